"""Regenerate doc_deposit_example.json from the reference's DOCUMENTATION.pdf.

The reference documents one worked single-particle deposition (printed pages
18-19: Figure 8 gives the triangle weights of a 3x3 staggered patch, Tables 1-3
the resulting m=0 current densities).  This script decompresses the PDF's
content streams, finds the three tables' text and prints the numbers; the
fixture JSON next to it holds those numbers (and the Figure 8 weights and the
stated parameters, transcribed from the same pages).  Run in a checkout that
has the reference at /root/reference; it reads nothing else.
"""
import re
import sys
import zlib

PDF = "/root/reference/DOCUMENTATION.pdf"


def page_texts(data):
    for m in re.finditer(rb"stream\r?\n", data):
        s = m.end()
        e = data.find(b"endstream", s)
        try:
            t = zlib.decompress(data[s:e])
        except zlib.error:
            continue
        out = []
        for tj in re.finditer(rb"\[(.*?)\]\s*TJ|\((.*?)\)\s*Tj", t):
            seg = tj.group(1) or tj.group(2)
            for p, num in re.findall(rb"\(((?:\\.|[^\\)])*)\)|(-?\d+\.?\d*)", seg):
                if p:
                    out.append(p.decode("latin-1"))
                elif num and float(num) < -200:
                    out.append(" ")
        yield "".join(out)


def main():
    data = open(PDF, "rb").read()
    for txt in page_texts(data):
        if "Table 1" in txt or "Table 2" in txt or "Table 3" in txt:
            for name in ("Table 1", "Table 2", "Table 3"):
                k = txt.find(name)
                if k >= 0:
                    print(txt[max(0, k - 200):k + 40])
                    print("----")


if __name__ == "__main__":
    sys.exit(main())
