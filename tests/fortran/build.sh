#!/bin/bash
# Builds the Fortran host driver (a shared library with one BIND(C) entry)
# against the in-tree libcylpic.so.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
OUT="$HERE/_build"
mkdir -p "$OUT"
cd "$OUT"
amdflang -O2 -fPIC -c "$ROOT/include/cylpic.f90" -o cylpic_mod.o
amdflang -O2 -fPIC -c "$HERE/driver.f90" -o driver.o
amdflang -O2 -shared -o libcylpic_fdriver.so driver.o cylpic_mod.o -L"$ROOT/cylindrical_epoch_amd" -lcylpic \
  -Wl,-rpath,'$ORIGIN/../../../cylindrical_epoch_amd'
echo "$OUT/libcylpic_fdriver.so"
