"""Known-answer test of the oracle's m=0 deposit against the worked example of
the reference documentation (printed pp. 18-19: Figure 8 weights, Tables 1-3).
The example pins the deposit formulas of particles.F90:633-662 and the
geometry tables of particles.F90:190-217."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
C_LIGHT = 2.99792458e8


def _printed_tol(v):
    """Half a unit in the last printed digit (3 significant figures)."""
    v = abs(v)
    if v == 0:
        return 0.005
    e = int(np.floor(np.log10(v)))
    return 0.5 * 10.0 ** (e - 2) * 1.02


def test_doc_tables(oracle):
    d = json.load(open(os.path.join(HERE, "golden", "doc_deposit_example.json")))
    p = d["params"]
    st = oracle.State(nx=120, ny=120, n_mode=1, dx=p["dx"], dy=p["dr"], dt=p["dt"])
    fx, fr = np.array(d["fig8_x"]), np.array(d["fig8_r"])
    # the code's triangle weights sum to 2 (gx.inc); the documented ones to 1
    gx, hx, gy, hy = (np.zeros(5) for _ in range(4))
    gx[1:4], hx[1:4] = 2 * fx[:, 0], 2 * fx[:, 2]
    gy[1:4], hy[1:4] = 2 * fr[:, 0], 2 * fr[:, 2]
    # staggered cell (0,0) centred at x0+2dx, r0+2dr -> index 102 with x_min = 0
    cell = int(round((p["x0"] + 2 * p["dx"]) / p["dx"]))
    assert cell == 102
    qwf = p["Q"] * p["weight"] * 0.25
    oracle.deposit_from_weights(st, gx, hx, gy, hy, cell, cell, 0, 0, qwf / p["dt"],
                                qwf / p["dt"], qwf * p["v_theta_over_c"] * C_LIGHT)
    jt = [[st.at("jt", ix, iy, 0).real / 1e5 for ix in (101, 102, 103)] for iy in (101, 102, 103)]
    jx = [[st.at("jx", ix, iy, 0).real / 1e7 for ix in (101, 102, 103, 104)] for iy in (101, 102, 103)]
    jr = [[st.at("jr", ix, iy, 0).real / 1e7 for ix in (101, 102, 103)] for iy in (101, 102, 103, 104)]
    for got, key in ((jt, "table1_jtheta_1e5"), (jx, "table2_jx_1e7"), (jr, "table3_jr_1e7")):
        want = np.array(d[key]["values"], float)
        got = np.array(got)
        for g_, w_ in zip(got.ravel(), want.ravel()):
            assert abs(g_ - w_) <= _printed_tol(w_), (key, got, want)
    # m=0 currents are real
    assert np.all(st.arr["jt"].imag == 0) and np.all(st.arr["jx"].imag == 0)


def test_doc_tables_detect_planted_mistake(oracle):
    """A transposed operand (hx<->hy) in the weights fails the table check."""
    d = json.load(open(os.path.join(HERE, "golden", "doc_deposit_example.json")))
    p = d["params"]
    st = oracle.State(nx=120, ny=120, n_mode=1, dx=p["dx"], dy=p["dr"], dt=p["dt"])
    fx, fr = np.array(d["fig8_x"]), np.array(d["fig8_r"])
    gx, hx, gy, hy = (np.zeros(5) for _ in range(4))
    gx[1:4], hx[1:4] = 2 * fx[:, 0], 2 * fr[:, 2]     # planted: hx from r
    gy[1:4], hy[1:4] = 2 * fr[:, 0], 2 * fx[:, 2]
    qwf = p["Q"] * 0.25
    oracle.deposit_from_weights(st, gx, hx, gy, hy, 102, 102, 0, 0, qwf / p["dt"],
                                qwf / p["dt"], qwf * 0.1 * C_LIGHT)
    jt = np.array([[st.at("jt", ix, iy, 0).real / 1e5 for ix in (101, 102, 103)]
                   for iy in (101, 102, 103)])
    want = np.array(d["table1_jtheta_1e5"]["values"])
    assert np.max(np.abs(jt - want)) > 1.0
