"""The element pairing of the laser boundaries' non-conformable terms.  No
GPU needed.

outflow_bcs_x_min divides the (1:ny) section bxm(1,1:ny,im) by the (0:ny)
array r_d_vals, and outflow_bcs_x_max adds the (0:ny) arrays source_t and
source_r to (1:ny) sections (src/laser.f90:490-493, 599-612).  The standard
leaves such expressions undefined; what the reference computes is what its
compiler emits.  tests/fortran/pairing.f90 evaluates the same expression
shapes (allocatable arrays, extents known only at run time) with amdflang
and reports which element each row used; the oracle must pair its rows the
same way."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
C = 2.99792458e8


def fortran_pairing(ny):
    fc = shutil.which("amdflang") or shutil.which("flang")
    if fc is None:
        pytest.skip("no Fortran compiler")
    out_dir = os.path.join(HERE, "fortran", "_build")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "pairing")
    subprocess.check_call([fc, "-O2", os.path.join(HERE, "fortran", "pairing.f90"), "-o", exe], cwd=out_dir)
    rd, src = {}, {}
    for line in subprocess.check_output([exe, str(ny)], text=True).split("\n"):
        if line.strip():
            tag, ir, k = line.split()
            (rd if tag == "rd" else src)[int(ir)] = int(k)
    return rd, src


def test_oracle_pairs_rows_as_the_fortran_compiler_does(oracle):
    ny, nx, M = 7, 6, 2
    rd, src = fortran_pairing(ny)
    assert rd == {ir: ir - 1 for ir in range(1, ny + 1)}
    assert src == {ir: ir - 1 for ir in range(1, ny + 1)}
    dx = dy = 1e-7
    dt = oracle.set_dt(dx, dy, 0.95)
    st = oracle.State(nx, ny, M, dx, dy, dt, bc_field=("simple_laser", "simple_laser", "other", "conduct"),
                      bc_particle=("open", "open", "other", "reflect"))
    g, ng = st.g, st.g.ng
    lx = dt * C * C / dx
    s = 1.0 / (lx + C)
    # x_min: only B_x on column 1, so btm(1, ir, 1) = sum * i m c^2 dt bx / r_d(row's element)
    st.arr["bx"][1][:, :] = 1.0
    st.call("orc_outflow_bcs_x_min", None)
    for ir in range(1, ny + 1):
        bt = st.arr["bt"][1][ir + ng - 1, 1 + ng - 1]
        r_used = s * C * C * dt / bt.imag
        k = rd[ir]
        assert r_used == pytest.approx(abs((k - 1) * dy + g.y_grid_min), rel=1e-12, abs=0), ir
    # x_max: source1(k) = k and nothing else, so btm(nx, ir, 1) = -4 sum source_t(row's element)
    st.arr["bx"][...] = 0.0
    source = np.concatenate([np.arange(ny + 1, dtype=float), np.zeros(ny + 1)])
    st.call("orc_outflow_bcs_x_max", source.ctypes.data)
    for ir in range(1, ny + 1):
        bt = st.arr["bt"][1][ir + ng - 1, nx + ng - 1]
        assert -bt.real / (4 * s) == pytest.approx(src[ir], rel=1e-12, abs=1e-12), ir
