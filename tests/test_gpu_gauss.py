"""Gauss's law for m = 0 (cylpic_gauss_residual_m0) through the C ABI.

* The entry's three maxima equal a numpy evaluation of the same quantity
  from the downloaded E and particles: the mean of each mobile particle's
  charges at x - v dt/2 and x + v dt/2 and the charge of each immobile one
  at x, with quadratic-spline fractions on the staggered grid, against eps0
  times the net flux of E_x, E_r out of each cell (the oracle's area
  tables), over the cells at least ng from every edge.
* A neutral start (electrons at rest on top of immobile ions, E = 0) has no
  residual beyond the round-off of the two species' sums, and a laser from
  x_min that drives the electrons keeps it at round-off of the electrons'
  per-cell charge while eps0 div E grows: the deposit conserves charge and
  the E update adds -dt J / eps0."""
import math

import numpy as np
import pytest

from cylindrical_epoch_amd import deck
from gpu_helpers import Q0, M0, C_LIGHT, log, make_pair, random_fields

pytestmark = pytest.mark.gpu
EPS0 = 8.854187817620389850536563031710750e-12


def _stagger(pos):
    c = np.floor(pos)
    f = c - pos + 0.5
    w = np.stack([0.25 + f * f + f, 1.5 - 2 * f * f, 0.25 + f * f - f]) / 2
    return c.astype(int), w          # first of three cells (Fortran index), weights


def numpy_gauss(st, ex, er, species):
    g, ng = st.g, st.g.ng
    q = np.zeros((g.ny + 2 * ng, g.nx + 2 * ng))
    for sp, charge, mass, immobile in species:
        u = np.stack([sp["px"], sp["py"], sp["pz"]]) / (mass * C_LIGHT)
        root = 0.0 if immobile else C_LIGHT * (g.dt / 2.0) / np.sqrt((u * u).sum(0) + 1.0)
        for vs in (-1.0, 1.0):
            x, y, z = sp["x"] + u[0] * vs * root, sp["y"] + u[1] * vs * root, sp["z"] + u[2] * vs * root
            cx, wx = _stagger((x - g.x_grid_min) / g.dx)
            cy, wy = _stagger((np.hypot(y, z) - g.y_grid_min) / g.dy)
            for a in range(3):
                for b in range(3):
                    np.add.at(q, (cy + b + ng - 1, cx + a + ng - 1), 0.5 * charge * sp["w"] * wx[a] * wy[b])
    inv_rt, inv_xt, _, _ = st.geometry()
    res = qq = ff = 0.0
    for cy in range(ng + 1, g.ny - ng + 1):
        for_cols = np.arange(ng + 1, g.nx - ng + 1)
        k = lambda i, j: (j + ng - 1, i + ng - 1)
        flux = EPS0 * ((ex[0][k(for_cols + 1, cy)].real - ex[0][k(for_cols, cy)].real) / inv_rt[cy + ng]
                       + er[0][k(for_cols, cy + 1)].real / inv_xt[cy + ng]
                       - er[0][k(for_cols, cy)].real / inv_xt[cy - 1 + ng])
        qc = q[k(for_cols, cy)]
        res, qq, ff = max(res, np.abs(flux - qc).max()), max(qq, np.abs(qc).max()), max(ff, np.abs(flux).max())
    return res, qq, ff


def test_gauss_entry_matches_numpy(oracle):
    nx, ny, dx, dy = 40, 24, 1e-7, 1e-7
    st, sim = make_pair(oracle, nx, ny, 2, dx, dy, tile=(8, 4))
    f = random_fields(st.shape, np.random.default_rng(5), e_amp=3e9)
    for n in ("ex", "er"):
        sim.upload_field(n, f[n])
    rng = np.random.default_rng(6)
    species = []
    for charge, mass, immobile, temp in ((-Q0, M0, False, 2e4), (Q0, 1836.15 * M0, True, 10.0)):
        p = deck.uniform_plasma(nx, ny, dx, dy, 6, 1e25, temp, mass=mass, seed=int(rng.integers(1 << 30)))
        sid = sim.set_species(charge, mass, immobile, capacity=len(p["x"]))
        sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
        species.append((p, charge, mass, immobile))
    got = sim.gauss_residual_m0()
    want = numpy_gauss(st, sim.download_field("ex"), sim.download_field("er"), species)
    log("gauss entry", got, want)
    for a, b in zip(got, want):
        assert a == pytest.approx(b, rel=1e-12, abs=0)
    sim.close()


@pytest.mark.parametrize("precision,bound", [("fp64", 1e-12), ("mixed", 1e-5)])
def test_gauss_kept_by_a_laser_driven_plasma(oracle, precision, bound):
    nx, ny, dx, dy = 128, 32, 0.05e-6, 0.05e-6
    bcf = ("simple_laser", "simple_outflow", "other", "conduct")
    bcp = ("open", "open", "other", "reflect")
    st, sim = make_pair(oracle, nx, ny, 2, dx, dy, bc_field=bcf, bc_particle=bcp, precision=precision)
    p = deck.uniform_plasma(nx, ny, dx, dy, 8, 5e26, 0.0, seed=9)
    for charge, mass, immobile in ((-Q0, M0, False), (Q0, 1836.15 * M0, True)):
        sid = sim.set_species(charge, mass, immobile, capacity=len(p["x"]))
        sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    sim.setup_field_boundaries()
    # the electrons' charge in the largest cell sets the scale: the net
    # charge of a neutral cell is round-off of the two species' sums
    scale = 5e26 * Q0 * 2 * math.pi * (ny * dy) * dx * dy
    r0, q0, f0 = sim.gauss_residual_m0()
    log("gauss start", precision, r0 / scale, q0 / scale, f0)
    assert r0 <= 1e-14 * scale and q0 <= 1e-14 * scale and f0 == 0.0
    lam = 0.8e-6
    omega = 2 * math.pi * C_LIGHT / lam
    amp = 1.0 * M0 * C_LIGHT * omega / Q0
    r = np.abs((np.arange(ny + 1) - 1) * dy + dy / 2)
    prof = np.exp(-(r / 0.6e-6) ** 2)
    worst = 0.0
    for n in range(160):
        t = (n + 1) * sim.dt
        src = np.concatenate([amp * min(1.0, t / 10e-15) * prof * math.sin(omega * t), np.zeros(ny + 1)])
        sim.step(src_xmin=src)
        if n % 20 == 19:
            res, q, fl = sim.gauss_residual_m0()
            worst = max(worst, res / scale)
            log(f"  step {n + 1}: residual {res / scale:.2e}, net charge {q / scale:.2e}, "
                f"eps0 div E {fl / scale:.2e} of the electron charge")
    assert worst <= bound
    assert fl / scale > 100 * bound       # the laser did separate charge
    sim.close()
