"""The benchmark table's smaller workloads at their full sizes, against the
oracle on the same inputs.

* thermal 256 x 64, 2 modes, 8 electrons per cell at 1 keV with an
  immobile ion species of the same density, periodic x, conducting r_max;
  five full steps.
* wakefield 2048 x 256, 2 modes, 16 cold electrons per cell (8.4M) at
  1e24 m^-3, the x_min laser (a0 = 2, 0.8 um, 10 um waist, 30 fs) with an
  outflow x_max and open r_max; three full steps.

FP64 must agree with the oracle to the multi-step bounds (m = 0 fields 1e-10,
m > 0 fields 1e-6 of each array's largest value, particles 1e-9); the mixed
path's per-mode field energies must agree to 1e-3 (E) and 1e-2 (B)."""
import math

import numpy as np
import pytest

from cylindrical_epoch_amd import deck
from gpu_helpers import by_id, log, make_pair, max_rel

pytestmark = pytest.mark.gpu

FIELD_NAMES = ("ex", "er", "et", "bx", "br", "bt", "jx", "jr", "jt")


class LaserSource:
    """x_min laser sources (laser.f90:444-461) with a Gaussian waist and a
    Gaussian envelope peaking at twice the FWHM."""

    def __init__(self, ny, dy, a0, lam, waist, fwhm):
        self.omega = 2 * math.pi * deck.C_LIGHT / lam
        self.amp = a0 * deck.M0 * deck.C_LIGHT * self.omega / deck.Q0
        r = np.abs((np.arange(ny + 1) - 1) * dy + dy / 2)
        self.profile = np.exp(-(r / waist) ** 2)
        self.tw = fwhm / (2 * math.sqrt(math.log(2)))
        self.t0 = 2 * fwhm

    def __call__(self, t):
        env = math.exp(-((t - self.t0) / self.tw) ** 2) * self.amp
        return np.concatenate([env * self.profile * math.sin(self.omega * t), np.zeros_like(self.profile)])


def _load(st, sim, w, seed, q, immobile=False):
    p = deck.uniform_plasma(w["nx"], w["ny"], w["dx"], w["dy"], w["ppc"], w["density"], w["temp_ev"],
                            seed=seed)
    st.add_species(q, deck.M0 if q < 0 else 1836.15267 * deck.M0, p["x"], p["y"], p["z"], p["px"], p["py"],
                   p["pz"], p["w"], immobile=immobile, ids=p["id"])
    m = deck.M0 if q < 0 else 1836.15267 * deck.M0
    sid = sim.set_species(q, m, immobile, capacity=len(p["x"]))
    sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])


def _run(oracle, name, precision, steps, bc_field, bc_particle, laser=None, ions=False):
    w = deck.WORKLOADS[name]
    st, sim = make_pair(oracle, w["nx"], w["ny"], w["n_mode"], w["dx"], w["dy"], bc_field=bc_field,
                        bc_particle=bc_particle, precision=precision)
    _load(st, sim, w, 7842432, -deck.Q0)
    if ions:
        _load(st, sim, w, 7842433, deck.Q0, immobile=True)
    st.setup_field_boundaries()
    sim.setup_field_boundaries()
    for n in range(steps):
        src = None if laser is None else laser((n + 1) * st.g.dt)
        st.step(src_xmin=src)
        sim.step(src_xmin=src)
    return st, sim


def _check_fp64(st, sim):
    for n in FIELD_NAMES:
        got = sim.download_field(n)
        for m in range(got.shape[0]):
            err = max_rel(got[m], st.arr[n][m])
            assert err < (1e-10 if m == 0 else 1e-6), (n, m, err)
    got = by_id(sim.download_particles(0))
    want = by_id({k: st.species[0][k] for k in ("x", "y", "z", "px", "py", "pz", "w", "id")})
    assert np.array_equal(got["id"], want["id"])
    for k in ("x", "y", "z", "px", "py", "pz"):
        assert max_rel(got[k], want[k]) < 1e-9, k


def _check_energies(st, sim):
    """Modes holding at least 1e-6 of the largest mode energy agree to the
    bounds; a mode below that (m = 0 under a pure m = 1 laser in a cold
    plasma holds only round-off) must stay below it on both paths."""
    we0, wb0 = st.mode_energy()
    we1, wb1 = sim.mode_energy()
    floor = 1e-6 * max(we0.max(), wb0.max())
    for m in range(len(we0)):
        log(f"  m={m} W_E {we0[m]:.4e} vs {we1[m]:.4e}, W_B {wb0[m]:.4e} vs {wb1[m]:.4e}")
        for a, b, tol in ((we0[m], we1[m], 1e-3), (wb0[m], wb1[m], 1e-2)):
            if a > floor:
                assert abs(b - a) <= tol * a
            else:
                assert b < floor


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_thermal_256x64(oracle, precision):
    log("thermal 256x64 m2 ppc8 + immobile ions", precision)
    st, sim = _run(oracle, "thermal_256x64_m2_ppc8", precision, 5, ("periodic", "periodic", "other", "conduct"),
                   ("periodic", "periodic", "other", "reflect"), ions=True)
    if precision == "fp64":
        _check_fp64(st, sim)
    else:
        _check_energies(st, sim)
    sim.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_wakefield_2048x256(oracle, precision):
    log("wakefield 2048x256 m2 ppc16 with the x_min laser", precision)
    w = deck.WORKLOADS["wakefield_2048x256_m2_ppc16"]
    laser = LaserSource(w["ny"], w["dy"], 2.0, 0.8e-6, 10e-6, 30e-15)
    st, sim = _run(oracle, "wakefield_2048x256_m2_ppc16", precision, 3,
                   ("simple_laser", "simple_outflow", "other", "open"), ("open", "open", "other", "open"),
                   laser=laser)
    if precision == "fp64":
        _check_fp64(st, sim)
    else:
        _check_energies(st, sim)
    sim.close()
