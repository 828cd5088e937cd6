"""GPU output diagnostics against the oracle run on the same particles (the
device's particles downloaded after a few steps, so removed particles and
unsorted state are exercised): number density modes
(calc_df.F90:588-661) and distribution functions (dist_fn.F90:118-715).
Cell indices, shape weights and bins are computed identically on both
sides; only the order of the FP64 atomic sums differs, hence the 1e-12
relative tolerance; automatic ranges (exact minima and maxima) must agree
exactly."""
import numpy as np
import pytest

from gpu_helpers import Q0, M0, add_plasma, log, make_pair

pytestmark = pytest.mark.gpu

NX, NY, M = 40, 16, 3
DX = DY = 1e-7
CASES = [
    (("periodic", "periodic", "other", "conduct"), ("periodic", "periodic", "other", "reflect")),
    (("simple_outflow", "simple_outflow", "other", "open"), ("open", "open", "other", "open")),
    (("conduct", "conduct", "other", "conduct"), ("reflect", "reflect", "other", "reflect")),
]


def _run(oracle, case, precision, steps=3, drift=(0.3, 0.1, -0.2), ions=False):
    bf, bp = CASES[case]
    st, sim = make_pair(oracle, NX, NY, M, dx=DX, dy=DY, bc_field=bf, bc_particle=bp, precision=precision)
    add_plasma(st, sim, NX, NY, DX, DY, 6, temp_ev=5000.0, drift=drift, seed=3 + case)
    if ions:
        add_plasma(st, sim, NX, NY, DX, DY, 3, temp_ev=10.0, seed=9, q=Q0, m=1836 * M0)
    sim.setup_field_boundaries()
    for _ in range(steps):
        sim.step()
    return sim


def _reference(oracle, sim, case):
    """A fresh oracle state holding the device's particles."""
    bf, bp = CASES[case]
    ref = oracle.State(NX, NY, M, DX, DY, sim.dt, bc_field=bf, bc_particle=bp)
    for sid, sp in enumerate(sim.species):
        p = sim.download_particles(sid, with_ids=False)
        ref.add_species(sp["charge"], sp["mass"], p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    return ref


def _close(a, b, rtol=1e-12):
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale) <= rtol


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_number_density_modes(oracle, case, precision):
    log("number density", case, precision)
    sim = _run(oracle, case, precision)
    got = sim.number_density_modes(0)  # straight after a push: removed particles still held
    want = _reference(oracle, sim, case).number_density_modes(0)
    assert _close(got, want), np.max(np.abs(got - want))
    sim.close()


def test_number_density_all_species(oracle):
    log("number density, all species")
    sim = _run(oracle, 0, "fp64", ions=True)
    got = sim.number_density_modes(-1)
    want = _reference(oracle, sim, 0).number_density_modes(-1)
    assert _close(got, want)
    assert _close(got, sim.number_density_modes(0) + sim.number_density_modes(1))
    sim.close()


DF_CASES = [
    dict(directions=["en"], resolution=[64]),
    dict(directions=["px"], ranges=[(-2e-22, 2e-22)], resolution=[50]),
    dict(directions=["x", "px"], resolution=[0, 120]),  # 40 x 120 bins: past the LDS histogram
    dict(directions=["x", "py"], ranges=[(1.3e-6, 3.1e-6), (0.0, 0.0)], resolution=[0, 16],
         restrictions={"pz": (0.0, 1e30)}),
    dict(directions=["y", "gamma_m1", "mod_p"], resolution=[0, 8, 6]),
    dict(directions=["xy_angle"], ranges=[(-np.pi, np.pi)], resolution=[36]),
    dict(directions=["zx_angle"], resolution=[24], restrictions={"en": (0.0, 4e-16)}),
]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("k", range(len(DF_CASES)))
def test_dist_fn(oracle, k, precision):
    log("dist_fn", k, precision)
    sim = _run(oracle, 1, precision)
    kw = DF_CASES[k]
    h1, r1, n1 = sim.dist_fn(0, **kw)
    h0, r0, n0 = _reference(oracle, sim, 1).dist_fn(0, **kw)
    assert np.array_equal(n1, n0)
    assert np.array_equal(r1, r0)
    assert h1.shape == h0.shape
    assert _close(h1, h0), np.max(np.abs(h1 - h0))
    assert h0.sum() > 0
    sim.close()


def test_dist_fn_empty_species_and_errors(oracle):
    log("dist_fn empty")
    st, sim = make_pair(oracle, NX, NY, M, dx=DX, dy=DY)
    sid = sim.set_species(-Q0, M0, False, capacity=16)
    assert sim.dist_fn(sid, ["en"], resolution=[10]) is None
    with pytest.raises(Exception):
        sim.dist_fn(sid, [12], resolution=[10])
    with pytest.raises(Exception):
        sim.dist_fn(sid, ["px"], resolution=[0])
    sim.close()


@pytest.mark.parametrize("precision,bound", [("fp64", 1e-11), ("mixed", 1e-5)])
def test_continuity_residual_m0(oracle, precision, bound):
    """The m = 0 continuity identity of the charge-conserving deposit over a
    few pushes of a drifting thermal plasma, relative to the largest
    per-cell charge change: at round-off in FP64 (measured 1.3e-14); the
    mixed path, with the charges of its own FP32 half step, is bounded by
    its FP32 currents and fixed-point sums (measured 2.7e-7)."""
    log("continuity", precision)
    bf, bp = CASES[0]
    st, sim = make_pair(oracle, NX, NY, M, dx=DX, dy=DY, bc_field=bf, bc_particle=bp, precision=precision)
    add_plasma(st, sim, NX, NY, DX, DY, 8, temp_ev=20000.0, drift=(0.5, 0.2, -0.3), seed=21)
    sim.setup_field_boundaries()
    worst = 0.0
    for _ in range(3):
        sim.update_eb_fields_half()
        sim.continuity_begin()
        sim.push_particles()
        res, dq, q = sim.continuity_end()
        sim.current_finish()
        sim.update_eb_fields_final()
        assert 0 < dq < q
        worst = max(worst, res / dq)
        log("  residual / max |dQ| =", res / dq, " / max |Q| =", res / q)
    assert worst < bound
    with pytest.raises(Exception):
        sim.continuity_end()  # not armed
    sim.close()
