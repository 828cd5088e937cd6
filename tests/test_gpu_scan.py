"""A seeded parameter scan of small runs in the manner of a batched scan of
uneven runs (grid size, mode count, particle load, boundaries, tile shape,
sort interval, current smoothing drawn per run), each stepped on the GPU and
by the oracle from the same inputs: FP64 to the multi-step tolerances
(m = 0 fields 1e-10, m > 0 1e-6, particles 1e-9), mixed to its mode-energy
tolerances (1e-3 for E, 1e-2 for B)."""
import numpy as np
import pytest

from gpu_helpers import add_plasma, by_id, log, make_pair, max_rel

pytestmark = pytest.mark.gpu

BCS = [
    (("periodic", "periodic", "other", "conduct"), ("periodic", "periodic", "other", "reflect")),
    (("simple_outflow", "simple_outflow", "other", "open"), ("open", "open", "other", "open")),
    (("conduct", "conduct", "other", "conduct"), ("reflect", "reflect", "other", "reflect")),
]
TILES_MIXED = [None, (16, 4), (8, 8), (8, 4), (16, 8), (32, 4)]
TILES_FP64 = [None, (16, 8), (8, 4), (4, 4)]


def _draw(k):
    rng = np.random.default_rng(4000 + k)
    run = dict(nx=int(rng.integers(24, 65)), ny=int(rng.integers(10, 25)), M=int(rng.integers(1, 4)),
               ppc=int(rng.integers(2, 7)), temp=float(rng.choice([100.0, 3e3, 5e4])),
               drift=tuple(float(v) for v in rng.normal(size=3) * 0.2), bc=int(rng.integers(0, len(BCS))),
               sort=int(rng.choice([1, 2, 4])), smooth=bool(rng.random() < 0.4), steps=4)
    run["tile_mixed"] = TILES_MIXED[int(rng.integers(0, len(TILES_MIXED)))]
    run["tile_fp64"] = TILES_FP64[int(rng.integers(0, len(TILES_FP64)))]
    return run


def _pair(oracle, run, precision):
    bf, bp = BCS[run["bc"]]
    tile = run["tile_" + precision]
    if tile is not None:
        tile = (min(tile[0], run["nx"]), min(tile[1], run["ny"]))
    st, sim = make_pair(oracle, run["nx"], run["ny"], run["M"], bc_field=bf, bc_particle=bp, precision=precision,
                        tile=tile, sort_interval=run["sort"])
    add_plasma(st, sim, run["nx"], run["ny"], 1e-7, 1e-7, run["ppc"], temp_ev=run["temp"], drift=run["drift"],
               seed=7)
    if run["smooth"]:
        st.smooth = (1, 1, (1, 2))
        sim.set_smoothing(1, 1, (1, 2))
    st.setup_field_boundaries()
    sim.setup_field_boundaries()
    return st, sim


@pytest.mark.parametrize("k", range(8))
def test_scan_fp64(oracle, k):
    run = _draw(k)
    log("scan fp64", k, run)
    st, sim = _pair(oracle, run, "fp64")
    for _ in range(run["steps"]):
        st.step()
        sim.step()
    for n in ("ex", "er", "et", "bx", "br", "bt", "jx", "jr", "jt"):
        got = sim.download_field(n)
        for m in range(run["M"]):
            err = max_rel(got[m], st.arr[n][m])
            assert err < (1e-10 if m == 0 else 1e-6), (n, m, err)
    got = by_id(sim.download_particles(0))
    want = by_id({kk: st.species[0][kk][: st.species[0]["n"]] for kk in ("x", "y", "z", "px", "py", "pz", "w", "id")})
    assert np.array_equal(got["id"], want["id"])
    for kk in ("x", "y", "z", "px", "py", "pz"):
        assert max_rel(got[kk], want[kk]) < 1e-9, kk
    sim.close()


@pytest.mark.parametrize("k", range(8))
def test_scan_mixed(oracle, k):
    run = _draw(k)
    log("scan mixed", k, run)
    st, sim = _pair(oracle, run, "mixed")
    for _ in range(run["steps"]):
        st.step()
        sim.step()
    we0, wb0 = st.mode_energy()
    we1, wb1 = sim.mode_energy()
    tot_e, tot_b = float(np.sum(we0)), float(np.sum(wb0))
    for m in range(run["M"]):
        assert abs(we1[m] - we0[m]) <= 1e-3 * tot_e + 1e-30, (m, we0[m], we1[m])
        assert abs(wb1[m] - wb0[m]) <= 1e-2 * tot_b + 1e-30, (m, wb0[m], wb1[m])
    n = sim.particle_count(0)
    assert abs(n - st.species[0]["n"]) <= max(2, n // 1000)
    sim.close()
