"""GPU checks of the mixed-precision path (int32 cell + FP32 offsets, FP32
fields and momenta, FP64 per-tile current sums, FP64 axis rows) against the
FP64 oracle on the same inputs.  Every comparison reports its error against
the oracle; the tolerances are the FP32 error budgets stated per test."""
import numpy as np
import pytest

from gpu_helpers import (add_plasma, by_id, load_fields, log, make_pair, max_rel,
                         random_fields)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M", [1, 3])
def test_mixed_fields_half(oracle, M):
    """One half E/B update in FP32: relative error of a few FP32 ulps of the
    largest field value."""
    log("mixed fields half M", M)
    st, sim = make_pair(oracle, 37, 13, M, precision="mixed")
    rng = np.random.default_rng(40 + M)
    load_fields(st, sim, random_fields(st.shape, rng))
    st.update_eb_fields_half()
    sim.update_eb_fields_half()
    for n in ("ex", "er", "et", "bx", "br", "bt", "bx_old", "br_old", "bt_old"):
        err = max_rel(sim.download_field(n), st.arr[n])
        log(" ", n, f"{err:.2e}")
        assert err < 2e-6, (n, err)
    sim.close()


@pytest.mark.parametrize("M,acc,gather_lds,tile", [(1, 3, 1, (8, 4)), (2, 3, 1, (8, 4)), (3, 3, 1, (8, 4)),
                                                    (3, 3, 0, (8, 4)), (4, 3, 0, (8, 4)), (3, 0, 1, (8, 4)),
                                                    (3, 1, 1, (8, 4)), (3, 2, 0, (8, 4)), (1, 4, 0, (8, 4)),
                                                    (3, 4, 1, (8, 4)), (2, 0, 0, (8, 4)), (3, 3, 2, (8, 4)),
                                                    (1, 3, 2, (8, 4)), (3, 3, 1, (16, 4)), (2, 3, 1, (8, 8)),
                                                    (3, 3, 2, (32, 4)), (3, 3, 1, (5, 3))])
def test_mixed_push(oracle, M, acc, gather_lds, tile):
    """One push of a warm plasma in given fields, for each current
    accumulator (FP32, FP32 + 2Sum, FP64, int64 fixed point, packed 2 x 32-bit
    fixed point) and gather source (LDS, L1/L2, or E from LDS and B from
    L1/L2), on tiles with compile-time shapes (16 x 4, 8 x 8, 32 x 4, whose
    staged field rows are padded to a bank-conflict-free pitch) and run-time
    ones (8 x 4, 5 x 3).
    Positions and momenta carry FP32 rounding (~1e-7 relative per
    operation); the current of a particle is a difference of O(1) shape
    weights, so FP32 positions put ~1e-5 relative error on slow particles'
    contributions."""
    log("mixed push M", M, "acc", acc, "gather_lds", gather_lds, "tile", tile)
    nx, ny, dx, dy = 40, 20, 1e-7, 1e-7
    st, sim = make_pair(oracle, nx, ny, M, dx, dy, tile=tile, precision="mixed")
    sim.tune(acc, gather_lds)
    rng = np.random.default_rng(M)
    f = random_fields(st.shape, rng, e_amp=3e10, b_amp=200.0, j_amp=0.0)
    load_fields(st, sim, f)
    add_plasma(st, sim, nx, ny, dx, dy, ppc=6, temp_ev=2.0e5, drift=(0.3, 0.0, 0.0), seed=M)
    st.push_particles()
    sim.push_particles()
    got = by_id(sim.download_particles(0))
    want = by_id({k: st.species[0][k] for k in ("x", "y", "z", "px", "py", "pz", "w", "id")})
    assert np.array_equal(got["id"], want["id"])
    pos_err = max(np.max(np.abs(got["x"] - want["x"])) / dx,
                  np.max(np.abs(got["y"] - want["y"])) / dy, np.max(np.abs(got["z"] - want["z"])) / dy)
    mom_err = max(max_rel(got[k], want[k]) for k in ("px", "py", "pz"))
    log(f"  position error {pos_err:.2e} cells, momentum error {mom_err:.2e}")
    assert pos_err < 1e-4
    assert mom_err < 1e-5
    for n in ("jx", "jr", "jt"):
        g = sim.download_field(n)
        for m in range(M):
            err = max_rel(g[m], st.arr[n][m])
            log(f"  {n} m={m} error {err:.2e}")
            assert err < 1e-3, (n, m, err)
    sim.close()


def test_mixed_multistep_energy(oracle):
    """Twenty steps of a thermal plasma: per-mode field energies track the
    FP64 oracle (the statistical comparison PIC calls for once trajectories
    decorrelate)."""
    log("mixed multi-step")
    nx, ny, dx, dy = 32, 16, 1e-7, 1e-7
    st, sim = make_pair(oracle, nx, ny, 2, dx, dy, tile=(8, 8), precision="mixed")
    add_plasma(st, sim, nx, ny, dx, dy, ppc=8, temp_ev=1000.0, seed=2)
    for _ in range(20):
        st.step()
        sim.step()
    we0, wb0 = st.mode_energy()
    we1, wb1 = sim.mode_energy()
    for m in range(2):
        log(f"  m={m} W_E {we0[m]:.4e} vs {we1[m]:.4e}, W_B {wb0[m]:.4e} vs {wb1[m]:.4e}")
        assert abs(we1[m] - we0[m]) <= 1e-3 * we0[m]
        assert abs(wb1[m] - wb0[m]) <= 1e-2 * max(wb0[m], 1e-30) + 1e-30
    sim.close()
