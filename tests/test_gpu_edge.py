"""Edge cases of the timestep against the oracle: a species with no
particles and an immobile species next to a mobile one, and grids smaller
than one tile (the tile shape is clipped to the grid); the size limits are
rejected with an error."""
import numpy as np
import pytest

from cylindrical_epoch_amd import deck
from gpu_helpers import Q0, M0, add_plasma, by_id, log, make_pair, max_rel

pytestmark = pytest.mark.gpu


def _compare(st, sim, M, species, tol_m0=1e-10, tol_m=1e-6, tol_p=1e-9):
    for n in ("ex", "er", "et", "bx", "br", "bt", "jx", "jr", "jt"):
        got = sim.download_field(n)
        for m in range(M):
            assert max_rel(got[m], st.arr[n][m]) < (tol_m0 if m == 0 else tol_m), (n, m)
    for sid in species:
        got = by_id(sim.download_particles(sid))
        sp = st.species[sid]
        want = by_id({k: sp[k][: sp["n"]] for k in ("x", "y", "z", "px", "py", "pz", "w", "id")})
        assert np.array_equal(got["id"], want["id"]), sid
        for k in ("x", "y", "z", "px", "py", "pz"):
            if len(want[k]):
                assert max_rel(got[k], want[k]) < tol_p, (sid, k)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_empty_and_immobile_species(oracle, precision):
    log("empty + immobile species", precision)
    nx, ny, M, dx, dy = 24, 12, 2, 1e-7, 1e-7
    st, sim = make_pair(oracle, nx, ny, M, dx, dy, precision=precision)
    add_plasma(st, sim, nx, ny, dx, dy, ppc=4, temp_ev=3000.0, seed=5)
    e = np.zeros(0)
    st.add_species(Q0, 1836 * M0, e, e, e, e, e, e, e, ids=np.zeros(0, np.int64))
    sid1 = sim.set_species(Q0, 1836 * M0, False, capacity=8)
    sim.upload_particles(sid1, e, e, e, e, e, e, e, ids=np.zeros(0, np.int64))
    p = deck.uniform_plasma(nx, ny, dx, dy, 2, 1e25, 10.0, mass=1836 * M0, seed=6)
    p["id"] = p["id"] + 10 ** 6
    st.add_species(Q0, 1836 * M0, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], immobile=True,
                   ids=p["id"])
    sid2 = sim.set_species(Q0, 1836 * M0, True, capacity=len(p["x"]))
    sim.upload_particles(sid2, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])
    for _ in range(3):
        st.step()
        sim.step()
    assert sim.particle_count(sid1) == 0
    if precision == "fp64":
        _compare(st, sim, M, (0, 1, 2))
    else:
        we0, wb0 = st.mode_energy()
        we1, wb1 = sim.mode_energy()
        assert np.allclose(we1, we0, rtol=1e-3, atol=1e-3 * we0.sum())
        got = by_id(sim.download_particles(sid2))  # immobile: unchanged
        assert np.array_equal(got["id"], np.sort(p["id"]))
        assert max_rel(got["x"], p["x"][np.argsort(p["id"])]) < 1e-6
    sim.close()


@pytest.mark.parametrize("nx,ny,M", [(10, 6, 2), (12, 6, 1), (13, 7, 3)])
def test_grid_smaller_than_a_tile(oracle, nx, ny, M):
    """The smallest grids the library accepts with a reflecting r_max
    (nx >= 2 ng, ny > ng) under the default 16 x 8 tile, clipped to the
    grid."""
    log("tiny grid", nx, ny, M)
    dx, dy = 1e-7, 1e-7
    st, sim = make_pair(oracle, nx, ny, M, dx, dy)
    add_plasma(st, sim, nx, ny, dx, dy, ppc=3, temp_ev=2000.0, seed=8, drift=(0.1, 0.1, 0.0))
    for _ in range(3):
        st.step()
        sim.step()
    _compare(st, sim, M, (0,))
    sim.close()


def test_size_limits_rejected(oracle):
    """nx < 2 ng, ny < 3, and a reflecting r_max with ny <= ng (its radius
    ratios divide by r_max - i dy = 0 at i = ny) are refused."""
    from cylindrical_epoch_amd._lib import CylpicError
    for nx, ny in ((9, 8), (16, 2), (16, 5)):
        with pytest.raises(CylpicError):
            make_pair(oracle, nx, ny, 1)
