"""The oracle's threaded push (the CPU baseline's timing path,
orc_push_particles_threads) against its serial push: particles bit-exact
(the push of a particle does not depend on the others), the current within
1e-13 of its largest value (only the order of the sums differs: per-slab
copies added in slab order), for 1, 3 and 8 threads, with periodic and
with open x boundaries (a hot plasma whose particles cross the slabs'
edges and, periodic, wrap)."""
import numpy as np
import pytest

from cylindrical_epoch_amd import deck


@pytest.mark.parametrize("threads", [1, 3, 8])
@pytest.mark.parametrize("bc", ["periodic", "open"])
def test_threaded_push_matches_serial(oracle, threads, bc):
    nx, ny, M, dx, dy = 40, 12, 2, 1e-7, 1e-7
    dt = deck.set_dt(dx, dy, 0.95)
    bf = ("periodic", "periodic", "other", "conduct") if bc == "periodic" else \
        ("simple_outflow", "simple_outflow", "other", "conduct")
    bp = ("periodic", "periodic", "other", "reflect") if bc == "periodic" else ("open", "open", "other", "reflect")
    p = deck.uniform_plasma(nx, ny, dx, dy, 6, 1e25, 2.0e5, seed=4)
    states = []
    for k in range(2):
        st = oracle.State(nx, ny, M, dx, dy, dt, bc_field=bf, bc_particle=bp)
        rng = np.random.default_rng(9)
        for n in ("ex", "er", "et", "bx", "br", "bt"):
            st.arr[n][:] = (3e10 if n[0] == "e" else 100.0) * (rng.normal(size=st.shape) + 1j * rng.normal(size=st.shape))
        st.add_species(-deck.Q0, deck.M0, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])
        if k == 0:
            st.push_particles()
        else:
            st.push_particles_threads(threads)
        states.append(st)
    a, b = (s.species[0] for s in states)
    assert a["n"] == b["n"]
    for key in ("id", "x", "y", "z", "px", "py", "pz", "w"):
        assert np.array_equal(a[key][: a["n"]], b[key][: b["n"]]), key
    for n in ("jx", "jr", "jt"):
        ja, jb = states[0].arr[n], states[1].arr[n]
        assert np.max(np.abs(ja - jb)) <= 1e-13 * np.max(np.abs(ja)), n
        assert np.max(np.abs(ja)) > 0
