"""cylpic_gauss_residual_m0_noalloc, the entry point that allocates no device
memory, against cylpic_gauss_residual_m0 through the C ABI: two calls in a
row on different handles, the second on a smaller grid with fewer
particles, both precisions, each within the FP64 atomic-order tolerance of
the existing entry's result (1e-12 relative: the per-cell charges are sums
of atomics in a run-dependent order); a grid beyond the scratch array's
1048576 cells is refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

from cylindrical_epoch_amd import deck
from cylindrical_epoch_amd._lib import lib
from cylindrical_epoch_amd.host import CylpicError, Simulation
from gpu_helpers import Q0, M0, log, random_fields

pytestmark = pytest.mark.gpu


def _sim(nx, ny, precision, ppc, seed):
    dx = dy = 1e-7
    sim = Simulation(nx, ny, 2, dx, dy, deck.set_dt(dx, dy, 0.95), precision=precision)
    f = random_fields((2, ny + 10, nx + 10), np.random.default_rng(seed), e_amp=3e9)
    for n in ("ex", "er"):
        sim.upload_field(n, f[n])
    rng = np.random.default_rng(seed + 1)
    for charge, mass, immobile, temp in ((-Q0, M0, False, 2e4), (Q0, 1836.15 * M0, True, 10.0)):
        p = deck.uniform_plasma(nx, ny, dx, dy, ppc, 1e25, temp, mass=mass, seed=int(rng.integers(1 << 30)))
        sid = sim.set_species(charge, mass, immobile, capacity=len(p["x"]))
        sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    return sim


def _noalloc(sim):
    r, q, f = C.c_double(), C.c_double(), C.c_double()
    rc = lib().cylpic_gauss_residual_m0_noalloc(sim.h, C.byref(r), C.byref(q), C.byref(f))
    return rc, (r.value, q.value, f.value)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_gauss_matches_the_allocating_entry(precision):
    for nx, ny, ppc, seed in ((96, 48, 6, 11), (40, 20, 3, 12)):
        log("no-alloc gauss", precision, nx, ny, ppc)
        sim = _sim(nx, ny, precision, ppc, seed)
        rc, got = _noalloc(sim)
        assert rc == 0, lib().cylpic_last_error()
        want = sim.gauss_residual_m0()
        log("  noalloc", got, "allocating", want)
        for a, b in zip(got, want):
            assert b > 0
            assert abs(a - b) <= 1e-12 * b
        sim.close()


def test_noalloc_gauss_refuses_a_grid_beyond_its_scratch():
    log("no-alloc gauss refuses 1100 x 1000")
    sim = Simulation(1100, 1000, 1, 1e-7, 1e-7, deck.set_dt(1e-7, 1e-7, 0.95), precision="mixed")
    rc, _ = _noalloc(sim)
    assert rc != 0
    assert b"exceed its scratch" in lib().cylpic_last_error()
    with pytest.raises(CylpicError, match="exceed its scratch"):
        sim.gauss_residual_m0_noalloc()
    sim.close()
