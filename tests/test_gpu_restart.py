"""Restart round trip: a run checkpointed after seven steps (the fifteen
field arrays, the particles with their ids, and cylpic_get_restart_state:
time, the window's position, the push count, the tile offset) and continued
on a new handle (boundary snapshots of the still-zero fields, then
cylpic_set_restart_state and the uploads, as the reference's after_deck_last
precedes restart_data, src/epoch2d.F90:113-117) equals the same run done in
one go.  Laser at x_min, outflow x_max, a moving window that injects a
column and moves one cell every second step, sorts every third push (so the
checkpoint falls mid-phase), an odd tile offset at the checkpoint.

FP64: the continued run matches the uninterrupted one to the multi-step
bounds (m = 0 fields 1e-10, m > 0 1e-6, particles 1e-9 of each array's
largest value): the reloaded particles are processed in the upload's order,
so the deposit's atomic sums add in another order.  Mixed: the same at
1e-4 (FP32 flush atomics)."""
import numpy as np
import pytest

from cylindrical_epoch_amd import deck
from cylindrical_epoch_amd.host import Simulation
from gpu_helpers import M0, Q0, by_id, log, max_rel
from test_gpu_batch import FIELD_NAMES, laser_sources

pytestmark = pytest.mark.gpu

NX, NY, M, DX, DY, PPC = 48, 16, 2, 1e-7, 1e-7, 4
BF = ("simple_laser", "simple_outflow", "other", "conduct")
BP = ("open", "open", "other", "reflect")
N1, N2 = 7, 8


def make(precision):
    dt = deck.set_dt(DX, DY, 0.95)
    sim = Simulation(NX, NY, M, DX, DY, dt, bc_field=BF, bc_particle=BP, precision=precision, tile=(8, 4),
                     sort_interval=3)
    sim.set_species(-Q0, M0, False, capacity=2 * NX * NY * PPC)
    return sim


def start(sim):
    p = deck.uniform_plasma(NX, NY, DX, DY, PPC, 1e25, 2.0e4, seed=5)
    sim.setup_field_boundaries()
    sim.upload_particles(0, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])


def advance(sim, k0, k1):
    """Steps k0 .. k1-1 of the run: each with the laser's sources; after the
    odd ones a column of new particles beyond x_max and a one-cell shift."""
    for k in range(k0, k1):
        sim.step(laser_sources(NY, DY, (k + 1) * sim.dt, a0=0.5, waist=6e-7))
        if k % 2 == 1:
            p = deck.uniform_plasma(NX, NY, DX, DY, PPC, 1e25, 2.0e4, x_min=sim.x_min, seed=100 + k, cells_x=[NX])
            p["id"] = p["id"] + 10 ** 6 + 10 ** 4 * k
            sim.inject_particles(0, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])
            sim.move_window(1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_restart_round_trip(precision):
    log("restart round trip", precision)
    whole = make(precision)
    start(whole)
    advance(whole, 0, N1 + N2)
    first = make(precision)
    start(first)
    advance(first, 0, N1)
    ckpt = first.checkpoint()
    first.close()
    st = ckpt["state"]
    log("  checkpoint state", st)
    assert st["push_count"] == N1 and st["push_count"] % 3 != 0
    assert st["tile_shift"] == (N1 // 2) % 8 and st["tile_shift"] != 0
    again = make(precision)
    again.restore(ckpt)
    assert again.restart_state() == st
    advance(again, N1, N1 + N2)
    a, b = whole.restart_state(), again.restart_state()
    log("  states", a, b)
    assert a == b
    ftol = {"fp64": (1e-10, 1e-6), "mixed": (1e-4, 1e-4)}[precision]
    for name in FIELD_NAMES:
        fa, fb = whole.download_field(name), again.download_field(name)
        for m in range(M):
            err = max_rel(fb[m], fa[m])
            log(f"  {name} m={m} {err:.1e}")
            assert err < ftol[m > 0], (name, m, err)
    pa, pb = by_id(whole.download_particles(0)), by_id(again.download_particles(0))
    assert np.array_equal(pa["id"], pb["id"])
    ptol = 1e-9 if precision == "fp64" else 1e-4
    for k in ("x", "y", "z", "px", "py", "pz", "w"):
        assert max_rel(pb[k], pa[k]) < ptol, k
    whole.close()
    again.close()


def test_restart_state_is_refused_after_an_upload():
    sim = make("fp64")
    start(sim)
    st = sim.restart_state()
    with pytest.raises(Exception, match="before uploading particles"):
        sim.set_restart_state(st)
    fresh = make("fp64")
    with pytest.raises(Exception, match="tile_shift"):
        fresh.set_restart_state(dict(st, tile_shift=8))
    sim.close()
    fresh.close()
