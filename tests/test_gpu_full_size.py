"""The benchmark's own workload at its full size (8192 x 512 cells, 3 modes,
32 particles per cell: 134,217,728 electrons), checked by size-independent
properties instead of the oracle, which cannot push that many particles in a
test: a 1 keV thermal load so that every particle moves, two full steps, and
around a third push the m = 0 continuity equation of the deposit, whose
residual must stay at round-off relative to the largest per-cell charge
(1e-11 for FP64, 1e-6 for the mixed path with its FP32 positions).  After
the push the particle count is unchanged (periodic x, reflecting r_max)."""
import pytest

from cylindrical_epoch_amd import deck
from cylindrical_epoch_amd.host import Simulation
from gpu_helpers import log

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,bound", [("fp64", 1e-11), ("mixed", 1e-6)])
def test_full_size_continuity(precision, bound):
    w = deck.WORKLOADS["wakefield_8192x512_m3_ppc32"]
    nx, ny, M, dx, dy = w["nx"], w["ny"], w["n_mode"], w["dx"], w["dy"]
    log("full size", precision, nx, ny, M, w["ppc"])
    p = deck.uniform_plasma(nx, ny, dx, dy, w["ppc"], w["density"], 1000.0, seed=5)
    n = len(p["x"])
    sim = Simulation(nx, ny, M, dx, dy, deck.set_dt(dx, dy, 0.95),
                     bc_field=("periodic", "periodic", "other", "conduct"),
                     bc_particle=("periodic", "periodic", "other", "reflect"), precision=precision)
    sid = sim.set_species(-deck.Q0, deck.M0, False, capacity=n)
    sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    del p
    log("  uploaded", n)
    for _ in range(2):
        sim.step()
    sim.update_eb_fields_half()
    sim.continuity_begin()
    sim.push_particles()
    res, dq, q = sim.continuity_end()
    log(f"  residual {res:.3e} C, max charge change {dq:.3e}, max charge {q:.3e}")
    assert q > 0 and dq > 0
    assert res <= bound * q
    assert sim.particle_count(sid) == n
    sim.close()
