"""A Fortran host drives the GPU path through include/cylpic.f90 (the
reference main loop's call order); its results match the same run driven
from Python through the same C ABI.

The two runs are separate executions of the same kernels: particle order,
positions and weights agree bit for bit, but the per-tile current sums of the
FP64 path are LDS floating-point atomics whose order between waves is not
fixed, so currents, and through them fields and momenta, may differ in the
last bits from run to run.  The tolerance (1e-11 of each array's largest
value after five steps) bounds that round-off, not physics."""
import os

import numpy as np
import pytest

from fortran_io import DRIVER_LIB, read_output, run_driver, write_input
from gpu_helpers import log, max_rel

pytestmark = pytest.mark.gpu


def test_fortran_host_matches_python_host(tmp_path):
    from cylindrical_epoch_amd import deck
    from cylindrical_epoch_amd.host import Simulation
    assert os.path.exists(DRIVER_LIB), "build the driver first (__graft_entry__.build())"
    nx, ny, M, dx, dy, nsteps = 32, 16, 2, 1e-7, 1e-7, 5
    dt = deck.set_dt(dx, dy, 0.95)
    p = deck.uniform_plasma(nx, ny, dx, dy, 4, 1e25, 1000.0, seed=5)
    bf, bp, tile = (1, 1, 2, 10), (1, 1, 2, 9), (8, 4)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(fin, nx, ny, M, nsteps, 0, bf, bp, tile, dx, dy, dt, 0.0, -deck.Q0, deck.M0, p)
    log("fortran driver")
    assert run_driver(fin, fout) == 0
    fparts, fwe, fwb, fex, fjr = read_output(fout, nx, ny, M)

    sim = Simulation(nx, ny, M, dx, dy, dt, bc_field=bf, bc_particle=bp, tile=tile)
    sid = sim.set_species(-deck.Q0, deck.M0, False, capacity=len(p["x"]))
    sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    sim.setup_field_boundaries()
    for _ in range(nsteps):
        sim.update_eb_fields_half()
        sim.push_particles()
        sim.current_finish()
        sim.update_eb_fields_final()
    got = sim.download_particles(sid, with_ids=False)
    we, wb = sim.mode_energy()
    assert np.array_equal(got["w"], fparts["w"])
    tol = 1e-11
    for k in ("x", "y", "z", "px", "py", "pz"):
        err = max_rel(got[k], fparts[k])
        log(f"  {k} {err:.1e}")
        assert err <= tol, (k, err)
    for name, a, b in (("w_e", we, fwe), ("w_b", wb, fwb), ("ex", sim.download_field("ex"), fex),
                       ("jr", sim.download_field("jr"), fjr)):
        err = max_rel(a, b)
        log(f"  {name} {err:.1e}")
        assert err <= tol, (name, err)
    sim.close()
