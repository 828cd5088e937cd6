"""The reference's current-density validation case (example_decks/
current_density_test.deck) through the C ABI: a uniform electron plasma
drifting with p = (-1.23e20, 1.23e20, 0) kg m/s puts J_x = n q v_x in mode 0
only and J_r = J_y, J_theta = -i J_y in mode 1 only (|J| = 3.40e-6 A m^-2).
FP64 runs the deck's drift; the mixed path rejects it (|p| / (m c) = 4.5e41
overflows FP32) and runs a drift of gamma ~ 4 instead."""
import numpy as np
import pytest

from cylindrical_epoch_amd import deck
from cylindrical_epoch_amd.host import CylpicError, Simulation
from gpu_helpers import log
from test_oracle_drift_modes import (DRIFT, DX, DY, M, NX, NY, check_axis_rows, check_mode_structure,
                                     drift_plasma, expected_current, interior_means)

pytestmark = pytest.mark.gpu

BC_FIELD = ("periodic", "periodic", "other", "conduct")
BC_PART = ("periodic", "periodic", "other", "reflect")


def _sim(precision):
    return Simulation(NX, NY, M, DX, DY, deck.set_dt(DX, DY, 0.95), bc_field=BC_FIELD, bc_particle=BC_PART,
                      precision=precision)


def _upload(sim, p):
    sid = sim.set_species(-deck.Q0, deck.M0, False, capacity=len(p["x"]))
    sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_drift_current_modes_gpu(precision):
    drift = DRIFT if precision == "fp64" else tuple(np.sign(DRIFT) * 3.0 * deck.M0 * deck.C_LIGHT)
    log("drift current modes", precision, drift)
    sim = _sim(precision)
    _upload(sim, drift_plasma(drift=drift))
    sim.push_particles()
    sim.current_finish()
    fields = {c: sim.download_field(c) for c in ("jx", "jr", "jt")}
    means = interior_means(lambda c: fields[c], 5)
    j = expected_current(drift)
    log("  J_x,0", means["jx"][0] / j[0], "J_r,1", means["jr"][1] / j[1], "J_t,1", means["jt"][1] / (-1j * j[1]))
    check_mode_structure(means, j)
    check_axis_rows(lambda c: fields[c], 5, j)
    sim.close()


def test_mixed_rejects_momenta_beyond_fp32():
    log("mixed rejects the deck's drift")
    sim = _sim("mixed")
    with pytest.raises(CylpicError, match="momentum"):
        _upload(sim, drift_plasma(ppc=4, drift=DRIFT))
    p = drift_plasma(ppc=4, drift=(0.0, 0.0, 0.0))
    p["pz"][7] = np.nan
    with pytest.raises(CylpicError, match="momentum"):
        sim.upload_particles(0, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    sim.close()
