"""The current folds at the axis and at a reflecting r_max through the C ABI,
against the same closed forms as the oracle (tests/test_oracle_axis_fold.py:
quiet loads, the cylinder's mirror geometry integrated by quadrature).

FP64 runs the oracle's cases at the oracle's bounds.  The mixed path
truncates each current contribution once to a fixed-point unit of 2^-30 of
its tile's bound (a one-cell move of the tile's heaviest particle), so a
particle moving 1e-4 cells per step with a few thousandths of the heaviest
weight (the quiet load's innermost radii) carries errors of a few percent;
its cases move the particles 30 to 100 times faster (drift 1e-2 c, rotation
omega dt = 2e-3), where the closed forms still hold to 3e-5, and meet 1e-4
of the row's unit, well inside the fold terms themselves (5e-3 of the row
for m = 1, 1e-3 for m = 0, 0.13 at r_max)."""
import pytest

from cylindrical_epoch_amd.host import Simulation
from gpu_helpers import log
from test_oracle_axis_fold import (BC_FIELD, BC_PART, DT, DX, DY, M, NX, NY, check_m0_rotation, check_m1_drift,
                                   check_rmax_fold)
from cylindrical_epoch_amd import deck

pytestmark = pytest.mark.gpu


def gpu_deposit(precision):
    def deposit(p, ny=NY):
        sim = Simulation(NX, ny, M, DX, DY, DT, bc_field=BC_FIELD, bc_particle=BC_PART, precision=precision)
        sid = sim.set_species(-deck.Q0, deck.M0, False, capacity=len(p["x"]))
        sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
        sim.push_particles()
        sim.current_finish()
        arr = {c: sim.download_field(c) for c in ("jt", "jr")}
        sim.close()
        return {c: (lambda m, j, c=c: arr[c][m][5 - 1 + j, 5:5 + NX].mean()) for c in ("jt", "jr")}
    return deposit


# (tolerance, speed) per case: m = 1 drift, m = 0 rotation, r_max drift
CASES = {"fp64": ((3e-5, 1e-4), (1e-5, 1e-4), (3e-5, 1e-4)),
         "mixed": ((1e-4, 1e-2), (1e-4, 2e-3), (1e-4, 1e-2))}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_jtheta_fold_m1_uniform_drift(precision):
    log("axis fold m=1 drift", precision)
    tol, beta = CASES[precision][0]
    check_m1_drift(gpu_deposit(precision), tol=tol, beta=beta)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_jtheta_fold_m0_rigid_rotation(precision):
    log("axis fold m=0 rotation", precision)
    tol, beta = CASES[precision][1]
    check_m0_rotation(gpu_deposit(precision), tol=tol, beta=beta)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_rmax_reflection_fold(precision):
    log("r_max fold m=1 drift", precision)
    tol, beta = CASES[precision][2]
    check_rmax_fold(gpu_deposit(precision), tol=tol, beta=beta)
