"""The vacuum-mode checks of the FDTD (exact discrete dispersion of the m = 1
plane wave, second-order convergence of TM/TE Bessel modes and m = 0 Bessel
beams, the m = 1 B_r axis row as the reference writes it) run on the device
through the C ABI, FP64 and mixed.  The mixed path keeps the interior fields
in FP32, so its round-off floor is ~1e-6 of the amplitude instead of
~1e-15."""
import pytest

from cylindrical_epoch_amd import deck
from cylindrical_epoch_amd.host import Simulation
from gpu_helpers import log
from test_oracle_fdtd_modes import (BC_FIELD, BC_PART, DX, DY, KNOWN, beam_errors, mode_errors,
                                    plane_wave_check, second_order_check)

pytestmark = pytest.mark.gpu

C = deck.C_LIGHT


class SimGrid:
    def __init__(self, nx, ny, M, precision):
        self.dt = deck.set_dt(DX, DY, 0.95)
        self.sim = Simulation(nx, ny, M, DX, DY, self.dt, bc_field=BC_FIELD, bc_particle=BC_PART,
                              precision=precision)
        self.ng, self.shape = self.sim.ng, self.sim.shape
        self.x_grid_min = self.sim.x_min + DX / 2

    def load(self, fields):
        for name, a in fields.items():
            self.sim.upload_field(name, a)
        self.sim.setup_field_boundaries()

    def step(self):
        self.sim.step()

    def get(self, name):
        return self.sim.download_field(name)


def _make(precision, grids):
    def make(nx, ny, M):
        g = SimGrid(nx, ny, M, precision)
        grids.append(g)
        return g
    return make


@pytest.fixture
def grids():
    made = []
    yield made
    for g in made:
        g.sim.close()


PRECISIONS = ["fp64", "mixed"]


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("pol", ["y", "z"])
def test_plane_wave_gpu(precision, pol, grids):
    log("m = 1 plane wave", precision, pol)
    plane_wave_check(_make(precision, grids), pol, tol=1e-13 if precision == "fp64" else 2e-5)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("kind,m", [("TM", 0), ("TM", 1), ("TE", 0), ("TE", 1), ("TM", 2)])
def test_bessel_modes_gpu(precision, kind, m, grids):
    log("Bessel mode", precision, kind, m)
    make = _make(precision, grids)
    skip = KNOWN.get((kind, m), ())
    coarse = mode_errors(make, kind, m, 0.15, skip=skip)[0]
    fine = mode_errors(make, kind, m, 0.075, skip=skip)[0]
    log("  coarse", {k: f"{v:.1e}" for k, v in coarse.items()}, "fine", {k: f"{v:.1e}" for k, v in fine.items()})
    second_order_check(coarse, fine)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("kind", ["TM", "TE"])
def test_bessel_beam_gpu(precision, kind, grids):
    log("m = 0 Bessel beam", precision, kind)
    make = _make(precision, grids)
    second_order_check(beam_errors(make, kind, 32, 0.15), beam_errors(make, kind, 64, 0.075), bound=1.5e-3)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_m1_axis_br_row_gpu(precision, grids):
    log("m = 1 B_r axis row", precision)
    _, st = mode_errors(_make(precision, grids), "TM", 1, 0.15)
    br = st.get("br")
    assert not br[1][st.ng - 1, :].any()
    assert abs(br[1][st.ng, st.ng]) * C / 1.0e9 > 0.05
