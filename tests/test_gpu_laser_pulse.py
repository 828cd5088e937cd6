"""The checks of the laser and outflow boundaries through the C ABI, FP64
and mixed: the pulse from the x_min laser enters at the source's amplitude
and polarisation, delayed by x / c, and leaves through the x_max outflow
with under 1 % reflected; the r_max outflow absorbs a cylindrical pulse
that a conducting r_max keeps."""
import pytest

from cylindrical_epoch_amd import deck
from cylindrical_epoch_amd.host import Simulation
from gpu_helpers import log
from test_oracle_laser import BC_FIELD, BC_PART, DX, DY, NX, NY, RMAX_CASES, check_pulse, rmax_energy_left

pytestmark = pytest.mark.gpu


class SimGrid:
    def __init__(self, precision):
        self.dt = deck.set_dt(DX, DY, 0.95)
        self.sim = Simulation(NX, NY, 2, DX, DY, self.dt, bc_field=BC_FIELD, bc_particle=BC_PART,
                              precision=precision)
        self.sim.setup_field_boundaries()
        self.ng = self.sim.ng
        self.x_grid_min = self.sim.x_min + DX / 2

    def step(self, src):
        self.sim.step(src_xmin=src)
        self._cache = {}

    def probe(self, name, row, col):
        if name not in self._cache:
            self._cache[name] = self.sim.download_field(name)
        return self._cache[name][1][self.ng - 1 + row, self.ng - 1 + col]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("pol", ["y", "z"])
def test_laser_pulse_gpu(precision, pol):
    log("laser pulse", precision, pol)
    grid = SimGrid(precision)
    check_pulse(grid, pol, phase_tol=1e-12 if precision == "fp64" else 1e-6)
    grid.sim.close()


class SimBox:
    def __init__(self, nx, ny, M, ybc, precision):
        dt = deck.set_dt(DX, DY, 0.95)
        self.sim = Simulation(nx, ny, M, DX, DY, dt, bc_field=("periodic", "periodic", "other", ybc),
                              bc_particle=("periodic", "periodic", "other", "reflect"), precision=precision)
        self.ng, self.shape = self.sim.ng, self.sim.shape

    def load(self, fields):
        for name, a in fields.items():
            self.sim.upload_field(name, a)
        self.sim.setup_field_boundaries()

    def step(self):
        self.sim.step()

    def energy(self):
        return self.sim.mode_energy()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("m,comp", RMAX_CASES)
def test_rmax_outflow_absorbs_gpu(precision, m, comp):
    log("r_max outflow", precision, m, comp)
    boxes = []

    def make(nx, ny, M, ybc):
        boxes.append(SimBox(nx, ny, M, ybc, precision))
        return boxes[-1]
    left_out = rmax_energy_left(make, m, comp, "simple_outflow")
    left_cond = rmax_energy_left(make, m, comp, "conduct")
    log(f"  energy left: outflow {left_out:.2e}, conduct {left_cond:.3f}")
    assert left_out < 0.02 and left_cond > 0.95
    for b in boxes:
        b.sim.close()
