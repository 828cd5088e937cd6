"""The focused Gaussian beam of example_decks/gaussian_pulse.deck through
the C ABI, FP64 and mixed: the on-axis peak field is the focal field
sqrt(2 I / (eps0 c)) = 8.68e10 V/m to 3 %, between 7 and 11 um."""
import pytest

from cylindrical_epoch_amd import deck
from cylindrical_epoch_amd.host import Simulation
from gpu_helpers import log
from test_oracle_gaussian_beam import BC_FIELD, BC_PART, DX, DY, M, NX, NY, axis_peak, check_focus

pytestmark = pytest.mark.gpu


class SimGrid:
    def __init__(self, precision):
        self.dt = deck.set_dt(DX, DY, 0.95)
        self.sim = Simulation(NX, NY, M, DX, DY, self.dt, bc_field=BC_FIELD, bc_particle=BC_PART,
                              precision=precision)
        self.sim.setup_field_boundaries()
        self.ng = self.sim.ng

    def step(self, src):
        self.sim.step(src_xmin=src)

    def et_axis(self):
        return self.sim.download_field("et")[1][self.ng - 1]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gaussian_beam_focus_gpu(precision):
    log("Gaussian beam", precision)
    g = SimGrid(precision)
    peak, x = axis_peak(g)
    i = int(peak.argmax())
    log(f"  peak {peak[i]:.4e} V/m at x = {x[i] * 1e6:.2f} um")
    check_focus(peak, x)
    g.sim.close()
