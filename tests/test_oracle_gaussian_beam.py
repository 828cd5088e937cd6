"""The reference's focused Gaussian beam (example_decks/gaussian_pulse.deck;
the user manual's validation reports a peak E_y of about 8.6e10 V/m).  A CW
beam of 1 um wavelength, focused 10 um from the x_min laser boundary with a
1.5 um intensity FWHM and 1e15 W/cm^2 peak intensity, is launched through
the boundary with the deck's curvature and Gouy phase on a 500 x 100 grid,
two modes, open boundaries elsewhere.  On the axis, |E_y| = |E_theta,1|; its
largest value once the beam has filled the box must be the focal field
sqrt(2 I / (eps0 c)) = 8.68e10 V/m to 3 %, reached between 7 and 11 um.
No GPU needed."""
import math

import numpy as np
import pytest

C = 2.99792458e8
EPS0 = 8.854187817620389850536563031710750e-12
NX, NY, M = 500, 100, 2
DX, DY = 20.0e-6 / NX, 5.0e-6 / NY
BC_FIELD = ("simple_laser", "simple_outflow", "other", "simple_outflow")
BC_PART = ("open", "open", "other", "open")
LAM, I_FWHM, I_PEAK, FOCUS = 1.0e-6, 1.5e-6, 1.0e15, 10.0e-6
T_END, T_SAMPLE = 70e-15, 55e-15


def focal_field():
    return math.sqrt(2.0 * I_PEAK * 1.0e4 / (EPS0 * C))


class Beam:
    """The deck's laser block evaluated on the boundary rows r = |y(ir)|,
    ir = 0..ny (laser.f90:444-461): amplitude from the boundary intensity,
    Gaussian profile of the boundary waist, curvature and Gouy phase."""

    def __init__(self):
        k = 2 * math.pi / LAM
        w0 = I_FWHM / math.sqrt(2.0 * math.log(2.0))
        zr = math.pi * w0 ** 2 / LAM
        wb = w0 * math.sqrt(1.0 + (FOCUS / zr) ** 2)
        rc = FOCUS * (1.0 + (zr / FOCUS) ** 2)
        gouy = math.atan(-FOCUS / rc)
        i_b = I_PEAK * (w0 / wb) ** 2
        self.amp = math.sqrt(2.0 * i_b * 1.0e4 / (EPS0 * C))
        self.omega = k * C
        r = np.abs((np.arange(NY + 1) - 1) * DY + DY / 2)
        self.profile = np.exp(-(r / wb) ** 2)
        self.phase = k * r ** 2 / (2.0 * rc) - gouy

    def sources(self, t):
        src = np.zeros(2 * (NY + 1))
        src[:NY + 1] = self.amp * self.profile * np.sin(self.omega * t + self.phase)
        return src


def axis_peak(grid):
    """Largest |E_theta,1| on the axis row after T_SAMPLE, per x column, and
    the x of each column."""
    beam = Beam()
    ng, dt = grid.ng, grid.dt
    peak = np.zeros(NX)
    for n in range(int(T_END / dt)):
        t = (n + 1) * dt
        grid.step(beam.sources(t))
        if t > T_SAMPLE:
            peak = np.maximum(peak, np.abs(grid.et_axis()[ng:ng + NX]))
    x = (np.arange(NX) + 0.5) * DX
    return peak, x


def check_focus(peak, x):
    i = int(np.argmax(peak))
    assert peak[i] == pytest.approx(focal_field(), rel=0.03, abs=0), peak[i]
    assert 7e-6 < x[i] < 11e-6, x[i]


class OracleGrid:
    def __init__(self, oracle):
        self.dt = oracle.set_dt(DX, DY, 0.95)
        self.st = oracle.State(NX, NY, M, DX, DY, self.dt, bc_field=BC_FIELD, bc_particle=BC_PART)
        self.st.setup_field_boundaries()
        self.ng = self.st.g.ng

    def step(self, src):
        self.st.step(src_xmin=src)

    def et_axis(self):
        return self.st.arr["et"][1][self.ng - 1]


def test_gaussian_beam_focus(oracle):
    peak, x = axis_peak(OracleGrid(oracle))
    check_focus(peak, x)
