"""The laser and outflow boundaries (outflow_bcs_x_min / _x_max,
src/laser.f90:411-633) against the plane wave they are meant to launch and
absorb.  A source E_y(t) (source1) or E_z(t) (source2), uniform in r, is a
Gaussian-enveloped pulse at x_min; x_max is an outflow boundary.

* Away from the wall, the pulse inside the box is the m = 1 uniform field
  E_r,1 = E_y, E_theta,1 = -i E_y (E_r,1 = i E_z, E_theta,1 = E_z for
  source2), delayed by x / c: the trace at x = 29.5 dx follows the source to
  the Yee scheme's dispersion (3 % of the peak at 16 cells per wavelength)
  once a constant delay below one step is fitted.
* After the pulse has left through x_max, what comes back is the outflow
  boundary's reflection: below 1 % of the peak.
* The r_max outflow boundary (outflow_bcs_r_max, laser.f90:637-690)
  absorbs a cylindrical pulse launched near the axis: after 300 steps at
  most 2 % of the mode's field energy is left, where a conducting r_max
  keeps all of it.
No GPU needed."""
import numpy as np
import pytest

C = 2.99792458e8
NX, NY, DX, DY = 60, 200, 1e-7, 1e-7
BC_FIELD = ("simple_laser", "simple_outflow", "other", "conduct")
BC_PART = ("open", "open", "other", "reflect")
E0 = 1.0e10
LAM = 16 * DX
OMEGA = 2 * np.pi * C / LAM
TAU = LAM / C
T0 = 3 * TAU
ROW, COL = 3, 30          # probe: r = 3 dr, x = 29.5 dx (cell 30)


def source(t):
    return E0 * np.exp(-((t - T0) / TAU) ** 2) * np.sin(OMEGA * (t - T0))


class OracleGrid:
    def __init__(self, oracle):
        dt = oracle.set_dt(DX, DY, 0.95)
        self.st = oracle.State(NX, NY, 2, DX, DY, dt, bc_field=BC_FIELD, bc_particle=BC_PART)
        self.st.setup_field_boundaries()
        self.ng, self.dt, self.x_grid_min = self.st.g.ng, dt, self.st.g.x_grid_min

    def step(self, src):
        self.st.step(src_xmin=src)

    def probe(self, name, row, col):
        return self.st.arr[name][1][self.ng - 1 + row, self.ng - 1 + col]


def pulse_traces(grid, pol):
    """E_r,1 and E_theta,1 at the probe, one value per step, with the times."""
    nsteps = int((6 * TAU + 2.2 * NX * DX / C) / grid.dt)
    k = 0 if pol == "y" else 1
    er, et = [], []
    for n in range(nsteps):
        src = np.zeros(2 * (NY + 1))
        src[k * (NY + 1):(k + 1) * (NY + 1)] = source((n + 1) * grid.dt)
        grid.step(src)
        er.append(grid.probe("er", ROW, COL))
        et.append(grid.probe("et", ROW, COL))
    return (np.arange(nsteps) + 1) * grid.dt, np.array(er), np.array(et)


def check_pulse(grid, pol, tol=0.03, phase_tol=1e-12):
    t, er, et = pulse_traces(grid, pol)
    dt = grid.dt
    x = grid.x_grid_min + (COL - 1) * DX
    # E_r,1 = pr * E_y (or E_z), E_theta,1 = pt * the same
    pr, pt = (1.0, -1j) if pol == "y" else (1j, 1.0)
    ey = (er / pr).real
    assert np.abs(er / pr - ey).max() < phase_tol * E0             # the polarisation's phase
    assert np.abs(et - pt * ey).max() < 0.01 * E0                  # E_theta follows E_r
    shifts = np.linspace(-dt, dt, 201)
    err, s = min((np.abs(ey - source(t - x / C - s)).max(), s) for s in shifts)
    assert err < tol * E0, (err / E0, s / dt)
    assert abs(np.abs(ey).max() / np.abs(source(t - x / C)).max() - 1) < 0.02
    gone = t > T0 + 3 * TAU + (2 * NX * DX - x) / C + 5 * dt
    assert gone.sum() > 20
    assert np.abs(ey[gone]).max() < 0.01 * E0


@pytest.mark.parametrize("pol", ["y", "z"])
def test_laser_pulse_enters_and_leaves(oracle, pol):
    check_pulse(OracleGrid(oracle), pol)


# ---------------------------------------------------------------- r_max outflow
RMAX_CASES = [(0, "ex"), (1, "ex"), (1, "bx"), (2, "ex")]


def rmax_energy_left(make, m, comp, ybc, steps=300):
    """Field energy of mode m left after `steps`, as a fraction of the start,
    for a ring pulse of `comp` (Gaussian in r at 8 dr, 3 dr wide, uniform in
    x) under an r_max boundary ybc."""
    nx, ny = 10, 60
    g = make(nx, ny, m + 1, ybc)
    j = np.arange(1 - g.ng, ny + g.ng + 1)
    r = (j - 0.5) * DY if comp in ("er", "bx", "bt") else j * DY
    prof = np.exp(-((r - 8 * DY) / (3 * DY)) ** 2) * (1e9 if comp[0] == "e" else 1e9 / C) * np.sign(r) ** m
    f = np.zeros(g.shape, complex)
    f[m][:, :] = prof[:, None]
    g.load({comp: f})
    e0 = sum(e[m] for e in g.energy())
    for _ in range(steps):
        g.step()
    return sum(e[m] for e in g.energy()) / e0


class OracleBox:
    def __init__(self, oracle, nx, ny, M, ybc):
        dt = oracle.set_dt(DX, DY, 0.95)
        self.st = oracle.State(nx, ny, M, DX, DY, dt, bc_field=("periodic", "periodic", "other", ybc),
                               bc_particle=("periodic", "periodic", "other", "reflect"))
        self.ng, self.shape = self.st.g.ng, self.st.shape

    def load(self, fields):
        for name, a in fields.items():
            self.st.arr[name][...] = a
        self.st.setup_field_boundaries()

    def step(self):
        self.st.step()

    def energy(self):
        return self.st.mode_energy()


@pytest.mark.parametrize("m,comp", RMAX_CASES)
def test_rmax_outflow_absorbs(oracle, m, comp):
    def make(nx, ny, M, ybc):
        return OracleBox(oracle, nx, ny, M, ybc)
    assert rmax_energy_left(make, m, comp, "simple_outflow") < 0.02
    assert rmax_energy_left(make, m, comp, "conduct") > 0.95
