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


# ------------------------------------------------- r_max outflow: static fields
def static_field_run(oracle, steps=200, ny=40):
    """Loads a uniform B_x,0 on every row (ghosts included) with an outflow
    r_max and periodic x, steps it, and returns the largest change of any
    component on rows 0..ny relative to the field."""
    dt = oracle.set_dt(DX, DY, 0.95)
    st = oracle.State(8, ny, 2, DX, DY, dt, bc_field=("periodic", "periodic", "other", "simple_outflow"),
                      bc_particle=("periodic", "periodic", "other", "open"))
    B0 = 1.0e9 / C
    st.arr["bx"][0][:, :] = B0
    st.setup_field_boundaries()
    start = {n: st.arr[n].copy() for n in ("ex", "er", "et", "bx", "br", "bt")}
    ng = st.g.ng
    rows, cols = slice(ng - 1, ng + ny), slice(ng, ng + 8)
    worst = 0.0
    for _ in range(steps):
        st.step()
        for n, a in start.items():
            scale = B0 if n[0] == "b" else B0 * C
            worst = max(worst, np.abs(st.arr[n][:, rows, cols] - a[:, rows, cols]).max() / scale)
    return worst


def test_rmax_outflow_keeps_a_uniform_axial_b(oracle):
    """A uniform B_x,0 (a solenoid's field) is a static vacuum solution, and
    the r_max outflow's B_x update (laser.f90:664-673) is a discretised
    one-way wave equation, so it must hold it: its four B_x terms sum to
    (c + ly) B0 = B0 / sum_x, and the E_theta, B_r, E_r and J terms vanish.
    Every row 0..ny of every component stays at its value to round-off over
    200 steps.  (A uniform transverse B is not static here: the clamp of
    the generic field boundary zeroes B_r on the r_max row, boundary.F90:
    1449-1476.)"""
    assert static_field_run(oracle) < 1e-13


def test_rmax_boundary_rows_hold_their_previous_value_when_updated(oracle):
    """update_b_field runs to row ny - 1 on a rank with r_max
    (fields.f90:199), so row ny of B_x and B_theta is written only by the
    boundaries.  When outflow_bcs_r_max runs, bxm(1:nx-1, ny) therefore
    still equals its *_old copy taken in update_eb_fields_half
    (fields.f90:316-337), and the reference's bxm_old(:, ny) term could read
    the live array with the same result.  For btm(1:nx, ny) that holds only
    while no x boundary writes the corner (nx, ny) first: with an x_max
    outflow it does (test_rmax_outflow_time_levels)."""
    dt = oracle.set_dt(DX, DY, 0.95)
    ny = 30
    st = oracle.State(8, ny, 2, DX, DY, dt, bc_field=("periodic", "periodic", "other", "simple_outflow"),
                      bc_particle=("periodic", "periodic", "other", "open"))
    rng = np.random.default_rng(3)
    for n in ("ex", "er", "et", "bx", "br", "bt"):
        st.arr[n][...] = (rng.standard_normal(st.shape) + 1j * rng.standard_normal(st.shape)) * (
            1e9 if n[0] == "e" else 1e9 / C)
    st.setup_field_boundaries()
    ng = st.g.ng
    cols = slice(ng, ng + 8)
    for _ in range(3):
        st.update_eb_fields_half()
        st.push_particles()
        st.current_finish()
        st.call("orc_update_b_field")          # the first part of update_eb_fields_final
        for n in ("bx", "bt"):
            assert np.array_equal(st.arr[n][:, ny + ng - 1, cols], st.arr[n + "_old"][:, ny + ng - 1, cols])
        st.call("orc_bfield_final_bcs", None, None)
        st.call("orc_update_e_field")
        st.call("orc_efield_bcs")


def test_rmax_outflow_time_levels(oracle):
    """The r_max outflow (laser.f90:637-690) is the one-way wave equation
    discretised between the start of the step (t^n: the *_old copies taken
    in update_eb_fields_half) and its end (t^{n+1}), with E at t^{n+1/2}.
    For m = 0 and no current, on every column it updates,

      (c + ly) Bx_N' + (c - ly) Bx_N-1' + (ly - c) Bx_N + (-c - ly) Bx_N-1
          + c dt/r E_theta,N-1 - lx/2 (dBr' + dBr) = 0,
      (c + ly + d) Bt_N' + (c - ly + d) Bt_N-1' + (ly - c + d) Bt_N
          + (-c - ly + d) Bt_N-1 + lx/(2c) dEr = 0,

    primes at t^{n+1}, d = dt c^2 / (4 r), dBr and dEr the x differences
    of B_r on row N-1 and of the two-row sum of E_r.  Checked at every
    column with outflow boundaries on x_min, x_max and r_max, so the corner
    (nx, N), which the x_max boundary writes first in the same call, must
    use its value from the start of the step as well."""
    nx, ny, dx, dy = 24, 20, 1e-7, 1e-7
    dt = oracle.set_dt(dx, dy, 0.95)
    st = oracle.State(nx, ny, 1, dx, dy, dt, bc_field=("simple_outflow", "simple_outflow", "other", "simple_outflow"),
                      bc_particle=("open", "open", "other", "open"))
    ng = st.g.ng
    j = np.arange(1 - ng, ny + ng + 1)
    i = np.arange(1 - ng, nx + ng + 1)
    r = j * dy
    st.arr["ex"][0][:, :] = 1e9 * np.exp(-((r[:, None] - 15 * dy) / (3 * dy)) ** 2) * np.cos(0.2 * i)[None, :]
    st.arr["bx"][0][:, :] = 1e9 / C * np.exp(-((r[:, None] - 14.5 * dy) / (3 * dy)) ** 2) * np.sin(0.3 * i)[None, :]
    st.setup_field_boundaries()
    c, lx, ly = C, dt * C * C / dx, dt * C * C / dy
    inv_r = 1.0 / ((ny - 1.5) * dy + st.g.y_grid_min)
    d = 0.25 * dt * C * C * inv_r
    N, Nm = ny + ng - 1, ny + ng - 2          # array rows of r-index ny and ny - 1
    col = lambda a, b: slice(a + ng - 1, b + ng)
    for _ in range(12):
        old = {n: st.arr[n][0].copy() for n in ("bx", "br", "bt")}
        st.update_eb_fields_half()
        half = {n: st.arr[n][0].copy() for n in ("er", "et")}
        st.push_particles()
        st.current_finish()
        st.update_eb_fields_final()
        new = {n: st.arr[n][0] for n in ("bx", "br", "bt")}
        bx = col(1, nx - 1)
        bxp = col(2, nx)
        res_x = ((c + ly) * new["bx"][N, bx] + (c - ly) * new["bx"][Nm, bx] + (ly - c) * old["bx"][N, bx]
                 + (-c - ly) * old["bx"][Nm, bx] + c * dt * inv_r * half["et"][Nm, bx]
                 - 0.5 * lx * (new["br"][Nm, bxp] - new["br"][Nm, bx] + old["br"][Nm, bxp] - old["br"][Nm, bx]))
        bt, btm1 = col(1, nx), col(0, nx - 1)
        res_t = ((c + ly + d) * new["bt"][N, bt] + (c - ly + d) * new["bt"][Nm, bt] + (ly - c + d) * old["bt"][N, bt]
                 + (-c - ly + d) * old["bt"][Nm, bt]
                 + 0.5 * lx / c * (half["er"][N, bt] + half["er"][Nm, bt] - half["er"][N, btm1] - half["er"][Nm, btm1]))
        scale = c * np.abs(new["bt"]).max()
        assert np.abs(res_x).max() < 1e-12 * scale and np.abs(res_t).max() < 1e-12 * scale
    for n in ("bx", "bt"):
        assert np.abs(st.arr[n][0][N, col(1, nx)]).max() > 1e-3 * np.abs(st.arr[n][0]).max()
