"""The FDTD of the oracle (update_e_field / update_b_field, src/fields.f90:
53-312, with their r = 0 rows) against vacuum solutions of Maxwell's
equations in modes (a field f(theta) = Re sum_m f_m e^{-i m theta}).

* A field uniform across the axis and standing along x is the m = 1 mode
  E_r,1 = E0, E_theta,1 = -i E0 (E along y).  For it the scheme reduces to
  the 1-D Yee scheme, whose standing wave E0 cos(k x) cos(w t_n) with
  cos(w dt) = 1 - 2 (c dt / dx)^2 sin^2(k dx / 2) is exact: the oracle must
  reproduce it to round-off on every row the r_max boundary has not reached,
  the r = 0 row of E_theta included.
* TM_m (E_x,m = E0 J_m(k r)) and TE_m (B_x,m = B0 J_m(k r)) standing modes
  with no x dependence are continuum solutions for all r; rows that the
  r_max boundary has not reached follow them to the scheme's O((k dr)^2)
  error, so halving k dr must cut the error about four-fold.

Two r = 0 rows of the reference do not converge to the continuum, and the
tests pin them as they are:
* the m = 1 B_r axis row (Lifschitz eq. 22, fields.f90:270-273) adds
  i / dy * exm(:, 0, 1), a row that update_e_field sets to zero for m = 1
  (fields.f90:140), so with no x dependence B_r,1 stays zero on the axis
  while the continuum value is i m / (r w) E_x,1 -> i E0 / (2 c) sin(w t);
* for m > 1, E_r's first row is set to a ninth of the second
  (fields.f90:165), the r^2 ratio of m = 3; an m = 2 field grows like r
  there, so TE_2 converges at first order near the axis.
No GPU needed."""
import numpy as np
import pytest
from scipy import special

C = 2.99792458e8
NX, NY, DX, DY = 10, 64, 1e-7, 1e-7
BC_FIELD = ("periodic", "periodic", "other", "conduct")
BC_PART = ("periodic", "periodic", "other", "reflect")
R_HALF = ("er", "bx", "bt")      # stored at r = (j - 1/2) dr; the others at r = j dr
ROWS = 21                        # rows 0..20: out of reach of r_max for 20 steps


class OracleGrid:
    """The field arrays and step of the oracle, behind the interface the
    checks below use (the GPU tests wrap the C ABI the same way)."""

    def __init__(self, oracle, nx, ny, M, dx=DX, dy=DY):
        self.st = oracle.State(nx, ny, M, dx, dy, oracle.set_dt(dx, dy, 0.95), bc_field=BC_FIELD,
                               bc_particle=BC_PART)
        self.ng, self.dt, self.shape = self.st.g.ng, self.st.g.dt, self.st.shape
        self.x_grid_min = self.st.g.x_grid_min

    def load(self, fields):
        for name, a in fields.items():
            self.st.arr[name][...] = a
        self.st.setup_field_boundaries()

    def step(self):
        self.st.step()

    def get(self, name):
        return self.st.arr[name]


def _oracle_grid(oracle):
    return lambda nx, ny, M: OracleGrid(oracle, nx, ny, M)


# ---------------------------------------------------------------- plane wave
def plane_wave_check(make, pol, tol=1e-13):
    """Runs the m = 1 standing plane wave for 15 steps; tol is relative to
    the amplitude."""
    nx, ny = 16, 48
    st = make(nx, ny, 2)
    ng, dt = st.ng, st.dt
    k = 2 * np.pi / (nx * DX)
    x = st.x_grid_min + (np.arange(1 - ng, nx + ng + 1) - 1) * DX
    E0 = 1.0e9
    # E along y: E_r = E0 cos(theta), E_theta = -E0 sin(theta); along z:
    # E_r = E0 sin(theta), E_theta = E0 cos(theta)
    pr, pt = (1.0, -1j) if pol == "y" else (1j, 1.0)
    init = {n: np.zeros(st.shape, complex) for n in ("ex", "er", "et", "bx", "br", "bt")}
    init["er"][1][:, :] = pr * E0 * np.cos(k * x)[None, :]
    init["et"][1][:, :] = pt * E0 * np.cos(k * x)[None, :]
    st.load(init)
    w = np.arccos(1.0 - 2.0 * (C * dt / DX) ** 2 * np.sin(k * DX / 2) ** 2) / dt
    cols = slice(ng, ng + nx)
    for n in range(1, 16):
        st.step()
        f = {name: st.get(name) for name in ("ex", "er", "et", "bx", "br", "bt")}
        want = E0 * np.cos(k * x[cols]) * np.cos(w * n * dt)
        er = f["er"][1][ng:ng - 1 + 14, cols]         # rows 1..13
        et = f["et"][1][ng - 1:ng - 1 + 14, cols]     # rows 0..13, the axis row first
        assert np.abs(er - pr * want).max() < tol * E0, n
        assert np.abs(et - pt * want).max() < tol * E0, n
        assert np.abs(f["ex"][1][ng - 1:ng - 1 + 14, cols]).max() < tol * E0
        assert not np.any(f["ex"][0]) and not np.any(f["bx"][0])
        # B is uniform across the axis too: rows 0..13 of B_r, B_theta (row 0
        # of B_theta at r = -dr/2) equal row 7; B_x stays zero
        for name in ("br", "bt"):
            b = f[name][1][ng - 1:ng - 1 + 14, cols]
            assert np.abs(b - b[7]).max() < tol * E0 / C, (n, name)
        assert np.abs(f["bx"][1][ng - 1:ng - 1 + 14, cols]).max() < tol * E0 / C


@pytest.mark.parametrize("pol", ["y", "z"])
def test_m1_plane_wave_exact_discrete_dispersion(oracle, pol):
    plane_wave_check(_oracle_grid(oracle), pol)


# ---------------------------------------------------------------- Bessel modes
def _mode(kind, m, k, r, t, A):
    """TM_m / TE_m standing vacuum mode with no x dependence (Maxwell in the
    reference's convention: d/dtheta -> -i m)."""
    w = C * k
    J, Jp = special.jv(m, k * r), special.jvp(m, k * r)
    safe = np.where(r == 0, 1.0, r)
    J_r = np.where(r == 0, k / 2 if m == 1 else 0.0, J / safe)    # J_m(k r) / r
    zero = np.zeros_like(r, dtype=complex)
    s, c = np.sin(w * t), np.cos(w * t)
    if kind == "TM":
        return dict(ex=A * J * c + zero, br=1j * m / w * A * J_r * s + zero, bt=A * k / w * Jp * s + zero,
                    er=zero, et=zero, bx=zero)
    return dict(bx=A * J * c + zero, er=-1j * m * C * C / w * A * J_r * s + zero, et=-C * A * Jp * s + zero,
                ex=zero, br=zero, bt=zero)


def mode_errors(make, kind, m, kdy, steps=20, skip=()):
    """Largest error per component over rows 0..20 and the steps, in units of
    the E amplitude (B scaled by c).  Row 0 of E_r, B_x and B_theta sits at
    r = -dr/2, where the mode's own parity (J_m(-x) = (-1)^m J_m(x)) gives
    the reflected value; skip lists (component, row) pairs left out."""
    st = make(NX, NY, m + 1)
    ng, dt = st.ng, st.dt
    k = kdy / DY
    A = 1.0e9 if kind == "TM" else 1.0e9 / C
    src = "ex" if kind == "TM" else "bx"
    j = np.arange(1 - ng, NY + ng + 1)
    r = (j - 0.5) * DY if src in R_HALF else j * DY
    init = {n: np.zeros(st.shape, complex) for n in ("ex", "er", "et", "bx", "br", "bt")}
    init[src][m][:, :] = _mode(kind, m, k, r, 0.0, A)[src][:, None]
    st.load(init)
    worst = {}
    for n in range(1, steps + 1):
        st.step()
        for name in ("ex", "er", "et", "bx", "br", "bt"):
            rows = np.arange(0, ROWS)
            rows = rows[~np.isin(rows, [rw for nm, rw in skip if nm == name])]
            r = (rows - 0.5) * DY if name in R_HALF else rows * DY
            want = _mode(kind, m, k, r, n * dt, A)[name]
            got = st.get(name)[m][ng - 1 + rows, ng]
            scale = 1.0 if name[0] == "e" else C
            worst[name] = max(worst.get(name, 0.0), np.abs(got - want).max() * scale / 1.0e9)
    return worst, st


SECOND_ORDER = [("TM", 0), ("TM", 1), ("TM", 2), ("TM", 3), ("TE", 0), ("TE", 1), ("TE", 3)]


# rows that do not follow the continuum (module docstring): the m = 1 B_r
# axis row and the B_theta row reflected from it; E_r's m = 1 row 0, which
# the reference sets from E_theta's axis value before updating that value
KNOWN = {("TM", 1): (("br", 0), ("bt", 0)), ("TE", 1): (("er", 0),)}


def second_order_check(coarse, fine, bound=8e-4):
    for name in coarse:
        assert fine[name] < bound, (name, fine[name])
        if coarse[name] > 1e-4:
            assert coarse[name] / fine[name] > 3.5, (name, coarse[name], fine[name])


@pytest.mark.parametrize("kind,m", SECOND_ORDER)
def test_bessel_mode_second_order(oracle, kind, m):
    skip = KNOWN.get((kind, m), ())
    make = _oracle_grid(oracle)
    second_order_check(mode_errors(make, kind, m, 0.15, skip=skip)[0], mode_errors(make, kind, m, 0.075, skip=skip)[0])


X_HALF = ("ex", "br", "bt")      # stored half a cell below the others in x


def _beam(kind, kx, kr, r, x, t, A):
    """m = 0 Bessel beam E_x (TM) or B_x (TE) = A J_0(kr r) cos(kx x - w t),
    w = c sqrt(kx^2 + kr^2): a vacuum solution for all r and x."""
    w = C * np.hypot(kx, kr)
    J0, J1 = special.jv(0, kr * r), special.jv(1, kr * r)
    zero = np.zeros(np.broadcast(r, x).shape, dtype=complex)
    ph = kx * x - w * t
    if kind == "TM":
        return dict(ex=A * J0 * np.cos(ph) + zero, er=kx / kr * A * J1 * np.sin(ph) + zero,
                    bt=w / (C * C * kr) * A * J1 * np.sin(ph) + zero, et=zero, bx=zero, br=zero)
    return dict(bx=A * J0 * np.cos(ph) + zero, br=kx / kr * A * J1 * np.sin(ph) + zero,
                et=-w / kr * A * J1 * np.sin(ph) + zero, ex=zero, er=zero, bt=zero)


def beam_errors(make, kind, nx, krdy, steps=20):
    st = make(nx, NY, 1)
    ng, dt = st.ng, st.dt
    kx, kr = 2 * np.pi / (nx * DX), krdy / DY
    A = 1.0e9 if kind == "TM" else 1.0e9 / C
    j = np.arange(1 - ng, NY + ng + 1)[:, None]
    xc = st.x_grid_min + (np.arange(1 - ng, nx + ng + 1) - 1)[None, :] * DX

    def at(name, rows, cols, t):
        r = (rows - 0.5) * DY if name in R_HALF else rows * DY
        x = xc[:, cols] - (0.5 * DX if name in X_HALF else 0.0)
        return _beam(kind, kx, kr, r, x, t, A)[name]

    all_cols = slice(None)
    init = {n: np.zeros(st.shape, complex) for n in ("ex", "er", "et", "bx", "br", "bt")}
    for name in init:
        init[name][0][:, :] = at(name, j, all_cols, 0.0)
    st.load(init)
    rows = np.arange(0, ROWS)[:, None]
    cols = slice(ng, ng + nx)
    worst = {}
    for n in range(1, steps + 1):
        st.step()
        for name in ("ex", "er", "et", "bx", "br", "bt"):
            got = st.get(name)[0][ng - 1 + rows[:, 0], cols]
            scale = 1.0 if name[0] == "e" else C
            err = np.abs(got - at(name, rows, cols, n * dt)).max() * scale / 1.0e9
            worst[name] = max(worst.get(name, 0.0), err)
    return worst


@pytest.mark.parametrize("kind", ["TM", "TE"])
def test_m0_bessel_beam_second_order(oracle, kind):
    """An m = 0 Bessel beam travelling along x: every component, with the x
    and r staggering, the r = 0 rows and the reflected row at r = -dr/2,
    converges at second order when dx and dr are halved together."""
    make = _oracle_grid(oracle)
    second_order_check(beam_errors(make, kind, 32, 0.15), beam_errors(make, kind, 64, 0.075), bound=1.5e-3)


def test_te2_first_row_rule(oracle):
    """TE_2: within 1e-2 and converging at about first order, the error
    coming from E_r's first row (a ninth of the second)."""
    make = _oracle_grid(oracle)
    coarse, _ = mode_errors(make, "TE", 2, 0.15)
    fine, _ = mode_errors(make, "TE", 2, 0.075)
    for name in ("er", "et", "bx"):
        assert fine[name] < 1e-2, (name, fine[name])
        assert coarse[name] / fine[name] > 2.0, (name, coarse[name], fine[name])


def test_te3_first_er_row(oracle):
    """TE_3: near the axis E_r,3 grows like r^2 (J_3(k r) / r), so setting
    E_r's first row (r = dr/2) to a ninth of the second (r = 3 dr / 2), the
    reference's rule for every m > 1 (fields.f90:165), is the continuum's
    own ratio there.  The row then follows the mode within 2e-5 of the E
    amplitude at k dr = 0.075, and its error falls about four-fold when k dr
    halves; a ratio of a quarter leaves 8e-5."""
    make = _oracle_grid(oracle)
    errs = []
    for kdy in (0.15, 0.075):
        st = make(NX, NY, 4)
        ng, dt = st.ng, st.dt
        k, A = kdy / DY, 1.0e9 / C
        j = np.arange(1 - ng, NY + ng + 1)
        init = {n: np.zeros(st.shape, complex) for n in ("ex", "er", "et", "bx", "br", "bt")}
        init["bx"][3][:, :] = _mode("TE", 3, k, (j - 0.5) * DY, 0.0, A)["bx"][:, None]
        st.load(init)
        worst = 0.0
        for n in range(1, 21):
            st.step()
            want = _mode("TE", 3, k, np.array([0.5 * DY]), n * dt, A)["er"][0]
            worst = max(worst, abs(st.get("er")[3][ng, ng] - want) / 1.0e9)
        errs.append(worst)
    assert errs[1] < 2e-5, errs
    assert errs[0] / errs[1] > 3.5, errs


def test_m1_er_reflected_row_lags_etheta_axis_row(oracle):
    """The m = 1 E_r row at r = -dr/2 is set from E_theta's axis value
    before that value is updated (fields.f90:144-147, in this order).  In the
    uniform standing wave E along y (E_r,1 = E_y, E_theta,1 = -i E_y, rows
    1.. exact, see the plane-wave test) the row therefore holds
    2 E_y(previous half update) - E_y(now) after each E update, where the
    other order would give E_y(now).  The two differ here by about 5 % of
    the amplitude per half update."""
    st = OracleGrid(oracle, 16, 48, 2)
    ng, nx = st.ng, 16
    k = 2 * np.pi / (nx * DX)
    x = st.x_grid_min + (np.arange(1 - ng, nx + ng + 1) - 1) * DX
    E0 = 1.0e9
    init = {n: np.zeros(st.shape, complex) for n in ("ex", "er", "et", "bx", "br", "bt")}
    init["er"][1][:, :] = E0 * np.cos(k * x)[None, :]
    init["et"][1][:, :] = -1j * E0 * np.cos(k * x)[None, :]
    st.load(init)
    cols = slice(ng, ng + nx)
    o = st.st
    prev = o.arr["er"][1][ng, cols].copy()          # row 1 = E_y before the first update
    lag = 0.0
    for n in range(12):
        for half in ("half", "final"):
            if half == "half":
                o.update_eb_fields_half()
            else:
                o.push_particles()
                o.current_finish()
                o.update_eb_fields_final()
            now = o.arr["er"][1][ng, cols].copy()
            row0 = o.arr["er"][1][ng - 1, cols]
            assert np.abs(row0 - (2 * prev - now)).max() < 1e-13 * E0, (n, half)
            lag = max(lag, np.abs(now - prev).max())
            prev = now
    assert lag > 0.03 * E0


def test_m1_axis_br_stays_zero_without_x_dependence(oracle):
    """TM_1 with no x dependence: B_r,1 on the axis row stays exactly zero
    (eq. 22 reads the zeroed E_x,1 axis row), although the continuum value
    reaches |E0| / (2 c) within the run."""
    _, st = mode_errors(_oracle_grid(oracle), "TM", 1, 0.15)
    ng, br = st.ng, st.get("br")
    assert not np.any(br[1][ng - 1, :])
    assert np.abs(br[1][ng, ng]) * C / 1.0e9 > 0.05     # row 1 is not zero
