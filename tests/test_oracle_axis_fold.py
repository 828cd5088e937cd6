"""The fold of the azimuthal current's r < 0 ghost rows
(current_bcs_r_min_final, src/boundary.F90:1917-1931), pinned by the
geometry of a cylinder on a quiet load.  No GPU needed.

The deposit spreads a particle's current over rows with a quadratic spline
along the signed radius through the particle.  A row at signed radius -a is
the point at radius a on the opposite side of the axis (angle theta + pi),
where the azimuthal unit vector is reversed and e^{i m theta} gains (-1)^m.
So row -a folds into row +a with the factor -(-1)^m for J_theta (as for
J_r), and the J_theta row j (radius j dr, ring volume 2 pi j dr^2 dx) of a
load of density n(r) holds, in the limit of many particles,

    J(j) = (1 / j) [ int_0^inf n(r) v(r) r S(r - j) dr
                     - (-1)^m int_0^inf n(r) v(r) r S(r + j) dr ] / n

in cell units, with S the quadratic spline of unit integral.  The two
integrals are evaluated here by adaptive quadrature of the spline itself,
not by the deposit's loops.  Two cases:

* m = 1, a uniform transverse drift v_y: J_theta,1 = -i n q v_y times the
  bracket with v = 1.  Row 1 gets 1.00521 (the fold adds 0.0026, and the
  part of row 1's spline at r < 0 is missing); a reversed fold sign gives
  1.00000.
* m = 0, a rigid rotation v_theta = omega r: row 1 gets 1.24896 n q omega dr;
  a reversed fold sign gives 1.25000.

The load is quiet: radii at the midpoints of dr/64 steps with weights
proportional to r (uniform density), eight equally spaced angles, two
positions per cell in x, so the deposit's sums reproduce the integrals to
about 1e-5.  Rows 2 and 3 (no fold) check the load against the same
integrals."""
import math

import numpy as np
import pytest
from scipy import integrate

from cylindrical_epoch_amd import deck

NX, NY, M = 10, 8, 2
DX = DY = 1.0e-7
DENSITY = 1.0e24
SUB_R = 64        # radial sub-steps per cell
N_THETA = 8
R_CELLS = 6       # load out to 6 dr


def spline(d):
    """Quadratic B-spline of unit integral and unit cell width."""
    a = abs(d)
    if a < 0.5:
        return 0.75 - a * a
    if a < 1.5:
        return 0.5 * (1.5 - a) ** 2
    return 0.0


def row_bracket(j, m, power):
    """(1/j) [int_0 r^power S(r - j) dr - (-1)^m int_0 r^power S(r + j) dr]
    over the spline's support, by adaptive quadrature with the spline's knots
    as break points."""
    def part(centre):
        lo, hi = max(0.0, centre - 1.5), max(0.0, centre + 1.5)
        if hi <= lo:
            return 0.0
        pts = [p for p in (centre - 0.5, centre + 0.5) if lo < p < hi]
        v, _ = integrate.quad(lambda r: r ** power * spline(r - centre), lo, hi, points=pts or None,
                              epsabs=1e-14, epsrel=1e-13)
        return v
    return (part(j) - (-1) ** m * part(-j)) / j


def quiet_load(velocity):
    """A uniform plasma (density DENSITY) on rows 0..R_CELLS: radii at the
    midpoints of dr/SUB_R steps, N_THETA equally spaced angles, two x
    positions per cell; velocity(r, theta) -> (vx, vy, vz) [m/s]."""
    rk = (np.arange(R_CELLS * SUB_R) + 0.5) * (DY / SUB_R)
    th = 2.0 * math.pi * (np.arange(N_THETA) + 0.25) / N_THETA
    xs = (np.arange(2 * NX) + 0.5) * (DX / 2)
    X, R, T = np.meshgrid(xs, rk, th, indexing="ij")
    X, R, T = X.ravel(), R.ravel(), T.ravel()
    # weight = density * volume represented: (dx/2) * (dr/SUB_R) * r * (2 pi / N_THETA)
    w = DENSITY * (DX / 2) * (DY / SUB_R) * R * (2.0 * math.pi / N_THETA)
    vx, vy, vz = velocity(R, T)
    v2 = vx * vx + vy * vy + vz * vz
    gam = 1.0 / np.sqrt(1.0 - v2 / deck.C_LIGHT ** 2)
    mc = deck.M0
    return dict(x=X, y=R * np.cos(T), z=R * np.sin(T), px=gam * mc * vx, py=gam * mc * vy,
                pz=gam * mc * vz, w=w)


BC_FIELD = ("periodic", "periodic", "other", "conduct")
BC_PART = ("periodic", "periodic", "other", "reflect")
DT = deck.set_dt(DX, DY, 0.95)


def oracle_deposit(oracle):
    """deposit(p, ny) -> jt(m, j), jr(m, j): one push and current_finish of
    the oracle; the getters return a row's mean over x."""
    def deposit(p, ny=NY):
        st = oracle.State(NX, ny, M, DX, DY, DT, bc_field=BC_FIELD, bc_particle=BC_PART)
        st.add_species(-deck.Q0, deck.M0, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
        st.push_particles()
        st.current_finish()
        ng = st.g.ng
        return {c: (lambda m, j, c=c: st.arr[c][m][ng - 1 + j, ng:ng + NX].mean()) for c in ("jt", "jr")}
    return deposit


def test_spline_brackets_closed_form():
    """The quadrature against hand integrals: row 1's fold term for v = 1
    is (1/2) int_0^1/2 r (1/2 - r)^2 dr = 1/384, and the missing r < 0 part
    of row 1's spline is the same; away from the axis the bracket is exactly
    1 (v = 1) and j + 1/(4 j) (v = r, the spline's variance 1/4)."""
    assert row_bracket(1, 1, 1) == pytest.approx(1.0 + 2.0 / 384.0, rel=1e-12, abs=0)
    assert row_bracket(1, 0, 2) == pytest.approx(1.25 - 2.0 / 1920.0, rel=1e-12, abs=0)
    for j in (2, 3):
        assert row_bracket(j, 1, 1) == pytest.approx(1.0, rel=1e-12, abs=0)
        assert row_bracket(j, 0, 2) == pytest.approx(j + 0.25 / j, rel=1e-12, abs=0)


def check_m1_drift(deposit, tol=3e-5, beta=1.0e-4):
    """J_theta,1 of a transverse drift v_y = beta c on rows 1..3 against the
    bracket; row 1 carries the folded r < 0 ghost row."""
    vy = beta * deck.C_LIGHT
    p = quiet_load(lambda r, t: (0.0 * r, np.full_like(r, vy), 0.0 * r))
    jt = deposit(p)["jt"]
    nqv = DENSITY * (-deck.Q0) * vy
    for j in (1, 2, 3):
        want = -1j * nqv * row_bracket(j, 1, 1)
        got = jt(1, j)
        assert abs(got - want) < tol * abs(nqv), (j, got / (-1j * nqv), row_bracket(j, 1, 1))


def check_m0_rotation(deposit, tol=1e-5, beta=1.0e-4):
    """J_theta,0 of a rigid rotation v_theta = omega r (omega dt = beta) on
    rows 1..3 against the bracket with r^2; row 1 carries the folded ghost
    row.  The free particles' radii grow by (omega dt)^2 / 2 in a step, which
    keeps beta small."""
    omega = beta / DT
    p = quiet_load(lambda r, t: (0.0 * r, -omega * r * np.sin(t), omega * r * np.cos(t)))
    jt = deposit(p)["jt"]
    unit = DENSITY * (-deck.Q0) * omega * DY
    for j in (1, 2, 3):
        want = unit * row_bracket(j, 0, 2)
        got = jt(0, j)
        assert abs(got - want) < tol * abs(unit), (j, got / unit, row_bracket(j, 0, 2))


def test_jtheta_fold_m1_uniform_drift(oracle):
    check_m1_drift(oracle_deposit(oracle))


def test_jtheta_fold_m0_rigid_rotation(oracle):
    check_m0_rotation(oracle_deposit(oracle))


# ------------------------------------------------------------------ r_max
def hat(d):
    """Linear B-spline of unit integral: the radial shape of the
    charge-conserving J_r (the derivative of the running sum of quadratic
    weights across a face, particles.F90:649-651)."""
    return max(0.0, 1.0 - abs(d))


def integral(shape, power, centre, lo, hi):
    lo, hi = max(lo, centre - 1.5), min(hi, centre + 1.5)
    if hi <= lo:
        return 0.0
    pts = [p for p in (centre - 1.0, centre - 0.5, centre, centre + 0.5, centre + 1.0) if lo < p < hi]
    v, _ = integrate.quad(lambda r: r ** power * shape(r - centre), lo, hi, points=pts or None,
                          epsabs=1e-14, epsrel=1e-13)
    return v


def check_rmax_fold(deposit, tol=3e-5, beta=1.0e-4):
    """The current rows beyond a reflecting r_max = R fold back
    (particle_reflection_bcs_complex, boundary.F90:986-1013): the face of J_r
    at R + a is the face at R - a of the mirror path, whose radial velocity
    is reversed, and the flux through the larger circle is spread over the
    smaller one, so the ghost row enters with the factor -(R + a)/(R - a);
    J_theta's ghost rows at R + a add to R - a as they are.  On a quiet
    uniform load drifting along y (R = 16 dr) the m = 1 rows then hold, in
    units of n q v_y (J_r,1) and -i n q v_y (J_theta,1),

        J_r(R - 1/2)  = [int_0^R r L(r - R + 1/2) - int_0^R r L(r - R - 1/2)] / (R - 1/2),
        J_th(R - 1)   = int_0^R r S(r - R + 1) / (R - 1) + int_0^R r S(r - R - 1) / (R + 1),

    with L the hat and S the quadratic spline; rows further in are 1."""
    ny = 16
    vy = beta * deck.C_LIGHT
    rk = (np.arange(ny * SUB_R) + 0.5) * (DY / SUB_R)
    th = 2.0 * math.pi * (np.arange(N_THETA) + 0.25) / N_THETA
    xs = (np.arange(2 * NX) + 0.5) * (DX / 2)
    X, R, T = (a.ravel() for a in np.meshgrid(xs, rk, th, indexing="ij"))
    w = DENSITY * (DX / 2) * (DY / SUB_R) * R * (2.0 * math.pi / N_THETA)
    gam = 1.0 / math.sqrt(1.0 - (vy / deck.C_LIGHT) ** 2)
    p = dict(x=X, y=R * np.cos(T), z=R * np.sin(T), px=0 * X, py=np.full_like(X, gam * deck.M0 * vy), pz=0 * X, w=w)
    rows = deposit(p, ny)
    nqv = DENSITY * (-deck.Q0) * vy
    Rm = float(ny)
    want_r = {ny: (integral(hat, 1, Rm - 0.5, 0, Rm) - integral(hat, 1, Rm + 0.5, 0, Rm)) / (Rm - 0.5),
              ny - 1: 1.0, ny - 2: 1.0}
    want_t = {ny - 1: integral(spline, 1, Rm - 1, 0, Rm) / (Rm - 1) + integral(spline, 1, Rm + 1, 0, Rm) / (Rm + 1),
              ny - 2: 1.0}
    for j, v in want_r.items():      # J_r row j sits at r = (j - 1/2) dr
        got = rows["jr"](1, j) / nqv
        assert abs(got - v) < tol, ("jr", j, got, v)
    for j, v in want_t.items():      # J_theta row j sits at r = j dr
        got = rows["jt"](1, j) / (-1j * nqv)
        assert abs(got - v) < tol, ("jt", j, got, v)
    # the part of the row's hat beyond R and the reversed fold each take
    # about 1/8 of the row
    assert 0.7 < want_r[ny] < 0.8


def test_rmax_reflection_folds_m1_uniform_drift(oracle):
    check_rmax_fold(oracle_deposit(oracle))
