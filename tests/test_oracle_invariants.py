"""CPU tests that pin the FP64 oracle with closed forms and conservation laws
(each chosen so that a dropped term, a wrong sign or a transposed operand
fails).  No GPU needed."""
import math

import numpy as np
import pytest

C_LIGHT = 2.99792458e8
EPS0 = 8.854187817620389850536563031710750e-12
Q0 = 1.602176565e-19
M0 = 9.10938291e-31


def _state(oracle, nx=24, ny=16, M=2, dx=1e-7, dy=1e-7, **kw):
    dt = oracle.set_dt(dx, dy, 0.95)
    return oracle.State(nx, ny, M, dx, dy, dt, **kw)


def _one_particle(st, x, r, theta, p=(0.0, 0.0, 0.0), q=-Q0, m=M0, w=1.0):
    y, z = r * math.cos(theta), r * math.sin(theta)
    return st.add_species(q, m, [x], [y], [z], [p[0]], [p[1]], [p[2]], [w])


# ---------------------------------------------------------------- set_dt
def test_set_dt_closed_form(oracle):
    dx, dy = 0.02e-6, 0.1e-6
    assert oracle.set_dt(dx, dy, 0.95) == pytest.approx(0.95 * 0.9 * dx / math.sqrt(2) / C_LIGHT, rel=1e-15, abs=0)
    assert oracle.set_dt(dx, dy, 1.0, dt_plasma=1e-18) == pytest.approx(1e-18, rel=1e-15, abs=0)
    assert oracle.set_dt(dx, dy, 1.0, dt_laser=2e-18) == pytest.approx(2e-18, rel=1e-15, abs=0)


# ---------------------------------------------------------------- push
@pytest.mark.parametrize("theta", [0.3, 2.0, -2.5])
def test_uniform_ex_mode0_kick(oracle, theta):
    """Uniform m=0 Ex: dp_x = q E dt exactly, nothing else moves."""
    st = _state(oracle)
    E0 = 3.7e9
    st.arr["ex"][0] = E0
    sp = _one_particle(st, 1.23e-6, 0.77e-6, theta)
    st.push_particles()
    assert sp["px"][0] == pytest.approx(-Q0 * E0 * st.g.dt, rel=1e-13, abs=0)
    assert abs(sp["py"][0]) < 1e-40 and abs(sp["pz"][0]) < 1e-40


@pytest.mark.parametrize("theta", [0.0, 0.9, -2.2, 3.0])
def test_mode1_ex_reconstruction(oracle, theta):
    """A uniform m=1 Ex of value a+ib acts as Re[(a+ib) e^{-i theta}] (mode sum
    of e_part.inc); a particle at rest samples theta at its own position."""
    st = _state(oracle)
    a, b = 2.0e9, -1.3e9
    st.arr["ex"][1] = a + 1j * b
    sp = _one_particle(st, 1.1e-6, 0.8e-6, theta)
    st.push_particles()
    want = -Q0 * st.g.dt * (a * math.cos(theta) + b * math.sin(theta))
    assert sp["px"][0] == pytest.approx(want, rel=1e-12, abs=1e-40)


@pytest.mark.parametrize("comp", ["er", "et"])
def test_radial_azimuthal_to_cartesian(oracle, comp):
    """Uniform m=0 Er -> (Ey,Ez) = Er (cos, sin); Etheta -> Et (-sin, cos)."""
    st = _state(oracle)
    E0 = 1.9e9
    st.arr[comp][0] = E0
    th = 0.61
    sp = _one_particle(st, 1.0e-6, 0.9e-6, th)
    st.push_particles()
    k = -Q0 * E0 * st.g.dt
    if comp == "er":
        want = (k * math.cos(th), k * math.sin(th))
    else:
        want = (-k * math.sin(th), k * math.cos(th))
    assert sp["py"][0] == pytest.approx(want[0], rel=1e-12, abs=0)
    assert sp["pz"][0] == pytest.approx(want[1], rel=1e-12, abs=0)


def test_boris_rotation_uniform_bx(oracle):
    """Pure magnetic rotation: |p| and p_x conserved, rotation angle
    2 atan(q B dt / (2 gamma m))."""
    st = _state(oracle)
    B0 = 2.0e4
    st.arr["bx"][0] = B0
    p0 = np.array([0.3, 1.1, -0.4]) * M0 * C_LIGHT
    sp = _one_particle(st, 1.2e-6, 0.8e-6, 0.2, p=p0)
    st.push_particles()
    p1 = np.array([sp["px"][0], sp["py"][0], sp["pz"][0]])
    assert np.linalg.norm(p1) == pytest.approx(np.linalg.norm(p0), rel=1e-14, abs=0)
    assert p1[0] == pytest.approx(p0[0], rel=1e-14, abs=0)
    gamma = math.sqrt(1 + (np.linalg.norm(p0) / (M0 * C_LIGHT)) ** 2)
    ang = 2 * math.atan(abs(Q0 * B0 * st.g.dt / (2 * gamma * M0)))
    got = math.atan2(p0[1] * p1[2] - p0[2] * p1[1], p0[1] * p1[1] + p0[2] * p1[2])
    # electron (q<0) in +Bx rotates counter-clockwise in (y,z)
    assert got == pytest.approx(ang, rel=1e-10, abs=0)


@pytest.mark.parametrize("vd", [0.05, 0.3])
def test_e_cross_b_drift(oracle, vd):
    """Uniform E along y and B along z (m = 1 fields: E_r,1 = E, E_theta,1 =
    -i E, B_r,1 = i B, B_theta,1 = B).  A particle starting at the drift
    velocity E x B / B^2 = (E / B) x feels no force in the continuum.  The
    relativistic Boris step (half kick, rotation with the gamma of the
    kicked momentum, half kick; particles.F90:404-451) holds it there up to
    a gyration whose amplitude, relative to the drift momentum, is eps^2,
    eps = q E dt / (2 m c) the half kick in u = p / (m c): the kick changes
    gamma by O(eps^2) before the rotation.  Over 200 pushes the largest
    relative momentum excursion is eps^2 to 5 %, and x advances by
    (E / B) t to eps^2."""
    st = _state(oracle, nx=24, ny=32, M=2, bc_particle=("periodic", "periodic", "other", "reflect"))
    B0 = 1.0e4
    E0 = vd * C_LIGHT * B0
    st.arr["er"][1] = E0
    st.arr["et"][1] = -1j * E0
    st.arr["br"][1] = 1j * B0
    st.arr["bt"][1] = B0
    gam = 1.0 / math.sqrt(1.0 - vd * vd)
    u0 = np.array([gam * vd, 0.0, 0.0])
    x0, y0 = 3.3 * st.g.dx, 10.4 * st.g.dy
    sp = _one_particle(st, x0, y0, 0.0, p=u0 * M0 * C_LIGHT)
    eps = Q0 * E0 * st.g.dt / (2 * M0 * C_LIGHT)
    n, dev = 200, 0.0
    for _ in range(n):
        st.push_particles()
        u = np.array([sp["px"][0], sp["py"][0], sp["pz"][0]]) / (M0 * C_LIGHT)
        dev = max(dev, np.abs(u - u0).max() / u0[0])
    assert dev == pytest.approx(eps * eps, rel=0.05, abs=0)
    L = st.g.x_max - st.g.x_min
    travelled = (sp["x"][0] - x0) % L
    dist = vd * C_LIGHT * n * st.g.dt
    assert travelled == pytest.approx(dist % L, abs=eps * eps * dist)
    assert sp["y"][0] == pytest.approx(y0, abs=eps * eps * dist)


def test_free_streaming(oracle):
    """No fields: x(t+dt) = x + v dt."""
    st = _state(oracle)
    p0 = np.array([0.5, -0.2, 0.7]) * M0 * C_LIGHT
    sp = _one_particle(st, 1.3e-6, 0.9e-6, 1.0, p=p0)
    x0 = np.array([sp["x"][0], sp["y"][0], sp["z"][0]])
    st.push_particles()
    u = p0 / (M0 * C_LIGHT)
    v = u * C_LIGHT / math.sqrt(1 + u @ u)
    x1 = np.array([sp["x"][0], sp["y"][0], sp["z"][0]])
    assert np.allclose(x1, x0 + v * st.g.dt, rtol=1e-14, atol=1e-22)


# ---------------------------------------------------------------- deposit
def _stagger_weights(pos_r):
    """Quadratic-spline charge fractions on the staggered grid (sum 1)."""
    c = math.floor(pos_r)
    f = c - pos_r + 0.5
    w = np.array([0.25 + f * f + f, 1.5 - 2 * f * f, 0.25 + f * f - f]) / 2
    return c + 1, w


@pytest.mark.parametrize("seed", range(6))
def test_continuity_mode0_finite_volume(oracle, seed):
    """Charge conservation of the m=0 deposit: for every staggered cell,
    Q(t+3dt/2) - Q(t+dt/2) + dt * (net outward current through its faces)
    vanishes to round-off (particles.F90:633-662 with the area tables of
    :190-217)."""
    rng = np.random.default_rng(seed)
    st = _state(oracle, nx=24, ny=20, M=1)
    g = st.g
    x = (8 + rng.random()) * g.dx
    r = (9 + rng.random()) * g.dy
    th = rng.uniform(-3, 3)
    u = rng.normal(size=3) * 2.0           # relativistic, crosses cells
    p = u * M0 * C_LIGHT
    q, w = -Q0, 1.7e3
    sp = _one_particle(st, x, r, th, p=p, w=w)
    gam = math.sqrt(1 + u @ u)
    dtco2 = C_LIGHT * g.dt / 2
    pos0 = np.array([sp["x"][0], sp["y"][0], sp["z"][0]])
    half = pos0 + u * dtco2 / gam
    end15 = pos0 + 3 * u * dtco2 / gam
    st.push_particles()
    inv_rt, inv_xt, inv_vol, _ = st.geometry()
    ng = g.ng

    def charge(pos):
        xr = (pos[0] - g.x_grid_min) / g.dx
        rr = (math.hypot(pos[1], pos[2]) - g.y_grid_min) / g.dy
        cx, wx = _stagger_weights(xr)
        cy, wy = _stagger_weights(rr)
        Q = {}
        for i in range(3):
            for j in range(3):
                Q[(cx - 1 + i, cy - 1 + j)] = q * w * wx[i] * wy[j]
        return Q

    Q0_, Q1_ = charge(half), charge(end15)
    cells = set(Q0_) | set(Q1_)
    scale = abs(q * w)
    for (cx, cy) in cells:
        jx1 = st.at("jx", cx + 1, cy, 0).real
        jx0 = st.at("jx", cx, cy, 0).real
        jr1 = st.at("jr", cx, cy + 1, 0).real
        jr0 = st.at("jr", cx, cy, 0).real
        a_rt = 1 / inv_rt[cy + ng]
        a_hi = 1 / inv_xt[cy + ng]
        a_lo = 1 / inv_xt[cy - 1 + ng]
        flux = (jx1 - jx0) * a_rt + jr1 * a_hi - jr0 * a_lo
        res = Q1_.get((cx, cy), 0.0) - Q0_.get((cx, cy), 0.0) + g.dt * flux
        assert abs(res) < 1e-12 * scale, (cx, cy, res)


@pytest.mark.parametrize("seed", range(8))
def test_continuity_mode0_after_axis_fold(oracle, seed):
    """A particle within 1.4 cells of the axis: after current_finish (the
    fold of the r < 0 ghost rows, boundary.F90:1917-1931), every cell
    with r >= dr conserves charge to round-off, its charge being the
    stencil plus the stencil's mirror (the node at -r folds onto +r).  The
    axis cell itself is left out: the reference replaces jxm(:,0,0) by
    (4 jxm(:,1,0) - jxm(:,2,0)) / 3."""
    rng = np.random.default_rng(100 + seed)
    st = _state(oracle, nx=24, ny=20, M=1)
    g = st.g
    x = (8 + rng.random()) * g.dx
    r = (0.05 + 1.35 * rng.random()) * g.dy
    th = rng.uniform(-3, 3)
    u = rng.normal(size=3) * 1.0
    p = u * M0 * C_LIGHT
    q, w = -Q0, 1.7e3
    sp = _one_particle(st, x, r, th, p=p, w=w)
    gam = math.sqrt(1 + u @ u)
    dtco2 = C_LIGHT * g.dt / 2
    pos0 = np.array([sp["x"][0], sp["y"][0], sp["z"][0]])
    st.push_particles()
    st.current_finish()
    inv_rt, inv_xt, _, _ = st.geometry()
    ng = g.ng

    def charge(pos):
        xr = (pos[0] - g.x_grid_min) / g.dx
        rr = (math.hypot(pos[1], pos[2]) - g.y_grid_min) / g.dy
        cx, wx = _stagger_weights(xr)
        cy, wy = _stagger_weights(rr)
        Q = {}
        for i in range(3):
            for j in range(3):
                key = (cx - 1 + i, abs(cy - 1 + j))
                Q[key] = Q.get(key, 0.0) + q * w * wx[i] * wy[j]
        return Q

    Q0_ = charge(pos0 + u * dtco2 / gam)
    Q1_ = charge(pos0 + 3 * u * dtco2 / gam)
    cells = {c for c in set(Q0_) | set(Q1_) if c[1] >= 1}
    assert cells
    scale = abs(q * w)
    for (cx, cy) in cells:
        jx1 = st.at("jx", cx + 1, cy, 0).real
        jx0 = st.at("jx", cx, cy, 0).real
        jr1 = st.at("jr", cx, cy + 1, 0).real
        jr0 = st.at("jr", cx, cy, 0).real
        flux = (jx1 - jx0) / inv_rt[cy + ng] + jr1 / inv_xt[cy + ng] - jr0 / inv_xt[cy - 1 + ng]
        res = Q1_.get((cx, cy), 0.0) - Q0_.get((cx, cy), 0.0) + g.dt * flux
        assert abs(res) < 1e-12 * scale, (cx, cy, res)


def _mode_charge(g, pos, q, w, M):
    """Charge modes on the staggered grid at a position: quadratic-spline
    fractions times 1 (m = 0) or 2 e^{+i m theta} (m > 0)."""
    xr = (pos[0] - g.x_grid_min) / g.dx
    rr = (math.hypot(pos[1], pos[2]) - g.y_grid_min) / g.dy
    th = math.atan2(pos[2], pos[1])
    cx, wx = _stagger_weights(xr)
    cy, wy = _stagger_weights(rr)
    Q = {}
    for m in range(M):
        f = 1.0 if m == 0 else 2 * np.exp(1j * m * th)
        for i in range(3):
            for j in range(3):
                Q[(cx - 1 + i, cy - 1 + j, m)] = q * w * wx[i] * wy[j] * f
    return Q


@pytest.mark.parametrize("seed", range(4))
def test_mode_continuity_structure(oracle, seed):
    """All modes of one push (particles.F90:508-665).  With Q_m the charge
    modes at x(t+dt/2) and x(t+3dt/2), R = Q_m(1) - Q_m(0) + dt (x and r
    face fluxes of Jx, Jr) must be exactly what the azimuthal current has
    to carry: R = i m dt K cy Jt in every cell and mode, with one K per
    particle.  m = 0 gives R = 0.  For m > 0 the same K in every cell and
    mode pins the m-factors of Jx and Jr against the exact phases
    e^{i m theta} of the charge; K differs between particles because the
    reference deposits Jt with v_theta(t+dt) rather than r dtheta/dt, so
    per-mode continuity for m > 0 is not a discrete identity of the
    reference's scheme."""
    rng = np.random.default_rng(100 + seed)
    M = 3
    st = _state(oracle, nx=24, ny=20, M=M)
    g = st.g
    x = (8 + rng.random()) * g.dx
    r = (9 + rng.random()) * g.dy
    th = rng.uniform(-3, 3)
    u = rng.normal(size=3) * 2.0
    q, w = -Q0, 1.7e3
    sp = _one_particle(st, x, r, th, p=u * M0 * C_LIGHT, w=w)
    gam = math.sqrt(1 + u @ u)
    pos0 = np.array([sp["x"][0], sp["y"][0], sp["z"][0]])
    half = pos0 + u * C_LIGHT * g.dt / 2 / gam
    st.push_particles()
    u1 = np.array([sp["px"][0], sp["py"][0], sp["pz"][0]]) / (M0 * C_LIGHT)
    pos1 = np.array([sp["x"][0], sp["y"][0], sp["z"][0]])
    end15 = pos1 + u1 * C_LIGHT * g.dt / 2 / math.sqrt(1 + u1 @ u1)
    inv_rt, inv_xt, inv_vol, _ = st.geometry()
    ng = g.ng
    Qa, Qb = _mode_charge(g, half, q, w, M), _mode_charge(g, end15, q, w, M)
    scale = abs(q * w)
    ks = []
    for (cx, cy, m) in sorted(set(Qa) | set(Qb)):
        flux = ((st.at("jx", cx + 1, cy, m) - st.at("jx", cx, cy, m)) / inv_rt[cy + ng]
                + st.at("jr", cx, cy + 1, m) / inv_xt[cy + ng] - st.at("jr", cx, cy, m) / inv_xt[cy - 1 + ng])
        R = Qb.get((cx, cy, m), 0.0) - Qa.get((cx, cy, m), 0.0) + g.dt * flux
        jt = st.at("jt", cx, cy, m)
        if m == 0:
            assert abs(R) < 1e-12 * scale, (cx, cy, R)
        elif abs(jt) * g.dt * g.dx * g.dy * cy > 1e-9 * scale:
            ks.append(R / (1j * m * g.dt * jt * g.dx * g.dy * cy))
        else:
            assert abs(R) < 1e-11 * scale
    ks = np.array(ks)
    assert len(ks) >= 12
    assert np.allclose(ks, ks[0], rtol=1e-8, atol=0), ks
    assert abs(ks[0].imag) < 1e-8 * abs(ks[0]) and 0.3 < ks[0].real < 1.5


def test_small_angle_switch_is_continuous(oracle):
    """The m>0 factors switch from the exact form to the small-angle Taylor
    form at |m dtheta| = 1e-4 (particles.F90:589-626); the deposited current
    must be continuous across the switch (a wrong Taylor coefficient jumps
    by O(1))."""
    out = []
    for dth in (1e-4 * (1 - 1e-7), 1e-4 * (1 + 1e-7)):
        st = oracle.State(16, 16, 2, 1e-7, 1e-7, 1e-16)
        gx = np.array([0, 0.04, 1.28, 0.68, 0])
        hx = np.array([0, 0.5, 0.12, -0.62, 0])
        gy = np.array([0, 0.6, 1.34, 0.06, 0])
        hy = np.array([0, -0.4, 0.14, 0.26, 0])
        e05 = (math.cos(0.4), math.sin(0.4))
        ed = (math.cos(dth), math.sin(dth))
        oracle.deposit_from_weights(st, gx, hx, gy, hy, 8, 8, 0, 0, 1.0, 1.0, 0.5, e05, ed, dth)
        out.append(np.concatenate([st.arr[c][1].ravel() for c in ("jx", "jr", "jt")]))
    scale = np.max(np.abs(out[0]))
    assert scale > 0
    assert np.max(np.abs(out[0] - out[1])) < 1e-3 * scale


# ---------------------------------------------------------------- FDTD
def _div_b(st, m):
    """Discrete divergence of B_m at the Btheta points (x centre, r centre),
    using dtheta -> -i m (fields are Re sum F_m exp(-i m theta))."""
    g, ng = st.g, st.g.ng
    bx, br, bt = st.arr["bx"][m], st.arr["br"][m], st.arr["bt"][m]
    out = {}
    for j in range(2, g.ny):
        r_d = (j - 0.5) * g.dy
        r_p, r_pm = j * g.dy, (j - 1) * g.dy
        for i in range(1, g.nx + 1):
            I, J = i + ng - 1, j + ng - 1
            out[(i, j)] = ((bx[J, I] - bx[J, I - 1]) / g.dx
                           + (r_p * br[J, I] - r_pm * br[J - 1, I]) / (r_d * g.dy)
                           - 1j * m * bt[J, I] / r_d)
    return out


def _div_e(st, m):
    """Discrete divergence of E_m at the Etheta points (x stagger, r edge)."""
    g, ng = st.g, st.g.ng
    ex, er, et = st.arr["ex"][m], st.arr["er"][m], st.arr["et"][m]
    out = {}
    for j in range(1, g.ny):
        r_p = j * g.dy
        r_d1, r_d0 = (j + 0.5) * g.dy, (j - 0.5) * g.dy
        for i in range(0, g.nx):
            I, J = i + ng - 1, j + ng - 1
            out[(i, j)] = ((ex[J, I + 1] - ex[J, I]) / g.dx
                           + (r_d1 * er[J + 1, I] - r_d0 * er[J, I]) / (r_p * g.dy)
                           - 1j * m * et[J, I] / r_p)
    return out


@pytest.mark.parametrize("m", [0, 1, 2])
def test_update_b_preserves_div_b(oracle, m):
    rng = np.random.default_rng(10 + m)
    st = _state(oracle, nx=12, ny=10, M=3)
    for c in ("ex", "er", "et", "bx", "br", "bt"):
        st.arr[c][:] = rng.normal(size=st.shape) + 1j * rng.normal(size=st.shape)
    before = _div_b(st, m)
    st.call("orc_update_b_field")
    after = _div_b(st, m)
    scale = max(abs(v) for v in before.values())
    for k in before:
        assert abs(after[k] - before[k]) < 1e-9 * scale, k


@pytest.mark.parametrize("m", [0, 1, 2])
def test_update_e_gauss_law(oracle, m):
    """d(div E)/dt = -div J / eps0 for the interior E update.  For m > 1 the
    reference overwrites Er on row 1 after the update (er(:,1) = er(:,2)/9,
    fields.f90:166), so row 1 is excluded there."""
    rng = np.random.default_rng(20 + m)
    st = _state(oracle, nx=12, ny=10, M=3)
    for c in ("ex", "er", "et", "bx", "br", "bt"):
        st.arr[c][:] = rng.normal(size=st.shape) + 1j * rng.normal(size=st.shape)
    for c in ("jx", "jr", "jt"):
        st.arr[c][:] = 1e6 * (rng.normal(size=st.shape) + 1j * rng.normal(size=st.shape))
    before = _div_e(st, m)
    # div J at the same points, same stencil as div E
    saved = {c: st.arr[c].copy() for c in ("ex", "er", "et")}
    for a, b in (("ex", "jx"), ("er", "jr"), ("et", "jt")):
        st.arr[a][:] = st.arr[b]
    div_j = _div_e(st, m)
    for c in ("ex", "er", "et"):
        st.arr[c][:] = saved[c]
    st.call("orc_update_e_field")
    after = _div_e(st, m)
    dt = st.g.dt
    # the curl-B terms cancel in div E; their size sets the round-off floor
    bmax = max(np.max(np.abs(st.arr[c])) for c in ("bx", "br", "bt"))
    floor = 64 * 2.2e-16 * C_LIGHT ** 2 * dt * bmax / (st.g.dx * st.g.dy) * 10
    jeff = max(0.5 * dt * abs(v) / EPS0 for v in div_j.values())
    assert floor < 1e-5 * jeff
    for k in before:
        if m > 1 and k[1] == 1:
            continue
        want = before[k] - 0.5 * dt * div_j[k] / EPS0
        assert abs(after[k] - want) < floor, k


def _div_e_axis(st, comp=("ex", "er")):
    """m = 0 divergence over the axis disk r < dy/2 around each Ex(x, r=0)
    point: d/dx of Ex on the axis plus the flux of Er through r = dy/2,
    2 pi (dy/2) Er / (pi (dy/2)^2) = 4 Er / dy."""
    g, ng = st.g, st.g.ng
    ex, er = st.arr[comp[0]][0], st.arr[comp[1]][0]
    J0 = ng - 1  # row iy = 0, r = 0
    return np.array([(ex[J0, i + ng] - ex[J0, i + ng - 1]) / g.dx + 4.0 * er[J0 + 1, i + ng - 1] / g.dy
                     for i in range(0, g.nx)])


def test_axis_gauss_law_m0(oracle):
    """The r = 0 row of update_e_field for m = 0 (fields.f90:116-134, Ex on
    the axis from 4 c^2 / dr Btheta(1)) keeps Gauss's law over the axis
    disk: its divergence changes by -dt/2 div J / eps0 (the curl-B terms
    cancel to round-off), as in the interior."""
    rng = np.random.default_rng(31)
    st = _state(oracle, nx=12, ny=10, M=3)
    for c in ("ex", "er", "et", "bx", "br", "bt"):
        st.arr[c][:] = rng.normal(size=st.shape) + 1j * rng.normal(size=st.shape)
    for c in ("jx", "jr", "jt"):
        st.arr[c][:] = 1e6 * (rng.normal(size=st.shape) + 1j * rng.normal(size=st.shape))
    before, div_j = _div_e_axis(st), _div_e_axis(st, ("jx", "jr"))
    st.call("orc_update_e_field")
    after = _div_e_axis(st)
    dt = st.g.dt
    bmax = max(np.max(np.abs(st.arr[c])) for c in ("bx", "br", "bt"))
    floor = 64 * 2.2e-16 * C_LIGHT ** 2 * dt * bmax / (st.g.dx * st.g.dy) * 10
    assert floor < 1e-5 * np.max(np.abs(0.5 * dt * div_j / EPS0))
    assert np.max(np.abs(after - (before - 0.5 * dt * div_j / EPS0))) < floor


@pytest.mark.parametrize("m", [0, 1, 2])
def test_first_cell_div_b(oracle, m):
    """div B over the first radial cell (0 < r < dy, the disk around the
    axis, whose inner face has zero area) is preserved by update_b_field's
    row-1 update (fields.f90:186-248), which reads the E values of the r = 0
    row."""
    rng = np.random.default_rng(40 + m)
    st = _state(oracle, nx=12, ny=10, M=3)
    for c in ("ex", "er", "et", "bx", "br", "bt"):
        st.arr[c][:] = rng.normal(size=st.shape) + 1j * rng.normal(size=st.shape)
    g, ng = st.g, st.g.ng

    def div1():
        bx, br, bt = st.arr["bx"][m], st.arr["br"][m], st.arr["bt"][m]
        J = ng  # row iy = 1
        return np.array([(bx[J, i + ng - 1] - bx[J, i + ng - 2]) / g.dx + 2.0 * br[J, i + ng - 1] / g.dy
                         - 1j * m * bt[J, i + ng - 1] / (0.5 * g.dy) for i in range(1, g.nx + 1)])

    before = div1()
    st.call("orc_update_b_field")
    after = div1()
    assert np.max(np.abs(after - before)) < 1e-9 * np.max(np.abs(before))


def test_vacuum_energy_bounded(oracle):
    """A vacuum m=1 pulse in a conducting box: the mode energy stays bounded
    (the leapfrog's discrete energy is conserved up to O(dt^2) oscillation)."""
    st = _state(oracle, nx=32, ny=16, M=2,
                bc_field=("conduct", "conduct", "other", "conduct"))
    g = st.g
    xs = (np.arange(g.nx + 2 * g.ng) - g.ng + 1) * g.dx
    rs = (np.arange(g.ny + 2 * g.ng) - g.ng + 1) * g.dy
    X, R = np.meshgrid(xs, rs)
    st.arr["er"][1] = 1e9 * np.exp(-((X - 16 * g.dx) / (4 * g.dx)) ** 2 - (R / (5 * g.dy)) ** 2)
    st.arr["et"][1] = 1j * st.arr["er"][1]
    e0 = None
    for n in range(60):
        st.update_eb_fields_half()
        st.update_eb_fields_final()
        we, wb = st.mode_energy()
        tot = we.sum() + wb.sum()
        if e0 is None:
            e0 = tot
        assert tot < 1.5 * e0
        assert np.isfinite(tot)


# ---------------------------------------------------------------- sort / key
def test_stable_sort_matches_numpy(oracle):
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 97, size=5000).astype(np.int32)
    perm = oracle.stable_sort(keys, 97)
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))


def test_cell_keys_brute_force(oracle):
    rng = np.random.default_rng(4)
    st = _state(oracle, nx=20, ny=12, M=1)
    g = st.g
    n = 4000
    x = rng.uniform(g.x_min, g.x_max, n)
    r = rng.uniform(0, g.y_max, n)
    th = rng.uniform(-math.pi, math.pi, n)
    y, z = r * np.cos(th), r * np.sin(th)
    # exact cell edges too
    x[:20] = g.x_min + np.arange(20) * g.dx
    keys = oracle.cell_keys(st, x, y, z, 4, 3)
    rr = np.sqrt(y * y + z * z)
    cx = np.floor((x - g.x_grid_min) / g.dx + 1.5).astype(int)
    cy = np.floor((rr - g.y_grid_min) / g.dy + 1.5).astype(int)
    lx, ly = cx - 1, cy - 1
    ntx = (g.nx + 3) // 4
    want = (((ly // 3) * ntx + lx // 4) * 3 + ly % 3) * 4 + lx % 4
    ok = (cx >= 1) & (cx <= g.nx) & (cy >= 1) & (cy <= g.ny)
    assert ok.sum() > n - 50
    assert np.array_equal(keys[ok], want[ok])
    assert np.all((keys >= 0) & (keys < ntx * ((g.ny + 2) // 3) * 12))


# ---------------------------------------------------------------- particle bcs
def test_particle_bcs_periodic_and_reflect(oracle):
    st = _state(oracle, nx=10, ny=8, M=1, bc_particle=("periodic", "periodic", "other", "reflect"))
    g = st.g
    L = g.x_max - g.x_min
    x = [g.x_min - 0.1 * g.dx, g.x_max + 0.2 * g.dx, 0.5 * (g.x_min + g.x_max)]
    y = [0.3 * g.dy, 0.3 * g.dy, g.y_max + 0.25 * g.dy]
    z = [0.0, 0.0, 0.0]
    p = np.array([1e-22, 2e-22, 3e-22])
    sp = st.add_species(-Q0, M0, x, y, z, p, p, p, [1, 1, 1])
    arr = st._species_array()
    oracle.lib().orc_particle_bcs(oracle.C.byref(st.g), arr, 1)
    st._species_back(arr)
    ids = list(sp["id"])
    # staying particle first, then x_min leavers, then x_max leavers
    assert ids == [2, 0, 1]
    assert sp["x"][1] == pytest.approx(x[0] + L, abs=0)
    assert sp["x"][2] == pytest.approx(x[1] - L, abs=0)
    # r_max reflection of particle 2 (at theta = 0, every momentum component
    # p[2]): radius mirrored about y_max, radial momentum (p_y) reversed,
    # azimuthal (p_z) and axial kept
    assert sp["y"][0] == pytest.approx(g.y_max - 0.25 * g.dy, rel=1e-14, abs=0)
    assert sp["py"][0] == pytest.approx(-p[2], rel=1e-14, abs=0)
    assert sp["pz"][0] == pytest.approx(p[2], rel=1e-14, abs=0)
    assert sp["px"][0] == p[2]


@pytest.mark.parametrize("theta", [0.7, 2.4, -1.9])
def test_rmax_reflection_at_any_angle(oracle, theta):
    """r_max reflection (boundary.F90:1776-1790): the radius is mirrored
    about y_max at the same angle; the radial momentum reverses, the
    azimuthal and axial momenta and |p| are kept."""
    st = _state(oracle, nx=10, ny=8, M=1, bc_particle=("periodic", "periodic", "other", "reflect"))
    g = st.g
    r = g.y_max + 0.3 * g.dy
    p = np.array([1.3e-22, -0.7e-22, 2.1e-22])
    c, s = math.cos(theta), math.sin(theta)
    pr, pt = p[1] * c + p[2] * s, -p[1] * s + p[2] * c
    sp = st.add_species(-Q0, M0, [0.5 * (g.x_min + g.x_max)], [r * c], [r * s], [p[0]], [p[1]], [p[2]], [1.0])
    arr = st._species_array()
    oracle.lib().orc_particle_bcs(oracle.C.byref(st.g), arr, 1)
    st._species_back(arr)
    y, z = sp["y"][0], sp["z"][0]
    assert math.hypot(y, z) == pytest.approx(g.y_max - 0.3 * g.dy, rel=1e-13, abs=0)
    assert math.atan2(z, y) == pytest.approx(theta, abs=1e-13)
    py, pz = sp["py"][0], sp["pz"][0]
    assert py * c + pz * s == pytest.approx(-pr, rel=1e-12, abs=0)
    assert -py * s + pz * c == pytest.approx(pt, rel=1e-12, abs=0)
    assert sp["px"][0] == p[0]
    assert math.hypot(py, pz) == pytest.approx(math.hypot(p[1], p[2]), rel=1e-14, abs=0)


def test_mode_factor_sensitivity(oracle):
    """The m>0 deposit's exact-branch factors are ill-conditioned near the
    small-angle switch: one ulp of dtheta (the difference two atan2
    implementations may show) moves the m=1 current by far more than one ulp
    of J.  This bounds what a GPU/CPU comparison of m>0 currents can expect."""
    rng = np.random.default_rng(17)
    worst = 0.0
    for _ in range(200):
        dth = 10 ** rng.uniform(-4, -2) * rng.choice([-1, 1])
        out = []
        for d in (dth, np.nextafter(dth, 1.0)):
            st = oracle.State(16, 16, 2, 1e-7, 1e-7, 1e-16)
            gx = np.array([0, 0.04, 1.28, 0.68, 0])
            hx = np.array([0, 0.5, 0.12, -0.62, 0])
            gy = np.array([0, 0.6, 1.34, 0.06, 0])
            hy = np.array([0, -0.4, 0.14, 0.26, 0])
            e05 = (math.cos(0.4), math.sin(0.4))
            ed = (math.cos(dth), math.sin(dth))     # phasor from positions: unchanged
            oracle.deposit_from_weights(st, gx, hx, gy, hy, 8, 8, 0, 0, 1.0, 1.0, 0.5, e05, ed, d)
            out.append(np.concatenate([st.arr[c][1].ravel() for c in ("jx", "jr", "jt")]))
        worst = max(worst, np.max(np.abs(out[0] - out[1])) / np.max(np.abs(out[0])))
    assert 1e-13 < worst < 1e-8


def test_oracle_refuses_multi_cell_moves(oracle):
    """The oracle flags a particle whose stencil it cannot place (here a
    non-finite momentum) instead of overrunning the reference's 5-point
    weight arrays or reading outside the field arrays."""
    st = _state(oracle, nx=16, ny=8, M=1)
    one = np.ones(1)
    st.add_species(-Q0, M0, one * 5e-7, one * 3e-7, one * 0, one * np.nan, one * 0, one * 0, one)
    with pytest.raises(FloatingPointError):
        st.push_particles()
