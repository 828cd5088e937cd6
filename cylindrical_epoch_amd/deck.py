"""Input-side helpers of the host: the timestep rule and synthetic plasma
loads for the benchmark and test workloads.

The reference reads these from its input deck (src/deck/) and loads
particles in src/user_interaction/helper.F90; here the same quantities are
produced directly for the named workloads, with numpy's seeded generator in
place of the reference's KISS stream (identical initial particles between a
CPU and a GPU run come from handing both the same arrays).
"""
import math

import numpy as np

C_LIGHT = 2.99792458e8
Q0 = 1.602176565e-19
M0 = 9.10938291e-31
KB = 1.3806488e-23
EPSILON0 = 8.854187817620389850536563031710750e-12
PI = math.pi
EV = Q0


def set_dt(dx, dy, dt_multiplier=1.0, plasma=(), laser_omegas=()):
    """set_dt of src/housekeeping/setup.F90:573-673.

    plasma: iterable of (number density [m^-3], temperature [K], mass [kg])
    for the Bohm-Gross clip; laser_omegas: angular frequencies [rad/s]."""
    dt = 0.9 * min(dx, dy) / math.sqrt(2.0) / C_LIGHT
    k_max = 2.0 * PI / min(dx, dy)
    dt_p = 1.0e6
    for n, temp, mass in plasma:
        fac1 = Q0 ** 2 / mass / EPSILON0
        fac2 = 3.0 * k_max ** 2 * KB / mass
        omega2 = fac1 * n + fac2 * temp
        if omega2 > 2.2250738585072014e-308:
            dt_p = min(dt_p, 2.0 * PI / math.sqrt(omega2))
    dt_p = dt_p / 2.0 if plasma else 0.0
    dt_l = min((2.0 * PI / w for w in laser_omegas), default=0.0) / 2.0
    if dt_p > 2.2250738585072014e-308:
        dt = min(dt, dt_p)
    if dt_l > 2.2250738585072014e-308:
        dt = min(dt, dt_l)
    return dt_multiplier * dt


def uniform_plasma(nx, ny, dx, dy, ppc, density, temp_ev=0.0, mass=M0, x_min=0.0, seed=7842432,
                   cells_x=None, cells_r=None):
    """ppc macro-particles per cell, x uniform in the cell, r uniform in
    [r_c - dr/2, r_c + dr/2], theta uniform (helper.F90:540-545); weight
    density * 2 pi dx dy r / ppc (helper.F90:785-795, partlist.F90:999-1013);
    Maxwellian momenta of temperature temp_ev, sigma = sqrt(m k T) per
    component.  Returns a dict of float64 arrays and int64 ids."""
    rng = np.random.default_rng(seed)
    cx = np.arange(nx) if cells_x is None else np.asarray(cells_x)
    cr = np.arange(ny) if cells_r is None else np.asarray(cells_r)
    ncell = len(cx) * len(cr)
    n = ncell * ppc
    ix = np.repeat(np.tile(cx, len(cr)), ppc)
    ir = np.repeat(np.repeat(cr, len(cx)), ppc)
    x = x_min + (ix + rng.random(n)) * dx
    r = (ir + rng.random(n)) * dy
    th = rng.uniform(-PI, PI, n)
    y = r * np.cos(th)
    z = r * np.sin(th)
    w = density * 2.0 * PI * dx * dy * r / ppc
    if temp_ev > 0:
        sig = math.sqrt(mass * temp_ev * EV)
        p = rng.normal(0.0, sig, (3, n))
    else:
        p = np.zeros((3, n))
    return dict(x=x, y=y, z=z, px=p[0], py=p[1], pz=p[2], w=w, id=np.arange(n, dtype=np.int64))


def ramp_profile(x_start, ramp):
    """A density profile 0 before x_start, rising linearly over `ramp`
    metres to 1, then flat (the wakefield deck's up-ramp)."""
    def f(x):
        return np.clip((np.asarray(x) - x_start) / ramp, 0.0, 1.0)
    return f


def profiled_plasma(nx, ny, dx, dy, ppc, density, profile, temp_ev=0.0, mass=M0, x_min=0.0, seed=7842432,
                    cells_x=None, id0=0):
    """uniform_plasma with each weight scaled by profile(x) (the deck's
    density function at the particle) and the particles of zero weight
    dropped; ids from id0."""
    p = uniform_plasma(nx, ny, dx, dy, ppc, density, temp_ev, mass=mass, x_min=x_min, seed=seed, cells_x=cells_x)
    f = profile(p["x"])
    keep = f > 0
    out = {k: v[keep] for k, v in p.items()}
    out["w"] = out["w"] * f[keep]
    out["id"] = np.arange(int(keep.sum()), dtype=np.int64) + id0
    return out


# The workloads of the benchmark table (grid, modes, load).  dx = dy where the
# table gives one spacing; the large single run uses the wakefield spacing.
WORKLOADS = {
    "thermal_256x64_m2_ppc8": dict(nx=256, ny=64, n_mode=2, dx=0.1e-6, dy=0.1e-6, ppc=8,
                                   density=1e25, temp_ev=1000.0, steps=100),
    "wakefield_2048x256_m2_ppc16": dict(nx=2048, ny=256, n_mode=2, dx=0.02e-6, dy=0.1e-6, ppc=16,
                                        density=1e24, temp_ev=0.0, steps=2000),
    "wakefield_8192x512_m3_ppc32": dict(nx=8192, ny=512, n_mode=3, dx=0.02e-6, dy=0.1e-6, ppc=32,
                                        density=5e23, temp_ev=0.0, steps=1000),
}
