"""The oracle's moving window (src/housekeeping/window.F90:62-153, 330-350)
checked by what a window shift must do: translate the fields one cell
towards x_min, keep a field that is uniform in x uniform, bring in the stored
x_max boundary columns, move the grid by dx per cell and drop exactly the
particles left of the new x_min, keeping the order of the others."""
import numpy as np
import pytest

FIELDS9 = ("ex", "er", "et", "bx", "br", "bt", "jx", "jr", "jt")


def _state(oracle, nx=24, ny=10, M=2):
    dx, dy = 1e-7, 1e-7
    bf = ("simple_outflow", "simple_outflow", "other", "open")
    bp = ("open", "open", "other", "open")
    return oracle.State(nx, ny, M, dx, dy, oracle.set_dt(dx, dy, 0.95), bc_field=bf, bc_particle=bp)


def test_shift_translates_interior(oracle):
    st = _state(oracle)
    rng = np.random.default_rng(1)
    for n in FIELDS9:
        st.arr[n][:] = rng.normal(size=st.shape) + 1j * rng.normal(size=st.shape)
    st.setup_field_boundaries()
    before = {n: st.arr[n].copy() for n in FIELDS9}
    k = 3
    st.shift_window(k)
    ng, nx = 5, st.g.nx
    # array column c holds Fortran ix = c + 1 - ng; after k shifts f(ix) = f_old(ix + k)
    # for every column untouched by the x_max fix (ix <= nx - 2 - ... for the last shift)
    for n in FIELDS9:
        a, b = st.arr[n], before[n]
        last = nx - 3 + ng - 1  # array column of ix = nx - 3
        assert np.array_equal(a[:, :, : last - k + 1], b[:, :, k: last + 1]), n


def test_uniform_field_stays_uniform(oracle):
    st = _state(oracle)
    rng = np.random.default_rng(2)
    prof = {n: rng.normal(size=(st.shape[0], st.shape[1], 1)) for n in FIELDS9}
    for n in FIELDS9:
        st.arr[n][:] = np.broadcast_to(prof[n], st.shape)
    st.setup_field_boundaries()
    st.shift_window(2)
    for n in FIELDS9:
        # every column equals the profile: the boundary columns stored from a
        # uniform field and their averages reproduce it
        assert np.allclose(st.arr[n], np.broadcast_to(prof[n], st.shape), rtol=0, atol=1e-15), n


def test_grid_moves_and_particles_left_of_x_min_go(oracle):
    st = _state(oracle)
    dx = st.g.dx
    x0_min, x0_grid = st.g.x_min, st.g.x_grid_min
    n = 400
    rng = np.random.default_rng(3)
    x = x0_min + rng.uniform(-0.5, 5.0, n) * dx
    y = rng.uniform(0.1, 5.0, n) * st.g.dy
    z = np.zeros(n)
    st.add_species(-1.6e-19, 9.1e-31, x, y, z, z, z, z, np.ones(n), ids=np.arange(n))
    st.shift_window(3)
    assert st.g.x_grid_min == x0_grid + dx + dx + dx
    assert st.g.x_min == x0_min + dx + dx + dx
    assert st.g.x_max == st.g.x_min + st.g.nx * dx
    keep = x >= st.g.x_min
    sp = st.species[0]
    assert sp["n"] == keep.sum()
    assert np.array_equal(sp["id"][: sp["n"]], np.arange(n)[keep])  # order kept


def test_x_max_fix_keeps_a_line_straight(oracle):
    """A field linear in x, whose stored x_max columns continue the line at
    the columns they refill, stays exactly that line through one shift: the
    averaged staggered neighbour of the incoming cell is the midpoint of the
    line's two values around it."""
    st = _state(oracle)
    rng = np.random.default_rng(7)
    ng, nx = 5, st.g.nx
    ix = np.arange(st.shape[2]) + 1 - ng  # Fortran column index per array column
    lines = {}
    for n in FIELDS9:
        a = rng.normal(size=(st.shape[0], st.shape[1], 1)) + 1j * rng.normal(size=(st.shape[0], st.shape[1], 1))
        b = rng.normal(size=(st.shape[0], st.shape[1], 1)) + 1j * rng.normal(size=(st.shape[0], st.shape[1], 1))
        lines[n] = (a, b)
        st.arr[n][:] = a + b * ix[None, None, :]
    # the column each stored boundary value refills, window.F90:101-128
    refill = dict(ex=nx + 1, er=nx, et=nx, bx=nx, br=nx + 1, bt=nx + 1)
    for k, n in enumerate(("ex", "er", "et", "bx", "br", "bt")):
        a, b = lines[n]
        st.bnd["xmax"][k][:] = (a + b * (refill[n] + 1))[:, :, 0]
    st.shift_window(1)
    for n in ("ex", "er", "et", "bx", "br", "bt"):
        a, b = lines[n]
        want = a + b * (ix[None, None, :] + 1)
        cols = slice(0, nx + 1 + ng)  # every column up to ix = nx + 1
        assert np.allclose(st.arr[n][:, :, cols], want[:, :, cols], rtol=0, atol=1e-12), n


def test_injection_and_removal_commute(oracle):
    """The reference inserts the new column (insert_particles) before the
    shift removes the particles left of x_min.  The inserted particles lie
    beyond x_max and the removed ones below the new x_min, so the two sets
    are disjoint and the order of the two operations cannot change the
    result: inserting after the shift gives the same particles in the same
    order."""
    rng = np.random.default_rng(4)
    runs = []
    for insert_first in (True, False):
        st = _state(oracle)
        dx, dy = st.g.dx, st.g.dy
        n = 300
        x = st.g.x_min + rng.uniform(-0.5, st.g.nx, n) * dx if insert_first else runs[0][1]
        y = rng.uniform(0.1, 5.0, n) * dy if insert_first else runs[0][2]
        z = np.zeros(n)
        st.add_species(-1.6e-19, 9.1e-31, x, y, z, z, z, z, np.ones(n), ids=np.arange(n))
        k = 2
        xi = st.g.x_max + (np.arange(40) % 20 + 0.5) / 20 * k * dx
        yi = (np.arange(40) // 20 + 0.5) * dy
        zi = np.zeros(40)
        args = (0, xi, yi, zi, zi, zi, zi, np.ones(40))
        if insert_first:
            st.append_particles(*args, ids=np.arange(n, n + 40))
            st.shift_window(k)
        else:
            st.shift_window(k)
            st.append_particles(*args, ids=np.arange(n, n + 40))
        sp = st.species[0]
        runs.append((sp["id"][: sp["n"]].copy(), x, y, sp["x"][: sp["n"]].copy()))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][3], runs[1][3])
    ids = runs[0][0]
    assert len(ids) < 340 and np.count_nonzero(ids >= 300) == 40    # some removed, all inserted kept
