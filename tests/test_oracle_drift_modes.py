"""The reference's current-density validation case (example_decks/
current_density_test.deck, the user manual's mode-current figure): a uniform
electron plasma of density 1e5 m^-3 drifting with p_x = -1.23e20, p_y =
+1.23e20 kg m/s, two modes.  At that momentum v = c p / |p|, so the current is
J = n q c (-1, 1, 0) / sqrt(2) with |J_x| = |J_y| = n |q| c / sqrt(2) =
3.40e-6 A m^-2.  In modes (a field f(theta) = Re sum_m f_m e^{-i m theta}):
J_x has only m = 0, equal to J_x; J_r = J_y cos(theta) and J_theta =
-J_y sin(theta) have only m = 1, with J_r,1 = J_y and J_theta,1 = -i J_y.
The plasma is loaded at random, so the checks are on the mean over interior
cells after the current boundaries (the periodic fold in x), to the
statistical error of the load (at 256 particles per cell: 0.1 % for m = 0,
0.4 % for m = 1, 0.7 % for a mode that should be empty).  No GPU
needed."""
import math

import numpy as np
import pytest

from cylindrical_epoch_amd import deck

DENSITY = 1.0e5
DRIFT = (-1.23e20, 1.23e20, 0.0)
NX, NY, M = 40, 20, 2
DX, DY = 20.0e-6 / 500, 5.0e-6 / 100


def drift_plasma(nx=NX, ny=NY, ppc=256, seed=11, drift=DRIFT):
    p = deck.uniform_plasma(nx, ny, DX, DY, ppc, DENSITY, 0.0, seed=seed)
    for k, d in zip(("px", "py", "pz"), drift):
        p[k] = np.full_like(p[k], d)
    return p


def expected_current(drift=DRIFT, q=-deck.Q0, m=deck.M0):
    """n q v for a uniform drift of momentum `drift` (SI)."""
    u = np.array(drift) / (m * deck.C_LIGHT)
    v = deck.C_LIGHT * u / math.sqrt(1.0 + float(u @ u))
    return DENSITY * q * v


def interior_means(get, ng, nx=NX, ny=NY, n_mode=M):
    """Mean of each mode of jx, jr, jt over cells i = 1..nx, j = 4..ny-4 (away
    from the axis fold and the r_max reflection).  get(name) returns the
    (n_mode, ny+2ng, nx+2ng) array."""
    rows = slice(ng - 1 + 4, ng - 1 + ny - 3)
    cols = slice(ng, ng + nx)
    return {c: [get(c)[m][rows, cols].mean() for m in range(n_mode)] for c in ("jx", "jr", "jt")}


def check_mode_structure(means, j, rel=0.012, leak=0.015):
    """J_x only in m = 0, J_r / J_theta only in m = 1, at n q v."""
    jx, jy = j[0], j[1]
    assert abs(means["jx"][0] - jx) < rel * abs(jx)
    assert abs(means["jr"][1] - jy) < rel * abs(jy)
    assert abs(means["jt"][1] - (-1j * jy)) < rel * abs(jy)
    for c, m in (("jx", 1), ("jr", 0), ("jt", 0)):
        assert abs(means[c][m]) < leak * abs(jy), (c, m, means[c][m])


def check_axis_rows(get, ng, j, nx=NX, rel=0.04, leak=0.03):
    """Rows 0..3, where the axis fold and the r = 0 rules of the current act
    (boundary.F90:1909-1959), at the same closed form: J_x,0 on the axis,
    J_theta,1 on the axis, and J_r,1 on the row reflected to r = -dr/2 (the
    uniform transverse current continues across the axis).  Rows near the
    axis hold few particles and see the largest angle changes per step, so
    the tolerance is looser than in the interior.  J_x,1 and J_theta,0 are
    zero on the axis."""
    jx, jy = j[0], j[1]
    cols = slice(ng, ng + nx)

    def row(c, m, r):
        return get(c)[m][ng - 1 + r, cols].mean()

    for r in range(4):
        assert abs(row("jx", 0, r) - jx) < rel * abs(jx), ("jx", 0, r)
        assert abs(row("jr", 1, r) - jy) < rel * abs(jy), ("jr", 1, r)
        assert abs(row("jt", 1, r) - (-1j * jy)) < rel * abs(jy), ("jt", 1, r)
        for c, m in (("jx", 1), ("jr", 0), ("jt", 0)):
            assert abs(row(c, m, r)) < leak * abs(jy), (c, m, r)
    assert not get("jx")[1][ng - 1, :].any() and not get("jt")[0][ng - 1, :].any()


def test_drift_current_magnitude_closed_form():
    """The manual's figure value: |J_x| = |J_y| = 3.40e-6 A m^-2."""
    j = expected_current()
    assert abs(j[0]) == pytest.approx(3.40e-6, abs=0.005e-6)
    assert j[0] == pytest.approx(-j[1], rel=1e-12, abs=0)
    assert j[2] == 0.0


def test_drift_current_modes(oracle):
    dt = oracle.set_dt(DX, DY, 0.95)
    st = oracle.State(NX, NY, M, DX, DY, dt, bc_field=("periodic", "periodic", "other", "conduct"),
                      bc_particle=("periodic", "periodic", "other", "reflect"))
    p = drift_plasma()
    st.add_species(-deck.Q0, deck.M0, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    st.push_particles()
    st.current_finish()
    means = interior_means(lambda c: st.arr[c], st.g.ng)
    check_mode_structure(means, expected_current())
    check_axis_rows(lambda c: st.arr[c], st.g.ng, expected_current())


def test_drift_check_catches_a_dropped_mode_factor(oracle):
    """The same check fails when the m = 1 current loses the factor 2 of the
    mode projection (halving the odd-mode currents, as a dropped factor would)."""
    dt = oracle.set_dt(DX, DY, 0.95)
    st = oracle.State(NX, NY, M, DX, DY, dt, bc_field=("periodic", "periodic", "other", "conduct"),
                      bc_particle=("periodic", "periodic", "other", "reflect"))
    p = drift_plasma()
    st.add_species(-deck.Q0, deck.M0, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    st.push_particles()
    st.current_finish()
    means = interior_means(lambda c: st.arr[c] * np.array([1.0, 0.5])[:, None, None], st.g.ng)
    with pytest.raises(AssertionError):
        check_mode_structure(means, expected_current())
