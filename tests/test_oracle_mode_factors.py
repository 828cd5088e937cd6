"""The m > 0 factors of the charge-conserving deposit (particles.F90:589-626,
the documentation's eqs. 98-105) against their closed forms.  No GPU needed.

With phi = m dtheta and e = exp(i m theta_05), the deposit's three factors
are the path integrals of exp(i m theta) along the angle change:

    f2 = 2 e (e^x - 1) / x,
    f3 = 2 e (e^x (1 - x) - 1) / x^2,
    f4 = 2 e (e^x (2 - 2x + x^2) - 2) / x^3,        x = i phi,

whose power series have the general terms 1/n!, (n-1)/n! and
(n-1)(n-2)/n! (shifted by 1, 2, 3 powers).  The tests evaluate 30 terms of
each series, which is the function itself to round-off for |phi| <= 2.

The factors are read out of the deposit: the J_theta contribution of one
stencil cell is fcz * inv_volume * w_xr with w_xr = gx (f2 gy + f3 hy) +
hx (f3 gy + f4 hy) for m > 0, so one nonzero weight per direction isolates
f2 (gx gy), f3 (gx hy) or f4 (hx hy).  Below |phi| = 1e-4 the reference
switches to three Taylor terms, which must agree with the series to
O(phi^3); above it, the exact forms lose about eps / phi^3 (f4) to
cancellation, so the comparison there starts at phi = 0.01."""
import math

import numpy as np
import pytest


def series(kind, phi, terms=30):
    x = 1j * phi
    if kind == 2:
        coef = [1.0 / math.factorial(n) for n in range(1, terms)]
    elif kind == 3:
        coef = [(n - 1) / math.factorial(n) for n in range(2, terms)]
    else:
        coef = [(n - 1) * (n - 2) / math.factorial(n) for n in range(3, terms)]
    return 2.0 * sum(c * x ** k for k, c in enumerate(coef))


def factors(oracle, phi, theta05=0.4, m=1):
    """(f2, f3, f4) / e of mode m as the oracle's deposit applies them."""
    out = []
    for kind in (2, 3, 4):
        st = oracle.State(16, 16, m + 1, 1e-7, 1e-7, 1e-16)
        z = np.zeros(5)
        gx, hx, gy, hy = z.copy(), z.copy(), z.copy(), z.copy()
        if kind == 2:
            gx[2] = gy[2] = 1.0
        elif kind == 3:
            gx[2] = hy[2] = 1.0
        else:
            hx[2] = hy[2] = 1.0
        dth = phi / m
        e05 = (math.cos(theta05), math.sin(theta05))
        ed = (math.cos(dth), math.sin(dth))
        fcz = 0.75
        oracle.deposit_from_weights(st, gx, hx, gy, hy, 8, 8, 0, 0, 0.0, 0.0, fcz, e05, ed, dth)
        _, _, inv_volume, _ = st.geometry()
        ng = st.g.ng
        jt = st.arr["jt"][m][8 + ng - 1, 8 + ng - 1]
        e = complex(math.cos(m * theta05), math.sin(m * theta05))
        out.append(jt / (fcz * inv_volume[8 + ng]) / e)
    return out


def test_series_reproduce_the_documented_taylor_terms():
    """The series' first three terms are the reference's small-angle forms:
    what is left is below 0.1 phi^3 (the phi^3 coefficients are 2/24, 2/30
    and 2/36 times i^3)."""
    for kind, first in ((2, (1, 0.5, -1 / 6)), (3, (0.5, 1 / 3, -0.125)), (4, (1 / 3, 0.25, -0.1))):
        phi = 1e-3
        taylor = 2.0 * complex(first[0] + first[2] * phi * phi, first[1] * phi)
        assert abs(series(kind, phi) - taylor) < 0.1 * phi ** 3, kind


@pytest.mark.parametrize("phi", [3e-5, -7e-5, 9.9e-5])
def test_small_angle_branch(oracle, phi):
    for kind, got in zip((2, 3, 4), factors(oracle, phi)):
        want = series(kind, phi)
        assert abs(got - want) < 1e-12 * abs(want), (kind, got, want)


@pytest.mark.parametrize("phi", [0.01, -0.05, 0.3, 1.1, -2.0])
def test_exact_branch(oracle, phi):
    for kind, got in zip((2, 3, 4), factors(oracle, phi)):
        want = series(kind, phi)
        tol = 1e-15 / abs(phi) ** (kind - 1) + 1e-13
        assert abs(got - want) < tol * abs(want), (kind, got, want)


def test_second_mode_uses_twice_the_angle(oracle):
    """Mode 2 applies the same forms at phi = 2 dtheta, with e = e^{2 i theta}."""
    for phi in (6e-5, 0.08):
        for kind, got in zip((2, 3, 4), factors(oracle, phi, m=2)):
            want = series(kind, phi)
            assert abs(got - want) < 1e-10 * abs(want), (phi, kind)
