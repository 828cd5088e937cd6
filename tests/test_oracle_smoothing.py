"""The oracle's current smoothing (smooth_current / smooth_mode_array,
src/housekeeping/current_smooth.F90:50-58, 145-196) checked by what the
stencil fixes: a Fourier mode exp(i(kx ix + ky iy)) is an eigenvector of one
pass with factor alpha + 2 beta (cos kx s + cos ky s), so n passes multiply
it by the product of the per-pass factors away from the ghost-fed rows; a
constant array is left unchanged by the alpha = 1/2, beta = 1/8 pass; alpha
switches to its/2 + 1 only after an iteration numbered above its."""
import numpy as np
import pytest

NX, NY, M, NG = 32, 40, 2, 5


def _state(oracle, bf=("periodic", "periodic", "other", "conduct")):
    dx = dy = 1e-7
    return oracle.State(NX, NY, M, dx, dy, oracle.set_dt(dx, dy, 0.95), bc_field=bf,
                        bc_particle=("periodic", "periodic", "other", "reflect"))


def _mode(kx, ky, shape):
    # array column c is Fortran ix = c + 1 - ng, row r is iy = r + 1 - ng
    ix = np.arange(shape[2]) + 1 - NG
    iy = np.arange(shape[1]) + 1 - NG
    ph = np.exp(1j * (kx * ix[None, :] + ky * iy[:, None]))
    return np.broadcast_to(ph, shape).copy()


def _alphas(its, comp, nstrides):
    out, alpha = [], 0.5
    for it in range(1, its + comp + 1):
        out += [alpha] * nstrides
        if it > its:
            alpha = its * 0.5 + 1.0
    return out


@pytest.mark.parametrize("its,comp,strides", [(1, 0, (1,)), (3, 0, (1,)), (2, 2, (1,)), (1, 3, (1,)),
                                              (2, 0, (1, 2)), (1, 1, (3,)), (2, 1, (1, 2, 4))])
def test_fourier_mode_damping(oracle, its, comp, strides):
    st = _state(oracle)
    kx = 2 * np.pi * 3 / NX  # periodic in x over nx cells
    ky = 0.37
    base = {n: _mode(kx, ky, st.shape) * (1 + 0.5 * j) for j, n in enumerate(("jx", "jr", "jt"))}
    for n, a in base.items():
        st.arr[n][:] = a
    st.smooth_current(its, comp, strides)
    beta = 0.125
    factor = 1.0
    for a, s in zip(_alphas(its, comp, len(strides)), list(strides) * (its + comp)):
        factor *= a + 2 * beta * (np.cos(kx * s) + np.cos(ky * s))
    reach = sum(strides) * (its + comp)  # rows the ghost-fed boundary rows can reach
    rows = slice(NG + reach, NG + NY - reach)
    cols = slice(NG, NG + NX)
    for n, a in base.items():
        got = st.arr[n][:, rows, cols]
        want = factor * a[:, rows, cols]
        assert np.allclose(got, want, rtol=0, atol=1e-12 * max(1.0, abs(factor))), n
        # ghost cells are never written
        assert np.array_equal(st.arr[n][:, :NG, :], a[:, :NG, :])
        assert np.array_equal(st.arr[n][:, :, :NG], a[:, :, :NG])


def test_alpha_switch_is_late(oracle):
    """its = 1, comp_its = 1 runs both iterations with alpha = 1/2: the
    switch to its/2 + 1 follows iteration 2, after which none is left."""
    st = _state(oracle)
    kx = 2 * np.pi * 5 / NX
    a = _mode(kx, 0.0, st.shape)
    for n in ("jx", "jr", "jt"):
        st.arr[n][:] = a
    st.smooth_current(1, 1, (1,))
    f = (0.5 + 0.25 * (np.cos(kx) + 1)) ** 2
    rows, cols = slice(NG + 2, NG + NY - 2), slice(NG, NG + NX)
    assert np.allclose(st.arr["jx"][:, rows, cols], f * a[:, rows, cols], atol=1e-13)


def test_constant_unchanged_and_off_is_identity(oracle):
    st = _state(oracle, bf=("conduct", "conduct", "other", "conduct"))
    rng = np.random.default_rng(4)
    c = complex(rng.normal(), rng.normal())
    for n in ("jx", "jr", "jt"):
        st.arr[n][:] = c
    st.smooth_current(3, 0, (1, 2))
    for n in ("jx", "jr", "jt"):
        assert np.array_equal(st.arr[n], np.full(st.shape, c)), n
    r = rng.normal(size=st.shape) + 1j * rng.normal(size=st.shape)
    st.arr["jr"][:] = r
    st.smooth_current(0, 0, (1,))
    assert np.array_equal(st.arr["jr"], r)


def test_periodic_halo_is_used(oracle):
    """With periodic x the pass at ix = 1 reads ix = 0 from the exchanged
    halo (ix = nx), not the stale ghost: a mode periodic over nx stays a
    mode in the first and last columns."""
    st = _state(oracle)
    kx = 2 * np.pi / NX
    a = _mode(kx, 0.0, st.shape)
    a[:, :, :NG] = 99.0  # stale ghosts on both sides
    a[:, :, NG + NX:] = -99.0
    for n in ("jx", "jr", "jt"):
        st.arr[n][:] = a
    st.smooth_current(1, 0, (2,))
    want = (0.75 + 0.25 * np.cos(2 * kx)) * _mode(kx, 0.0, st.shape)
    rows, cols = slice(NG + 2, NG + NY - 2), slice(NG, NG + NX)
    assert np.allclose(st.arr["jt"][:, rows, cols], want[:, rows, cols], atol=1e-13)
