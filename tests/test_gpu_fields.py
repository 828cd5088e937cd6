"""GPU parity of the per-mode FDTD solver and the field boundaries against
the FP64 oracle: same random inputs, bit-identical FP64 outputs (both sides
evaluate the reference's expressions in the same order without fused
multiply-add)."""
import numpy as np
import pytest

from gpu_helpers import load_fields, log, make_pair, random_fields

pytestmark = pytest.mark.gpu

CASES = [
    # (bc_field, bc_particle)
    (("periodic", "periodic", "other", "conduct"), ("periodic", "periodic", "other", "reflect")),
    (("conduct", "conduct", "other", "conduct"), ("reflect", "reflect", "other", "reflect")),
    (("simple_laser", "simple_outflow", "other", "open"), ("open", "open", "other", "open")),
    (("clamp", "zero_gradient", "other", "zero_b"), ("open", "open", "other", "open")),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("M", [1, 2, 4])
def test_update_eb_fields_half(oracle, case, M):
    log("fields_half case", case, "M", M)
    bf, bp = CASES[case]
    nx, ny = 37, 13
    st, sim = make_pair(oracle, nx, ny, M, bc_field=bf, bc_particle=bp)
    rng = np.random.default_rng(100 + 7 * case + M)
    load_fields(st, sim, random_fields(st.shape, rng))
    st.update_eb_fields_half()
    sim.update_eb_fields_half()
    for n in st.arr:
        got = sim.download_field(n)
        assert np.array_equal(got, st.arr[n]), (n, np.max(np.abs(got - st.arr[n])))
    sim.close()


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("M", [1, 3])
def test_update_eb_fields_final_with_laser(oracle, case, M):
    log("fields_final case", case, "M", M)
    bf, bp = CASES[case]
    nx, ny = 29, 11
    st, sim = make_pair(oracle, nx, ny, M, bc_field=bf, bc_particle=bp)
    rng = np.random.default_rng(300 + 7 * case + M)
    load_fields(st, sim, random_fields(st.shape, rng))
    st.setup_field_boundaries()
    sim.setup_field_boundaries()
    # change the live fields after the snapshot so the snapshot matters
    f2 = random_fields(st.shape, rng)
    load_fields(st, sim, {k: v for k, v in f2.items() if not k.endswith("_old")})
    src_min = rng.normal(size=2 * (ny + 1)) * 1e12
    src_max = rng.normal(size=2 * (ny + 1)) * 1e12
    st.update_eb_fields_final(src_min, src_max)
    sim.update_eb_fields_final(src_min, src_max)
    for n in st.arr:
        got = sim.download_field(n)
        assert np.array_equal(got, st.arr[n]), (n, np.max(np.abs(got - st.arr[n])))
    sim.close()


def test_current_finish(oracle):
    log("current_finish")
    for bf, bp in CASES:
        st, sim = make_pair(oracle, 24, 12, 3, bc_field=bf, bc_particle=bp)
        rng = np.random.default_rng(5)
        load_fields(st, sim, random_fields(st.shape, rng))
        st.current_finish()
        sim.current_finish()
        for n in ("jx", "jr", "jt"):
            got = sim.download_field(n)
            assert np.array_equal(got, st.arr[n]), (bf, n)
        sim.close()


SMOOTH = [(1, 0, (1,)), (2, 1, (1, 2)), (1, 3, (3,)), (3, 0, (1, 2, 5))]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("sm", range(len(SMOOTH)))
def test_current_finish_with_smoothing(oracle, case, sm):
    """smooth_current after the current bcs: FP64 bit-identical."""
    log("current_finish smoothing case", case, "smooth", SMOOTH[sm])
    bf, bp = CASES[case]
    st, sim = make_pair(oracle, 27, 14, 3, bc_field=bf, bc_particle=bp)
    rng = np.random.default_rng(50 + case + 10 * sm)
    load_fields(st, sim, random_fields(st.shape, rng))
    its, comp, strides = SMOOTH[sm]
    sim.set_smoothing(its, comp, strides)
    for _ in range(2):  # the second call reuses the work buffer
        st.current_finish(smooth=SMOOTH[sm])
        sim.current_finish()
    for n in ("jx", "jr", "jt"):
        got = sim.download_field(n)
        assert np.array_equal(got, st.arr[n]), (bf, n, np.max(np.abs(got - st.arr[n])))
    sim.close()


def test_current_finish_with_smoothing_mixed(oracle):
    """The FP32 field path: the same passes in FP32, relative error of a
    few FP32 ulps of the largest current."""
    log("current_finish smoothing mixed")
    bf, bp = CASES[2]
    st, sim = make_pair(oracle, 27, 14, 3, bc_field=bf, bc_particle=bp, precision="mixed")
    rng = np.random.default_rng(77)
    f = random_fields(st.shape, rng)
    load_fields(st, sim, f)
    for n in f:  # the oracle starts from the FP32-rounded arrays the device holds
        st.arr[n][:] = sim.download_field(n)
    sim.set_smoothing(2, 1, (1, 2))
    st.current_finish(smooth=(2, 1, (1, 2)))
    sim.current_finish()
    for n in ("jx", "jr", "jt"):
        got = sim.download_field(n)
        scale = np.max(np.abs(st.arr[n]))
        assert np.max(np.abs(got - st.arr[n])) <= 1e-6 * scale, n
    sim.close()


def test_smoothing_rejects_bad_strides(oracle):
    st, sim = make_pair(oracle, 16, 8, 1)
    with pytest.raises(Exception):
        sim.set_smoothing(1, 0, (6,))  # beyond ng = 5
    with pytest.raises(Exception):
        sim.set_smoothing(-1, 0, (1,))
    sim.close()


def test_mode_energy(oracle):
    log("mode_energy")
    st, sim = make_pair(oracle, 40, 20, 3)
    rng = np.random.default_rng(9)
    load_fields(st, sim, random_fields(st.shape, rng))
    we0, wb0 = st.mode_energy()
    we1, wb1 = sim.mode_energy()
    assert np.allclose(we0, we1, rtol=1e-12) and np.allclose(wb0, wb1, rtol=1e-12)
    sim.close()
