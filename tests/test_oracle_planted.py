"""Planted mistakes: each entry of PLANTED is a plausible slip in one term of
the oracle (oracle/cylpic_oracle.c, the `PLANT(k, ...)` sites, built with
-DORACLE_PLANT=k by `make -C oracle plant`), and each must fail the
committed test named beside it.  This is what makes the oracle's checks
checks: a restatement error of that kind cannot pass the suite.

EQUIVALENT lists mistakes that no test can catch because they change
nothing the program computes; for those the test runs a laser-driven plasma
with both builds and requires bit-identical fields and particles.  No GPU
needed."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
F, L, I, A, W, D, S, M = ("test_oracle_fdtd_modes.py", "test_oracle_laser.py", "test_oracle_invariants.py",
                          "test_oracle_axis_fold.py", "test_oracle_window.py", "test_oracle_diagnostics.py",
                          "test_oracle_smoothing.py", "test_oracle_mode_factors.py")

# k: (the mistake, the reference lines it breaks, the test that must fail)
PLANTED = {
    1: ("m = 0 E_x axis row: factor 4 c^2/dr -> 2 c^2/dr", "fields.f90:121-123",
        I + "::test_axis_gauss_law_m0"),
    2: ("m = 1 E_theta axis row: -i/8 -> +i/8", "fields.f90:146-147",
        F + "::test_m1_plane_wave_exact_discrete_dispersion[y]"),
    3: ("m = 1 E_r row at -dr/2 set after E_theta's axis row", "fields.f90:144-147",
        F + "::test_m1_er_reflected_row_lags_etheta_axis_row"),
    4: ("m = 0 E_r row at -dr/2: -E_r(1) -> +E_r(1)", "fields.f90:127",
        F + "::test_m0_bessel_beam_second_order[TM]"),
    5: ("E_x update: c^2/r_p -> c^2/r_d", "fields.f90:77-79",
        F + "::test_bessel_mode_second_order[TM-0]"),
    6: ("E_r update: i m c^2/r_d -> i m c^2/r_p", "fields.f90:80",
        F + "::test_bessel_mode_second_order[TE-1]"),
    7: ("m > 1 first E_r row: E_r(2)/9 -> E_r(2)/4", "fields.f90:165",
        F + "::test_te3_first_er_row"),
    8: ("m = 0 B_x row at -dr/2: +B_x(1) -> -B_x(1)", "fields.f90:258",
        F + "::test_bessel_mode_second_order[TE-0]"),
    9: ("m = 1 B_theta axis row: -2i B_r - B_theta(1) -> -2i B_r + B_theta(1)", "fields.f90:276-277",
        F + "::test_m1_plane_wave_exact_discrete_dispersion[y]"),
    10: ("m = 1 B_r axis row (eq. 22): sign of d/dx E_theta", "fields.f90:270-273",
         F + "::test_m1_plane_wave_exact_discrete_dispersion[z]"),
    11: ("J_theta fold of r < 0 rows: -mode_sign -> +mode_sign", "boundary.F90:1925",
         A + "::test_jtheta_fold_m1_uniform_drift"),
    12: ("J_x fold of r < 0 rows: +mode_sign -> -mode_sign", "boundary.F90:1923",
         I + "::test_continuity_mode0_after_axis_fold[0]"),
    13: ("J_r fold of row 0 into row 1 dropped", "boundary.F90:1929",
         I + "::test_continuity_mode0_after_axis_fold[0]"),
    14: ("m = 0 J_x axis rule: divisor 3 -> 4", "boundary.F90:1942",
         "test_oracle_drift_modes.py::test_drift_current_modes"),
    15: ("m = 1 J_theta axis row: -i -> +i", "boundary.F90:1947",
         "test_oracle_drift_modes.py::test_drift_current_modes"),
    16: ("m = 1 J_r axis row: 2i J_theta - J_r(1) -> 2i J_theta + J_r(1)", "boundary.F90:1949",
         "test_oracle_drift_modes.py::test_drift_current_modes"),
    17: ("axis fold's mode sign does not alternate", "boundary.F90:1932",
         A + "::test_jtheta_fold_m1_uniform_drift"),
    18: ("periodic current ghost sum at x_min dropped", "boundary.F90:1133-1245",
         A + "::test_jtheta_fold_m0_rigid_rotation"),
    19: ("r_max J_r fold: radius ratio inverted", "boundary.F90:986-1013",
         A + "::test_rmax_reflection_folds_m1_uniform_drift"),
    20: ("x_min laser source: 4 source_t -> 2 source_t", "laser.f90:490",
         L + "::test_laser_pulse_enters_and_leaves[y]"),
    21: ("source_t = source1 + i source2 -> source1 - i source2", "laser.f90:483",
         L + "::test_laser_pulse_enters_and_leaves[z]"),
    22: ("x_min B_theta: sign of the diff term", "laser.f90:495",
         L + "::test_laser_pulse_enters_and_leaves[y]"),
    23: ("x_min B_r: etm(1) -> etm(2)", "laser.f90:511",
         L + "::test_laser_pulse_enters_and_leaves[y]"),
    24: ("x_max sum: 1/(lx + c) -> 1/lx", "laser.f90:541",
         L + "::test_laser_pulse_enters_and_leaves[y]"),
    25: ("r_max B_x: sum_x -> sum_t", "laser.f90:664",
         L + "::test_rmax_outflow_keeps_a_uniform_axial_b"),
    26: ("r_max B_theta: sign of the E_r term", "laser.f90:678-679",
         L + "::test_rmax_outflow_absorbs[1-bx]"),
    28: ("r_max B_x: sign of the bxm_old(ny-1) term", "laser.f90:666",
         L + "::test_rmax_outflow_keeps_a_uniform_axial_b"),
    30: ("r_max B_theta reads btm(ny) for btm_old(ny)", "laser.f90:676",
         L + "::test_rmax_outflow_time_levels"),
    29: ("x_min r_d_vals: row ir paired with r_d_vals(ir), not r_d_vals(ir-1)", "laser.f90:490-493",
         "test_oracle_laser_pairing.py::test_oracle_pairs_rows_as_the_fortran_compiler_does"),
    31: ("Boris: gamma taken before the first half kick", "particles.F90:424",
         I + "::test_e_cross_b_drift[0.3]"),
    32: ("Boris: second half kick subtracted", "particles.F90:449",
         I + "::test_uniform_ex_mode0_kick[0.3]"),
    33: ("Boris: sign of tau_z in the rotation", "particles.F90:438-439",
         I + "::test_e_cross_b_drift[0.3]"),
    34: ("mode projection factor 2 dropped", "particles.F90:611, 619",
         "test_oracle_drift_modes.py::test_drift_current_modes"),
    35: ("mode phase e^{i m theta} conjugated", "particles.F90:607",
         I + "::test_mode_continuity_structure[0]"),
    36: ("small-angle m_fac_3: i mdth / 3 -> i mdth / 4", "particles.F90:614",
         M + "::test_small_angle_branch[3e-05]"),
    37: ("gather: sign of E_theta in E_y", "include/triangle/e_part.inc:43",
         I + "::test_radial_azimuthal_to_cartesian[et]"),
    38: ("deposit: sign of the J_r increment", "particles.F90:658",
         "test_oracle_doc_example.py::test_doc_tables"),
    39: ("r_max particle reflection keeps p_r", "boundary.F90:1776-1790",
         I + "::test_rmax_reflection_at_any_angle[0.7]"),
    40: ("cell key: + 1.5 -> + 1.0", "split_particle.F90:61",
         I + "::test_cell_keys_brute_force"),
    41: ("moving window x_max fix: E_x(nx) = E_x(nx-1)", "window.F90:120",
         W + "::test_x_max_fix_keeps_a_line_straight"),
    43: ("current smoothing: alpha switched one iteration early", "current_smooth.F90:187-188",
         S + "::test_alpha_switch_is_late"),
    44: ("number density r < dy fold radius -> dy/2", "calc_df.F90:588-661",
         D + "::test_density_axis_fold[0.5]"),
    46: ("mode energy: m = 0 weight 2 pi -> pi", "(the boundary's own diagnostic)",
         D + "::test_mode_energy_of_a_uniform_field[ex0]"),
}

EQUIVALENT = {
    27: ("r_max B_x reads bxm(ny) for bxm_old(ny)", "laser.f90:665",
         "row ny is written only by outflow_bcs_r_max on the columns its B_x loop covers (1..nx-1), so at "
         "that call it still equals its *_old copy "
         "(test_oracle_laser.py::test_rmax_boundary_rows_hold_their_previous_value_when_updated)"),
}


@pytest.fixture(scope="module")
def plants():
    ks = sorted(PLANTED) + sorted(EQUIVALENT)
    subprocess.check_call(["make", "-s", "-j8", "-C", os.path.join(ROOT, "oracle"), "all", "plant",
                           "PLANTS=" + " ".join(str(k) for k in ks)])
    return ks


def _env(k):
    env = dict(os.environ)
    env["CYLPIC_ORACLE_PLANT"] = str(k)
    env["PYTHONDONTWRITEBYTECODE"] = "1"
    return env


@pytest.mark.parametrize("k", sorted(PLANTED))
def test_planted_mistake_fails_its_check(plants, k):
    what, cite, node = PLANTED[k]
    if "pairing" in node and not (shutil.which("amdflang") or shutil.which("flang")):
        pytest.skip("the pairing check needs a Fortran compiler")
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(HERE, node), "-q", "-p", "no:cacheprovider",
                        "-p", "no:randomly"], cwd=ROOT, env=_env(k), capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 1 and " failed" in out, f"planted mistake {k} ({what}, {cite}) passed {node}:\n{out[-2000:]}"


RUN = r"""
import sys, numpy as np
sys.path.insert(0, {root!r})
from oracle import oracle as O
from cylindrical_epoch_amd import deck
nx, ny, M, dx, dy = 40, 24, 2, 1e-7, 1e-7
dt = O.set_dt(dx, dy, 0.95)
st = O.State(nx, ny, M, dx, dy, dt, bc_field=("simple_laser", "simple_outflow", "other", "simple_outflow"),
             bc_particle=("open", "open", "other", "open"))
p = deck.uniform_plasma(nx, ny, dx, dy, 4, 1e25, 500.0, seed=5)
st.add_species(-deck.Q0, deck.M0, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
st.setup_field_boundaries()
r = np.abs((np.arange(ny + 1) - 1) * dy + dy / 2)
for n in range(30):
    s = 3e11 * np.sin(0.3 * n) * np.exp(-(r / (8 * dy)) ** 2)
    st.step(src_xmin=np.concatenate([s, 0.5 * s]))
np.savez(sys.argv[1], **st.arr, **{{k: st.species[0][k] for k in ("x", "y", "z", "px", "py", "pz")}})
"""


@pytest.mark.parametrize("k", sorted(EQUIVALENT))
def test_equivalent_mistake_changes_nothing(plants, k, tmp_path):
    script = tmp_path / "run.py"
    script.write_text(RUN.format(root=ROOT))
    res = {}
    for kk in (0, k):
        out = tmp_path / f"out{kk}.npz"
        subprocess.check_call([sys.executable, str(script), str(out)], cwd=ROOT, env=_env(kk), timeout=600)
        res[kk] = np.load(out)
    for name in res[0].files:
        assert np.array_equal(res[0][name], res[k][name]), name
    assert np.abs(res[0]["bt"]).max() > 0 and np.abs(res[0]["bx"]).max() > 0
