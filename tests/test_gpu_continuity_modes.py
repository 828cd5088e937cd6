"""Continuity of every azimuthal mode across a push, on the device
(cylpic_continuity_begin / _end / _modes).

* The x-r identity: per cell and mode, Q_m(t+3dt/2) - Q_m(t+dt/2) + dt (net
  Jx, Jr out of the cell) equals the azimuthal transport Z_m of the
  reference's m-factors (f_k = int_0^1 s^(k-2) e^{i m dtheta s} ds), so the
  residual stays at the level of the arithmetic in every mode (a wrong
  m-factor, phase or flux sign leaves O(1) of the change).  FP64: below
  1e-7 of the largest per-cell charge change.  m = 0 is at 1e-14; for
  m > 0 the deposit evaluates f3 = (e^{ix}(1 - ix) - 1) / x^2 in closed form
  down to |x| = 1e-4 (particles.F90:589-626), losing about 2 eps / x^2
  relative (2e-8 there), and f3 carries the weight-change part of Jx, Jr
  (measured 1e-9).  Mixed: 1e-4 (its FP32 half steps and currents; the
  mixed m = 0 check measures 2.7e-7 on the same load).
* The J_theta part: one particle, so |Z - Zp| / |Z| is the same in every
  cell and equals |1 - v_theta dt / (r dtheta)| from its own half-step
  positions (to 1e-9).  Away from the theta = pi cut, r dtheta /
  (v_theta dt) - 1 matches the leading-order (dt / r)^2 (3 v_r^2 -
  v_theta^2) / 12 to 5 %.  A particle crossing the cut has dtheta off by
  2 pi (the difference of two ATAN2 values, particles.F90:321-322, 522-524):
  its Jx, Jr carry the longer turn, its J_theta does not, and the ratio is
  about 1; on a plasma these few particles dominate |Z - Zp|."""
import math

import numpy as np
import pytest

from cylindrical_epoch_amd import deck
from cylindrical_epoch_amd.host import Simulation
from gpu_helpers import C_LIGHT, M0, Q0, add_plasma, load_fields, log, make_pair, random_fields

pytestmark = pytest.mark.gpu

NX, NY, M, DX, DY = 48, 24, 3, 1e-7, 1e-7


@pytest.mark.parametrize("precision,bound,tile", [("fp64", 1e-7, None), ("mixed", 1e-4, None),
                                                  ("fp64", 1e-7, (32, 8))])
def test_mode_identity(oracle, precision, bound, tile):
    """(tile None: the library's default, whose check sums per tile in LDS;
    32 x 8 with 3 modes needs more than 64 KB for that, and the check adds
    each word to global memory instead)"""
    log("continuity of every mode", precision, tile)
    st, sim = make_pair(oracle, NX, NY, M, DX, DY, precision=precision, tile=tile)
    rng = np.random.default_rng(3)
    load_fields(st, sim, random_fields(st.shape, rng, e_amp=1e10, b_amp=50.0, j_amp=0.0))
    add_plasma(st, sim, NX, NY, DX, DY, 6, temp_ev=20000.0, drift=(0.3, 0.2, -0.4), seed=13)
    sim.setup_field_boundaries()
    for k in range(3):
        sim.update_eb_fields_half()
        sim.continuity_begin()
        sim.push_particles()
        res0, dq0, _ = sim.continuity_end()
        sim.current_finish()
        sim.update_eb_fields_final()
        cm = sim.continuity_modes()
        log(f"  push {k}: residual / max |dQ| per mode", cm["residual"] / cm["dq"],
            " |Z - Zp| / max |dQ|", cm["jtheta"][1:] / cm["dq"][1:])
        # m = 0 is the m = 0 check's own charge change (its charges are summed
        # in another order, and the mixed one weights in FP32)
        assert cm["dq"][0] == pytest.approx(dq0, rel=1e-9 if precision == "fp64" else 1e-4, abs=0)
        assert cm["residual"][0] < bound * dq0 and res0 < bound * dq0
        assert np.all(cm["dq"] > 0) and np.all(cm["z"][1:] > 0) and cm["z"][0] == 0
        assert np.all(cm["residual"] < bound * cm["dq"])
        # the azimuthal term is a large part of the change for m > 0
        assert np.all(cm["z"][1:] > 1e-3 * cm["dq"][1:])
    sim.close()


@pytest.mark.parametrize("precision,bound", [("fp64", 1e-7), ("mixed", 1e-4)])
def test_mode_identity_with_stale_slots_and_injection(precision, bound):
    """The same identity where a sort decided on the device has left stale
    copies behind the kept particles (open boundaries, a hot plasma that
    loses particles, a sort every second push) and a window shift has
    appended a column: only the particles count (the check follows the
    tile ranges and the appended ranges, not every slot up to the count)."""
    log("continuity of every mode, stale slots", precision)
    bf, bp = ("simple_outflow", "simple_outflow", "other", "conduct"), ("open", "open", "other", "open")
    dt = deck.set_dt(DX, DY, 0.95)
    sim = Simulation(NX, NY, M, DX, DY, dt, bc_field=bf, bc_particle=bp, precision=precision, tile=(8, 4),
                     sort_interval=2)
    p = deck.uniform_plasma(NX, NY, DX, DY, 6, 1e25, 2.0e5, seed=17)
    sid = sim.set_species(-Q0, M0, False, capacity=2 * len(p["x"]))
    sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])
    sim.setup_field_boundaries()
    n0 = len(p["x"])
    for k in range(5):
        sim.step()
    col = deck.uniform_plasma(NX, NY, DX, DY, 6, 1e25, 2.0e5, x_min=sim.x_min, seed=99, cells_x=[NX])
    sim.inject_particles(sid, col["x"], col["y"], col["z"], col["px"], col["py"], col["pz"], col["w"],
                         ids=col["id"] + 10 ** 6)
    sim.move_window(1)
    # push 6: sorts in its push_end, so the after pass sees the new ranges
    sim.update_eb_fields_half()
    sim.continuity_begin()
    sim.push_particles()
    sim.continuity_end()
    cm = sim.continuity_modes()
    log("  residual / max |dQ| per mode", cm["residual"] / cm["dq"])
    assert np.all(cm["dq"] > 0)
    assert np.all(cm["residual"] < bound * cm["dq"])
    n = sim.particle_count(sid)
    log(f"  particles {n0} + {len(col['x'])} injected -> {n}")
    assert n < n0 + len(col["x"])
    sim.close()


@pytest.mark.parametrize("crossing", [False, True])
def test_jtheta_part_single_particle(crossing):
    log("J_theta part, one particle, crossing the cut:", crossing)
    dt = deck.set_dt(DX, DY, 0.95)
    sim = Simulation(NX, NY, M, DX, DY, dt, precision="fp64")
    # near theta = pi with v_theta > 0 the half steps straddle the cut
    u, th0 = (np.array([0.4, -0.3, -1.3]), math.pi - 0.03) if crossing else (np.array([0.4, 1.3, 0.9]), 0.7)
    r0, x0 = 11.3 * DY, 24.4 * DX
    p = [np.array([v]) for v in (x0, r0 * math.cos(th0), r0 * math.sin(th0), *(u * M0 * C_LIGHT), 1e3)]
    sid = sim.set_species(-Q0, M0, False, capacity=1)
    sim.upload_particles(sid, *p)
    sim.continuity_begin()
    sim.push_particles()
    sim.continuity_end()
    cm = sim.continuity_modes()
    got = sim.download_particles(sid)
    x1 = np.array([got["x"][0], got["y"][0], got["z"][0]])
    u1 = np.array([got["px"][0], got["py"][0], got["pz"][0]]) / (M0 * C_LIGHT)
    gam = math.sqrt(1 + u1 @ u1)
    d = u1 * C_LIGHT * dt / 2 / gam
    a, b = x1 - d, x1 + d
    dth = math.atan2(b[2], b[1]) - math.atan2(a[2], a[1])
    assert (abs(dth) > 6) == crossing
    r1 = math.hypot(x1[1], x1[2])
    v = u1 * C_LIGHT / gam
    vt = (-v[1] * x1[2] + v[2] * x1[1]) / r1
    vr = (v[1] * x1[1] + v[2] * x1[2]) / r1
    want = abs(1 - vt * dt / (r1 * dth))
    for m in range(1, M):
        ratio = cm["jtheta"][m] / cm["z"][m]
        log(f"  m={m}: |Z - Zp| / |Z| {ratio:.6e}, from the particle {want:.6e}")
        assert ratio == pytest.approx(want, rel=1e-9, abs=0)
        assert cm["residual"][m] < 1e-7 * cm["dq"][m]
    if not crossing:
        lead = (dt / r1) ** 2 * (3 * vr ** 2 - vt ** 2) / 12
        k = r1 * dth / (vt * dt) - 1
        log(f"  r dtheta / (v_theta dt) - 1 = {k:.6e}, leading order {lead:.6e}")
        assert k == pytest.approx(lead, rel=0.05, abs=0)
    sim.close()
