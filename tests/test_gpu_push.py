"""GPU parity of push_particles (Boris push + charge-conserving mode deposit
+ axis fold + particle boundaries) and of the tile sort, against the FP64
oracle.

Particle positions and momenta: bit-identical (same operation order, no
fused multiply-add, correctly rounded sqrt and division on both sides).
Currents: the GPU adds each particle's contributions with atomics in a
different order and atan2 comes from a different math library.  For m = 0
the tolerance is |J_gpu - J_cpu| <= 1e-12 * max|J_0|.  For m > 0 the
reference's exact-branch factors (particles.F90:619-624) lose ~12 eps /
(m dtheta)^3 of relative accuracy near the |m dtheta| = 1e-4 switch, so an
ulp of difference in atan2 moves those currents by up to ~1e-9 of max|J_m|
(tests/test_oracle_invariants.py::test_mode_factor_sensitivity measures
this on the CPU); the tolerance there is 1e-7 * max|J_m|."""
import numpy as np
import pytest

from cylindrical_epoch_amd.host import Simulation
from gpu_helpers import (M0, Q0, add_plasma, by_id, load_fields, log, make_pair, max_rel,
                         random_fields)

pytestmark = pytest.mark.gpu
J_TOL_M0 = 1e-12
J_TOL_M = 1e-7


def _compare_particles(st, sim, sid=0):
    got = by_id(sim.download_particles(sid))
    want = by_id({k: st.species[sid][k] for k in ("x", "y", "z", "px", "py", "pz", "w", "id")})
    assert np.array_equal(got["id"], want["id"])
    for k in ("x", "y", "z", "px", "py", "pz", "w"):
        assert np.array_equal(got[k], want[k]), (k, max_rel(got[k], want[k]))


def _compare_currents(st, sim):
    for n in ("jx", "jr", "jt", "jx_old", "jr_old", "jt_old"):
        got = sim.download_field(n)
        for m in range(got.shape[0]):
            err = max_rel(got[m], st.arr[n][m])
            assert err <= (J_TOL_M0 if m == 0 else J_TOL_M), (n, m, err)


@pytest.mark.parametrize("M,gather_lds", [(1, 0), (2, 0), (3, 0), (2, 1), (3, 1)])
@pytest.mark.parametrize("bcs", ["periodic", "open"])
def test_push_particles_parity(oracle, M, gather_lds, bcs):
    """Particles bit-exact and currents within the stated tolerances, with
    the fields gathered from L1/L2 or from the LDS-staged tile region."""
    log("push parity M", M, "gather_lds", gather_lds, bcs)
    nx, ny, dx, dy = 40, 20, 1e-7, 1e-7
    if bcs == "periodic":
        bf = ("periodic", "periodic", "other", "conduct")
        bp = ("periodic", "periodic", "other", "reflect")
    else:
        bf = ("simple_outflow", "simple_outflow", "other", "open")
        bp = ("open", "open", "other", "open")
    st, sim = make_pair(oracle, nx, ny, M, dx, dy, bc_field=bf, bc_particle=bp, tile=(8, 4))
    sim.tune(gather_lds=gather_lds)
    rng = np.random.default_rng(M)
    f = random_fields(st.shape, rng, e_amp=3e10, b_amp=200.0, j_amp=0.0)
    load_fields(st, sim, f)
    # relativistic plasma: many particles cross cells, the axis and the edges
    add_plasma(st, sim, nx, ny, dx, dy, ppc=6, temp_ev=2.0e5, drift=(0.3, 0.0, 0.0), seed=M)
    st.push_particles()
    sim.push_particles()
    _compare_particles(st, sim)
    _compare_currents(st, sim)
    sim.close()


@pytest.mark.parametrize("M,gather_lds", [(1, 0), (3, 0), (3, 1)])
def test_unsorted_pushes_parity(oracle, M, gather_lds):
    """Four pushes of a relativistic plasma without a sort, on 4 x 2 tiles:
    particles that left their tile's neighbourhood (their stencil can reach
    past the tile's current box) take the FP64 push's second pass, which
    deposits straight into global J.  After every push the particles are
    bit-exact and the currents within the push tolerances."""
    log("unsorted pushes M", M, "gather_lds", gather_lds)
    nx, ny, dx, dy = 32, 16, 1e-7, 1e-7
    bf = ("periodic", "periodic", "other", "conduct")
    bp = ("periodic", "periodic", "other", "reflect")
    st, sim = make_pair(oracle, nx, ny, M, dx, dy, bc_field=bf, bc_particle=bp, tile=(4, 2), sort_interval=1000)
    sim.tune(gather_lds=gather_lds)
    rng = np.random.default_rng(10 + M)
    load_fields(st, sim, random_fields(st.shape, rng, e_amp=3e10, b_amp=200.0, j_amp=0.0))
    add_plasma(st, sim, nx, ny, dx, dy, ppc=6, temp_ev=2.0e5, drift=(0.4, 0.0, 0.0), seed=20 + M)
    sp = st.species[0]
    c0 = np.floor(sp["x"] / dx), np.floor(np.hypot(sp["y"], sp["z"]) / dy)
    for k in range(4):
        st.push_particles()
        sim.push_particles()
        _compare_particles(st, sim)
        _compare_currents(st, sim)
    moved = (np.abs(np.floor(sp["x"] / dx) - c0[0]) >= 2) | (np.abs(np.floor(np.hypot(sp["y"], sp["z"]) / dy) - c0[1]) >= 2)
    log(f"  {int(moved.sum())} of {len(moved)} particles moved two cells or more without a sort")
    assert moved.mean() > 0.05
    sim.close()


def test_current_finish_after_push(oracle):
    log("push + current_finish")
    nx, ny, dx, dy = 32, 16, 1e-7, 1e-7
    st, sim = make_pair(oracle, nx, ny, 2, dx, dy, tile=(16, 8))
    add_plasma(st, sim, nx, ny, dx, dy, ppc=8, temp_ev=5e4, seed=4)
    st.push_particles()
    sim.push_particles()
    st.current_finish()
    sim.current_finish()
    _compare_currents(st, sim)
    sim.close()


def interleave_within_tiles(order, keys, cells, ntiles):
    """The per-tile rounds interleave: `order` in the stable order of the
    sort keys `keys` (4 x tile-major cell key + quadrant; removed particles
    keyed ntiles * cells * 4, last); each tile's particles are laid out as
    rounds, round r holding the r-th particle of every cell that has more
    than r, in cell order.  Written as loops over tiles, rounds and cells."""
    out = []
    n = len(order)
    cell = keys // 4
    i = 0
    while i < n and keys[i] < ntiles * cells * 4:
        t = keys[i] // (cells * 4)
        members = {}
        while i < n and keys[i] // (cells * 4) == t:
            members.setdefault(int(cell[i]), []).append(order[i])
            i += 1
        r = 0
        while any(len(v) > r for v in members.values()):
            for c in sorted(members):
                if len(members[c]) > r:
                    out.append(members[c][r])
            r += 1
    out.extend(order[i:])
    return np.array(out, dtype=order.dtype)


def half_cell_quadrant(st, x, y, z):
    """The upload's sub-key below the cell: 2 [r in the lower half of its
    cell] + [x in the lower half], from the key's own expressions
    (split_particle.F90:61-62 before the FLOOR)."""
    fx = (x - st.g.x_grid_min) / st.g.dx + 1.5
    fy = (np.sqrt(y * y + z * z) - st.g.y_grid_min) / st.g.dy + 1.5
    return 2 * (fy - np.floor(fy) < 0.5).astype(np.int32) + (fx - np.floor(fx) < 0.5).astype(np.int32)


@pytest.mark.parametrize("tile", [(4, 4), (16, 8), (7, 3)])
def test_sort_matches_stable_cpu_sort(oracle, tile):
    """The upload sorts particles by their tile-major cell key and, within a
    cell, by half-cell quadrant (stable), then interleaves each tile's
    particles; the downloaded order must equal that permutation of the
    uploaded order computed on the CPU (bit-exact)."""
    log("sort tile", tile)
    nx, ny, dx, dy = 48, 20, 1e-7, 1e-7
    st, sim = make_pair(oracle, nx, ny, 1, dx, dy, tile=tile)
    p = add_plasma(st, sim, nx, ny, dx, dy, ppc=40, temp_ev=100.0, seed=11, shuffle=True)
    cells = tile[0] * tile[1]
    keys = 4 * oracle.cell_keys(st, p["x"], p["y"], p["z"], tile[0], tile[1]) + half_cell_quadrant(
        st, p["x"], p["y"], p["z"])
    ntiles = ((nx + tile[0] - 1) // tile[0]) * ((ny + tile[1] - 1) // tile[1])
    perm = oracle.stable_sort(keys, ntiles * cells * 4)
    want = interleave_within_tiles(p["id"][perm], keys[perm], cells, ntiles)
    got = sim.download_particles(0)
    assert np.array_equal(got["id"], want)
    # a second, plain sort keeps the order (no particle moved)
    sim.sort_particles()
    got2 = sim.download_particles(0)
    assert np.array_equal(got2["id"], want)
    sim.close()


def test_multi_step_thermal(oracle):
    """Ten full steps of a small thermal plasma: fields and particles stay
    within round-off growth of the oracle."""
    log("multi-step thermal")
    nx, ny, dx, dy = 32, 16, 1e-7, 1e-7
    st, sim = make_pair(oracle, nx, ny, 2, dx, dy, tile=(8, 8))
    add_plasma(st, sim, nx, ny, dx, dy, ppc=8, temp_ev=1000.0, seed=2)
    for _ in range(10):
        st.step()
        sim.step()
    for n in ("ex", "er", "et", "bx", "br", "bt", "jx", "jr", "jt"):
        got = sim.download_field(n)
        for m in range(got.shape[0]):
            err = max_rel(got[m], st.arr[n][m])
            assert err < (1e-10 if m == 0 else 1e-6), (n, m, err)
    got = by_id(sim.download_particles(0))
    want = by_id({k: st.species[0][k] for k in ("x", "y", "z", "px", "py", "pz", "w", "id")})
    for k in ("x", "y", "z", "px", "py", "pz"):
        assert max_rel(got[k], want[k]) < 1e-9, k
    sim.close()


@pytest.mark.parametrize("precision,tile,bc,temp_ev,pushes,method", [
    ("fp64", (8, 4), "periodic", 100.0, 1, "migrate"),
    ("fp64", (1, 1), "open", 2.0e5, 1, "migrate"),
    ("fp64", (16, 8), "periodic", 2.0e5, 1, "migrate"),
    ("fp64", (2, 1), "periodic", 2.0e5, 1, "migrate"),
    ("fp64", (1, 1), "periodic", 2.0e5, 5, "radix"),
    ("mixed", (8, 4), "periodic", 2.0e5, 1, "migrate"),
    ("mixed", (4, 2), "open", 2.0e5, 1, "migrate"),
    ("fp64", (8, 4), "periodic", 100.0, 1, "cells"),
    ("fp64", (4, 4), "open", 2.0e5, 3, "cells"),
    ("fp64", (16, 8), "periodic", 2.0e5, 2, "cells"),
    ("mixed", (8, 4), "open", 2.0e5, 2, "cells"),
])
def test_sort_after_push_matches_stable_cpu_sort(oracle, precision, tile, bc, temp_ev, pushes, method):
    """After `pushes` pushes without a sort, the next sort puts the particles
    in the stable order of their new tile keys, removed ones dropped
    (bit-exact against the CPU stable sort of the keys).  One push moves a
    particle at most one tile, so the neighbour migration sort applies; five
    pushes of a hot plasma on 1 x 1 tiles move some farther, and the sort
    falls back to the radix passes.  The cell method re-keys by cell and
    half-cell quadrant and lays each tile out in rounds, as the upload does."""
    log("sort after push", precision, tile, bc, temp_ev, pushes)
    nx, ny, dx, dy = 24, 12, 1e-7, 1e-7
    bcf = (bc, bc, "other", "conduct") if bc == "periodic" else ("simple_outflow", "simple_outflow", "other",
                                                                   "conduct")
    bcp = (bc, bc, "other", "reflect") if bc == "periodic" else ("open", "open", "other", "open")
    st, sim = make_pair(oracle, nx, ny, 1, dx, dy, bc_field=bcf, bc_particle=bcp, precision=precision,
                        tile=tile, sort_interval=1000)
    if method == "cells":
        sim.set_sort_method("cells")
    sim_inputs = add_plasma(st, sim, nx, ny, dx, dy, ppc=6, temp_ev=temp_ev, seed=31, shuffle=True)
    before = sim.download_particles(0)["id"]
    c0 = sim.sort_counts()
    for _ in range(pushes):
        sim.push_particles()
        if precision == "fp64":
            st.push_particles()
    got = sim.download_particles(0)
    c1 = sim.sort_counts()
    counted = "radix" if method == "cells" else method
    assert c1[counted] == c0[counted] + 1, (c0, c1)
    ntiles = ((nx + tile[0] - 1) // tile[0]) * ((ny + tile[1] - 1) // tile[1])
    cells = tile[0] * tile[1]
    if precision == "fp64" and method == "cells":
        sp = st.species[0]
        pos = {i: k for k, i in enumerate(sp["id"])}
        ck = 4 * oracle.cell_keys(st, sp["x"], sp["y"], sp["z"], tile[0], tile[1]) + half_cell_quadrant(
            st, sp["x"], sp["y"], sp["z"])
        keys = np.array([ck[pos[i]] if i in pos else ntiles * cells * 4 for i in before], np.int64)
        perm = oracle.stable_sort(keys, ntiles * cells * 4 + 1)
        want = interleave_within_tiles(before[perm], keys[perm], cells, ntiles)[: int(np.sum(keys < ntiles * cells * 4))]
        log(f"  {len(before)} -> {len(want)} particles, sort counts {c1}")
        assert np.array_equal(got["id"], want)
        sim.close()
        return
    if precision == "fp64":
        # positions are bit-exact with the oracle: keys from the oracle's particles
        sp = st.species[0]
        pos = {i: k for k, i in enumerate(sp["id"])}
        keys_alive = oracle.cell_keys(st, sp["x"], sp["y"], sp["z"], tile[0], tile[1]) // (tile[0] * tile[1])
        keys = np.array([keys_alive[pos[i]] if i in pos else ntiles for i in before], np.int32)
    else:
        # the mixed path's keys come from its FP32 state: compare with the same
        # run sorted by the radix passes (the push of a particle does not
        # depend on the others, so both runs hold identical keys)
        ref = Simulation(nx, ny, 1, dx, dy, st.g.dt, bc_field=bcf, bc_particle=bcp, precision=precision,
                         tile=tile, sort_interval=1000)
        ref.set_sort_method("radix")
        sid = ref.set_species(-Q0, M0, False, capacity=len(before))
        p = sim_inputs
        ref.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])
        for _ in range(pushes):
            ref.push_particles()
        want = ref.download_particles(0)
        assert ref.sort_counts()["migrate"] == 0
        ref.close()
        if method == "cells":
            # another order of the same particles: compare them by id
            got, want = by_id(got), by_id(want)
        for k in ("id", "x", "y", "z", "px", "py", "pz", "w"):
            assert np.array_equal(got[k], want[k]), k
        sim.close()
        return
    perm = oracle.stable_sort(keys, ntiles + 1)
    want = before[perm][: int(np.sum(keys < ntiles))]
    log(f"  {len(before)} -> {len(want)} particles, sort counts {c1}")
    assert np.array_equal(got["id"], want)
    sim.close()


def test_device_decided_sorts_in_sequence(oracle):
    """The sort is chosen on the device (migration, or radix after a move
    beyond the tile neighbourhood) and the kept count stays there: a
    migration sort that drops particles leaves stale slots behind the kept
    ones, and a later radix fallback over the host's upper bound must not
    bring them back.  Open boundaries, 4 x 1 tiles: one push and a sort
    (migration; a window shift of two cells, within a tile's width, removes
    the first two columns), five pushes and a sort (radix: some particles
    move two or more tiles in r), against the CPU stable sort of the
    oracle's keys after each.  The keys are those the last push left: a
    window shift removes the particles left of the new x_min and keeps the
    others' keys (a particle that the push found beyond x_max, where the
    cell clamps to nx, keeps that tile until the next push).  A twin run,
    identical up to the first sort, is downloaded there (a download reads
    the count back, which the run under test must not do before its second
    sort)."""
    log("device-decided sorts in sequence")
    nx, ny, dx, dy, tile = 24, 12, 1e-7, 1e-7, (4, 1)
    bcf = ("simple_outflow", "simple_outflow", "other", "conduct")
    bcp = ("open", "open", "other", "open")
    st, sim = make_pair(oracle, nx, ny, 1, dx, dy, bc_field=bcf, bc_particle=bcp, precision="fp64", tile=tile,
                        sort_interval=1000)
    _, twin = make_pair(oracle, nx, ny, 1, dx, dy, bc_field=bcf, bc_particle=bcp, precision="fp64", tile=tile,
                        sort_interval=1000)
    p = add_plasma(st, sim, nx, ny, dx, dy, ppc=6, temp_ev=2.0e5, seed=33, shuffle=True)
    sid = twin.set_species(-Q0, M0, False, capacity=len(p["x"]))
    twin.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])
    ntiles = 1 << 16                                 # beyond every tile key of the offset tiling
    order = sim.download_particles(0)["id"]          # the GPU's order after the upload sort
    assert np.array_equal(twin.download_particles(0)["id"], order)
    shifted = 0
    c0 = sim.sort_counts()
    for stage, (pushes, shift) in enumerate(((1, 2), (5, 0))):
        for _ in range(pushes):
            sim.push_particles()
            st.push_particles()
            if stage == 0:
                twin.push_particles()
        # the keys of the last push (the tiles move with the plasma: a window
        # cell is a tile offset of 1)
        sp = st.species[0]
        pushed = oracle.cell_keys(st, sp["x"], sp["y"], sp["z"], *tile, shift=shifted) // (tile[0] * tile[1])
        key_of = dict(zip(sp["id"].tolist(), pushed.tolist()))
        if shift:
            sim.move_window(shift)
            st.shift_window(shift)
            shifted += shift
            if stage == 0:
                twin.move_window(shift)
        sim.sort_particles()
        alive = set(st.species[0]["id"].tolist())
        key_of = {i: k for i, k in key_of.items() if i in alive}
        keys = np.array([key_of.get(int(i), ntiles) for i in order], np.int32)
        perm = oracle.stable_sort(keys, ntiles + 1)
        kept = order[perm][: int(np.sum(keys < ntiles))]
        log(f"  after {pushes} push(es), shift {shift}: {len(order)} -> {len(kept)} particles")
        assert len(kept) < len(order)
        if stage == 0:
            twin.sort_particles()
            c = twin.sort_counts()
            got = twin.download_particles(0)["id"]
            log(f"  twin after the first sort: counts {c}")
            report_order(got, kept, key_of)
            twin.close()
        order = kept
    c1 = sim.sort_counts()
    got = sim.download_particles(0)["id"]
    log(f"  counts {c0} -> {c1}")
    report_order(got, order, key_of)
    assert (c1["migrate"] - c0["migrate"], c1["radix"] - c0["radix"]) == (1, 1), (c0, c1)
    sim.close()


def report_order(got, want, key_of):
    """Asserts equal id orders; on a mismatch logs where and the keys there."""
    if len(got) == len(want) and np.array_equal(got, want):
        return
    n = min(len(got), len(want))
    bad = np.nonzero(got[:n] != want[:n])[0]
    k = int(bad[0]) if len(bad) else n
    lo, hi = max(0, k - 3), min(n, k + 6)
    log(f"  order differs: {len(got)} vs {len(want)} ids, {len(bad)} positions differ, first at {k}")
    log("  got  ", [(int(i), int(key_of.get(int(i), -1))) for i in got[lo:hi]])
    log("  want ", [(int(i), int(key_of.get(int(i), -1))) for i in want[lo:hi]])
    assert False, "sorted order differs from the stable sort of the oracle's keys"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotune_restores_state(oracle, precision):
    """The autotuner times pushes under each candidate tile shape and puts
    the particles and currents back: the steps after it equal the oracle's
    as without it (to the multi-step tolerances), and the choice is one of
    the candidates."""
    log("autotune", precision)
    nx, ny, dx, dy = 48, 24, 1e-7, 1e-7
    st, sim = make_pair(oracle, nx, ny, 2, dx, dy, precision=precision)
    add_plasma(st, sim, nx, ny, dx, dy, ppc=6, temp_ev=5000.0, seed=12, drift=(0.2, 0.0, 0.1))
    st.step()
    sim.step()
    tx, tr, gl, ms = sim.autotune(pushes=2)
    log(f"  chose {tx}x{tr} gather_lds={gl}, {ms:.4f} ms/push")
    assert (tx, tr) in {(16, 8), (16, 4), (8, 8), (8, 4), (32, 4)} and gl in (0, 1) and ms > 0
    for _ in range(2):
        st.step()
        sim.step()
    got = by_id(sim.download_particles(0))
    want = by_id({k: st.species[0][k][: st.species[0]["n"]] for k in ("x", "y", "z", "px", "py", "pz", "w", "id")})
    assert np.array_equal(got["id"], want["id"])
    tol = 1e-9 if precision == "fp64" else 1e-4  # the multi-step tolerances
    for k in ("x", "y", "z", "px", "py", "pz"):
        assert max_rel(got[k], want[k]) < tol, k
    jm = sim.download_field("jx")
    assert max_rel(jm[0], st.arr["jx"][0]) < (1e-10 if precision == "fp64" else 1e-3)
    sim.close()
