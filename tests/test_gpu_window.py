"""GPU moving window against the oracle: each step the same column of new
particles is injected beyond x_max and the window moves one cell
(src/housekeeping/window.F90:62-153, 157-350)."""
import numpy as np
import pytest

from cylindrical_epoch_amd import deck
from gpu_helpers import add_plasma, by_id, log, make_pair, max_rel, random_fields, load_fields

pytestmark = pytest.mark.gpu

BF = ("simple_outflow", "simple_outflow", "other", "open")
BP = ("open", "open", "other", "open")


def _column(st, nx, ny, dx, dy, ppc, seed, id0):
    """insert_particles' batch: ppc particles in each r cell of the column
    beyond x_max (window.F90:222-238), from a seeded generator."""
    p = deck.uniform_plasma(nx, ny, dx, dy, ppc, 1e25, 2.0e4, x_min=st.g.x_min, seed=seed, cells_x=[nx])
    p["id"] = p["id"] + id0
    return p


def _window_run(oracle, precision, M, steps, tile):
    nx, ny, dx, dy = 32, 12, 1e-7, 1e-7
    st, sim = make_pair(oracle, nx, ny, M, dx, dy, bc_field=BF, bc_particle=BP, precision=precision, tile=tile)
    rng = np.random.default_rng(7)
    load_fields(st, sim, random_fields(st.shape, rng, e_amp=1e9, b_amp=5.0, j_amp=0.0))
    add_plasma(st, sim, nx, ny, dx, dy, ppc=4, temp_ev=2.0e4, seed=5, capacity=4 * nx * ny * 4)
    st.setup_field_boundaries()
    sim.setup_field_boundaries()
    next_id = nx * ny * 4
    for k in range(steps):
        st.step()
        sim.step()
        p = _column(st, nx, ny, dx, dy, 4, 100 + k, next_id)
        next_id += len(p["x"])
        st.append_particles(0, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])
        sim.inject_particles(0, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])
        st.shift_window(1)
        sim.move_window(1)
    return st, sim


def test_window_fp64_matches_oracle(oracle):
    """FP64: after six steps with injection and a one-cell shift each, fields
    and particles track the oracle as in the plain multi-step test (fields
    m = 0 1e-10, m > 0 1e-6 of max; particles 1e-9), and the same particles
    were removed and kept."""
    log("window fp64")
    st, sim = _window_run(oracle, "fp64", 2, 6, (8, 4))
    assert sim.x_min == st.g.x_min and sim.x_max == st.g.x_max
    for n in ("ex", "er", "et", "bx", "br", "bt", "jx", "jr", "jt"):
        got = sim.download_field(n)
        for m in range(got.shape[0]):
            err = max_rel(got[m], st.arr[n][m])
            log(f"  {n} m={m} {err:.1e}")
            assert err < (1e-10 if m == 0 else 1e-6), (n, m, err)
    got = by_id(sim.download_particles(0))
    want = by_id({k: st.species[0][k][: st.species[0]["n"]] for k in ("x", "y", "z", "px", "py", "pz", "w", "id")})
    assert np.array_equal(got["id"], want["id"])
    for k in ("x", "y", "z", "px", "py", "pz"):
        assert max_rel(got[k], want[k]) < 1e-9, k
    sim.close()


def test_window_mixed_energies(oracle):
    """Mixed precision: the same run's mode energies and particle count
    track the oracle (FP32 error budget of the multi-step mixed test)."""
    log("window mixed")
    st, sim = _window_run(oracle, "mixed", 2, 6, (16, 4))
    we0, wb0 = st.mode_energy()
    we1, wb1 = sim.mode_energy()
    for m in range(2):
        log(f"  m={m} W_E {we0[m]:.4e} vs {we1[m]:.4e}, W_B {wb0[m]:.4e} vs {wb1[m]:.4e}")
        assert abs(we1[m] - we0[m]) <= 1e-3 * we0[m]
        assert abs(wb1[m] - wb0[m]) <= 1e-2 * max(wb0[m], 1e-30) + 1e-30
    assert abs(sim.particle_count(0) - st.species[0]["n"]) <= 2
    sim.close()


def _inject_push(oracle, precision, method, nbatch=1):
    nx, ny, dx, dy = 32, 12, 1e-7, 1e-7
    tile = (8, 4)
    st, sim = make_pair(oracle, nx, ny, 1, dx, dy, bc_field=BF, bc_particle=BP, precision=precision, tile=tile,
                        sort_interval=1000)
    sim.set_sort_method(method)
    add_plasma(st, sim, nx, ny, dx, dy, ppc=4, temp_ev=2.0e4, seed=9, capacity=4 * nx * ny * 4)
    before = sim.download_particles(0)["id"]
    ntiles = (nx // tile[0]) * (ny // tile[1])
    cells = tile[0] * tile[1]
    ids, keys = [], []
    for b in range(nbatch):  # batches since the last sort: one region, stable-sorted by key at each injection
        p = _column(st, nx, ny, dx, dy, 4, 11 + b, 10 ** 6 * (b + 1))
        p["x"] = p["x"] + b * dx
        sim.inject_particles(0, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])
        ids.append(p["id"])
        keys.append(oracle.cell_keys(st, p["x"], p["y"], p["z"], *tile) // cells)
    ids, keys = np.concatenate(ids), np.concatenate(keys)
    batch = ids[oracle.stable_sort(keys.astype(np.int32), ntiles)]
    sim.move_window(nbatch)
    st.shift_window(nbatch)  # the oracle's grid follows (its own particles are not used)
    sim.push_particles()
    c0 = sim.sort_counts()
    got = sim.download_particles(0)
    c1 = sim.sort_counts()
    sim.close()
    return st, got, np.concatenate([before, batch]), ntiles, cells, tile, c0, c1


@pytest.mark.parametrize("nbatch", [1, 2, 3])
def test_injected_batch_sorts_as_stable_cpu_sort(oracle, nbatch):
    """FP64: injected particles join the tile ranges without a full re-sort;
    after the next push the order equals the stable CPU sort by tile key of
    the previous order followed by the batches (together in stable tile-key
    order), removed particles dropped (bit-exact)."""
    log("window sort fp64", nbatch)
    st, got, order, ntiles, cells, tile, c0, c1 = _inject_push(oracle, "fp64", "migrate", nbatch)
    assert c1["migrate"] == c0["migrate"] + 1
    pos = {i: k for k, i in enumerate(got["id"])}
    # the tiles moved with the plasma: one window cell = tile offset 1
    keys_alive = oracle.cell_keys(st, got["x"], got["y"], got["z"], *tile, shift=nbatch) // cells
    removed = 1 << 16  # beyond every tile key of the offset tiling
    keys = np.array([keys_alive[pos[i]] if i in pos else removed for i in order], np.int32)
    want = order[oracle.stable_sort(keys, removed + 1)][: len(got["id"])]
    log(f"  {len(order)} -> {len(want)} particles")
    assert np.array_equal(got["id"], want)


@pytest.mark.parametrize("nbatch", [1, 2])
def test_injected_batch_mixed_migrate_equals_radix(oracle, nbatch):
    """Mixed precision: the migration merge of injected batches gives the
    same order and state as the full radix sort of the same run."""
    log("window sort mixed", nbatch)
    _, got, *_ , c0, c1 = _inject_push(oracle, "mixed", "migrate", nbatch)
    assert c1["migrate"] == c0["migrate"] + 1
    _, ref, *_ = _inject_push(oracle, "mixed", "radix", nbatch)
    for k in ("id", "x", "y", "z", "px", "py", "pz", "w"):
        assert np.array_equal(got[k], ref[k]), k
