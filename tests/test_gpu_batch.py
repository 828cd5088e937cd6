"""Batches of runs stepped through captured HIP graphs, with the mixed
pushes of all runs of a mode count in one persistent launch
(cylpic_batch_*), against the same runs stepped one call at a time and
against the oracle.

* Three small runs of different shapes and mode counts, one with a laser,
  sorted every second step: two single batch steps and then 18 steps in
  kept-and-relaunched chunks equal twenty ordinary steps of identical copies
  of the runs to the bounds of the multi-step FP64 comparisons (m = 0 fields
  1e-10, m > 0 1e-6, particles 1e-9 of each array's largest value: the
  deposit's atomic sums add in a run-dependent order, and the steps carry
  that round-off on) and to 1e-4 for the mixed path (FP32 flush atomics).
  The chunks run eagerly until one ends in the state it started from; the
  next is captured into a graph, and the one after relaunches it.
* A seeded 32-run parameter scan at the scan's sizes (nx 512-2048, ny
  64-256, 1-3 modes, 8-32 particles per cell, densities 5e23-3e24 m^-3,
  a0 1-4 at x_min): the FP64 batch matches the oracle on the four smallest
  runs over two steps (the multi-step FP64 bounds), and after ten steps the
  mixed batch's mode energies match the FP64 batch's on every run (1e-3 of
  the electric energy, 1e-2 of the magnetic, as for the single-run mixed
  path)."""
import math

import numpy as np
import pytest

from cylindrical_epoch_amd import deck
from cylindrical_epoch_amd.host import Batch, Simulation
from gpu_helpers import log, max_rel

pytestmark = pytest.mark.gpu

FIELD_NAMES = ("ex", "er", "et", "bx", "br", "bt", "jx", "jr", "jt")


def small_runs():
    return [dict(nx=48, ny=20, M=1, ppc=6, laser=False, seed=1),
            dict(nx=40, ny=16, M=2, ppc=4, laser=True, seed=2),
            dict(nx=64, ny=24, M=3, ppc=5, laser=False, seed=3)]


def laser_sources(ny, dy, t, a0=1.0, lam=0.8e-6, waist=3e-6):
    om = 2 * math.pi * deck.C_LIGHT / lam
    amp = a0 * deck.M0 * deck.C_LIGHT * om / deck.Q0
    r = np.abs((np.arange(ny + 1) - 1) * dy + dy / 2)
    env = amp * math.exp(-((t - 20e-15) / 8e-15) ** 2) * math.sin(om * t)
    s1 = env * np.exp(-(r / waist) ** 2)
    return np.concatenate([s1, 0.0 * s1])


def make_run(w, precision, dx=1e-7, dy=1e-7, sort_interval=2, temp_ev=2000.0, density=1e25, parts=None):
    dt = deck.set_dt(dx, dy, 0.95)
    bf = ("simple_laser", "simple_outflow", "other", "conduct") if w.get("laser") else \
        ("periodic", "periodic", "other", "conduct")
    bp = ("open", "open", "other", "reflect") if w.get("laser") else ("periodic", "periodic", "other", "reflect")
    sim = Simulation(w["nx"], w["ny"], w["M"], dx, dy, dt, bc_field=bf, bc_particle=bp, precision=precision,
                     sort_interval=sort_interval)
    p = parts if parts is not None else deck.uniform_plasma(w["nx"], w["ny"], dx, dy, w["ppc"], density, temp_ev,
                                                            seed=w["seed"])
    sid = sim.set_species(-deck.Q0, deck.M0, False, capacity=len(p["x"]))
    sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])
    sim.setup_field_boundaries()
    return sim, p, (bf, bp, dt)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_single_steps(precision):
    log("batch vs single steps", precision)
    runs = small_runs()
    eager = [make_run(w, precision)[0] for w in runs]
    batched = [make_run(w, precision)[0] for w in runs]
    batch = Batch(batched)
    steps, single = 20, 2
    dt = eager[0].dt

    def srcs(n):
        return [laser_sources(w["ny"], 1e-7, (n + 1) * dt) if w["laser"] else None for w in runs]

    for n in range(steps):
        for sim, s in zip(eager, srcs(n)):
            sim.step(s)
    for n in range(single):
        batch.step(src_xmin=srcs(n))
    batch.run(steps - single, src_xmin=lambda j: srcs(single + j))
    batch.synchronize()
    st = batch.stats()
    log("  batch stats", st)
    assert st["steps"] == steps
    # 18 steps in chunks of 4: once a chunk ends where it started, the next
    # is captured and the one after relaunches it
    assert st["replays"] >= 1 and st["instantiations"] <= 2
    assert st["push_groups"] == (3 if precision == "mixed" else 0)
    ptol = 1e-9 if precision == "fp64" else 1e-4
    for w, a, b in zip(runs, eager, batched):
        for name in FIELD_NAMES:
            fa, fb = a.download_field(name), b.download_field(name)
            for m in range(w["M"]):
                ftol = (1e-10 if m == 0 else 1e-6) if precision == "fp64" else 1e-4
                err = max_rel(fb[m], fa[m])
                assert err < ftol, (w, name, m, err)
        pa, pb = a.download_particles(0), b.download_particles(0)
        oa, ob = np.argsort(pa["id"]), np.argsort(pb["id"])
        assert np.array_equal(pa["id"][oa], pb["id"][ob])
        for k in ("x", "y", "z", "px", "py", "pz"):
            assert max_rel(pb[k][ob], pa[k][oa]) < ptol, (w, k)
        assert abs(a.time - b.time) < 1e-30
    batch.close()
    for s in eager + batched:
        s.close()


def scan_draw(n=32, seed=20231):
    """The parameter scan of the batched benchmark: a seeded draw within the
    ranges nx 512-2048, ny 64-256, 1-3 modes, 8-32 particles per cell,
    n_e 5e23-3e24 m^-3, a0 1-4 (dx = 0.02 um, dy = 0.1 um)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        out.append(dict(nx=int(rng.integers(512, 2049)), ny=int(rng.integers(64, 257)), M=int(rng.integers(1, 4)),
                        ppc=int(rng.integers(8, 33)), density=float(rng.uniform(5e23, 3e24)),
                        a0=float(rng.uniform(1.0, 4.0)), laser=True, seed=100 + k))
    return out


def test_batch_parameter_scan_against_oracle_and_fp64(oracle):
    runs = scan_draw()
    dx, dy = 0.02e-6, 0.1e-6
    log("scan draw:", sum(w["nx"] * w["ny"] * w["ppc"] for w in runs), "particles in", len(runs), "runs")
    order = sorted(range(len(runs)), key=lambda k: runs[k]["nx"] * runs[k]["ny"] * runs[k]["ppc"])
    check = set(order[:4])
    f64, mix, orc = [], [], {}
    for k, w in enumerate(runs):
        p = deck.uniform_plasma(w["nx"], w["ny"], dx, dy, w["ppc"], w["density"], 100.0, seed=w["seed"])
        a, _, (bf, bp, dt) = make_run(w, "fp64", dx, dy, sort_interval=4, parts=p)
        b, _, _ = make_run(w, "mixed", dx, dy, sort_interval=4, parts=p)
        f64.append(a)
        mix.append(b)
        if k in check:
            st = oracle.State(w["nx"], w["ny"], w["M"], dx, dy, dt, bc_field=bf, bc_particle=bp)
            st.add_species(-deck.Q0, deck.M0, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"],
                           ids=p["id"])
            st.setup_field_boundaries()
            orc[k] = st
    log("  uploaded")
    bf64, bmix = Batch(f64), Batch(mix)

    def sources(t):
        return [laser_sources(w["ny"], dy, t, a0=w["a0"], waist=8e-6) for w in runs]

    dt = f64[0].dt
    for n in range(2):
        src = sources((n + 1) * dt)
        bf64.step(src_xmin=src)
        bmix.step(src_xmin=src)
        for k, st in orc.items():
            st.step(src_xmin=src[k])
        if n == 1:
            bf64.synchronize()
            for k, st in orc.items():
                for name in FIELD_NAMES:
                    g = f64[k].download_field(name)
                    for m in range(runs[k]["M"]):
                        tol = 1e-10 if m == 0 else 1e-6       # the FP64 multi-step bounds
                        err = max_rel(g[m], st.arr[name][m])
                        assert err < tol, (k, name, m, err)
                got = f64[k].download_particles(0)
                sp = st.species[0]
                og, ow = np.argsort(got["id"]), np.argsort(sp["id"][: sp["n"]])
                assert np.array_equal(got["id"][og], sp["id"][: sp["n"]][ow])
                for key in ("x", "y", "z", "px", "py", "pz"):
                    assert max_rel(got[key][og], sp[key][: sp["n"]][ow]) < 1e-9, (k, key)
            log("  FP64 batch equals the oracle on runs", sorted(check))
    bf64.run(8, src_xmin=lambda j: sources((j + 3) * dt))
    bmix.run(8, src_xmin=lambda j: sources((j + 3) * dt))
    bf64.synchronize()
    bmix.synchronize()
    log("  stats", bf64.stats(), bmix.stats())
    assert bmix.stats()["push_groups"] == len({w["M"] for w in runs})
    for k, w in enumerate(runs):
        we0, wb0 = f64[k].mode_energy()
        we1, wb1 = mix[k].mode_energy()
        for m in range(w["M"]):
            assert abs(we1[m] - we0[m]) <= 1e-3 * max(we0) , (k, m, we0[m], we1[m])
            assert abs(wb1[m] - wb0[m]) <= 1e-2 * max(wb0), (k, m, wb0[m], wb1[m])
    bf64.close()
    bmix.close()
    for s in f64 + mix:
        s.close()
