"""Throughput of the quasi-3D PIC timestep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|mixed|both]

A step is one full timestep of the reference's main loop (update_eb_fields_half,
push_particles, current_finish, update_eb_fields_final, moving_window;
src/epoch2d.F90:213-265) over the whole particle load, with every input
resident in HBM.  Workload: the 8192 x 512 cell, 3-mode, 32 particles-per-cell
single run (134,217,728 electrons, n_e = 5e23 m^-3 plateau, dx = 0.02 um,
dy = 0.1 um) with a Gaussian laser (lambda 0.8 um, a0 = 3, waist 15 um, 25 fs)
entering through x_min, open particle boundaries and a window moving at c
(a column of 16,384 new particles injected per cell moved).  Synthetic data:
particles from a seeded generator with a 100 eV thermal spread.

Prints one JSON line.  The CPU baseline is the FP64 oracle timed on the host's
cores (OMP_NUM_THREADS, else all): whole steps of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from cylindrical_epoch_amd import deck  # noqa: E402
from cylindrical_epoch_amd.host import Simulation  # noqa: E402

METRIC = "particle-steps/s (push+deposit+FDTD), mixed vs FP64, 1 MI355X; % HBM peak"
HBM_PEAK_GBS = 8000.0
WORKLOAD = "wakefield_8192x512_m3_ppc32"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Laser:
    """Host-side laser source of one boundary (laser.f90:444-461 with a
    Gaussian transverse profile and Gaussian temporal envelope)."""

    def __init__(self, ny, dy, a0=3.0, lam=0.8e-6, waist=15e-6, fwhm=25e-15, pol=0.0):
        self.omega = 2 * math.pi * deck.C_LIGHT / lam
        self.amp = a0 * deck.M0 * deck.C_LIGHT * self.omega / deck.Q0
        r = np.abs((np.arange(ny + 1) - 1) * dy + dy / 2)
        self.profile = np.exp(-(r / waist) ** 2)
        self.tw = fwhm / (2 * math.sqrt(math.log(2)))
        self.t0 = 2 * fwhm
        self.pol = pol
        self.ny = ny

    def sources(self, t):
        env = math.exp(-((t - self.t0) / self.tw) ** 2) * self.amp
        base = env * self.profile * math.sin(self.omega * t)
        return np.concatenate([base * math.cos(self.pol), base * math.sin(self.pol)])


class Window:
    """The reference's moving_window (window.F90:353-376) at window_v_x = c
    from t = 0: the shift fraction accumulates dt v / dx, and each whole cell
    injects a column of new particles beyond x_max (insert_particles,
    window.F90:157-325) and shifts the window.  The columns are drawn ahead
    of time from the same generator as the initial load (synthetic data)."""

    def __init__(self, w, dt, max_cells, seed=11):
        self.w, self.dt, self.fraction = w, dt, 0.0
        self.v = deck.C_LIGHT
        self.batches = []
        for k in range(max_cells):
            p = deck.uniform_plasma(1, w["ny"], w["dx"], w["dy"], w["ppc"], w["density"], 100.0,
                                    seed=seed + k, cells_x=[0])
            self.batches.append({key: p[key] for key in ("x", "y", "z", "px", "py", "pz", "w")})
        self.used = 0
        self.injected = 0

    @staticmethod
    def cells_for(w, dt, steps):
        return int(math.ceil(steps * dt * deck.C_LIGHT / w["dx"])) + 2

    def advance(self, sim, sid):
        self.fraction += self.dt * self.v / self.w["dx"]
        k = int(math.floor(self.fraction))
        if k <= 0:
            return
        for c in range(k):
            p = self.batches[self.used]
            self.used += 1
            sim.inject_particles(sid, p["x"] + (sim.x_max + c * self.w["dx"]), p["y"], p["z"], p["px"],
                                 p["py"], p["pz"], p["w"])
            self.injected += len(p["x"])
        sim.move_window(k)
        self.fraction -= k


def make_sim(w, precision, particles, tune=None, extra_capacity=0):
    dx, dy = w["dx"], w["dy"]
    laser = Laser(w["ny"], dy)
    dt = deck.set_dt(dx, dy, 0.95, plasma=[(w["density"], 100.0 * deck.EV / deck.KB, deck.M0)],
                     laser_omegas=[laser.omega])
    tune = tune or {}
    mixed = precision == "mixed"
    sim = Simulation(w["nx"], w["ny"], w["n_mode"], dx, dy, dt,
                     bc_field=("simple_laser", "simple_outflow", "other", "open"),
                     bc_particle=("open", "open", "other", "open"), precision=precision,
                     tile=tune.get("tile_" + precision),
                     acc_mode=tune.get("acc_mode") if mixed else None,
                     gather_lds=tune.get("gather_lds") if mixed else None)
    if tune.get("sort_method"):
        sim.set_sort_method(tune["sort_method"])
    if tune.get("sort_interval"):
        sim.set_sort_interval(tune["sort_interval"])
    n = len(particles["x"])
    sid = sim.set_species(-deck.Q0, deck.M0, False, capacity=n + extra_capacity)
    sim.upload_particles(sid, particles["x"], particles["y"], particles["z"], particles["px"],
                         particles["py"], particles["pz"], particles["w"])
    sim.setup_field_boundaries()
    return sim, laser, dt


def run(sim, laser, dt, steps, warmup, window=None):
    t = 0.0
    for _ in range(warmup):
        t += dt
        sim.step(laser.sources(t))
        if window:
            window.advance(sim, 0)
    sim.synchronize()
    sim.profile(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        t += dt
        sim.step(laser.sources(t))
        if window:
            window.advance(sim, 0)
    sim.synchronize()
    t1 = time.perf_counter()
    prof = {k: sim.kernel_time(k) for k in ("push_deposit", "update_e", "update_b", "field_bcs",
                                            "zero_j", "current_axis", "particle_bcs",
                                            "current_finish", "sort", "window", "inject")}
    sim.profile(False)
    return (t1 - t0) / steps, prof


def accuracy_step(sim, laser, t, window=None):
    """One more step, outside the timed region, in the main loop's order
    with the m = 0 continuity check around the push; returns the residual
    relative to the largest per-cell charge change and to the largest
    per-cell charge, and the on-axis Ex (Re exm(1:nx, 0, 0)) after it."""
    sim.update_eb_fields_half()
    sim.continuity_begin()
    sim.push_particles()
    res, dq, q = sim.continuity_end()
    cm = sim.continuity_modes()
    sim.current_finish()
    sim.update_eb_fields_final(laser.sources(t))
    if window:
        window.advance(sim, 0)
    ng = 5
    ex = sim.download_field("ex")
    axis = ex[0, ng - 1, ng:ng + sim.nx].real.copy()  # iy = 0 is r = 0
    rel = lambda a, b: [float(x / y) if y > 0 else None for x, y in zip(a, b)]
    return {"residual_over_max_dq": res / dq if dq > 0 else None,
            "residual_over_max_q": res / q if q > 0 else None,
            # every mode: the x-r identity's residual, and the part of the
            # charge change the deposited J_theta misses, both over max |dQ_m|
            "modes_xr_identity_residual_over_max_dq": rel(cm["residual"], cm["dq"]),
            "modes_jtheta_part_over_max_dq": rel(cm["jtheta"], cm["dq"])}, axis


def wake_accuracy(steps, every=100, sort_interval=8):
    """The wakefield deck at its definition, mixed against FP64 on the GPU
    from the same particles: 2048 x 256 cells (41 um x 25.6 um), 2 modes, 16
    cold electrons per cell, n_e = 1e24 m^-3 reached over a 50 um linear
    up-ramp from x = 0, the x_min laser (a0 = 2, 0.8 um, 10 um waist, 30 fs
    FWHM, peak entering at 40 fs), open particle boundaries, and a window
    that starts at 80 fs (the pulse's peak 12 um into the box) and then
    moves at c, injecting the ramp's columns.  After `steps` steps (3000:
    121 fs, the pulse in the plasma with its wake behind it) compares the
    on-axis Ez (Re exm(1:nx, r = 0, m = 0)), the per-mode field energies
    sampled every `every` steps, and the electrons' kinetic-energy spectrum
    (cylpic_dist_fn over "en", 64 bins on the FP64 run's range)."""
    from cylindrical_epoch_amd.window import MovingWindow, apply_to_sim
    w = deck.WORKLOADS["wakefield_2048x256_m2_ppc16"]
    nx, ny, M, dx, dy = w["nx"], w["ny"], w["n_mode"], w["dx"], w["dy"]
    profile = deck.ramp_profile(0.0, 50e-6)
    parts = deck.profiled_plasma(nx, ny, dx, dy, w["ppc"], w["density"], profile, seed=20240)
    laser = Laser(ny, dy, a0=2.0, waist=10e-6, fwhm=30e-15)
    laser.t0 = 40e-15
    dt = deck.set_dt(dx, dy, 0.95)
    cols = int(math.ceil(steps * dt * deck.C_LIGHT / dx)) + 2
    res = {}
    # "fp64_reordered": FP64 with another tile shape, so that only the order
    # of the deposit's atomic sums differs from "fp64": how far two FP64 runs
    # drift apart by round-off alone (the trajectories are chaotic)
    for run in ("fp64", "mixed", "fp64_reordered"):
        prec = "mixed" if run == "mixed" else "fp64"
        sim = Simulation(nx, ny, M, dx, dy, dt, bc_field=("simple_laser", "simple_outflow", "other", "open"),
                         bc_particle=("open", "open", "other", "open"), precision=prec, sort_interval=sort_interval,
                         tile=(8, 8) if run == "fp64_reordered" else None)
        sid = sim.set_species(-deck.Q0, deck.M0, False, capacity=len(parts["x"]) + cols * ny * w["ppc"])
        sim.upload_particles(sid, parts["x"], parts["y"], parts["z"], parts["px"], parts["py"], parts["pz"],
                             parts["w"], ids=parts["id"])
        sim.setup_field_boundaries()
        win = MovingWindow(ny, dx, dy, dt, w["ppc"], w["density"], profile, start=80e-15)
        hist, axes = [], []
        t0 = time.perf_counter()
        for k in range(steps):
            sim.step(laser.sources((k + 1) * dt))
            apply_to_sim(sim, sid, *win.advance(sim.x_max, (k + 1) * dt))
            if (k + 1) % every == 0:
                hist.append(np.concatenate(sim.mode_energy()))
            if (k + 1) % (5 * every) == 0 or k + 1 == steps:
                axes.append(sim.download_field("ex")[0, 4, 5:5 + nx].real.copy())
        spec = sim.dist_fn(sid, ["en"], ranges=res["fp64"]["range"] if run != "fp64" else [(0.0, 0.0)],
                           resolution=[64])
        res[run] = dict(axes=axes, hist=np.array(hist), spec=spec[0].ravel(), range=spec[1].tolist(),
                        n=sim.particle_count(sid), s=time.perf_counter() - t0)
        log(f"wake accuracy {run}: {steps} steps in {res[run]['s']:.1f} s, {res[run]['n']} electrons, "
            f"window moved {win.columns} cells")
        sim.close()
    a = res["fp64"]
    floor = 1e-6 * a["hist"].max()
    big = a["hist"] > floor

    def against(b):
        ez = [float(np.max(np.abs(y - x)) / np.max(np.abs(x))) for x, y in zip(a["axes"], b["axes"])]
        return {"on_axis_ez_rel_err": ez[-1], "on_axis_ez_rel_err_every_500_steps": ez,
                "field_energy_history_max_rel_err": float(np.max(np.abs(b["hist"] - a["hist"])[big] / a["hist"][big])),
                "electron_spectrum_l1_rel_diff": float(np.sum(np.abs(b["spec"] - a["spec"])) / np.sum(a["spec"])),
                "electrons": b["n"]}
    return {"workload": "wakefield_2048x256_m2_ppc16, 50 um up-ramp from x = 0, a0 = 2 peaking at 40 fs, "
                        "window at c from 80 fs",
            "steps": steps, "max_on_axis_ez_v_per_m": float(np.max(np.abs(a["axes"][-1]))), "electrons": a["n"],
            "field_energy_samples": int(len(a["hist"])),
            "mixed_vs_fp64": against(res["mixed"]), "fp64_reordered_vs_fp64": against(res["fp64_reordered"])}


def _profile_summary(precision):
    path = os.path.join(ROOT, "profiles", f"push_deposit_{precision}_{WORKLOAD}_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def measured_traffic(precision):
    """HBM bytes per push_deposit launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    or None when no summary for this kernel and workload exists."""
    return _profile_summary(precision).get("hbm_bytes_per_launch")


# ds_add_u64 lane-operations per second over the 256 CUs, random addresses in
# a 64 KB box (the per-tile current box's atomic), as measured on gfx950
LDS_ATOMIC_PEAK = 3208e9
# VALU lane-operations per second: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz,
# one FP32 or FP64 operation per lane and cycle (not counting packed FP32)
VALU_PEAK = 256 * 4 * 16 * 2.4e9


def roofline(w, n, prof_push, precision):
    """Algorithmic bytes of one push_deposit launch / its average duration."""
    ms, cnt = prof_push
    if cnt == 0:
        return None
    cells = w["nx"] * w["ny"]
    if precision == "fp64":
        per_particle = 104.0                       # read 7 doubles, write 6
        per_cell_mode = 16.0 * (6 + 6)             # gather 6 comps, J 3 comps read+write
    else:
        per_particle = 60.0
        per_cell_mode = 8.0 * (6 + 6)
    bytes_ = per_particle * n + per_cell_mode * w["n_mode"] * cells
    avg_s = ms / cnt / 1e3
    achieved = bytes_ / avg_s / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(precision),
           "kernel": "push_deposit", "avg_launch_ms": round(ms / cnt, 4),
           "algorithmic_bytes_per_launch": bytes_}
    # the deposit's LDS atomics (the kernel's measured bound): lane-operations
    # per launch from the committed SQ_INSTS_LDS_ATOMIC reading (x 64 lanes)
    # over this run's launch time, against the ds_add_u64 rate
    ops = _profile_summary(precision).get("lds_atomic_lane_ops_per_launch")
    if ops:
        rate = ops / avg_s
        out["lds_atomic"] = {"achieved": round(rate / 1e9, 1), "peak": LDS_ATOMIC_PEAK / 1e9, "unit": "Gops/s",
                             "frac": round(rate / LDS_ATOMIC_PEAK, 4), "ops_per_launch": ops,
                             "ops_per_particle": round(ops / n, 1)}
    # the VALU issue rate: the committed SQ_INSTS_VALU reading (wave
    # instructions, x 64 lanes) over this run's launch time
    valu = _profile_summary(precision).get("valu_wave_instructions_per_launch")
    if valu:
        rate = valu * 64 / avg_s
        out["valu"] = {"achieved": round(rate / 1e12, 2), "peak": round(VALU_PEAK / 1e12, 2),
                       "unit": "T lane-ops/s", "frac": round(rate / VALU_PEAK, 4),
                       "instructions_per_particle": round(valu * 64 / n, 1)}
    return out


def cpu_baseline(w, sample_cols, cores=None):
    """The FP64 oracle on the host's cores, timing whole steps: field
    updates (rows in parallel), the push with the particles cut into one x
    slab per core (private currents added in slab order) and the current
    boundaries, over the particles of the first `sample_cols` x-columns (all
    of them by default); the second of two steps is timed, and a partial
    load is scaled to the whole one by its particle count."""
    try:
        from oracle import oracle as O
    except Exception as e:  # pragma: no cover
        return {"error": f"oracle unavailable: {e}"}
    cores = cores or int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    dx, dy = w["dx"], w["dy"]
    dt = deck.set_dt(dx, dy, 0.95)
    st = O.State(w["nx"], w["ny"], w["n_mode"], dx, dy, dt,
                 bc_field=("simple_laser", "simple_outflow", "other", "open"),
                 bc_particle=("open", "open", "other", "open"))
    cols = min(sample_cols, w["nx"])
    p = deck.uniform_plasma(w["nx"], w["ny"], dx, dy, w["ppc"], w["density"], 100.0,
                            seed=1, cells_x=np.arange(cols))
    ns = len(p["x"])
    st.add_species(-deck.Q0, deck.M0, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    del p
    st.setup_field_boundaries()
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        st.update_eb_fields_half()
        t1 = time.perf_counter()
        st.push_particles_threads(cores)
        t2 = time.perf_counter()
        st.current_finish()
        st.update_eb_fields_final()
        t3 = time.perf_counter()
        times.append(((t1 - t0) + (t3 - t2), t2 - t1))
    t_fields, t_push = times[-1]
    n_full = w["nx"] * w["ny"] * w["ppc"]
    value = n_full / (t_fields + t_push * n_full / ns)
    what = "all" if cols == w["nx"] else f"the {ns} in x-columns 1..{cols} (push time scaled to {n_full})"
    return {"value": round(value, 1), "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": (f"the second of two whole steps of the FP64 oracle on {cores} host threads, {what} "
                       f"particles of the {w['nx']}x{w['ny']}x{w['n_mode']} workload: field updates and current "
                       f"boundaries {t_fields:.2f} s, push {t_push:.2f} s"),
            "push_particle_steps_per_s": round(ns / t_push, 1)}


def scan_runs(n=32, seed=20231):
    """The parameter scan: uneven wakefield runs drawn from a seeded
    generator, nx in [512, 2048], ny in [64, 256], 1-3 modes, 8-32 particles
    per cell, n_e in [5e23, 3e24] m^-3, a0 in [1, 4] (dx = 0.02 um, dy =
    0.1 um)."""
    rng = np.random.default_rng(seed)
    return [dict(nx=int(rng.integers(512, 2049)), ny=int(rng.integers(64, 257)), n_mode=int(rng.integers(1, 4)),
                 ppc=int(rng.integers(8, 33)), density=float(rng.uniform(5e23, 3e24)),
                 a0=float(rng.uniform(1.0, 4.0)), dx=0.02e-6, dy=0.1e-6) for _ in range(n)]


def scan(precision, steps, warmup, sort_interval=4, legs=("batch", "interleaved")):
    """Steps every run of the scan: (a) all runs in one batch, stepped in
    chunks of captured HIP graphs that are kept and relaunched, with the
    pushes of each mode count in one persistent launch (the headline); (b)
    each run in its own library handle and stream, interleaved step by step
    from the host.  Each run has its laser at x_min (waist 8 um, 25 fs, its
    own a0), an outflow x_max, open particle boundaries in x, a conducting
    r_max, a 100 eV plasma; the window stays fixed.  `legs` picks which of
    the two are run (each on its own copies of the runs)."""
    from cylindrical_epoch_amd.host import Batch
    runs = scan_runs()
    sims = {leg: [] for leg in legs}
    lasers, dts, total = [], [], 0
    for k, w in enumerate(runs):
        lasers.append(Laser(w["ny"], w["dy"], a0=w["a0"], waist=8e-6, fwhm=25e-15))
        dts.append(deck.set_dt(w["dx"], w["dy"], 0.95))
        p = deck.uniform_plasma(w["nx"], w["ny"], w["dx"], w["dy"], w["ppc"], w["density"], 100.0, seed=100 + k)
        for leg in legs:
            sim = Simulation(w["nx"], w["ny"], w["n_mode"], w["dx"], w["dy"], dts[-1],
                             bc_field=("simple_laser", "simple_outflow", "other", "conduct"),
                             bc_particle=("open", "open", "other", "reflect"), precision=precision,
                             sort_interval=sort_interval)
            sid = sim.set_species(-deck.Q0, deck.M0, False, capacity=len(p["x"]))
            sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
            sim.setup_field_boundaries()
            sims[leg].append(sim)
        total += len(p["x"])

    def sources(n):
        return [laser.sources(n * dt) for laser, dt in zip(lasers, dts)]

    # three sort periods: batch chunks run eagerly until one ends where it
    # started, the next is captured, and from then on the graph is relaunched
    warmup = max(warmup, 3 * 2 * sort_interval)
    # the sources of the timed steps are tabulated before the clock starts,
    # as a caller that knows its laser ahead would do
    table = [sources(warmup + j + 1) for j in range(steps)]
    out = dict(runs=len(runs), particles=total, cells=sum(w["nx"] * w["ny"] * w["n_mode"] for w in runs),
               warmup=warmup)
    if "batch" in legs:
        batch = Batch(sims["batch"])
        batch.run(warmup, src_xmin=lambda j: sources(j + 1))
        batch.synchronize()
        t0 = time.perf_counter()
        batch.run(steps, src_xmin=lambda j: table[j])
        batch.synchronize()
        t = time.perf_counter() - t0
        out.update(batch_s=t, value_batch=total * steps / t, stats=batch.stats())
        batch.close()
    if "interleaved" in legs:
        eager = sims["interleaved"]
        for n in range(warmup):
            for sim, src in zip(eager, sources(n + 1)):
                sim.step(src)
        for sim in eager:
            sim.synchronize()
        t0 = time.perf_counter()
        for j in range(steps):
            for sim, src in zip(eager, table[j]):
                sim.step(src)
        for sim in eager:
            sim.synchronize()
        t = time.perf_counter() - t0
        out.update(int_s=t, value_int=total * steps / t)
    for leg in legs:
        for sim in sims[leg]:
            sim.close()
    return out


def graph_single(w, steps, warmup):
    """The small workload stepped on its own (cylpic_step per step) and as a
    batch of one (kept graphs of a sort period's steps), per-step wall time."""
    from cylindrical_epoch_amd.host import Batch
    out = {}
    for mode in ("eager", "graph"):
        p = deck.uniform_plasma(w["nx"], w["ny"], w["dx"], w["dy"], w["ppc"], w["density"], w.get("temp_ev", 0.0),
                                seed=7)
        dt = deck.set_dt(w["dx"], w["dy"], 0.95)
        sim = Simulation(w["nx"], w["ny"], w["n_mode"], w["dx"], w["dy"], dt,
                         bc_field=("periodic", "periodic", "other", "conduct"),
                         bc_particle=("periodic", "periodic", "other", "reflect"), precision="mixed",
                         sort_interval=8)
        sid = sim.set_species(-deck.Q0, deck.M0, False, capacity=len(p["x"]))
        sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
        sim.setup_field_boundaries()
        stepper = Batch([sim]) if mode == "graph" else None

        def advance(n):
            if stepper:
                stepper.run(n)
            else:
                for _ in range(n):
                    sim.step()

        advance(max(warmup, 3 * 2 * 8))     # three sort periods: the kept graph exists
        (stepper or sim).synchronize()
        t0 = time.perf_counter()
        advance(steps)
        (stepper or sim).synchronize()
        out[mode] = (time.perf_counter() - t0) / steps * 1e3
        if stepper:
            stepper.close()
        sim.close()
    return out


def copy_calibration():
    """A plain device copy of 2^28 float32 (1 GiB read + 1 GiB written per
    launch), three launches."""
    import torch
    x = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    x.fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": 1 << 30, "copy_bytes_written_per_launch": 1 << 30,
                      "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64,
                    help="timed steps (default 64: two sort periods, long enough for the particle order's "
                         "drift between sorts to show in the push time)")
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--workload", default=WORKLOAD, choices=sorted(deck.WORKLOADS))
    ap.add_argument("--cpu-sample-cols", type=int, default=1 << 30,
                    help="x-columns of particles the CPU baseline loads (default: all of them; a part is scaled to "
                         "the whole load)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tile-mixed", default=None, help="tile of the mixed path, e.g. 32x4 (library default)")
    ap.add_argument("--tile-fp64", default=None, help="tile of the FP64 path, e.g. 16x8 (library default 16x4)")
    ap.add_argument("--acc-mode", type=int, default=None, help="mixed current accumulator 0..4")
    ap.add_argument("--gather-lds", type=int, default=None, help="mixed gather from LDS (1) or L1/L2 (0)")
    ap.add_argument("--sort-method", default="cells", choices=["radix", "migrate", "cells"],
                    help="library sort method (default cells: every sort re-keys by cell and lays each tile out "
                         "in rounds, so the push's wavefronts keep spanning distinct cells; the migration sort "
                         "keeps the upload's order while it decays: with it every 8 steps the mixed push "
                         "averaged 11.5 ms over 64 steps against 9.6 ms)")
    ap.add_argument("--no-window", action="store_true", help="keep the window fixed (no injection)")
    ap.add_argument("--autotune", action="store_true",
                    help="let the library's autotuner pick the tile shape and gather source before timing")
    ap.add_argument("--sort-interval", type=int, default=32,
                    help="sort every n pushes (library default 1; 32 measured fastest with the cells method on "
                         "this workload, 12.0 ms per mixed step against 12.4 at 16 and 13.5 at 8: a radix sort "
                         "has no neighbour limit, and the particles that drift out of their tile's current box "
                         "between sorts deposit through the push's global path)")
    ap.add_argument("--scan", action="store_true",
                    help="instead of the single run: the 32-run wakefield parameter scan (nx 512-2048, ny 64-256, "
                         "1-3 modes, 8-32 ppc, n_e 5e23-3e24, a0 1-4) as one graph batch, and as interleaved handles")
    ap.add_argument("--scan-leg", choices=("both", "batch", "interleaved"), default="both",
                    help="with --scan: run only one of its two ways of stepping (for profiling)")
    ap.add_argument("--graph-single", action="store_true",
                    help="only time one run of --workload stepped eagerly and as a one-run graph batch")
    ap.add_argument("--wake-steps", type=int, default=3000,
                    help="steps of the wakefield accuracy study (mixed vs FP64 with the pulse in the plasma) "
                         "reported in the accuracy block; 0 skips it")
    ap.add_argument("--wake-only", action="store_true", help="only run the wakefield accuracy study")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy three times (the reference kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.scan:
        prec = "mixed" if a.precision == "both" else a.precision
        legs = ("batch", "interleaved") if a.scan_leg == "both" else (a.scan_leg,)
        r = scan(prec, a.steps, a.warmup, legs=legs)
        if a.scan_leg != "both":
            log(f"scan {prec}, {a.scan_leg} only: {r}")
            return
        log(f"scan {prec}: {r['runs']} runs, {r['particles']} particles; batch (graph) {r['batch_s']:.3f} s, "
            f"interleaved handles {r['int_s']:.3f} s; {r['stats']}")
        print(json.dumps({
            "metric": "particle-steps/s, 32-run parameter scan (aggregate), 1 MI355X",
            "value": round(r["value_batch"], 1), "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
            "warmup": r["warmup"], "ms_per_step": round(r["batch_s"] / a.steps * 1e3, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == "fp64" else "f32+f64",
            "data": "synthetic",
            "config": {"workload": "scan_32_wakefield_runs_nx512-2048_ny64-256_m1-3_ppc8-32", "runs": r["runs"],
                       "particles": r["particles"], "cell_modes": r["cells"], "precision": prec,
                       "parallelism": "single-gpu, one graph per step, persistent push per mode count"},
            "batch_stats": r["stats"],
            "interleaved_handles_particle_steps_per_s": round(r["value_int"], 1)}), flush=True)
        return
    if a.wake_only:
        print(json.dumps({"metric": "wakefield accuracy, mixed vs FP64", **wake_accuracy(a.wake_steps)}), flush=True)
        return
    if a.graph_single:
        w = deck.WORKLOADS[a.workload]
        r = graph_single(w, a.steps, a.warmup)
        log(f"{a.workload}: eager {r['eager']:.4f} ms/step, graph {r['graph']:.4f} ms/step")
        n = w["nx"] * w["ny"] * w["ppc"]
        print(json.dumps({"metric": "ms/step, one small run, eager calls vs a captured graph", "workload": a.workload,
                          "particles": n, "eager_ms_per_step": round(r["eager"], 4),
                          "graph_ms_per_step": round(r["graph"], 4)}), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the library has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)
    w = deck.WORKLOADS[a.workload]
    log("generating particles", w)
    t0 = time.perf_counter()
    parts = deck.uniform_plasma(w["nx"], w["ny"], w["dx"], w["dy"], w["ppc"], w["density"], 100.0, seed=7)
    n = len(parts["x"])
    log(f"{n} particles in {time.perf_counter() - t0:.1f} s")
    precs = ["fp64", "mixed"] if a.precision == "both" else [a.precision]
    tune = {"acc_mode": a.acc_mode, "gather_lds": a.gather_lds, "sort_method": a.sort_method,
            "sort_interval": a.sort_interval}
    for k in ("mixed", "fp64"):
        t = getattr(a, "tile_" + k)
        tune["tile_" + k] = tuple(int(v) for v in t.split("x")) if t else None
    results = {}
    for prec in precs:
        win = None
        extra = 0
        if not a.no_window:
            dt0 = deck.set_dt(w["dx"], w["dy"], 0.95)
            cells = Window.cells_for(w, dt0, a.steps + a.warmup)
            extra = cells * w["ny"] * w["ppc"]
        sim, laser, dt = make_sim(w, prec, parts, tune, extra_capacity=extra)
        if not a.no_window:
            win = Window(w, dt, Window.cells_for(w, dt, a.steps + a.warmup))
        log(f"{prec}: uploaded, dt = {dt:.4e}")
        tuned = None
        if a.autotune:
            tuned = sim.autotune(pushes=2)
            log(f"{prec}: autotune chose tile {tuned[0]}x{tuned[1]}, gather_lds={tuned[2]} ({tuned[3]:.3f} ms/push)")
        ms, prof = run(sim, laser, dt, a.steps, a.warmup, win)
        sorts = sim.sort_counts()
        cont, axis = accuracy_step(sim, laser, (a.warmup + a.steps + 1) * dt, win)
        if win:
            log(f"{prec}: window moved {win.used} cells, {win.injected} particles injected")
        log(f"{prec}: m = 0 continuity residual {cont}")
        sim.close()
        results[prec] = dict(ms=ms * 1e3, value=n / ms, prof=prof, sorts=sorts, cont=cont, axis=axis, tuned=tuned)
        log(f"{prec}: {ms * 1e3:.3f} ms/step, {n / ms:.4e} particle-steps/s")
        for k, (kms, cnt) in prof.items():
            if cnt:
                log(f"   {k:15s} {kms / a.steps:8.3f} ms/step ({cnt} launches)")
    del parts
    main_prec = "mixed" if "mixed" in results else "fp64"
    r = results[main_prec]
    cells = w["nx"] * w["ny"] * w["n_mode"]
    out = {
        "metric": METRIC, "value": round(r["value"], 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(r["ms"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if main_prec == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": a.workload, "nx": w["nx"], "ny": w["ny"], "n_mode": w["n_mode"],
                   "ppc": w["ppc"], "particles": n, "precision": main_prec, "parallelism": "single-gpu",
                   "laser": "x_min, a0=3, lambda 0.8 um, waist 15 um, 25 fs",
                   "moving_window": not a.no_window, "window_v_x": "c" if not a.no_window else None,
                   "sort_method": a.sort_method, "sort_interval": a.sort_interval},
        "field_cell_updates_per_s": round(cells / (r["ms"] / 1e3), 1),
        "per_precision": {k: {"ms_per_step": round(v["ms"], 4), "particle_steps_per_s": round(v["value"], 1),
                              "kernel_ms_per_step": {kk: round(kms / a.steps, 4) for kk, (kms, cnt) in
                                                     v["prof"].items() if cnt},
                              "sorts": v["sorts"],
                              "autotuned": None if v["tuned"] is None else
                              {"tile": f"{v['tuned'][0]}x{v['tuned'][1]}", "gather_lds": v["tuned"][2],
                               "push_ms": round(v["tuned"][3], 4)}}
                          for k, v in results.items()},
        "roofline": roofline(w, n, r["prof"]["push_deposit"], main_prec),
    }
    acc = {"after_steps": a.warmup + a.steps + 1,
           "continuity": {k: v["cont"] for k, v in results.items()}}
    if "fp64" in results and "mixed" in results:
        out["mixed_over_fp64"] = round(results["mixed"]["value"] / results["fp64"]["value"], 3)
        ref = results["fp64"]["axis"]
        acc["on_axis_ex_rel_err_mixed_vs_fp64"] = float(np.max(np.abs(results["mixed"]["axis"] - ref)) /
                                                        max(np.max(np.abs(ref)), 1e-300))
    if a.wake_steps > 0 and "fp64" in results and "mixed" in results:
        acc["wakefield"] = wake_accuracy(a.wake_steps)
    out["accuracy"] = acc
    if not a.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_sample_cols)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
