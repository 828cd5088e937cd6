"""Host-side mirror of cylindrical EPOCH's hot-path interface.

The reference's main loop (src/epoch2d.F90:213-265) calls argument-less
subroutines on module globals: update_eb_fields_half, push_particles,
current_finish, update_eb_fields_final.  `Simulation` exposes the same
routines, by the same names, over a handle of the HIP library
(include/cylpic.h); the state (mode fields, particles) lives on the GPU.

Field arrays cross the boundary in the reference's Fortran layout
COMPLEX(8) f(1-ng:nx+ng, 1-ng:ny+ng, 0:n_mode-1) (src/shared_data.F90:
474-478), here numpy complex128 arrays of shape (n_mode, ny+2ng, nx+2ng).
"""
import ctypes as C

import numpy as np

from ._lib import CylpicError, check, lib

# c_bc_* codes, src/constants.F90:56-72
BC = dict(periodic=1, other=2, simple_laser=3, simple_outflow=4, open=5,
          zero_gradient=7, clamp=8, reflect=9, conduct=10, thermal=11,
          cpml_laser=12, cpml_outflow=13, zero_b=16)

FIELDS = ("ex", "er", "et", "bx", "br", "bt", "jx", "jr", "jt",
          "bx_old", "br_old", "bt_old", "jx_old", "jr_old", "jt_old")
PRECISION = {"fp64": 0, "mixed": 1}
# distribution function directions, c_dir_* of constants.F90:231-247
DIRS = dict(x=1, y=2, px=3, py=4, pz=5, en=6, gamma_m1=7, xy_angle=8, yz_angle=9, zx_angle=10, mod_p=11)


def _code(b):
    return BC[b] if isinstance(b, str) else int(b)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Simulation:
    """One rank's share of the domain (the r=0 axis and r_max are local)."""

    def __init__(self, nx, ny, n_mode, dx, dy, dt, x_min=0.0, y_max=None,
                 bc_field=("periodic", "periodic", "other", "conduct"),
                 bc_particle=("periodic", "periodic", "other", "reflect"),
                 precision="fp64", device=0, ng=5, tile=None, sort_interval=1, acc_mode=None,
                 gather_lds=None):
        self.nx, self.ny, self.n_mode, self.ng = nx, ny, n_mode, ng
        self.dx, self.dy, self.dt = dx, dy, dt
        self.x_min, self.x_max = x_min, x_min + nx * dx
        self.y_min = 0.0
        self.y_max = ny * dy if y_max is None else y_max
        self.shape = (n_mode, ny + 2 * ng, nx + 2 * ng)
        self.precision = precision
        self.time = 0.0
        h = C.c_void_p()
        bf = (C.c_int * 4)(*[_code(b) for b in bc_field])
        bp = (C.c_int * 4)(*[_code(b) for b in bc_particle])
        check(lib().cylpic_init(C.byref(h), nx, ny, n_mode, ng, dx, dy, dt, x_min, self.x_max,
                                self.y_min, self.y_max, C.cast(bf, C.c_void_p), C.cast(bp, C.c_void_p),
                                PRECISION[precision], device))
        self.h = h
        if acc_mode is not None or gather_lds is not None:
            self.tune(3 if acc_mode is None else int(acc_mode), gather_lds)
        if tile is not None:
            check(lib().cylpic_set_tile(self.h, int(tile[0]), int(tile[1])))
        if sort_interval != 1:
            check(lib().cylpic_set_sort_interval(self.h, int(sort_interval)))
        self.species = []

    # ------------------------------------------------------------ lifetime
    def close(self):
        if getattr(self, "h", None):
            check(lib().cylpic_finalize(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------ species
    def set_species(self, charge, mass, immobile=False, capacity=0):
        sid = len(self.species)
        check(lib().cylpic_set_species(self.h, sid, charge, mass, int(bool(immobile)), int(capacity)))
        self.species.append(dict(charge=charge, mass=mass, immobile=immobile, capacity=capacity))
        return sid

    def upload_particles(self, sid, x, y, z, px, py, pz, w, ids=None):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, px, py, pz, w)]
        n = len(arrs[0])
        idarr = None if ids is None else np.ascontiguousarray(ids, np.int64)
        check(lib().cylpic_upload_particles(self.h, sid, n, *[_ptr(a) for a in arrs], _ptr(idarr)))

    def inject_particles(self, sid, x, y, z, px, py, pz, w, ids=None):
        """Append particles (the moving window's insert_particles batch)."""
        arrs = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, px, py, pz, w)]
        n = len(arrs[0])
        idarr = None if ids is None else np.ascontiguousarray(ids, np.int64)
        check(lib().cylpic_inject_particles(self.h, sid, n, *[_ptr(a) for a in arrs], _ptr(idarr)))

    def move_window(self, ncells):
        """shift_window (window.F90:62-97) for ncells cells."""
        check(lib().cylpic_move_window(self.h, int(ncells)))
        for _ in range(int(ncells)):
            self.x_min = self.x_min + self.dx
            self.x_max = self.x_min + self.nx * self.dx

    def particle_count(self, sid):
        n = C.c_int64()
        check(lib().cylpic_particle_count(self.h, sid, C.byref(n)))
        return n.value

    def download_particles(self, sid, with_ids=True):
        cap = max(1, self.species[sid]["capacity"])
        out = {k: np.empty(cap, np.float64) for k in ("x", "y", "z", "px", "py", "pz", "w")}
        ids = np.empty(cap, np.int64) if with_ids else None
        n = C.c_int64()
        check(lib().cylpic_download_particles(self.h, sid, cap, C.byref(n),
                                              *[_ptr(out[k]) for k in ("x", "y", "z", "px", "py", "pz", "w")],
                                              _ptr(ids)))
        res = {k: v[:n.value].copy() for k, v in out.items()}
        if with_ids:
            res["id"] = ids[:n.value].copy()
        return res

    # ------------------------------------------------------------ fields
    def upload_field(self, name, arr):
        a = np.ascontiguousarray(arr, np.complex128)
        if a.shape != self.shape:
            raise CylpicError(f"field {name}: shape {a.shape} != {self.shape}")
        check(lib().cylpic_upload_field(self.h, FIELDS.index(name), _ptr(a)))

    def download_field(self, name):
        a = np.empty(self.shape, np.complex128)
        check(lib().cylpic_download_field(self.h, FIELDS.index(name), _ptr(a)))
        return a

    def setup_field_boundaries(self):
        check(lib().cylpic_setup_field_boundaries(self.h))

    # ------------------------------------------------------------ restart
    def restart_state(self):
        """time, the window's position (x_min, x_grid_min), the push count
        (the sort phase) and the tile columns' offset."""
        t, x0, xg = C.c_double(), C.c_double(), C.c_double()
        pc, sh = C.c_int64(), C.c_int()
        check(lib().cylpic_get_restart_state(self.h, C.byref(t), C.byref(x0), C.byref(xg), C.byref(pc),
                                             C.byref(sh)))
        return {"time": t.value, "x_min": x0.value, "x_grid_min": xg.value, "push_count": pc.value,
                "tile_shift": sh.value}

    def set_restart_state(self, st):
        check(lib().cylpic_set_restart_state(self.h, float(st["time"]), float(st["x_min"]),
                                             float(st["x_grid_min"]), int(st["push_count"]),
                                             int(st["tile_shift"])))
        self.time = st["time"]
        self.x_min = st["x_min"]
        self.x_max = self.x_min + self.nx * self.dx

    def checkpoint(self):
        """What a restart dump holds: the fifteen field arrays (E, B, J and
        the *_old copies), every species' particles with their ids, and
        restart_state()."""
        return {"fields": {n: self.download_field(n) for n in FIELDS},
                "particles": [self.download_particles(k) for k in range(len(self.species))],
                "state": self.restart_state()}

    def restore(self, ckpt):
        """Continue from checkpoint() on a new handle made with the same grid,
        boundaries, tile and tuning, and with its species registered: the
        boundary snapshots are taken from the still-zero fields (the
        reference's after_deck_last runs before restart_data reads the
        dump), then the state, fields and particles are loaded."""
        self.setup_field_boundaries()
        self.set_restart_state(ckpt["state"])
        for n, a in ckpt["fields"].items():
            self.upload_field(n, a)
        for k, p in enumerate(ckpt["particles"]):
            self.upload_particles(k, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])

    # ------------------------------------------------------------ hot path
    def update_eb_fields_half(self):
        check(lib().cylpic_update_eb_fields_half(self.h))

    def push_particles(self):
        check(lib().cylpic_push_particles(self.h))

    def current_finish(self):
        check(lib().cylpic_current_finish(self.h))

    # ------------------------------------------------------------ diagnostics
    def number_density_modes(self, species=-1):
        """number_density_mode output (calc_number_density_modes,
        calc_df.F90:588-661); species < 0 sums every species."""
        a = np.empty(self.shape, np.complex128)
        check(lib().cylpic_number_density_modes(self.h, int(species), _ptr(a)))
        return a

    def dist_fn(self, species, directions, ranges=None, resolution=None, restrictions=None):
        """A distribution function (general_dist_fn, dist_fn.F90:118-715).
        directions: 1-3 of DIRS (names or c_dir_* numbers); ranges: (lo, hi)
        per direction, equal values for the automatic range; resolution:
        bins per momentum direction; restrictions: {direction: (lo, hi)}.
        Returns (bins with the first direction fastest, used ranges, used
        resolution), or None for an empty species."""
        nd = len(directions)
        d = np.zeros(3, np.int32)
        d[:nd] = [DIRS[k] if isinstance(k, str) else k for k in directions]
        rg = np.zeros(6)
        if ranges is not None:
            rg[: 2 * nd] = np.ravel(ranges)
        res = np.ones(3, np.int32)
        if resolution is not None:
            res[:nd] = resolution
        restr = np.zeros(22)
        use = np.zeros(11, np.int32)
        for k, (lo, hi) in (restrictions or {}).items():
            i = (DIRS[k] if isinstance(k, str) else k) - 1
            restr[2 * i: 2 * i + 2] = lo, hi
            use[i] = 1
        nb = C.c_int64(0)
        check(lib().cylpic_dist_fn(self.h, int(species), nd, _ptr(d), _ptr(rg), _ptr(res), _ptr(restr), _ptr(use),
                                   None, 0, C.byref(nb)))
        if nb.value == 0:
            return None
        out = np.zeros(nb.value)
        rg2 = np.zeros(6)
        if ranges is not None:
            rg2[: 2 * nd] = np.ravel(ranges)
        res2 = np.ones(3, np.int32)
        if resolution is not None:
            res2[:nd] = resolution
        check(lib().cylpic_dist_fn(self.h, int(species), nd, _ptr(d), _ptr(rg2), _ptr(res2), _ptr(restr),
                                   _ptr(use), _ptr(out), nb.value, C.byref(nb)))
        return out.reshape(res2[2], res2[1], res2[0]), rg2[: 2 * nd].reshape(nd, 2), res2[:nd].copy()

    def continuity_begin(self):
        """Store the m = 0 charge at x + v dt/2 before a push."""
        check(lib().cylpic_continuity_begin(self.h))

    def continuity_end(self):
        """After the push (before current_finish): (max |residual|, max
        |charge change|, max |charge|) of the m = 0 continuity equation over
        the cells at least ng from every edge, in coulombs."""
        r, dq, q = C.c_double(), C.c_double(), C.c_double()
        check(lib().cylpic_continuity_end(self.h, C.byref(r), C.byref(dq), C.byref(q)))
        return r.value, dq.value, q.value

    def continuity_modes(self):
        """Per-mode maxima of the last continuity check: the x-r identity's
        residual, the charge change, the J_theta difference |Z - Zp| and |Z|
        (arrays of n_mode)."""
        out = [np.zeros(self.n_mode) for _ in range(4)]
        check(lib().cylpic_continuity_modes(self.h, *[_ptr(a) for a in out]))
        return dict(zip(("residual", "dq", "jtheta", "z"), out))

    def gauss_residual_m0(self):
        """Between steps: (max |eps0 net flux of Ex, Er - Q|, max |Q|, max
        |eps0 net flux|) of the m = 0 Gauss law over the cells at least ng
        from every edge, in coulombs (Q: the mean of the mobile species'
        charges at x - v dt/2 and x + v dt/2, immobile ones at x)."""
        r, q, f = C.c_double(), C.c_double(), C.c_double()
        check(lib().cylpic_gauss_residual_m0(self.h, C.byref(r), C.byref(q), C.byref(f)))
        return r.value, q.value, f.value

    def gauss_residual_m0_noalloc(self):
        """gauss_residual_m0 through the entry point that allocates no
        device memory (grids of at most 1048576 cells with guards)."""
        r, q, f = C.c_double(), C.c_double(), C.c_double()
        check(lib().cylpic_gauss_residual_m0_noalloc(self.h, C.byref(r), C.byref(q), C.byref(f)))
        return r.value, q.value, f.value

    def autotune(self, pushes=2):
        """Pick the fastest tile shape and gather source for the current
        particles (timed pushes, state restored); returns (tile_x, tile_r,
        gather_lds, ms per push)."""
        tx, tr, gl, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        check(lib().cylpic_autotune(self.h, int(pushes), C.byref(tx), C.byref(tr), C.byref(gl), C.byref(ms)))
        return tx.value, tr.value, gl.value, ms.value

    def set_sort_interval(self, n):
        """Sort the particles every n pushes."""
        check(lib().cylpic_set_sort_interval(self.h, int(n)))

    def set_smoothing(self, its, comp_its=0, strides=(1,)):
        """smooth_currents of the deck: smooth_its, smooth_comp_its,
        smooth_strides (current_smooth.F90:50-58, 145-196)."""
        st = np.ascontiguousarray(strides, np.int32)
        check(lib().cylpic_set_smoothing(self.h, int(its), int(comp_its), _ptr(st), len(st)))

    def update_eb_fields_final(self, src_xmin=None, src_xmax=None):
        a = None if src_xmin is None else np.ascontiguousarray(src_xmin, np.float64)
        b = None if src_xmax is None else np.ascontiguousarray(src_xmax, np.float64)
        for s in (a, b):
            if s is not None and s.size != 2 * (self.ny + 1):
                raise CylpicError("laser sources hold source1(0:ny) then source2(0:ny)")
        check(lib().cylpic_update_eb_fields_final(self.h, _ptr(a), _ptr(b)))

    def step(self, src_xmin=None, src_xmax=None):
        """One timestep in the reference's main-loop order."""
        a = None if src_xmin is None else np.ascontiguousarray(src_xmin, np.float64)
        b = None if src_xmax is None else np.ascontiguousarray(src_xmax, np.float64)
        check(lib().cylpic_step(self.h, _ptr(a), _ptr(b)))
        self.time += self.dt

    def tune(self, acc_mode=3, gather_lds=None):
        """Current accumulator of the mixed path (0 FP32, 1 FP32 + 2Sum,
        2 FP64, 3 int64 fixed point, 4 packed 2 x 32-bit fixed point) and
        gather source of both paths (1 LDS-staged tile, 0 L1/L2; default:
        1 for mixed, 0 for FP64)."""
        if gather_lds is None:
            gather_lds = 1 if self.precision == "mixed" else 0
        check(lib().cylpic_tune(self.h, int(acc_mode), int(gather_lds)))

    def sort_particles(self):
        check(lib().cylpic_sort_particles(self.h))

    def set_sort_method(self, method):
        """"radix" (full LSD radix sort every time), "migrate" (neighbour
        migration when it applies, radix otherwise) or "cells" (the upload's
        cell-keyed sort and per-tile rounds interleave every time)."""
        check(lib().cylpic_set_sort_method(self.h, {"radix": 0, "migrate": 1, "cells": 2}[method]))

    def sort_counts(self):
        r, m = C.c_int64(), C.c_int64()
        check(lib().cylpic_sort_counts(self.h, C.byref(r), C.byref(m)))
        return {"radix": r.value, "migrate": m.value}

    def synchronize(self):
        check(lib().cylpic_synchronize(self.h))

    # ------------------------------------------------------------ diagnostics
    def mode_energy(self):
        we = np.zeros(self.n_mode)
        wb = np.zeros(self.n_mode)
        check(lib().cylpic_mode_energy(self.h, _ptr(we), _ptr(wb)))
        return we, wb

    def profile(self, enable=True):
        check(lib().cylpic_profile(self.h, int(bool(enable))))

    def kernel_time(self, name):
        ms = C.c_double()
        cnt = C.c_int64()
        check(lib().cylpic_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


class Batch:
    """Several runs stepped together through captured HIP graphs, their
    mixed pushes sharing persistent launches (cylpic_batch_*).  run() steps
    in chunks whose graphs are kept and relaunched while the runs' state
    repeats; step() captures a single step each call.  The runs stay
    ordinary Simulations for set-up, downloads and diagnostics; between
    batch steps call synchronize() before touching them directly."""

    def __init__(self, sims):
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.h for s in self.sims])
        self._arr = arr
        b = C.c_void_p()
        check(lib().cylpic_batch_create(C.byref(b), C.cast(arr, C.c_void_p), len(self.sims)))
        self.b = b

    def step(self, src_xmin=None, src_xmax=None):
        """One timestep of every run; src_xmin / src_xmax: per-run laser
        sources (each None or 2 (ny + 1) doubles), or None for no laser."""
        keep = []

        def table(srcs):
            if srcs is None:
                return None
            t = (C.c_void_p * len(self.sims))()
            for k, v in enumerate(srcs):
                if v is not None:
                    a = np.ascontiguousarray(v, np.float64)
                    keep.append(a)
                    t[k] = a.ctypes.data
            keep.append(t)
            return C.cast(t, C.c_void_p)

        check(lib().cylpic_batch_step(self.b, table(src_xmin), table(src_xmax)))
        for s in self.sims:
            s.time += s.dt

    def run(self, nsteps, src_xmin=None, src_xmax=None):
        """nsteps timesteps of every run.  src_xmin / src_xmax: None, or a
        function of the step index j (0 .. nsteps - 1) returning the per-run
        laser sources of that step (a list with None or 2 (ny + 1) doubles
        per run)."""
        src = None
        if src_xmin is not None or src_xmax is not None:
            widths = [2 * (s.ny + 1) for s in self.sims]
            block = 2 * sum(widths)
            src = np.zeros((nsteps, block), np.float64)
            for j in range(nsteps):
                off = 0
                lo = src_xmin(j) if src_xmin is not None else None
                hi = src_xmax(j) if src_xmax is not None else None
                for r, wdt in enumerate(widths):
                    if lo is not None and lo[r] is not None:
                        src[j, off:off + wdt] = lo[r]
                    if hi is not None and hi[r] is not None:
                        src[j, off + wdt:off + 2 * wdt] = hi[r]
                    off += 2 * wdt
        ptr = src.ctypes.data if src is not None else None
        check(lib().cylpic_batch_steps(self.b, int(nsteps), ptr))
        for s in self.sims:
            s.time += nsteps * s.dt

    def synchronize(self):
        check(lib().cylpic_batch_synchronize(self.b))

    def stats(self):
        steps, inst, rep, groups = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int()
        check(lib().cylpic_batch_stats(self.b, C.byref(steps), C.byref(inst), C.byref(rep), C.byref(groups)))
        return {"steps": steps.value, "instantiations": inst.value, "replays": rep.value,
                "push_groups": groups.value}

    def close(self):
        if self.b:
            check(lib().cylpic_batch_destroy(self.b))
            self.b = None
