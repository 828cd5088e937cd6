"""ctypes binding of the FP64 CPU oracle (oracle/cylpic_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg as the parity checker.  The product library
(cylindrical_epoch_amd) never imports this module.

Arrays follow the reference's Fortran layout COMPLEX(8) f(1-ng:nx+ng,
1-ng:ny+ng, 0:n_mode-1) (src/shared_data.F90:474-478), held here as numpy
complex128 arrays of shape (n_mode, ny+2ng, nx+2ng): element [m, j+ng-1,
i+ng-1] is Fortran f(i, j, m).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CYLPIC_ORACLE_PLANT=k loads the oracle built with planted mistake k
# (oracle/Makefile `plant`), for tests/test_oracle_planted.py, which checks
# that every planted mistake fails a committed test.
PLANT = int(os.environ.get("CYLPIC_ORACLE_PLANT", "0") or 0)
_LIB_PATH = (os.path.join(_HERE, "build", "plant", f"libcylpic_oracle_{PLANT}.so") if PLANT
             else os.path.join(_HERE, "build", "libcylpic_oracle.so"))
_lib = None

BC = dict(periodic=1, other=2, simple_laser=3, simple_outflow=4, open=5,
          zero_gradient=7, clamp=8, reflect=9, conduct=10, thermal=11,
          zero_b=16)


def build():
    if PLANT:
        subprocess.check_call(["make", "-s", "-C", _HERE, "plant", f"PLANTS={PLANT}"])
    else:
        subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("n_mode", C.c_int), ("ng", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dt", C.c_double),
                ("x_min", C.c_double), ("x_max", C.c_double),
                ("y_min", C.c_double), ("y_max", C.c_double),
                ("x_grid_min", C.c_double), ("y_grid_min", C.c_double),
                ("bc_field", C.c_int * 4), ("bc_particle", C.c_int * 4),
                ("time", C.c_double)]


_P = C.c_void_p


class OCplx(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class Fields(C.Structure):
    _fields_ = [(n, _P) for n in ("ex", "er", "et", "bx", "br", "bt", "jx", "jr", "jt",
                                   "bx_old", "br_old", "bt_old", "jx_old", "jr_old", "jt_old")] + [
        ("bnd_xmin", _P * 6), ("bnd_xmax", _P * 6)]


class Species(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(n, _P) for n in ("x", "y", "z", "px", "py", "pz", "w", "id")] + [
        ("q", C.c_double), ("m", C.c_double), ("immobile", C.c_int)]


def _declare(L):
    gp, fp, sp = C.POINTER(Grid), C.POINTER(Fields), C.POINTER(Species)
    dp = C.c_void_p
    sig = {
        "orc_update_e_field": [gp, fp], "orc_update_b_field": [gp, fp],
        "orc_update_eb_fields_half": [gp, fp],
        "orc_update_eb_fields_final": [gp, fp, dp, dp],
        "orc_efield_bcs": [gp, fp], "orc_bfield_bcs": [gp, fp, C.c_int],
        "orc_bfield_final_bcs": [gp, fp, dp, dp],
        "orc_outflow_bcs_x_min": [gp, fp, dp], "orc_outflow_bcs_x_max": [gp, fp, dp],
        "orc_outflow_bcs_r_max": [gp, fp], "orc_setup_field_boundaries": [gp, fp],
        "orc_push_particles": [gp, fp, sp, C.c_int],
        "orc_push_particles_threads": [gp, fp, sp, C.c_int, C.c_int],
        "orc_current_bcs": [gp, fp], "orc_current_bcs_r_min_final": [gp, fp],
        "orc_current_finish": [gp, fp], "orc_particle_bcs": [gp, sp, C.c_int],
        "orc_shift_window": [gp, fp, sp, C.c_int, C.c_int],
        "orc_smooth_current": [gp, fp, C.c_int, C.c_int, C.c_void_p, C.c_int],
        "orc_geometry": [gp, dp, dp, dp, dp],
        "orc_deposit_from_weights": [gp, fp, dp, dp, dp, dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_double, C.c_double, C.c_double,
                                     OCplx, OCplx, C.c_double,
                                     dp, dp, dp, dp],
        "orc_cell_keys": [gp, C.c_int64, dp, dp, dp, C.c_int, C.c_int, C.c_int, dp],
        "orc_stable_sort_keys": [C.c_int64, dp, C.c_int32, dp],
        "orc_mode_energy": [gp, fp, dp, dp],
        "orc_number_density_modes": [gp, sp, C.c_int, C.c_int, dp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = None
    L.orc_dist_fn.argtypes = [gp, sp, C.c_int, dp, dp, dp, dp, dp, dp]
    L.orc_dist_fn.restype = C.c_int
    L.orc_take_error.argtypes = []
    L.orc_take_error.restype = C.c_int
    L.orc_set_dt.argtypes = [C.c_double] * 5
    L.orc_set_dt.restype = C.c_double


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


FIELD_NAMES = ("ex", "er", "et", "bx", "br", "bt", "jx", "jr", "jt",
               "bx_old", "br_old", "bt_old", "jx_old", "jr_old", "jt_old")


class State:
    """Grid + fields + species for the oracle, numpy-owned."""

    def __init__(self, nx, ny, n_mode, dx, dy, dt, x_min=0.0, y_max=None,
                 bc_field=("periodic", "periodic", "other", "conduct"),
                 bc_particle=("periodic", "periodic", "other", "reflect"), ng=5):
        self.g = Grid()
        g = self.g
        g.nx, g.ny, g.n_mode, g.ng = nx, ny, n_mode, ng
        g.dx, g.dy, g.dt = dx, dy, dt
        g.x_min = x_min
        g.x_max = x_min + nx * dx
        g.y_min = 0.0
        g.y_max = ny * dy if y_max is None else y_max
        g.x_grid_min = g.x_min + dx / 2.0
        g.y_grid_min = g.y_min + dy / 2.0
        g.bc_field[:] = [map_field_bc(b) for b in bc_field]
        g.bc_particle[:] = [map_particle_bc(b) for b in bc_particle]
        g.time = 0.0
        shape = (n_mode, ny + 2 * ng, nx + 2 * ng)
        self.shape = shape
        self.arr = {n: np.zeros(shape, np.complex128) for n in FIELD_NAMES}
        self.bnd = {s: [np.zeros((n_mode, ny + 2 * ng), np.complex128) for _ in range(6)]
                    for s in ("xmin", "xmax")}
        self.f = Fields()
        for n in FIELD_NAMES:
            setattr(self.f, n, self.arr[n].ctypes.data)
        for k in range(6):
            self.f.bnd_xmin[k] = self.bnd["xmin"][k].ctypes.data
            self.f.bnd_xmax[k] = self.bnd["xmax"][k].ctypes.data
        self.species = []
        self.smooth = None  # (its, comp_its, strides) of smooth_current, or None

    # Fortran-index view helpers
    def at(self, name, i, j, m):
        ng = self.g.ng
        return self.arr[name][m, j + ng - 1, i + ng - 1]

    def add_species(self, q, m, x, y, z, px, py, pz, w, immobile=False, ids=None):
        n = len(x)
        d = dict(x=x, y=y, z=z, px=px, py=py, pz=pz, w=w)
        d = {k: np.ascontiguousarray(v, dtype=np.float64).copy() for k, v in d.items()}
        d["id"] = (np.arange(n, dtype=np.int64) if ids is None
                   else np.ascontiguousarray(ids, dtype=np.int64).copy())
        sp = dict(q=q, m=m, immobile=bool(immobile), n=n, **d)
        self.species.append(sp)
        return sp

    def _species_array(self):
        arr = (Species * max(1, len(self.species)))()
        for k, s in enumerate(self.species):
            a = arr[k]
            a.n = s["n"]
            for nm in ("x", "y", "z", "px", "py", "pz", "w", "id"):
                setattr(a, nm, s[nm].ctypes.data)
            a.q, a.m, a.immobile = s["q"], s["m"], int(s["immobile"])
        return arr

    def _species_back(self, arr):
        for k, s in enumerate(self.species):
            n = arr[k].n
            s["n"] = n
            for nm in ("x", "y", "z", "px", "py", "pz", "w", "id"):
                s[nm] = s[nm][:n].copy()

    # reference routines
    def update_eb_fields_half(self):
        lib().orc_update_eb_fields_half(C.byref(self.g), C.byref(self.f))

    def update_eb_fields_final(self, src_xmin=None, src_xmax=None):
        a = None if src_xmin is None else np.ascontiguousarray(src_xmin, np.float64)
        b = None if src_xmax is None else np.ascontiguousarray(src_xmax, np.float64)
        lib().orc_update_eb_fields_final(C.byref(self.g), C.byref(self.f),
                                         None if a is None else _ptr(a),
                                         None if b is None else _ptr(b))

    def push_particles(self):
        arr = self._species_array()
        lib().orc_push_particles(C.byref(self.g), C.byref(self.f), arr, len(self.species))
        self._species_back(arr)
        if lib().orc_take_error():
            raise FloatingPointError("oracle push: a particle moved more than one cell in a step "
                                     "(or its position is not finite)")

    def push_particles_threads(self, nthreads):
        """The push on nthreads host threads (x slabs, private currents added
        in slab order): the CPU baseline's timing path."""
        arr = self._species_array()
        lib().orc_push_particles_threads(C.byref(self.g), C.byref(self.f), arr, len(self.species), int(nthreads))
        self._species_back(arr)
        if lib().orc_take_error():
            raise FloatingPointError("oracle push: a particle moved more than one cell in a step "
                                     "(or its position is not finite)")

    def current_finish(self, smooth=None):
        """current_finish (current_smooth.F90:29-46); smooth = (its,
        comp_its, strides) applies smooth_current after the current bcs."""
        lib().orc_current_finish(C.byref(self.g), C.byref(self.f))
        if smooth:
            self.smooth_current(*smooth)

    def smooth_current(self, its, comp_its=0, strides=(1,)):
        """smooth_current (current_smooth.F90:50-58, 145-196) on jx, jr, jt."""
        st = np.ascontiguousarray(strides, np.int32)
        lib().orc_smooth_current(C.byref(self.g), C.byref(self.f), int(its), int(comp_its),
                                 st.ctypes.data_as(C.c_void_p), len(st))

    def setup_field_boundaries(self):
        lib().orc_setup_field_boundaries(C.byref(self.g), C.byref(self.f))

    def append_particles(self, species, x, y, z, px, py, pz, w, ids=None):
        """Append particles to a species (the window's injection)."""
        sp = self.species[species]
        d = dict(x=x, y=y, z=z, px=px, py=py, pz=pz, w=w)
        n = len(x)
        for k, v in d.items():
            sp[k] = np.concatenate([sp[k][:sp["n"]], np.asarray(v, np.float64)])
        new_ids = (np.arange(n, dtype=np.int64) if ids is None else np.asarray(ids, np.int64))
        sp["id"] = np.concatenate([sp["id"][:sp["n"]], new_ids])
        sp["n"] = sp["n"] + n

    def shift_window(self, ncells):
        """shift_window (window.F90:62-97) for ncells cells."""
        arr = self._species_array()
        lib().orc_shift_window(C.byref(self.g), C.byref(self.f), arr, len(self.species), int(ncells))
        self._species_back(arr)

    def step(self, src_xmin=None, src_xmax=None, threads=0):
        """One timestep of the reference main loop, epoch2d.F90:213-265
        (no moving window, no optional physics); threads > 0 pushes on that
        many host threads (push_particles_threads)."""
        self.update_eb_fields_half()
        if threads > 0:
            self.push_particles_threads(threads)
        else:
            self.push_particles()
        self.current_finish(smooth=self.smooth)
        self.g.time += self.g.dt
        self.update_eb_fields_final(src_xmin, src_xmax)

    def call(self, name, *args):
        fn = getattr(lib(), name)
        fn(C.byref(self.g), C.byref(self.f), *args)

    def mode_energy(self):
        M = self.g.n_mode
        we, wb = np.zeros(M), np.zeros(M)
        lib().orc_mode_energy(C.byref(self.g), C.byref(self.f), _ptr(we), _ptr(wb))
        return we, wb

    def number_density_modes(self, species=-1):
        """calc_number_density_modes (calc_df.F90:588-661); species < 0 sums
        every species."""
        out = np.zeros(self.shape, np.complex128)
        arr = self._species_array()
        lib().orc_number_density_modes(C.byref(self.g), arr, len(self.species), int(species), _ptr(out))
        return out

    def dist_fn(self, species, directions, ranges=None, resolution=None, restrictions=None):
        """general_dist_fn (dist_fn.F90:118-715) on one rank.  directions:
        1-3 of DIRS; ranges: per direction (lo, hi), equal values meaning
        the automatic range; resolution: bins per momentum direction;
        restrictions: {direction: (lo, hi)}.  Returns (histogram with the
        first direction fastest, used ranges, used resolution) or None for
        an empty species."""
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
        bound = 1
        for k in range(nd):
            bound *= {1: self.g.nx + 1, 2: self.g.ny + 1}.get(int(d[k]), int(res[k]))
        out = np.zeros(max(1, bound))
        arr = self._species_array()
        sp = C.cast(C.byref(arr[species]), C.POINTER(Species))
        rc = lib().orc_dist_fn(C.byref(self.g), sp, nd, _ptr(d), _ptr(rg), _ptr(res), _ptr(restr), _ptr(use),
                               _ptr(out))
        if rc == 2:
            return None
        if rc:
            raise ValueError("bad distribution function direction")
        n = int(res[0]) * int(res[1]) * int(res[2])
        return out[:n].reshape(res[2], res[1], res[0]), rg[: 2 * nd].reshape(nd, 2), res[:nd].copy()

    def geometry(self):
        n = self.g.ny + 2 * self.g.ng + 1
        a = [np.zeros(n) for _ in range(4)]
        lib().orc_geometry(C.byref(self.g), *[_ptr(x) for x in a])
        return a


# c_dir_* of constants.F90:231-247
DIRS = dict(x=1, y=2, px=3, py=4, pz=5, en=6, gamma_m1=7, xy_angle=8, yz_angle=9, zx_angle=10, mod_p=11)


def map_field_bc(name):
    """setup_boundaries' field mapping, boundary.F90:30-68."""
    b = BC[name] if isinstance(name, str) else int(name)
    if b == BC["other"]:
        b = BC["clamp"]
    if b == BC["reflect"]:
        b = BC["clamp"]
    if b == BC["open"]:
        b = BC["simple_outflow"]
    return b


def map_particle_bc(name):
    """setup_particle_boundary, boundary.F90:98-131."""
    b = BC[name] if isinstance(name, str) else int(name)
    if b in (BC["other"], BC["conduct"]):
        b = BC["reflect"]
    if b in (BC["simple_laser"], BC["simple_outflow"]):
        b = BC["open"]
    return b


def set_dt(dx, dy, dt_multiplier=1.0, dt_plasma=0.0, dt_laser=0.0):
    return lib().orc_set_dt(dx, dy, dt_multiplier, dt_plasma, dt_laser)


def cell_keys(state, x, y, z, tile_x, tile_r, shift=0):
    n = len(x)
    keys = np.zeros(n, np.int32)
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    lib().orc_cell_keys(C.byref(state.g), n, _ptr(x), _ptr(y), _ptr(z), tile_x, tile_r, shift, _ptr(keys))
    return keys


def stable_sort(keys, nkeys):
    keys = np.ascontiguousarray(keys, np.int32)
    perm = np.zeros(len(keys), np.int64)
    lib().orc_stable_sort_keys(len(keys), _ptr(keys), nkeys, _ptr(perm))
    return perm


def deposit_from_weights(state, gx, hx, gy, hy, cell_x2, cell_y2, dcellx, dcelly,
                         fcx, fcy, fcz, exp_itheta_05=(1.0, 0.0), exp_idtheta=(1.0, 0.0),
                         dtheta=0.0):
    geo = state.geometry()
    arrs = [np.ascontiguousarray(a, np.float64) for a in (gx, hx, gy, hy)]
    lib().orc_deposit_from_weights(C.byref(state.g), C.byref(state.f),
                                   *[_ptr(a) for a in arrs], cell_x2, cell_y2, dcellx, dcelly,
                                   fcx, fcy, fcz, OCplx(*exp_itheta_05),
                                   OCplx(*exp_idtheta), dtheta,
                                   *[_ptr(a) for a in geo])
