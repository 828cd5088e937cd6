"""Stream-file format shared by tests/fortran/driver.f90 and its tests, and
the ctypes call of the Fortran driver library."""
import ctypes
import os

import numpy as np

DRIVER_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "_build",
                          "libcylpic_fdriver.so")


def run_driver(fin, fout):
    """Run the Fortran host (one call of cylpic_fortran_driver); returns its
    status (0 = ok, else the failing cylpic call's return code)."""
    L = ctypes.CDLL(DRIVER_LIB)
    L.cylpic_fortran_driver.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    L.cylpic_fortran_driver.restype = ctypes.c_int
    return L.cylpic_fortran_driver(fin.encode(), fout.encode())


def write_input(path, nx, ny, n_mode, nsteps, precision, bc_field, bc_particle, tile, dx, dy, dt, x_min,
                charge, mass, p):
    with open(path, "wb") as f:
        np.array([nx, ny, n_mode, nsteps, precision, *bc_field, *bc_particle, *tile], np.int32).tofile(f)
        np.array([dx, dy, dt, x_min, charge, mass], np.float64).tofile(f)
        np.array([len(p["x"])], np.int64).tofile(f)
        for k in ("x", "y", "z", "px", "py", "pz", "w"):
            np.ascontiguousarray(p[k], np.float64).tofile(f)


def read_output(path, nx, ny, n_mode, ng=5):
    raw = open(path, "rb").read()
    off = 0
    n = int(np.frombuffer(raw, np.int64, 1, off)[0])
    off += 8
    parts = {}
    for k in ("x", "y", "z", "px", "py", "pz", "w"):
        parts[k] = np.frombuffer(raw, np.float64, n, off).copy()
        off += 8 * n
    we = np.frombuffer(raw, np.float64, n_mode, off).copy()
    off += 8 * n_mode
    wb = np.frombuffer(raw, np.float64, n_mode, off).copy()
    off += 8 * n_mode
    shape = (n_mode, ny + 2 * ng, nx + 2 * ng)
    cnt = shape[0] * shape[1] * shape[2]
    ex = np.frombuffer(raw, np.complex128, cnt, off).reshape(shape).copy()
    off += 16 * cnt
    jr = np.frombuffer(raw, np.complex128, cnt, off).reshape(shape).copy()
    return parts, we, wb, ex, jr
