"""Shared set-up for the GPU parity tests: the same seeded inputs handed to
the oracle (CPU, FP64) and to the HIP library through its C ABI."""
import sys

import numpy as np

from cylindrical_epoch_amd import deck
from cylindrical_epoch_amd.host import FIELDS, Simulation

Q0, M0, C_LIGHT = deck.Q0, deck.M0, deck.C_LIGHT


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_fields(shape, rng, e_amp=1e9, b_amp=1.0, j_amp=1e12):
    out = {}
    for n in ("ex", "er", "et"):
        out[n] = e_amp * (rng.normal(size=shape) + 1j * rng.normal(size=shape))
    for n in ("bx", "br", "bt"):
        out[n] = b_amp * (rng.normal(size=shape) + 1j * rng.normal(size=shape))
    for n in ("jx", "jr", "jt"):
        out[n] = j_amp * (rng.normal(size=shape) + 1j * rng.normal(size=shape))
    for n in ("bx_old", "br_old", "bt_old"):
        out[n] = out[n[:2]] * (1 + 0.01 * rng.normal())
    for n in ("jx_old", "jr_old", "jt_old"):
        out[n] = out[n[:2]] * 0.5
    return out


def make_pair(oracle, nx, ny, M, dx=1e-7, dy=1e-7, dt=None, bc_field=None, bc_particle=None,
              precision="fp64", tile=None, sort_interval=1):
    bc_field = bc_field or ("periodic", "periodic", "other", "conduct")
    bc_particle = bc_particle or ("periodic", "periodic", "other", "reflect")
    if dt is None:
        dt = deck.set_dt(dx, dy, 0.95)
    st = oracle.State(nx, ny, M, dx, dy, dt, bc_field=bc_field, bc_particle=bc_particle)
    sim = Simulation(nx, ny, M, dx, dy, dt, bc_field=bc_field, bc_particle=bc_particle,
                     precision=precision, tile=tile, sort_interval=sort_interval)
    return st, sim


def load_fields(st, sim, fields):
    for n, a in fields.items():
        st.arr[n][:] = a
        sim.upload_field(n, a)


def add_plasma(st, sim, nx, ny, dx, dy, ppc, density=1e25, temp_ev=1000.0, drift=(0, 0, 0),
               seed=1, q=-Q0, m=M0, shuffle=False, capacity=None):
    p = deck.uniform_plasma(nx, ny, dx, dy, ppc, density, temp_ev, mass=m, seed=seed)
    for k, d in zip(("px", "py", "pz"), drift):
        p[k] = p[k] + d * m * C_LIGHT
    if shuffle:
        perm = np.random.default_rng(seed + 1).permutation(len(p["x"]))
        p = {k: v[perm] for k, v in p.items()}
    n = len(p["x"])
    st.add_species(q, m, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])
    sid = sim.set_species(q, m, False, capacity=n if capacity is None else capacity)
    sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])
    return p


def by_id(d):
    order = np.argsort(d["id"], kind="stable")
    return {k: v[order] for k, v in d.items()}


def max_rel(a, b):
    scale = max(np.max(np.abs(a)), np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def all_field_names():
    return FIELDS
