"""Shared runs of the plasma-wave checks, for the oracle (CPU) and for the
library through the C ABI (GPU).  `make(nx, ny, M, dx, dy, dt, bc_field,
bc_particle)` returns an object with the host's Simulation methods used here
(set_species, upload_particles, upload_field, download_field,
setup_field_boundaries, step) and the attributes ng, shape and x_min."""
import math

import numpy as np

from cylindrical_epoch_amd import deck

EPS0 = 8.854187817620389850536563031710750e-12
PERIODIC_FIELD = ("periodic", "periodic", "other", "conduct")
PERIODIC_PART = ("periodic", "periodic", "other", "reflect")


def _zero_crossings(ts, sig):
    return [ts[i] - sig[i] * (ts[i + 1] - ts[i]) / (sig[i + 1] - sig[i])
            for i in range(len(sig) - 1) if sig[i] * sig[i + 1] < 0]


def langmuir(make, nx, ny, ppc, nwaves, dx=1e-6, n=2.3e21, periods=3.2, every=10):
    """Cold electrons on immobile ions (neutral, E = 0), u_x = 1e-3 sin(k x),
    k = 2 pi nwaves / (nx dx); returns (measured frequency, w_p, half periods
    seen) from the sign changes of the E_x Fourier amplitude at k."""
    C, Q0, M0 = deck.C_LIGHT, deck.Q0, deck.M0
    kx = 2 * math.pi * nwaves / (nx * dx)
    w_p = math.sqrt(n * Q0 ** 2 / (EPS0 * M0))
    dt = deck.set_dt(dx, dx, 0.95)
    sim = make(nx, ny, 1, dx, dx, dt, PERIODIC_FIELD, PERIODIC_PART)
    p = deck.uniform_plasma(nx, ny, dx, dx, ppc, n, 0.0, seed=21)
    e = dict(p)
    e["px"] = 1e-3 * M0 * C * np.sin(kx * p["x"])
    for sp, q, m, still in ((e, -Q0, M0, False), (p, Q0, 1836.15 * M0, True)):
        sid = sim.set_species(q, m, still, capacity=len(sp["x"]))
        sim.upload_particles(sid, sp["x"], sp["y"], sp["z"], sp["px"], sp["py"], sp["pz"], sp["w"])
    ng, ts, sig, c0 = sim.ng, [], [], None
    row = min(10, ny - 1)
    for step in range(int(periods * 2 * math.pi / (w_p * dt))):
        sim.step()
        if step % every == every - 1:
            c = np.fft.rfft(sim.download_field("ex")[0][ng - 1 + row, ng:ng + nx].real)[nwaves]
            c0 = c if c0 is None else c0
            sig.append((c * np.conj(c0)).real)
            ts.append((step + 1) * dt)
    cross = _zero_crossings(np.array(ts), np.array(sig))
    return (math.pi / np.diff(cross).mean() if len(cross) > 1 else 0.0), w_p, len(cross) - 1


def em_wave(make, nx=16, ny=96, dx=1e-7, steps=95, row=3):
    """m = 1 standing wave, E along y, uniform across the axis, in a cold
    plasma with w_p = c k, k = 2 pi / (nx dx); returns (measured frequency,
    the discrete-dispersion frequency, half periods seen)."""
    C, Q0, M0 = deck.C_LIGHT, deck.Q0, deck.M0
    k = 2 * math.pi / (nx * dx)
    w_p = C * k
    n = w_p ** 2 * EPS0 * M0 / Q0 ** 2
    dt = deck.set_dt(dx, dx, 0.95)
    kd = 2.0 / dx * math.sin(k * dx / 2)
    want = 2.0 / dt * math.asin(math.sqrt((w_p ** 2 + (C * kd) ** 2) / ((2.0 / dt) ** 2 + w_p ** 2)))
    sim = make(nx, ny, 2, dx, dx, dt, PERIODIC_FIELD, PERIODIC_PART)
    p = deck.uniform_plasma(nx, ny, dx, dx, 8, n, 0.0, seed=23)
    for q, m, still in ((-Q0, M0, False), (Q0, 1836.15 * M0, True)):
        sid = sim.set_species(q, m, still, capacity=len(p["x"]))
        sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    ng = sim.ng
    x = sim.x_min + dx / 2 + (np.arange(1 - ng, nx + ng + 1) - 1) * dx
    E0 = 1.0e8
    for name, ph in (("er", 1.0), ("et", -1j)):
        f = np.zeros(sim.shape, complex)
        f[1][:, :] = ph * E0 * np.cos(k * x)[None, :]
        sim.upload_field(name, f)
    sim.setup_field_boundaries()
    ts, sig, c0 = [], [], None
    for step in range(steps):
        sim.step()
        c = np.fft.fft(sim.download_field("er")[1][ng - 1 + row, ng:ng + nx].real)[1]
        c0 = c if c0 is None else c0
        sig.append((c * np.conj(c0)).real)
        ts.append((step + 1) * dt)
    cross = _zero_crossings(np.array(ts), np.array(sig))
    return (math.pi / np.diff(cross).mean() if len(cross) > 1 else 0.0), want, len(cross) - 1
