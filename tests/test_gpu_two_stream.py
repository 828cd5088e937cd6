"""Two-stream instability through the whole timestep (the reference's
example_decks/two_stream_instability.deck, rescaled so that it runs in a few
thousand steps): two cold electron beams of density n_b each drift at +-0.1 c
along x in a periodic box of 400 x 64 cells, one mode.  A 1e-3 wavelength
displacement seeds the wavenumber k = 2 pi / (80 dx), chosen as the fastest
growing one, k v = (sqrt(3) / 2) w_b.  For a perturbation uniform in r (the
conducting r_max clamps E_r, so it is an eigenmode) the cold-beam dispersion
relation 1 = w_e^2 [(w - k v)^-2 + (w + k v)^-2], with the relativistic
longitudinal mass w_e = w_b / gamma^(3/2), gives the growth rate
Gamma^2 = w_e sqrt(w_e^2 + 4 k^2 v^2) - k^2 v^2 - w_e^2, 0.496 w_b here.
The amplitude of E_x at that k must grow at Gamma to 10 % over the linear
phase, 5 < w_b t < 8.5.

It runs on the device only: about 2.9e3 steps of 8e5 particles is beyond
the CPU oracle's reach in a test; the oracle parity of each step is what the
other tests establish."""
import math

import numpy as np
import pytest

from cylindrical_epoch_amd import deck
from cylindrical_epoch_amd.host import Simulation
from gpu_helpers import log
from plasma_waves import em_wave, langmuir

pytestmark = pytest.mark.gpu

EPS0 = 8.854187817620389850536563031710750e-12
NX, NY, DX = 400, 64, 1e-6


def growth_rate(w_e, kv):
    return math.sqrt(w_e * math.sqrt(w_e ** 2 + 4 * kv ** 2) - kv ** 2 - w_e ** 2)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_two_stream_growth_rate(precision):
    C, Q0, M0 = deck.C_LIGHT, deck.Q0, deck.M0
    v = 0.1 * C
    gam = 1.0 / math.sqrt(1.0 - (v / C) ** 2)
    kx = 2 * math.pi / (80 * DX)
    w_b = kx * v / (math.sqrt(3) / 2)
    n_b = w_b ** 2 * EPS0 * M0 / Q0 ** 2
    want = growth_rate(w_b / gam ** 1.5, kx * v) / w_b
    dt = deck.set_dt(DX, DX, 0.95)
    log("two-stream", precision, f"w_b dt = {w_b * dt:.2e}, predicted growth {want:.3f} w_b")
    sim = Simulation(NX, NY, 1, DX, DX, dt, bc_field=("periodic", "periodic", "other", "conduct"),
                     bc_particle=("periodic", "periodic", "other", "reflect"), precision=precision)
    for s, sign in enumerate((1, -1)):
        p = deck.uniform_plasma(NX, NY, DX, DX, 16, n_b, 0.0, seed=100 + s)
        p["x"] = (p["x"] + 1e-3 * (80 * DX) * np.sin(kx * p["x"])) % (NX * DX)
        p["px"][:] = sign * gam * v * M0
        sid = sim.set_species(-Q0, M0, False, capacity=len(p["x"]))
        sim.upload_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"])
    ng, amps, ts = sim.ng, [], []
    for n in range(int(8.6 / (w_b * dt))):
        sim.step()
        if n % 20 == 19:
            ex = sim.download_field("ex")[0][ng - 1 + 10, ng:ng + NX].real
            amps.append(2.0 * np.abs(np.fft.rfft(ex))[5] / NX)
            ts.append((n + 1) * dt * w_b)
    ts, amps = np.array(ts), np.array(amps)
    sel = (ts > 5.0) & (ts < 8.5)
    got = np.polyfit(ts[sel], np.log(amps[sel]), 1)[0]
    log(f"  measured growth {got:.3f} w_b")
    assert got == pytest.approx(want, rel=0.10, abs=0)
    sim.close()


def _device(precision):
    def make(nx, ny, M, dx, dy, dt, bc_field, bc_particle):
        sim = Simulation(nx, ny, M, dx, dy, dt, bc_field=bc_field, bc_particle=bc_particle, precision=precision)
        _made.append(sim)
        return sim
    return make


_made = []


@pytest.fixture(autouse=True)
def _close_sims():
    yield
    while _made:
        _made.pop().close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_langmuir_frequency(precision):
    """Cold electrons on immobile ions (neutral, E = 0), u_x = 1e-3 sin(k x)
    with k = 2 pi / (80 dx) on 400 x 32 cells, oscillate at the plasma
    frequency w_p = sqrt(n e^2 / (eps0 m)) to 1 % (measured 0.998; the r_max
    boundary and the axis pull it down by about 1 / ny^2).  A wrong factor in
    the current, the charge or eps0 moves the frequency by its square root."""
    got, w_p, halves = langmuir(_device(precision), nx=NX, ny=32, ppc=16, nwaves=5)
    log(f"Langmuir {precision}: measured {got / w_p:.4f} w_p over {halves} half periods")
    assert halves >= 5
    assert got == pytest.approx(w_p, rel=0.01, abs=0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_em_wave_in_plasma(precision):
    """An electromagnetic standing wave in a cold plasma (electrons on
    immobile ions, w_p = c k): the m = 1 field uniform across the axis, E
    along y, with k = 2 pi / (16 dx), drives the m = 1 currents of the
    deposit and their axis rows.  It oscillates at
    (2 / dt)^2 sin^2(w dt / 2) = w_p^2 cos^2(w dt / 2) + c^2 k_d^2,
    k_d = (2 / dx) sin(k dx / 2) the Yee wavenumber; the cos^2 comes from the
    step applying the previous and the new current for half a step each
    (without it the frequency would be 0.7 % higher here).  Bound 0.2 %, on a
    row that r_max has not reached.  Without the plasma's current the wave
    would oscillate at c k_d, 30 % slower."""
    got, want, halves = em_wave(_device(precision))
    log(f"EM wave {precision}: measured {got / want:.4f} of the expected frequency ({halves} half periods)")
    assert halves >= 6
    assert got == pytest.approx(want, rel=0.002, abs=0)
