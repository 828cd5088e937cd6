"""The plasma-wave checks on the oracle (the whole step: push, deposit,
current boundaries, FDTD), at sizes the CPU runs in seconds.

* Langmuir oscillation: cold electrons on immobile ions, seeded with
  u_x = 1e-3 sin(k x) (one wavelength over 80 cells), oscillate at the
  plasma frequency to 1 % with 32 radial cells.  The r_max boundary and the
  axis pull the frequency down by about 1 / ny^2: 2 % at ny = 8, 0.9 % at
  16, 0.2 % at 32.
* Electromagnetic wave in a plasma: an m = 1 standing wave uniform across
  the axis, w_p = c k, oscillates at the scheme's discrete dispersion
  relation (2 / dt)^2 sin^2(w dt / 2) = w_p^2 cos^2(w dt / 2) + c^2 k_d^2 to
  0.2 %; the cos^2 comes from the step applying the previous and the new
  current for half a step each.
No GPU needed."""
import pytest

from plasma_waves import em_wave, langmuir


class OracleSim:
    """The oracle behind the host Simulation methods the runs use."""

    def __init__(self, oracle, nx, ny, M, dx, dy, dt, bc_field, bc_particle):
        self.st = oracle.State(nx, ny, M, dx, dy, dt, bc_field=bc_field, bc_particle=bc_particle)
        self.ng, self.shape, self.x_min = self.st.g.ng, self.st.shape, self.st.g.x_min
        self.pending = []

    def set_species(self, q, m, immobile, capacity=0):
        self.pending.append((q, m, immobile))
        return len(self.pending) - 1

    def upload_particles(self, sid, x, y, z, px, py, pz, w):
        q, m, immobile = self.pending[sid]
        self.st.add_species(q, m, x, y, z, px, py, pz, w, immobile=immobile)

    def upload_field(self, name, a):
        self.st.arr[name][...] = a

    def download_field(self, name):
        return self.st.arr[name].copy()

    def setup_field_boundaries(self):
        self.st.setup_field_boundaries()

    def step(self):
        self.st.step()


def _make(oracle):
    return lambda *a: OracleSim(oracle, *a)


def test_langmuir_oracle(oracle):
    got, w_p, halves = langmuir(_make(oracle), nx=80, ny=32, ppc=4, nwaves=1)
    assert halves >= 5
    assert got == pytest.approx(w_p, rel=0.01, abs=0)


def test_em_wave_in_plasma_oracle(oracle):
    got, want, halves = em_wave(_make(oracle))
    assert halves >= 6
    assert got == pytest.approx(want, rel=0.002, abs=0)
