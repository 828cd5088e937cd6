"""Host side of the moving window: the reference's moving_window bookkeeping
(src/housekeeping/window.F90:353-376: the shift fraction accumulates
window_v_x dt / dx and each whole cell moves the window) and insert_particles
(window.F90:157-325: a column of new particles beyond x_max per cell, from
the deck's density), producing the columns and shifts once so that any
number of targets (library handles, cylindrical_epoch_amd.host.Simulation,
or a CPU state a test compares with) see identical columns."""
import math

from . import deck


class MovingWindow:
    """Window moving at v (default c) once the time reaches `start`
    (window_start_time).  Columns are drawn from a seeded generator per
    column index with `ppc` particles per r cell, weights scaled by
    profile(x) (None: uniform), temperature temp_ev."""

    def __init__(self, ny, dx, dy, dt, ppc, density, profile=None, temp_ev=0.0, v=deck.C_LIGHT, seed=11,
                 id0=10 ** 9, start=0.0):
        self.ny, self.dx, self.dy, self.dt = ny, dx, dy, dt
        self.ppc, self.density, self.temp_ev = ppc, density, temp_ev
        self.profile = profile or (lambda x: 1.0 + 0.0 * x)
        self.v, self.seed, self.start = v, seed, start
        self.fraction = 0.0
        self.columns = 0
        self.next_id = id0
        self.injected = 0

    def column(self, x_max):
        """The particles of the next column beyond x_max."""
        p = deck.profiled_plasma(1, self.ny, self.dx, self.dy, self.ppc, self.density, self.profile,
                                 self.temp_ev, x_min=x_max, seed=self.seed + self.columns, cells_x=[0],
                                 id0=self.next_id)
        self.columns += 1
        self.next_id += len(p["x"])
        self.injected += len(p["x"])
        return p

    def advance(self, x_max, time=None):
        """After the step that ends at `time` (None: the window has started):
        the columns to inject (beyond x_max, one per cell) and the number of
        cells to move."""
        if time is not None and time < self.start:
            return [], 0
        self.fraction += self.dt * self.v / self.dx
        k = int(math.floor(self.fraction))
        if k <= 0:
            return [], 0
        self.fraction -= k
        return [self.column(x_max + c * self.dx) for c in range(k)], k


def apply_to_sim(sim, sid, cols, k):
    """Inject the columns into species `sid` of a library handle, then move
    its window k cells."""
    for p in cols:
        sim.inject_particles(sid, p["x"], p["y"], p["z"], p["px"], p["py"], p["pz"], p["w"], ids=p["id"])
    if k:
        sim.move_window(k)
