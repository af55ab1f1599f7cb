"""MD tuners (HOOMD's ``md.tune``): the neighbor-list buffer tuner."""

import time

import numpy as np

from ... import _azp
from .. import trigger as _trigger

torch = _azp.torch


class NeighborListBuffer:
    """Chooses a neighbor list's ``buffer`` (r_buff) by measured speed
    (the role of HOOMD's ``md.tune.NeighborListBuffer``, grid solver).

    A larger buffer rebuilds the list less often but lists more pairs; the
    fastest value depends on the system and the hardware.  The tuner visits
    ``n_bins`` buffers evenly spaced in [minimum_buffer, maximum_buffer],
    times the steps between two of its triggers at each (the GPU is
    synchronized at each trigger, so the window covers the steps' work),
    repeats the grid ``n_rounds`` times around the best value with half the
    spacing, and leaves the buffer at the fastest.  ``tuned`` turns true then;
    ``best_buffer`` and ``results`` (buffer -> timesteps/s) record the search.

    Use ``with_grid``; the trigger's period should cover several list
    rebuilds (a few hundred steps for liquids).
    """

    def __init__(self, trigger, nlist, minimum_buffer, maximum_buffer, n_bins=5, n_rounds=1):
        if not 0.0 <= minimum_buffer < maximum_buffer:
            raise ValueError("need 0 <= minimum_buffer < maximum_buffer")
        if int(n_bins) < 2:
            raise ValueError("n_bins must be at least 2")
        self.trigger = _trigger.as_trigger(trigger)
        self.nlist = nlist
        self.minimum_buffer = float(minimum_buffer)
        self.maximum_buffer = float(maximum_buffer)
        self.n_bins = int(n_bins)
        self.n_rounds = int(n_rounds)
        self.results = {}
        self.best_buffer = None
        self._sim = None
        self._grid = list(np.linspace(self.minimum_buffer, self.maximum_buffer, self.n_bins))
        self._round = 0
        self._current = None
        self._t0 = None
        self._step0 = None

    @classmethod
    def with_grid(cls, trigger, nlist, maximum_buffer, minimum_buffer=0.0, n_bins=5, n_rounds=1):
        return cls(trigger, nlist, minimum_buffer, maximum_buffer, n_bins, n_rounds)

    @property
    def tuned(self):
        return self.best_buffer is not None

    def _attach(self, sim):
        self._sim = sim

    def _start(self, timestep, buffer):
        self.nlist.buffer = float(buffer)
        self._current = float(buffer)
        torch.cuda.synchronize()
        self._t0 = time.perf_counter()
        self._step0 = int(timestep)

    def _act(self, timestep):
        if self.tuned or not self.trigger(timestep):
            return
        if self._current is not None and timestep > self._step0:
            torch.cuda.synchronize()
            dt = time.perf_counter() - self._t0
            self.results[self._current] = (timestep - self._step0) / dt
        pending = [b for b in self._grid if b not in self.results]
        if pending:
            self._start(timestep, pending[0])
            return
        best = max(self._grid, key=lambda b: self.results[b])
        if self._round < self.n_rounds:
            self._round += 1
            step = (self._grid[1] - self._grid[0]) / 2.0
            lo = max(self.minimum_buffer, best - step)
            hi = min(self.maximum_buffer, best + step)
            self._grid = sorted(set([lo, best, hi]))
            pending = [b for b in self._grid if b not in self.results]
            if pending:
                self._start(timestep, pending[0])
                return
        self.best_buffer = best
        self.nlist.buffer = float(best)
        self._current = None
