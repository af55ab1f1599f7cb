"""Tuners (HOOMD's ``hoomd.tune``): the particle sorter."""

from .. import _azp
from . import trigger as _trigger

torch = _azp.torch


class ParticleSorter:
    """Reorders the particles in HBM along a space-filling curve.

    The role of HOOMD's ``tune.ParticleSorter`` (SFCPackTuner): particles
    close in space end up close in memory, which keeps the neighbor gathers
    of the pair kernels and the integrator's streams local as particles
    diffuse.  The order is the Morton curve over a grid of about ``grid``
    bins along the box's longest axis (coarsened to at most ~4 bins per
    particle), each bin's particles by their current index; the library's
    ``azp_sorter_order`` computes it and ``azp_gather_rows`` applies it to
    every per-particle array.  Tags travel with their particles, so outputs
    read by tag (``forces``, ``energies``, snapshots) do not change; neighbor
    lists rebuild after a sort.

    Args:
        trigger: timesteps on which to sort (an int means every `trigger` steps).
        grid: bins along the longest box axis; None means 256.
    """

    def __init__(self, trigger=200, grid=None):
        self.trigger = _trigger.as_trigger(trigger)
        self.grid = 256 if grid is None else int(grid)
        if not 1 <= self.grid <= 1024:
            raise ValueError("grid must be in [1, 1024]")
        self._sim = None
        self._handle = None
        self._order = None
        self.num_sorts = 0

    def _attach(self, sim):
        if self._handle is None:
            h = _azp.ctypes.c_void_p()
            _azp.call("azp_sorter_create", _azp.ctypes.byref(h))
            self._handle = h
        self._sim = sim

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                _azp.lib().azp_sorter_destroy(h)
            except Exception:
                pass

    def _act(self, timestep):
        if self.trigger(timestep):
            self.sort()

    def sort(self):
        """Sort now (also what the trigger calls)."""
        st = self._sim.state
        if st.N < 2:
            return
        if self._order is None or self._order.shape[0] != st.N:
            self._order = torch.empty(st.N, dtype=torch.int32, device=st.pos.device)
        _azp.call("azp_sorter_order", self._handle, _azp.ctypes.byref(st.particles()), self.grid,
                  _azp.dptr(self._order), _azp.stream_ptr())
        self._sim._reorder(self._order)
        self.num_sorts += 1
