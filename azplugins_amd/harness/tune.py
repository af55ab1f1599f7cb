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
        grid: bins along the longest box axis.  None (the default) bins at
            the cell size of the attached neighbor lists (the largest
            r_cut + r_buff), the order their cell lists walk in, which
            measured fastest for both the plain and the tiled pair kernels;
            256 when no neighbor list is attached.
    """

    def __init__(self, trigger=200, grid=None):
        self.trigger = _trigger.as_trigger(trigger)
        self.grid = None if grid is None else int(grid)
        if self.grid is not None and not 1 <= self.grid <= 1024:
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

    def _grid(self):
        if self.grid is not None:
            return self.grid
        integ = self._sim.operations.integrator
        r_list = 0.0
        for f in (integ.forces if integ is not None else []):
            nl = getattr(f, "nlist", None)
            if nl is None or getattr(f, "_cpp", None) is None:
                continue
            r_list = max(r_list, max(f._r_cut_matrix(), default=0.0) + nl.buffer)
        if r_list <= 0.0:
            return 256
        L = max(self._sim.state.box.L)
        return int(min(1024, max(1, L // r_list)))

    def sort(self):
        """Sort now (also what the trigger calls)."""
        st = self._sim.state
        if st.N < 2:
            return
        if self._order is None or self._order.shape[0] != st.N:
            self._order = torch.empty(st.N, dtype=torch.int32, device=st.pos.device)
        _azp.call("azp_sorter_order", self._handle, _azp.ctypes.byref(st.particles()), self._grid(),
                  _azp.dptr(self._order), _azp.stream_ptr())
        self._sim._reorder(self._order)
        self.num_sorts += 1
