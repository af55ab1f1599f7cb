"""Neighbor lists (``md.nlist.Cell``) backed by the library's cell list."""

from ... import _azplugins


class NeighborList:
    """Verlet-buffered full neighbor list shared by the pair potentials.

    ``buffer`` is r_buff: the list holds pairs within r_cut(ti,tj) + r_buff
    and is rebuilt when a particle has moved more than r_buff / 2.
    """

    def __init__(self, buffer, exclusions=("bond",), rebuild_check_delay=1, check_dist=True):
        self.buffer = float(buffer)
        self.exclusions = tuple(exclusions)
        self.rebuild_check_delay = int(rebuild_check_delay)
        self.check_dist = bool(check_dist)
        self._cpp = None
        self._sim = None
        self._users = []
        self._last = None

    @property
    def buffer(self):
        return self._buffer

    @buffer.setter
    def buffer(self, value):
        self._buffer = float(value)
        self._pushed_for = None  # push the settings again at the next update

    def _attach(self, sim):
        if self._cpp is None or self._sim is not sim:
            self._cpp = _azplugins.NeighborListGPU(self.buffer)
            self._sim = sim
            self._users = []
            self._last = None
            self._tiled = False

    def _invalidate(self):
        """Particles were reordered: the next update rebuilds."""
        self._force_next = True
        self._last = None

    def _register(self, pair):
        if pair not in self._users:
            self._users.append(pair)

    @property
    def handle(self):
        return self._cpp.handle

    def _r_cut_matrix(self):
        nt = len(self._sim.state.types)
        m = [0.0] * (nt * nt)
        for u in self._users:
            for k, v in enumerate(u._r_cut_matrix()):
                m[k] = max(m[k], v)
        return nt, m

    def _wants_tiles(self):
        """The tiled list serves isotropic pair forces in both forms and DPD."""
        from ... import _azp

        return any(u._cpp is not None
                   and u._cpp.driver in (_azp.DRIVER_PAIR, _azp.DRIVER_CONSERVATIVE, _azp.DRIVER_DPD_THERMO)
                   for u in self._users)

    def _needs_update(self, timestep):
        """Whether this step's update (rebuild check) is still due; pushes the
        settings first when they may have changed (once per run() otherwise)."""
        if self._last == timestep:
            return False
        run_id = getattr(self._sim.state, "_view", None)
        if run_id is None or run_id is not getattr(self, "_pushed_for", None):
            nt, m = self._r_cut_matrix()
            self._cpp.set_r_buff(self.buffer)
            self._cpp.set_r_cut(nt, m)
            tiled = self._wants_tiles()
            if tiled != getattr(self, "_tiled", False):
                self._cpp.set_tiled(tiled)
                self._tiled = tiled
            self._pushed_for = run_id
        return True

    def _take_force_flag(self):
        f = getattr(self, "_force_next", False)
        self._force_next = False
        return f

    def _update(self, timestep):
        if not self._needs_update(timestep):
            return
        self._cpp.update(self._sim.state.particles(), force_rebuild=self._take_force_flag())
        self._last = timestep

    @property
    def num_builds(self):
        return self._cpp.num_builds() if self._cpp is not None else 0


class Cell(NeighborList):
    """Cell-list-accelerated neighbor list (HOOMD's ``md.nlist.Cell``)."""
