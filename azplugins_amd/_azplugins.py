"""Stand-in for the reference's pybind11 module ``_azplugins`` (pair part).

The reference registers one C++ class per evaluator and device
(``module.cc:142-150,171-179``): ``PotentialPair<Name>``,
``PotentialPair<Name>GPU``, ``PotentialPairDPDThermo<Name>[GPU]``,
``PotentialPairConservative<Name>`` and ``AnisoPotentialPair<Name>[GPU]``.  The
Python classes in ``pair.py`` select one by ``_cpp_class_name`` (+ "GPU").
Here every name resolves to the same gfx950 implementation behind the C ABI
(``azp_pair_*``); there is no CPU runtime path.
"""

import ctypes

from . import _azp


class PotentialPairGPUBase:
    """Owns one ``azp_pair`` handle (a PotentialPair*GPU object)."""

    cpp_class_name = None

    def __init__(self, ntypes):
        self._h = ctypes.c_void_p()
        _azp.call("azp_pair_create", ctypes.byref(self._h), self.cpp_class_name.encode(), int(ntypes))
        ev = ctypes.c_int()
        drv = ctypes.c_int()
        _azp.call("azp_class_lookup", self.cpp_class_name.encode(), ctypes.byref(ev), ctypes.byref(drv))
        self.evaluator = ev.value
        self.driver = drv.value
        self.ntypes = int(ntypes)
        self.nvalues = _azp.lib().azp_param_nvalues(self.evaluator)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _azp.lib().azp_pair_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def takes_step(self):
        """Whether update_compute_step serves this force: the isotropic
        drivers (DPD and the anisotropic driver integrate otherwise)."""
        return self.driver in (0, 2)

    def setParams(self, ti, tj, values):
        arr = _azp.dbl_array(values)
        _azp.call("azp_pair_set_params", self._h, ti, tj, arr)

    def getParams(self, ti, tj):
        arr = (ctypes.c_double * self.nvalues)()
        _azp.call("azp_pair_get_params", self._h, ti, tj, arr)
        return list(arr)

    def setRCut(self, ti, tj, r_cut):
        _azp.call("azp_pair_set_r_cut", self._h, ti, tj, float(r_cut))

    def setROn(self, ti, tj, r_on):
        _azp.call("azp_pair_set_r_on", self._h, ti, tj, float(r_on))

    def setShiftMode(self, mode):
        _azp.call("azp_pair_set_shift_mode", self._h, _azp.SHIFT_MODES[mode])

    def setPrecision(self, precision):
        _azp.call("azp_pair_set_precision", self._h, _azp.PRECISIONS[precision])

    def setThermostat(self, seed, dt, kT):
        _azp.call("azp_pair_set_thermostat", self._h, int(seed) & 0xFFFF, float(dt), float(kT))

    def getRCutMatrix(self):
        arr = (ctypes.c_double * (self.ntypes * self.ntypes))()
        _azp.call("azp_pair_get_r_cut", self._h, arr)
        return list(arr)

    def update_compute(self, nlist_handle, particles, outputs, timestep, compute_virial=True, force_rebuild=False,
                       stream=None):
        """azp_pair_update_compute: the neighbor list's update, then the force."""
        rebuilt = ctypes.c_int()
        _azp.call("azp_pair_update_compute", self._h, nlist_handle, ctypes.byref(particles), ctypes.byref(outputs),
                  int(timestep), 1 if compute_virial else 0, 1 if force_rebuild else 0, ctypes.byref(rebuilt),
                  _azp.stream_ptr(stream))
        return bool(rebuilt.value)

    def update_compute_step(self, nlist_handle, particles, outputs, timestep, step, compute_virial=True,
                            force_rebuild=False, stream=None):
        """azp_pair_update_compute_step: update, force, then the MD step
        `step` (an _azp.StepArgs) with the next update's rebuild check."""
        rebuilt = ctypes.c_int()
        _azp.call("azp_pair_update_compute_step", self._h, nlist_handle, ctypes.byref(particles),
                  ctypes.byref(outputs), int(timestep), 1 if compute_virial else 0, 1 if force_rebuild else 0,
                  ctypes.byref(rebuilt), ctypes.byref(step), _azp.stream_ptr(stream))
        return bool(rebuilt.value)

    def compute(self, nlist_handle, particles, outputs, timestep, compute_virial=True, stream=None):
        _azp.call(
            "azp_pair_compute",
            self._h,
            nlist_handle,
            ctypes.byref(particles),
            ctypes.byref(outputs),
            int(timestep),
            1 if compute_virial else 0,
            _azp.stream_ptr(stream),
        )


_NAMES = (
    "PotentialPairColloid",
    "PotentialPairExpandedYukawa",
    "PotentialPairHertz",
    "PotentialPairPerturbedLennardJones",
    "PotentialPairDPDThermoGeneralWeight",
    "PotentialPairConservativeGeneralWeight",
    "AnisoPotentialPairTwoPatchMorse",
)

for _n in _NAMES:
    for _suffix in ("", "GPU"):
        _cls = type(_n + _suffix, (PotentialPairGPUBase,), {"cpp_class_name": _n + _suffix})
        globals()[_n + _suffix] = _cls
del _n, _suffix, _cls


class WallsPotentialGPUBase:
    """Owns one ``azp_wall`` handle (the reference's WallsPotential<Name>[GPU],
    export_PotentialExternalWall.cc.inc)."""

    cpp_class_name = None

    def __init__(self, ntypes):
        self._h = ctypes.c_void_p()
        _azp.call("azp_wall_create", ctypes.byref(self._h), self.cpp_class_name.encode(), int(ntypes))
        self.ntypes = int(ntypes)
        self.nvalues = _azp.lib().azp_wall_param_nvalues(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _azp.lib().azp_wall_destroy(h)
            except Exception:
                pass
            self._h = None

    def setParams(self, t, values):
        _azp.call("azp_wall_set_params", self._h, int(t), _azp.dbl_array(values))

    def getParams(self, t):
        arr = (ctypes.c_double * self.nvalues)()
        _azp.call("azp_wall_get_params", self._h, int(t), arr)
        return list(arr)

    def setGeometry(self, spheres, cylinders, planes):
        """Rows as in include/azp.h (sphere 6, cylinder 9, plane 7 values)."""
        def flat(rows, width):
            vals = [float(x) for r in rows for x in r]
            assert len(vals) == width * len(rows)
            return len(rows), (_azp.dbl_array(vals) if vals else None)

        ns, sp = flat(spheres, 6)
        nc, cy = flat(cylinders, 9)
        np_, pl = flat(planes, 7)
        _azp.call("azp_wall_set_geometry", self._h, ns, sp, nc, cy, np_, pl)

    def compute(self, particles, outputs, compute_virial=True, stream=None):
        _azp.call("azp_wall_compute", self._h, ctypes.byref(particles), ctypes.byref(outputs),
                  1 if compute_virial else 0, _azp.stream_ptr(stream))

    def check(self, stream=None):
        """Waits for the stream and raises the type error a compute flagged."""
        _azp.call("azp_wall_check", self._h, _azp.stream_ptr(stream))


for _n in ("WallsPotentialLJ93", "WallsPotentialColloid"):
    for _suffix in ("", "GPU"):
        globals()[_n + _suffix] = type(_n + _suffix, (WallsPotentialGPUBase,), {"cpp_class_name": _n + _suffix})
del _n, _suffix


class NeighborListGPU:
    """Owns one ``azp_nlist`` handle (cell list + full neighbor list)."""

    def __init__(self, r_buff=0.4):
        self._h = ctypes.c_void_p()
        _azp.call("azp_nlist_create", ctypes.byref(self._h))
        self.set_r_buff(r_buff)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _azp.lib().azp_nlist_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def set_r_buff(self, r_buff):
        _azp.call("azp_nlist_set_r_buff", self._h, float(r_buff))

    def set_r_cut(self, ntypes, r_cut_matrix):
        arr = _azp.dbl_array(r_cut_matrix)
        _azp.call("azp_nlist_set_r_cut", self._h, int(ntypes), arr)

    def update(self, particles, force_rebuild=False, stream=None):
        rebuilt = ctypes.c_int(0)
        _azp.call(
            "azp_nlist_update",
            self._h,
            ctypes.byref(particles),
            1 if force_rebuild else 0,
            ctypes.byref(rebuilt),
            _azp.stream_ptr(stream),
        )
        return bool(rebuilt.value)

    def get(self):
        nn = ctypes.c_void_p()
        nl = ctypes.c_void_p()
        pitch = ctypes.c_uint32()
        mx = ctypes.c_uint32()
        _azp.call("azp_nlist_get", self._h, ctypes.byref(nn), ctypes.byref(nl), ctypes.byref(pitch), ctypes.byref(mx))
        return nn.value, nl.value, pitch.value, mx.value

    def set_tiled(self, mode=True):
        """Also build the tiled list (16-bit entries per block of particles) for the mixed form.

        mode: False / 0 off, True / 1 blocks of 512 else 256, or 256 / 512 for that size only.
        """
        m = int(mode) if not isinstance(mode, bool) else (1 if mode else 0)
        _azp.call("azp_nlist_set_tiled", self._h, m)

    def get_tiles(self):
        """(d_tile_count, d_tile_jset, d_tile_list, tile_size, tile_slots); tile_size 0 when not available."""
        c = ctypes.c_void_p()
        j = ctypes.c_void_p()
        t = ctypes.c_void_p()
        size = ctypes.c_uint32()
        slots = ctypes.c_uint32()
        _azp.call("azp_nlist_get_tiles", self._h, ctypes.byref(c), ctypes.byref(j), ctypes.byref(t),
                  ctypes.byref(size), ctypes.byref(slots))
        return c.value, j.value, t.value, size.value, slots.value

    def get_tile_order(self):
        """(d_tile_order, d_tile_nn) for slot-ordered tiles, (None, None) otherwise."""
        o = ctypes.c_void_p()
        n = ctypes.c_void_p()
        _azp.call("azp_nlist_get_tile_order", self._h, ctypes.byref(o), ctypes.byref(n))
        return o.value, n.value

    def walk_kind(self):
        """1 when the last build wrote the tiled list from the per-block walk, 0 for the plain walk."""
        k = ctypes.c_int()
        _azp.call("azp_nlist_walk_kind", self._h, ctypes.byref(k))
        return k.value

    def tile_stats(self):
        """(blocks, blocks walked in FP32 offsets) of the last per-block walk."""
        n = ctypes.c_uint32()
        r = ctypes.c_uint32()
        _azp.call("azp_nlist_tile_stats", self._h, ctypes.byref(n), ctypes.byref(r))
        return n.value, r.value

    def download(self, N):
        """(n_neigh[N], list[max_neighbors, pitch]) as numpy arrays.  They view
        pinned host memory, so each array arrives in one DMA copy."""
        import numpy as np

        torch = _azp.torch
        _, _, pitch, mx = self.get()
        nn_t = torch.empty(max(N, 1), dtype=torch.int32, pin_memory=True)
        lst_t = torch.empty(max(mx * pitch, 1), dtype=torch.int32, pin_memory=True)
        _azp.call("azp_nlist_download", self._h, int(N), nn_t.data_ptr(), lst_t.data_ptr(), lst_t.numel())
        nn = nn_t.numpy().view(np.uint32)
        lst = lst_t.numpy().view(np.uint32)
        return nn[:N], lst[: mx * pitch].reshape(mx, pitch) if pitch else lst[:0].reshape(0, 0)

    def num_builds(self):
        n = ctypes.c_uint64()
        _azp.call("azp_nlist_num_builds", self._h, ctypes.byref(n))
        return n.value
