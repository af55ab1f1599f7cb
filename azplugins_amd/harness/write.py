"""Writers: operations that save the state (HOOMD's ``hoomd.write``).

``GSD`` appends frames of the ``hoomd`` schema to a GSD file.  Writers act
at the end of a step, on the step count that step ends at, with the
state's per-particle arrays complete (no half step pending), in tag order.
Fields are stored as the schema's 32-bit types, so a frame read back holds
the state rounded to float32.
"""

from . import filter as _filter
from . import gsd as _gsd
from .trigger import as_trigger

# chunk names by HOOMD's dynamic categories
_CATEGORIES = {
    "property": ("configuration/box", "particles/N", "particles/position", "particles/orientation"),
    "momentum": ("particles/velocity", "particles/angmom", "particles/image"),
    "attribute": ("particles/types", "particles/typeid", "particles/mass", "particles/charge",
                  "particles/diameter", "particles/body", "particles/moment_inertia"),
    "topology": (),
}


def _dynamic_chunks(dynamic):
    names = set()
    for d in dynamic:
        if d in _CATEGORIES:
            names.update(_CATEGORIES[d])
        elif d in _gsd.SCHEMA or d == "particles/types":
            names.add(d)
        else:
            raise ValueError(f"unknown dynamic quantity {d!r}")
    # position and orientation are always written, as in HOOMD
    names.update(_CATEGORIES["property"])
    return names


class GSD:
    """Writes the state to `filename` on the steps `trigger` selects.

    mode: 'ab' appends (creating the file), 'wb' truncates, 'xb' refuses an
    existing file.  dynamic: the categories ('property', 'momentum',
    'attribute') or chunk names written every frame; everything else is
    written in the file's first frame only and read back from it.
    Only ``filter.All()`` is supported."""

    def __init__(self, trigger, filename, filter=None, mode="ab", truncate=False, dynamic=None):
        self.trigger = as_trigger(trigger)
        self.filename = str(filename)
        self.filter = filter if filter is not None else _filter.All()
        if not isinstance(self.filter, _filter.All):
            raise ValueError("only filter.All() is supported")
        if mode not in ("ab", "wb", "xb"):
            raise ValueError("mode must be 'ab', 'wb' or 'xb'")
        self.mode = mode
        self.truncate = bool(truncate)
        self.dynamic = list(dynamic) if dynamic is not None else ["property"]
        self._only = _dynamic_chunks(self.dynamic)
        self._traj = None
        self._opened = False
        self._sim = None

    def _attach(self, sim):
        self._sim = sim
        self._open()

    def _open(self):
        # the mode applies to the first opening; a closed writer reopens to append
        if self._traj is None:
            self._traj = _gsd.open(self.filename, self.mode if not self._opened else "ab")
            self._opened = True

    def _act(self, timestep):
        self._open()
        if self.truncate and len(self._traj):
            self._traj.close()
            self._traj = _gsd.open(self.filename, "wb")
        frame = self._sim.state.get_snapshot()._gsd_frame(step=timestep)
        self._traj.append(frame, only=self._only)

    def flush(self):
        if self._traj is not None:
            self._traj.flush()

    def close(self):
        if self._traj is not None:
            self._traj.close()
            self._traj = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def write(state, filename, filter=None, mode="wb"):
        """Writes the current state as one frame."""
        if filter is not None and not isinstance(filter, _filter.All):
            raise ValueError("only filter.All() is supported")
        with _gsd.open(filename, mode) as traj:
            traj.append(state.get_snapshot()._gsd_frame(step=state._sim.timestep))
