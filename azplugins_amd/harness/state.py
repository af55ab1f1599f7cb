"""Box, host snapshot and the device-resident particle state."""

import numpy as np

from .. import _azp

torch = _azp.torch


def permute_rows(t, order, N, scratch):
    """t[k] = t[order[k]] for k < N, in place (through `scratch`, a byte
    tensor of at least N * row bytes); rows of 4, 8, 16 or 32 bytes."""
    rows = t[:N]
    row_bytes = rows[0].numel() * rows.element_size()
    tmp = scratch[:N * row_bytes].view(rows.dtype).view(rows.shape)
    _azp.call("azp_gather_rows", _azp.dptr(rows), _azp.dptr(tmp), _azp.dptr(order), N, row_bytes,
              _azp.stream_ptr())
    rows.copy_(tmp)


class Box:
    """Orthorhombic periodic box (tilt factors must be zero)."""

    def __init__(self, Lx, Ly, Lz, xy=0.0, xz=0.0, yz=0.0):
        if xy or xz or yz:
            raise ValueError("only orthorhombic boxes are supported")
        self.Lx, self.Ly, self.Lz = float(Lx), float(Ly), float(Lz)

    @classmethod
    def cube(cls, L):
        return cls(L, L, L)

    @classmethod
    def from_box(cls, box):
        if isinstance(box, Box):
            return cls(box.Lx, box.Ly, box.Lz)
        b = list(box) + [0.0] * (6 - len(box))
        return cls(*b[:6])

    @property
    def L(self):
        return np.array([self.Lx, self.Ly, self.Lz])

    @property
    def volume(self):
        return self.Lx * self.Ly * self.Lz

    def to_c(self):
        return _azp.Box(self.Lx, self.Ly, self.Lz)

    def __eq__(self, other):
        return isinstance(other, Box) and np.array_equal(self.L, other.L)


class _Configuration:
    def __init__(self):
        self.box = [1.0, 1.0, 1.0, 0.0, 0.0, 0.0]
        self.dimensions = 3
        self.step = 0


class _Particles:
    _fields = {
        "position": (3, np.float64, 0.0),
        "velocity": (3, np.float64, 0.0),
        "mass": (None, np.float64, 1.0),
        "typeid": (None, np.uint32, 0),
        "orientation": (4, np.float64, None),
        "moment_inertia": (3, np.float64, 0.0),
        "angmom": (4, np.float64, 0.0),
        "image": (3, np.int32, 0),
        "diameter": (None, np.float64, 1.0),
        "charge": (None, np.float64, 0.0),
        "body": (None, np.int32, -1),
    }

    def __init__(self):
        self._N = 0
        self.types = ["A"]
        for name, (w, dt, fill) in self._fields.items():
            setattr(self, name, np.zeros((0,) if w is None else (0, w), dtype=dt))

    @property
    def N(self):
        return self._N

    @N.setter
    def N(self, n):
        self._N = int(n)
        for name, (w, dt, fill) in self._fields.items():
            shape = (self._N,) if w is None else (self._N, w)
            if name == "orientation":
                a = np.zeros(shape, dtype=dt)
                a[:, 0] = 1.0
            else:
                a = np.full(shape, fill, dtype=dt)
            setattr(self, name, a)


_GSD_FIELDS = ("position", "velocity", "mass", "typeid", "orientation", "moment_inertia", "angmom", "image",
               "diameter", "charge", "body")


class Snapshot:
    """Host-side system description (HOOMD's Snapshot layout)."""

    def __init__(self, communicator=None):
        self.configuration = _Configuration()
        self.particles = _Particles()
        self.communicator = communicator if communicator is not None else _Comm()

    @classmethod
    def from_gsd_frame(cls, frame, communicator=None):
        """A snapshot holding one frame of a GSD file (hoomd schema)."""
        frame.validate()
        snap = cls(communicator)
        c, p = frame.configuration, frame.particles
        snap.configuration.box = [float(v) for v in c.box]
        snap.configuration.dimensions = int(c.dimensions)
        snap.configuration.step = int(c.step)
        snap.particles.types = list(p.types)
        snap.particles.N = p.N
        sp = snap.particles
        for name in _GSD_FIELDS:
            getattr(sp, name)[:] = getattr(p, name)
        return snap

    def _gsd_frame(self, step=0):
        """This snapshot as a GSD frame (hoomd schema, 32-bit fields)."""
        from .gsd import Frame

        fr = Frame()
        c, p = fr.configuration, fr.particles
        c.step = int(step)
        c.dimensions = int(self.configuration.dimensions)
        c.box = self.configuration.box
        sp = self.particles
        p.N = sp.N
        p.types = list(sp.types)
        for name in _GSD_FIELDS:
            setattr(p, name, getattr(sp, name))
        return fr.validate()


class _Comm:
    rank = 0
    num_ranks = 1


class State:
    """Particle data in HBM (torch tensors used as device allocations).

    pos          (N, 4) float64  x, y, z, type
    vel          (N, 4) float64  vx, vy, vz, mass
    tag          (N,)   int32    particle tags (uint32 semantics)
    orientation  (N, 4) float64  quaternion (s, x, y, z)
    angmom       (N, 4) float64  space-frame angular momentum (Lx, Ly, Lz, 0)
    inertia      (N,)   float64  isotropic moment of inertia
    image        (N, 4) int32    periodic images (x, y, z, 0)
    accel        (N, 4) float64  accelerations from the last second half step
    """

    def __init__(self, simulation, snapshot):
        self._sim = simulation
        p = snapshot.particles
        self.types = list(p.types)
        self.box = Box.from_box(snapshot.configuration.box)
        N = p.N
        self.N = N
        dev = torch.device("cuda")
        pos = np.zeros((N, 4))
        pos[:, :3] = p.position
        pos[:, 3] = p.typeid
        self._wrap_host(pos)
        vel = np.zeros((N, 4))
        vel[:, :3] = p.velocity
        vel[:, 3] = p.mass
        mi = np.asarray(p.moment_inertia, dtype=np.float64).reshape(N, 3)
        if N and not np.allclose(mi, mi[:, :1]):
            raise ValueError("only isotropic moments of inertia (spherical tops) are supported")
        ang = np.zeros((N, 4))
        ang[:, :3] = np.asarray(p.angmom, dtype=np.float64).reshape(N, 4)[:, 1:]
        img = np.zeros((N, 4), dtype=np.int32)
        img[:, :3] = p.image
        self.pos = torch.tensor(pos, dtype=torch.float64, device=dev)
        self.vel = torch.tensor(vel, dtype=torch.float64, device=dev)
        self.tag = torch.arange(N, dtype=torch.int32, device=dev)
        self.orientation = torch.tensor(np.asarray(p.orientation, dtype=np.float64), device=dev).reshape(N, 4)
        self.angmom = torch.tensor(ang, dtype=torch.float64, device=dev)
        self.inertia = torch.tensor(mi[:, 0].copy(), dtype=torch.float64, device=dev)
        self.image = torch.tensor(img, device=dev)
        self.accel = torch.zeros((N, 4), dtype=torch.float64, device=dev)
        self.pos_fixed = None
        # per-tag fields the force path does not read, kept for snapshots
        self._diameter = np.array(p.diameter, dtype=np.float64).reshape(N)
        self._charge = np.array(p.charge, dtype=np.float64).reshape(N)
        self._body = np.array(p.body, dtype=np.int32).reshape(N)

    # ---- particle order (a sorter moves particles; tags travel with them) ----
    def _scratch(self, nbytes):
        sc = getattr(self, "_scratch_buf", None)
        if sc is None or sc.numel() < nbytes:
            sc = torch.empty(nbytes, dtype=torch.uint8, device=self.pos.device)
            self._scratch_buf = sc
        return sc

    def _reorder(self, order):
        """Moves particle order[k] to index k in every per-particle array."""
        N = self.N
        sc = self._scratch(32 * N)
        for t in (self.pos, self.vel, self.tag, self.orientation, self.angmom, self.inertia, self.image,
                  self.accel):
            permute_rows(t, order, N, sc)
        self._tag_host = None
        self._sorted = True
        return sc

    def _tags(self):
        """Host copy of the tags, or None while index == tag."""
        if not getattr(self, "_sorted", False):
            return None
        if getattr(self, "_tag_host", None) is None:
            self._tag_host = self.tag.cpu().numpy().astype(np.int64)
        return self._tag_host

    def _by_tag(self, a):
        """Rows of a per-index host array rearranged by tag."""
        tags = self._tags()
        if tags is None:
            return a
        out = np.empty_like(a)
        out[tags] = a
        return out

    def _wrap_host(self, pos):
        L = self.box.L
        for a in range(3):
            pos[:, a] -= L[a] * np.floor(pos[:, a] / L[a] + 0.5)

    @property
    def particle_types(self):
        return self.types

    def _freeze_view(self):
        """Inside Simulation.run: reuse one azp_particles view."""
        self._view = None
        self._view = self.particles()

    def _thaw_view(self):
        self._view = None

    def particles(self):
        """azp_particles view of the device arrays (valid while tensors live)."""
        if getattr(self, "_view", None) is not None:
            return self._view
        s = _azp.Particles()
        s.N = self.N
        s.box = self.box.to_c()
        s.d_pos = self.pos.data_ptr() if self.N else None
        s.d_pos_fixed = self.pos_fixed.data_ptr() if self.pos_fixed is not None else None
        s.d_vel = self.vel.data_ptr() if self.N else None
        s.d_tag = self.tag.data_ptr() if self.N else None
        s.d_orientation = self.orientation.data_ptr() if self.N else None
        return s

    def thermalize_particle_momenta(self, filter, kT):
        """Maxwell-Boltzmann velocities at kT with zero total momentum."""
        rng = np.random.default_rng(int(self._sim.seed) * 7919 + int(self._sim.timestep))
        m = self.vel[:, 3].cpu().numpy()
        v = rng.normal(size=(self.N, 3)) * np.sqrt(kT / m)[:, None]
        if self.N > 1:
            v -= (m[:, None] * v).sum(axis=0) / m.sum()
        vel = self.vel.cpu().numpy()
        vel[:, :3] = v
        self.vel.copy_(torch.tensor(vel, device=self.vel.device))
        inertia = self.inertia.cpu().numpy()
        if np.any(inertia > 0):
            L = rng.normal(size=(self.N, 3)) * np.sqrt(kT * inertia)[:, None]
            ang = np.zeros((self.N, 4))
            ang[:, :3] = np.where(inertia[:, None] > 0, L, 0.0)
            self.angmom.copy_(torch.tensor(ang, device=self.angmom.device))

    def get_snapshot(self):
        snap = Snapshot()
        snap.configuration.box = [self.box.Lx, self.box.Ly, self.box.Lz, 0.0, 0.0, 0.0]
        snap.particles.types = list(self.types)
        snap.particles.N = self.N
        bt = self._by_tag
        pos = bt(self.pos.cpu().numpy())
        vel = bt(self.vel.cpu().numpy())
        snap.particles.position[:] = pos[:, :3]
        snap.particles.typeid[:] = pos[:, 3].astype(np.uint32)
        snap.particles.velocity[:] = vel[:, :3]
        snap.particles.mass[:] = vel[:, 3]
        snap.particles.orientation[:] = bt(self.orientation.cpu().numpy())
        snap.particles.image[:] = bt(self.image.cpu().numpy())[:, :3]
        ang = np.zeros((self.N, 4))
        ang[:, 1:] = bt(self.angmom.cpu().numpy())[:, :3]
        snap.particles.angmom[:] = ang
        snap.particles.moment_inertia[:] = bt(self.inertia.cpu().numpy())[:, None]
        snap.particles.diameter[:] = self._diameter
        snap.particles.charge[:] = self._charge
        snap.particles.body[:] = self._body
        return snap
