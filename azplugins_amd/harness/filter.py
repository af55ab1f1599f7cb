"""Particle filters (only ``All`` is needed by the pair path)."""


class ParticleFilter:
    def indices(self, N):
        raise NotImplementedError


class All(ParticleFilter):
    """Selects every particle."""

    def __eq__(self, other):
        return isinstance(other, All)

    def __hash__(self):
        return hash("All")

    def indices(self, N):
        return slice(0, N)
