"""Triggers: on which timesteps an operation acts (HOOMD's ``hoomd.trigger``)."""


class Trigger:
    def __call__(self, timestep):
        raise NotImplementedError


class Periodic(Trigger):
    """Acts when (timestep - phase) is a non-negative multiple of period."""

    def __init__(self, period, phase=0):
        if int(period) < 1:
            raise ValueError("period must be at least 1")
        self.period = int(period)
        self.phase = int(phase)

    def __call__(self, timestep):
        d = int(timestep) - self.phase
        return d >= 0 and d % self.period == 0


def as_trigger(value):
    """An int means Periodic(int), as in HOOMD."""
    if isinstance(value, Trigger):
        return value
    return Periodic(int(value))
