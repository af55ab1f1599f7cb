"""Time-dependent scalars (``Constant`` and ``Ramp``)."""


class Variant:
    def __call__(self, timestep):
        raise NotImplementedError


class Constant(Variant):
    def __init__(self, value):
        self.value = float(value)

    def __call__(self, timestep):
        return self.value


class Ramp(Variant):
    """Linear ramp from A at t_start to B at t_start + t_ramp."""

    def __init__(self, A, B, t_start, t_ramp):
        self.A, self.B = float(A), float(B)
        self.t_start, self.t_ramp = int(t_start), int(t_ramp)

    def __call__(self, timestep):
        if timestep < self.t_start:
            return self.A
        if timestep >= self.t_start + self.t_ramp:
            return self.B
        f = (timestep - self.t_start) / self.t_ramp
        return self.A + f * (self.B - self.A)


def as_variant(v):
    return v if isinstance(v, Variant) else Constant(v)
