"""Per-type-pair parameter containers with HOOMD's TypeParameter semantics.

``params[("A", "B")] = dict(...)`` validates the keys against the schema of the
potential (the reference's ``TypeParameterDict(..., len_keys=2)``, pair.py),
treats ("A", "B") and ("B", "A") as one entry, and, once the potential is
attached, reads values back from the device object (the reference reads back
through ``asDict()``, so derived fields round-trip exactly as there).
"""

import numbers


def _canon(key):
    a, b = key
    return (a, b) if a <= b else (b, a)


def _expand(key):
    """Accept ("A","B") or a list of such keys or (["A","B"], ["C"])."""
    if isinstance(key, list):
        out = []
        for k in key:
            out.extend(_expand(k))
        return out
    if isinstance(key, tuple) and len(key) == 2:
        a, b = key
        aa = a if isinstance(a, (list, tuple)) else [a]
        bb = b if isinstance(b, (list, tuple)) else [b]
        return [(x, y) for x in aa for y in bb]
    raise KeyError(f"type-pair key expected, got {key!r}")


def _expand1(key):
    """Accept "A" or a list of type names."""
    if isinstance(key, (list, tuple)):
        out = []
        for k in key:
            out.extend(_expand1(k))
        return out
    if isinstance(key, str):
        return [key]
    raise KeyError(f"type key expected, got {key!r}")


class TypeParameterDict:
    """Schema-checked mapping from type pairs (len_keys=2) or single types
    (len_keys=1) to parameter dicts.  A schema entry is a type (required) or
    a default value (optional, of that value's type), as in HOOMD."""

    def __init__(self, len_keys=2, **schema):
        if len_keys not in (1, 2):
            raise ValueError("len_keys must be 1 or 2")
        self._len_keys = len_keys
        self._schema = {k: (v if isinstance(v, type) else type(v)) for k, v in schema.items()}
        self._defaults = {k: v for k, v in schema.items() if not isinstance(v, type)}
        self._data = {}

    def _keys(self, key):
        if self._len_keys == 1:
            return _expand1(key)
        return [_canon(k) for k in _expand(key)]

    def _convert(self, value):
        if not isinstance(value, dict):
            raise TypeError(f"parameters must be a dict, got {type(value).__name__}")
        extra = set(value) - set(self._schema)
        if extra:
            raise KeyError(f"unexpected parameter keys {sorted(extra)}")
        value = {**self._defaults, **value}
        missing = [k for k in self._schema if k not in value]
        if missing:
            raise KeyError(f"missing required parameter keys {missing}")
        out = {}
        for k, typ in self._schema.items():
            v = value[k]
            if typ is bool:
                if not isinstance(v, (bool, numbers.Integral)):
                    raise TypeError(f"{k} must be a bool")
                out[k] = bool(v)
            else:
                if isinstance(v, bool) or not isinstance(v, numbers.Real):
                    raise TypeError(f"{k} must be a number")
                out[k] = float(v)
        return out

    def keys(self):
        return self._data.keys()

    def __contains__(self, key):
        return (key if self._len_keys == 1 else _canon(key)) in self._data

    def __setitem__(self, key, value):
        for k in self._keys(key):
            self._data[k] = self._convert(value)

    def __getitem__(self, key):
        return dict(self._data[key if self._len_keys == 1 else _canon(key)])

    def items(self):
        return self._data.items()


class TypeParameter:
    """Named TypeParameterDict bound to an owner (the pair potential)."""

    def __init__(self, name, type_kind, param_dict):
        self.name = name
        self.type_kind = type_kind
        self.param_dict = param_dict
        self._owner = None

    def __setitem__(self, key, value):
        self.param_dict[key] = value
        if self._owner is not None:
            self._owner._typeparam_changed(self.name, key)

    def __getitem__(self, key):
        if self._owner is not None and self._owner._attached:
            return self._owner._typeparam_readback(self.name, key)
        return self.param_dict[key]

    def __contains__(self, key):
        return key in self.param_dict

    def keys(self):
        return self.param_dict.keys()


class RCutDict:
    """r_cut / r_on per type pair with a default for unset pairs."""

    def __init__(self, default=None):
        self.default = default
        self._data = {}
        self._owner = None

    def __setitem__(self, key, value):
        for k in _expand(key):
            self._data[_canon(k)] = float(value)
        if self._owner is not None:
            self._owner._rcut_changed()

    def __getitem__(self, key):
        k = _canon(key)
        if k in self._data:
            return self._data[k]
        if self.default is None:
            raise KeyError(f"no value for {key!r} and no default")
        return float(self.default)

    def get(self, key, fallback=0.0):
        try:
            return self[key]
        except KeyError:
            return fallback
