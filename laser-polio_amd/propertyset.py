"""Parameter container with attribute and item access.

Mirrors the subset of ``laser_core.propertyset.PropertySet`` the reference uses
(``pars.x``, ``pars["x"]``, ``"x" in pars``, ``pars <<= other`` to override
existing keys, ``pars += other`` to add keys, ``to_dict()``); laser_core is not
vendored by the reference (pyproject.toml:45), so this is written from how
src/laser_polio/model.py:58-99 and the reference tests call it.
"""

from copy import deepcopy


class PropertySet:
    def __init__(self, *dicts, **kwargs):
        for d in dicts:
            self._merge(d, allow_new=True)
        if kwargs:
            self._merge(kwargs, allow_new=True)

    def _items_of(self, other):
        if isinstance(other, PropertySet):
            return other.to_dict().items()
        return dict(other).items()

    def _merge(self, other, allow_new, allow_existing=True):
        for k, v in self._items_of(other):
            if k in self.__dict__ and not allow_existing:
                raise KeyError(f"PropertySet: key {k!r} already present")
            if k not in self.__dict__ and not allow_new:
                raise KeyError(f"PropertySet: key {k!r} not present (use += to add keys)")
            self.__dict__[k] = v
        return self

    def __ilshift__(self, other):
        return self._merge(other, allow_new=False)

    def __lshift__(self, other):
        return deepcopy(self)._merge(other, allow_new=False)

    def __iadd__(self, other):
        return self._merge(other, allow_new=True)

    def __add__(self, other):
        return deepcopy(self)._merge(other, allow_new=True)

    def __or__(self, other):
        return deepcopy(self)._merge(other, allow_new=True)

    def __getitem__(self, key):
        return self.__dict__[key]

    def __setitem__(self, key, value):
        self.__dict__[key] = value

    def __contains__(self, key):
        return key in self.__dict__

    def __iter__(self):
        return iter(self.__dict__)

    def __len__(self):
        return len(self.__dict__)

    def get(self, key, default=None):
        return self.__dict__.get(key, default)

    def keys(self):
        return self.__dict__.keys()

    def items(self):
        return self.__dict__.items()

    def values(self):
        return self.__dict__.values()

    def to_dict(self):
        return dict(self.__dict__)

    def __repr__(self):
        return f"PropertySet({self.__dict__!r})"
