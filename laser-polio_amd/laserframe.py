"""Struct-of-arrays containers for agents and results.

``LaserFrame`` mirrors the part of laser_core.laserframe.LaserFrame the
reference uses (count, capacity, add_scalar_property, add_array_property,
add_vector_property; model.py:139-172, 547-559).  Columns are numpy arrays on
the host.  While a simulation runs they live on the GPU; an optional sync hook
pulls a column back before it is read on the host and records that it may
have been written, so it is pushed again before the next GPU step.

``PeopleFrame`` adds the node-segmented layout: node n owns the slot range
[seg_begin[n], seg_end[n]) (a multiple of 1024 slots), its initial agents fill
the start of that range and its births are appended at live_end[n].

Snapshots (``PeopleFrame.save_snapshot`` / ``load_snapshot``, the calls of
laser_core's LaserFrame that run_sim.py:388-431 uses) are numpy ``.npz`` archives
(h5py is not available here): per node the count of handed-out slots, every
column's handed-out slots in node order with its dtype and default, optionally
results.R, and the parameters as JSON.
"""

import json

import numpy as np

from .utils import calc_capacity


class LaserFrame:
    def __init__(self, capacity, initial_count=0):
        object.__setattr__(self, "_cols", {})
        object.__setattr__(self, "_defaults", {})
        object.__setattr__(self, "_sync_hook", None)
        object.__setattr__(self, "_capacity", int(capacity))
        object.__setattr__(self, "_count", int(initial_count))

    # -- columns ---------------------------------------------------------------
    def add_scalar_property(self, name, dtype=np.uint32, default=0):
        self._cols[name] = np.full(self._capacity, default, dtype=dtype)
        self._defaults[name] = default

    def add_vector_property(self, name, length, dtype=np.uint32, default=0):
        self._cols[name] = np.full((length, self._capacity), default, dtype=dtype)

    def add_array_property(self, name, shape, dtype=np.uint32, default=0):
        self._cols[name] = np.full(shape, default, dtype=dtype)

    def has(self, name):
        return name in self._cols

    def column_names(self):
        return list(self._cols)

    def raw(self, name):
        """Host array without triggering a device pull (engine use)."""
        return self._cols[name]

    def __getattr__(self, name):
        cols = self.__dict__.get("_cols")
        if cols is not None and name in cols:
            hook = self.__dict__.get("_sync_hook")
            if hook is not None:
                hook(self, name)
            return cols[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        cols = self.__dict__.get("_cols")
        if cols is not None and name in cols:
            hook = self.__dict__.get("_sync_hook")
            if hook is not None:
                hook(self, name)
            arr = cols[name]
            arr[...] = value
            return
        object.__setattr__(self, name, value)

    def __contains__(self, name):
        return name in self._cols

    # -- sizes -----------------------------------------------------------------
    @property
    def capacity(self):
        return self._capacity

    @property
    def count(self):
        return self._count

    def add(self, count):
        if self._count + count > self._capacity:
            raise ValueError(f"LaserFrame.add: capacity exceeded ({self._count} + {count} > {self._capacity})")
        start = self._count
        object.__setattr__(self, "_count", self._count + count)
        return start, self._count


class PeopleFrame(LaserFrame):
    def __init__(self, init_pop, node_capacity, tile=1024, lazy=False):
        """lazy: columns are allocated without being filled (their values are written on
        the GPU and pulled before the host reads them)."""
        init_pop = np.asarray(init_pop, dtype=np.int64)
        cap_n = np.asarray(node_capacity, dtype=np.int64)
        cap_n = ((np.maximum(cap_n, init_pop) + tile - 1) // tile) * tile
        seg_begin = np.zeros(len(init_pop), dtype=np.int64)
        seg_begin[1:] = np.cumsum(cap_n)[:-1]
        super().__init__(int(cap_n.sum()), int(init_pop.sum()))
        object.__setattr__(self, "init_pop", init_pop)
        object.__setattr__(self, "seg_begin", seg_begin)
        object.__setattr__(self, "seg_end", seg_begin + cap_n)
        object.__setattr__(self, "live_end", seg_begin + init_pop)
        object.__setattr__(self, "_init_cum", np.concatenate([[0], np.cumsum(init_pop)]))
        object.__setattr__(self, "lazy", bool(lazy))

    def add_scalar_property(self, name, dtype=np.uint32, default=0):
        if self.lazy:
            self._cols[name] = np.empty(self._capacity, dtype=dtype)
            self._defaults[name] = default
        else:
            super().add_scalar_property(name, dtype=dtype, default=default)

    # -- snapshots -----------------------------------------------------------------
    def save_snapshot(self, path, results_r=None, pars=None):
        """Write every handed-out slot of every column (node order), results.R and the
        parameters to `path` (an .npz archive whatever the file name)."""
        count = (self.live_end - self.seg_begin).astype(np.int64)
        meta = {"format": "laser_polio_amd.snapshot/1", "columns": {},
                "pars": _jsonable(pars) if pars is not None else None}
        arrays = {"node_count": count}
        for name in self.column_names():
            col = getattr(self, name)          # pulls a device-newer column first
            arrays["col_" + name] = np.concatenate([col[b:e] for b, e in zip(self.seg_begin, self.live_end, strict=True)])
            d = self._defaults.get(name, 0)
            meta["columns"][name] = [col.dtype.str, d.item() if isinstance(d, np.generic) else d]
        if results_r is not None:
            arrays["results_R"] = np.asarray(results_r)
        arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
        with open(path, "wb") as fh:
            np.savez(fh, **arrays)

    @classmethod
    def load_snapshot(cls, path, n_ppl=None, cbr=None, nt=None):
        """(people, results_R or None, saved parameters or None) from a snapshot.  Node
        capacities add the births calc_capacity expects of `n_ppl` (per node; default the
        saved counts) over `nt` days at `cbr`, with the 4-sigma margin of a new model."""
        with np.load(path, allow_pickle=False) as z:
            meta = json.loads(bytes(z["meta"]).decode())
            count = z["node_count"].astype(np.int64)
            N = len(count)
            base = count if n_ppl is None else np.broadcast_to(np.atleast_1d(np.asarray(n_ppl)).astype(np.int64), (N,))
            cap = count.copy()
            if cbr is not None and nt is not None:
                c = np.broadcast_to(np.atleast_1d(np.asarray(cbr, dtype=np.float64)), (N,))
                for n in range(N):
                    eb = calc_capacity(int(base[n]), int(nt), float(c[n])) - int(base[n])
                    if eb > 0:
                        cap[n] = int((1 + 4 / np.sqrt(eb)) * (count[n] + eb)) + 16
            people = cls(count, cap)
            off = np.concatenate([[0], np.cumsum(count)])
            for name, (dts, default) in meta["columns"].items():
                src = z["col_" + name]
                people.add_scalar_property(name, dtype=np.dtype(dts), default=default)
                dst = people.raw(name)
                for n in range(N):
                    b = int(people.seg_begin[n])
                    dst[b: b + count[n]] = src[off[n]:off[n + 1]]
            results_r = np.array(z["results_R"]) if "results_R" in z.files else None
        return people, results_r, meta.get("pars")

    @property
    def count(self):
        """Agent slots handed out so far (initial agents + births, dead included)."""
        return int((self.live_end - self.seg_begin).sum())

    @property
    def n_nodes(self):
        return len(self.init_pop)

    def node_slice(self, n):
        return slice(int(self.seg_begin[n]), int(self.live_end[n]))

    def initial_slice(self, n):
        b = int(self.seg_begin[n])
        return slice(b, b + int(self.init_pop[n]))

    def initial_to_slot(self, idx):
        """Map indices 0..sum(init_pop)-1 (initial agents in node order) to slots."""
        idx = np.asarray(idx, dtype=np.int64)
        node = np.searchsorted(self._init_cum, idx, side="right") - 1
        return self.seg_begin[node] + (idx - self._init_cum[node])

    def live_slots(self):
        """All handed-out slot indices (node order)."""
        return np.concatenate([np.arange(b, e, dtype=np.int64) for b, e in zip(self.seg_begin, self.live_end, strict=True)])

    def add(self, count):
        raise NotImplementedError("births are appended per node by VitalDynamics_ABM on the GPU")


def _jsonable(pars):
    """The JSON-representable entries of a parameter set (arrays as lists, dict keys as
    strings); others (distributions, data frames) are stored as their repr."""
    items = pars.items() if hasattr(pars, "items") else dict(pars).items()

    def conv(v):
        if isinstance(v, np.ndarray):
            return v.tolist()
        if isinstance(v, np.generic):
            return v.item()
        if isinstance(v, dict):
            return {str(k): conv(x) for k, x in v.items()}
        if isinstance(v, (list, tuple)):
            return [conv(x) for x in v]
        if v is None or isinstance(v, (bool, int, float, str)):
            return v
        return repr(v)

    return {str(k): conv(v) for k, v in items}
