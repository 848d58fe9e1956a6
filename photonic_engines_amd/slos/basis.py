"""Mask-pruned Fock basis served by lexicographic rank / unrank — no stored states or edges.

Drop-in for the layer object of photonic_engines/slos/basis.py:92-301 (`FockLayers`,
`get_layers`, `clear_layers_cache`). A k-photon state is the sorted tuple of its photon
positions; layer k lists them in lexicographic order, exactly the order FockLayers builds
(basis.py:120-148). Here that order is a closed-form bijection (the counting tables of the
native plan), so a layer costs nothing until it is read, and `states[k]` is produced on demand.

What is served lazily, and how far:
  size(k), output_size()              always (closed form, 64-bit)
  rank / unrank / index_of / state_range   always (any range of any layer)
  states[k], layers[k], final_states, final_occupations(), prob_norms(k), output_prob_norms(),
  edges[k], edges_by_mode(k)          materialised on first access, only for layers of at most
                                      MATERIALIZE_LIMIT states (MemoryError beyond, with the
                                      ranged alternative in the message): at n=11, m=32 an
                                      [S, m] int64 occupation array would be 1.1 TB.
"""
from __future__ import annotations

from collections.abc import Sequence
from functools import cache

import numpy as np

from ..masks import Mask

MATERIALIZE_LIMIT = 1 << 25


def mode_dtype(m: int):
    """int8 positions up to 127 modes, int16 beyond (basis.py:85-89)."""
    if m > 32767:
        raise ValueError(f"m={m}: at most 32767 modes")
    return np.int8 if m <= 127 else np.int16


def positions_to_occupation(t: Sequence[int], m: int) -> tuple[int, ...]:
    """Sorted position tuple -> occupation vector (basis.py:60-65)."""
    occ = [0] * m
    for p in t:
        occ[int(p)] += 1
    return tuple(occ)


class _TupleView:
    """Lazy sequence of position tuples of one layer (basis.py:68-82)."""

    def __init__(self, layers: "RankedLayers", k: int):
        self._L, self._k = layers, k

    def __len__(self):
        return self._L.size(self._k)

    def __getitem__(self, i):
        S = len(self)
        i = int(i)
        if i < 0:
            i += S
        if not 0 <= i < S:
            raise IndexError(i)
        return tuple(int(x) for x in self._L.unrank(self._k, [i])[0])

    def __iter__(self):
        for start in range(0, len(self), 1 << 16):
            for row in self._L.state_range(self._k, start, min(len(self), start + (1 << 16))).tolist():
                yield tuple(row)


class _States:
    """`states[k]`: the [S_k, k] position array of layer k, materialised on first access."""

    def __init__(self, layers: "RankedLayers"):
        self._L = layers
        self._cache: dict[int, np.ndarray] = {}

    def __len__(self):
        return self._L.n + 1

    def __getitem__(self, k: int) -> np.ndarray:
        k = int(k)
        if k < 0:
            k += self._L.n + 1
        if k not in self._cache:
            self._L._check_materialize(k, "states[k]")
            self._cache[k] = self._L.state_range(k, 0, self._L.size(k))
        return self._cache[k]

    def __iter__(self):
        for k in range(len(self)):
            yield self[k]


class RankedLayers:
    """Layers 0..n of the mask-pruned Fock basis (see the module docstring)."""

    def __init__(self, n: int, m: int, mask: str = Mask.FOCK.value):
        from ._plan import Plan
        self.n, self.m, self.mask = int(n), int(m), Mask(mask).value
        self.mode_dtype = mode_dtype(m)
        self._plan = Plan(self.n, self.m, self.mask, [0] * self.n) if self.n > 0 else None
        self.states = _States(self)
        self.layers = [_TupleView(self, k) for k in range(self.n + 1)]
        self._edges_cache: dict[int, tuple] = {}

    # ---- sizes ------------------------------------------------------------------------------
    def size(self, k: int) -> int:
        """Number of states in layer k (basis.py:231-233)."""
        if k == 0:
            return 1
        return self._plan.sizes[k]

    def output_size(self) -> int:
        """Number of states in the final layer (basis.py:235-237)."""
        return self.size(self.n)

    # ---- rank / unrank ------------------------------------------------------------------------
    def rank(self, k: int, positions) -> np.ndarray:
        """Sorted positions [count, k] -> lex indices (uint64)."""
        if k == 0:
            return np.zeros(len(positions), dtype=np.uint64)
        return self._plan.rank(k, positions)

    def unrank(self, k: int, idx) -> np.ndarray:
        """Lex indices -> sorted positions [count, k] (mode dtype)."""
        if k == 0:
            return np.zeros((len(idx), 0), dtype=self.mode_dtype)
        return self._plan.unrank(k, idx).astype(self.mode_dtype)

    def state_range(self, k: int, start: int, stop: int) -> np.ndarray:
        """Positions of states start .. stop-1 of layer k, [stop-start, k]."""
        return self.unrank(k, np.arange(start, stop, dtype=np.uint64))

    def index_of(self, occupation) -> int:
        """Lex index in layer sum(occupation) of an occupation vector (replaces the
        final_states dict of adapters/perceval/backends.py:105-111)."""
        occ = [int(c) for c in occupation]
        if len(occ) != self.m:
            raise ValueError(f"occupation has {len(occ)} modes, expected {self.m}")
        pos = [j for j, c in enumerate(occ) for _ in range(c)]
        return int(self.rank(len(pos), [pos])[0])

    # ---- materialised views (bounded) -------------------------------------------------------
    def _check_materialize(self, k: int, what: str) -> None:
        if self.size(k) > MATERIALIZE_LIMIT:
            raise MemoryError(f"{what}: layer {k} has {self.size(k):,} states (> {MATERIALIZE_LIMIT:,}); "
                              f"use state_range / unrank / rank / index_of on ranges instead")

    @property
    def final_states(self) -> list[tuple[int, ...]]:
        """Output basis as occupation tuples in layer order (basis.py:239-242)."""
        return [tuple(r) for r in self.final_occupations().tolist()]

    def final_occupations(self, dtype=np.int64) -> np.ndarray:
        """Output basis as an [S, m] occupation array (basis.py:244-252)."""
        return self.occupations(self.n, dtype=dtype)

    def occupations(self, k: int, dtype=np.int64) -> np.ndarray:
        """Layer k as an [S_k, m] occupation array."""
        t = self.states[k]
        S = t.shape[0]
        occ = np.zeros((S, self.m), dtype=dtype)
        ar = np.arange(S)
        for j in range(k):
            np.add.at(occ, (ar, t[:, j].astype(np.int64)), 1)
        return occ

    def output_prob_norms(self) -> np.ndarray:
        """prod_j t_j! per final state, int64 (basis.py:254-258)."""
        return self.prob_norms(self.n)

    def prob_norms(self, k: int) -> np.ndarray:
        """prod_j t_j! per state of layer k, int64 (basis.py:260-272)."""
        if k == 0:
            return np.ones(1, dtype=np.int64)
        self._check_materialize(k, "prob_norms")
        return self._plan.prob_norms(k).astype(np.int64)

    def prob_norms_range(self, k: int, start: int, stop: int) -> np.ndarray:
        """prod_j t_j! of states start .. stop-1 of layer k (float64, any layer size)."""
        return self._plan.prob_norms(k, start, stop - start)

    # ---- edges (compatibility; the engine never stores them) --------------------------------
    @property
    def edges(self):
        return _EdgeList(self)

    def edge_arrays(self, k: int):
        """(parent_pos, child_pos, mode, factor) of the k -> k+1 raising edges, child-major: one
        edge per (child, distinct mode), factor = sqrt(count of the mode in the child)
        (basis.py:160-193). Built from ranks; bounded by MATERIALIZE_LIMIT."""
        if k not in self._edges_cache:
            ch = self.states[k + 1].astype(np.int64)
            S1, k1 = ch.shape
            first = np.ones((S1, k1), dtype=bool)
            if k1 > 1:
                first[:, 1:] = ch[:, 1:] != ch[:, :-1]
            rows, cols = np.nonzero(first)
            E = rows.shape[0]
            parent = np.zeros(E, dtype=np.int64)
            if k:
                keep = np.array([[j for j in range(k1) if j != c] for c in range(k1)], dtype=np.int64)
                par_pos = ch[rows[:, None], keep[cols]]
                parent = self.rank(k, par_pos).astype(np.int64)
            mode = ch[rows, cols].astype(self.mode_dtype)
            factor = np.sqrt((ch[rows] == ch[rows, cols][:, None]).sum(axis=1)).astype(np.float32)
            idx_dtype = np.int32 if max(S1, self.size(k)) < 2 ** 31 else np.int64
            self._edges_cache[k] = (parent.astype(idx_dtype), rows.astype(idx_dtype), mode, factor)
        return self._edges_cache[k]

    def edges_by_mode(self, k: int):
        """The k -> k+1 edges grouped by raised mode: (parent_pos, child_pos, offsets[m+1])
        (basis.py:216-228)."""
        parent, child, mode, _ = self.edge_arrays(k)
        order = np.argsort(mode, kind="stable")
        offsets = np.searchsorted(mode[order], np.arange(self.m + 1, dtype=mode.dtype)).astype(np.int64)
        return parent[order], child[order], offsets


class _EdgeList:
    def __init__(self, layers: RankedLayers):
        self._L = layers

    def __len__(self):
        return self._L.n

    def __getitem__(self, k):
        return self._L.edge_arrays(int(k))


FockLayers = RankedLayers


def get_layers(n: int, m: int, mask: str = Mask.FOCK.value, device: str = "cpu", backend: str = "auto"):
    """Cached basis (basis.py:275-292). The layers are host-side closed forms whatever the device,
    so `device` and `backend` only canonicalise the cache key as in the reference."""
    return _get_layers_cached(int(n), int(m), Mask(mask).value)


@cache
def _get_layers_cached(n: int, m: int, mask: str) -> RankedLayers:
    return RankedLayers(n, m, mask)


def clear_layers_cache() -> None:
    """Drop every cached basis (basis.py:295-301)."""
    _get_layers_cached.cache_clear()


__all__ = ["FockLayers", "RankedLayers", "clear_layers_cache", "get_layers", "mode_dtype",
           "positions_to_occupation", "MATERIALIZE_LIMIT"]
