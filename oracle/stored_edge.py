"""Torch-CPU restatement of the reference's stored-edge SLOS forward, used as a CPU baseline (test
infrastructure: never imported by the product).

It follows the reference algorithm step by step, in code of its own:
- the lex basis layer by layer: every parent row followed by every admissible mode >= its last
  position, row-major (photonic_engines/slos/basis.py:120-148; fock mask only here);
- the raising edges: one per (child, distinct mode), the parent found by a binary search of the
  child-minus-one-photon key among the parents' packed keys, whose numeric order is the lex order
  (basis.py:151-193);
- the edges regrouped by mode into slabs (basis.py:216-228);
- the forward in the "mode" layout: per layer and per mode slab, gather the parent amplitudes,
  multiply by U[:, j, i_k], scatter-add into the children (slos/kernel.py:216-230), then the
  readout |amp|^2 * prod t_j! (kernel.py:262-283).
Setup (basis and edges) and the forward are timed separately, as in the reference's bench
protocol (bench/bench_product_slos.py:44-57).
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch


def _layers(n: int, m: int) -> list[np.ndarray]:
    """Fock layers 0..n as int8/int16 [S_k, k] arrays of sorted positions, in lex order."""
    dt = np.int8 if m <= 127 else np.int16
    out = [np.zeros((1, 0), dtype=dt)]
    for k in range(n):
        par = out[-1]
        first = par[:, -1].astype(np.int64) if k else np.zeros(1, dtype=np.int64)
        counts = m - first                               # modes a >= last (fock: no cap)
        rows = np.repeat(np.arange(par.shape[0]), counts)
        starts = np.repeat(np.cumsum(counts) - counts, counts)
        a = first[rows] + (np.arange(rows.shape[0]) - starts)
        child = np.empty((rows.shape[0], k + 1), dtype=dt)
        child[:, :k] = par[rows]
        child[:, k] = a
        out.append(child)
    return out


def _keys(states: np.ndarray, bits: int) -> np.ndarray:
    key = np.zeros(states.shape[0], dtype=np.uint64)
    for j in range(states.shape[1]):
        key = (key << np.uint64(bits)) | states[:, j].astype(np.uint64)
    return key


def _edges_by_mode(parents: np.ndarray, children: np.ndarray, m: int, bits: int):
    """(parent, child, offsets): the edges of one layer sorted by mode, slab j = [off[j], off[j+1])."""
    S1, k1 = children.shape
    pkeys = _keys(parents, bits)
    ckeys = _keys(children, bits)
    distinct = np.ones((S1, k1), dtype=bool)
    distinct[:, 1:] = children[:, 1:] != children[:, :-1]
    cidx, col = np.nonzero(distinct)
    mode = children[cidx, col].astype(np.int64)
    # the parent's key: the child's key with digit `col` taken out (digits after it move up)
    low_bits = ((k1 - 1 - col) * bits).astype(np.uint64)
    key = ckeys[cidx]
    high = key >> (low_bits + np.uint64(bits))
    low = key & ((np.uint64(1) << low_bits) - np.uint64(1))
    par = np.searchsorted(pkeys, (high << low_bits) | low).astype(np.int64) if k1 > 1 else np.zeros(cidx.shape[0], np.int64)
    order = np.argsort(mode, kind="stable")
    off = np.searchsorted(mode[order], np.arange(m + 1))
    return (torch.from_numpy(par[order]), torch.from_numpy(cidx[order].astype(np.int64)),
            off.astype(np.int64).tolist())


class StoredEdgeSLOS:
    """The reference's mode-slab forward on the CPU in torch (complex64 by default)."""

    def __init__(self, n: int, m: int, occ, dtype=torch.complex64):
        t0 = time.perf_counter()
        self.n, self.m, self.dtype = n, m, dtype
        self.modes = [j for j, c in enumerate(occ) for _ in range(int(c))]
        bits = max(1, math.ceil(math.log2(max(m, 2))))
        lay = _layers(n, m)
        self.sizes = [a.shape[0] for a in lay]
        self.slabs = [_edges_by_mode(lay[k], lay[k + 1], m, bits) for k in range(n)]
        last = lay[n].astype(np.int64)
        run = np.ones(last.shape, dtype=np.int64)
        for j in range(1, n):
            run[:, j] = np.where(last[:, j] == last[:, j - 1], run[:, j - 1] + 1, 1)
        self.norm = torch.from_numpy(np.prod(run, axis=1).astype(np.float32))   # prod t_j! from run lengths
        self.input_norm = 1.0 / math.sqrt(math.prod(math.factorial(int(c)) for c in occ))
        self.setup_s = time.perf_counter() - t0

    def probs(self, U: torch.Tensor) -> torch.Tensor:
        U = U.to(self.dtype)
        psi = torch.ones(1, dtype=self.dtype)
        for k, i in enumerate(self.modes):
            par, chi, off = self.slabs[k]
            nxt = torch.zeros(self.sizes[k + 1], dtype=self.dtype)
            for j in range(self.m):
                a, b = off[j], off[j + 1]
                if b > a:
                    nxt.scatter_add_(0, chi[a:b], U[j, i] * psi[par[a:b]])
            psi = nxt
        amp = psi * self.input_norm
        return (amp.real ** 2 + amp.imag ** 2) * self.norm
