"""Strong simulation under partial distinguishability on MI355X: the FockSLOSKernel part of
photonic_engines/slos/species.py (`layer_probs` :84-101, `classical_raise` :37-56,
`convolve_layers` :64-81, `species_probs` :104-120, `x_mixture` :123-146).

A shot is a species partition of the input photons: the photons of one species interfere among
themselves, different species do not, and the detector sees the sum of the species' patterns.
On the layered basis of the kernel:

  * a species of g >= 2 photons is the SLOS recurrence on its columns stopped on layer g of the
    kernel's basis, squared with that layer's deferred normalisation (one slos_forward call with
    the species' columns, `layer_probs`);
  * a photon of a species of its own is a probability-level raise along the k -> k+1 edges,
    q_{k+1}(t + e_j) += q_k(t) |U[j, col]|^2, with no bosonic factor (slos_raise: each child
    gathers its parents; `classical_raise`);
  * two multi-photon species combine by the explicit convolution of their layers (small n,
    `convolve_layers`).

`x_mixture` is the x-model: the sum over the interfering subsets G of the input photons with
weights x^|G| (1 - x)^(n - |G|), each a `species_probs`; `order=b` keeps at most b
distinguishable photons (the dropped mass is `imperfections.truncation_bound`). A tensor x gets
a gradient (the weights are torch operations). U gets one too ("Differentiable in q and U",
species.py:37-41): with U requiring a gradient, the interfering layers run as SLOSRaw (the raw
amplitudes with the hand-written adjoint) squared by torch, and every raise runs as a LayerStep
(the complex layer step with the column |U[j, col]|^2, whose adjoint gives the gradients of the
layer below and of that column).
"""
from __future__ import annotations

import math
from collections.abc import Sequence
from itertools import combinations

import numpy as np
import torch

from .. import _native as nat
from ..imperfections import spec


def _needs_grad(U: torch.Tensor) -> bool:
    return U.requires_grad and torch.is_grad_enabled()


def _batched_u(kernel, U: torch.Tensor):
    """U [m,m] | [B,m,m] -> (contiguous complex64 [B,m,m] on the compute GPU, was batched); the
    result stays attached to U's graph when U requires a gradient."""
    return kernel._prepare(U if _needs_grad(U) else U.detach())


def _layer_norm(kernel, g: int, device) -> torch.Tensor:
    """prod_j t_j! of the states of layer g as float32 [S_g] (1 when no state can bunch)."""
    if kernel._output_norm_is_one:
        return torch.ones(1, dtype=torch.float32, device=device)
    return torch.as_tensor(kernel.layers.prob_norms(g), dtype=torch.float32, device=device)


def _raw_probs(kernel, Ub: torch.Tensor, modes: Sequence[int], g: int, scale: float) -> torch.Tensor:
    """|scale psi_g|^2 prod t! of layer g for the columns `modes` with U-gradients (SLOSRaw)."""
    from .autograd import SLOSRaw
    raw = SLOSRaw.apply(Ub, Ub.detach().contiguous(), kernel, list(modes), g)
    return (raw.real ** 2 + raw.imag ** 2) * (scale * scale) * _layer_norm(kernel, g, Ub.device)


def _raise_steps(kernel, q: torch.Tensor, k: int, Wcols: Sequence[torch.Tensor]) -> torch.Tensor:
    """Raises of q [B', S_k] (float32) by the coefficient columns Wcols[t] ([B', m], real, >= 0)
    as complex layer steps with their adjoint (LayerStep): [B', S_{k + len(Wcols)}]."""
    from .autograd import LayerStep
    Bp, m = q.shape[0], kernel.m
    qc = torch.complex(q, torch.zeros_like(q))
    for w in Wcols:
        W = torch.zeros((Bp, m, m), dtype=torch.complex64, device=q.device)
        W = W.index_copy(2, torch.zeros(1, dtype=torch.int64, device=q.device),
                         w.to(torch.complex64).unsqueeze(2))
        qc = LayerStep.apply(W, qc, kernel._plan, k + 1, 0)
        k += 1
    return qc.real


def layer_probs(kernel, Ub: torch.Tensor, cols: Sequence[int]) -> torch.Tensor:
    """|<t| U |cols>|^2 on layer g = len(cols) of the kernel's basis, [B, S_g] float32: one forward
    over the columns `cols` with the input norm 1/sqrt(prod s_i!) and the layer's prod t_j!.
    g = 0 is the vacuum (probability 1)."""
    g = len(cols)
    B = Ub.shape[0]
    if g == 0:
        return torch.ones((B, 1), dtype=torch.float32, device=Ub.device)
    counts: dict = {}
    for c in cols:
        counts[c] = counts.get(c, 0) + 1
    s_norm = math.prod(math.factorial(v) for v in counts.values())
    if _needs_grad(Ub):
        return _raw_probs(kernel, Ub, [int(c) for c in cols], g, 1.0 / math.sqrt(s_norm))
    flags = nat.FLAG_COMPENSATED if kernel.compensated else 0
    return kernel._plan.forward(Ub, modes=[int(c) for c in cols], layer=g, out_kind=nat.OUT_PROBS,
                                scale=1.0 / math.sqrt(s_norm), flags=flags)


def classical_raise(kernel, q: torch.Tensor, k: int, col: int, Ub: torch.Tensor) -> torch.Tensor:
    """One photon that interferes with nothing enters column `col`: the layer-k distribution
    q [B, S_k] becomes the layer-(k+1) one, q'(t + e_j) = sum q(t) |U[j, col]|^2."""
    w = (Ub[:, :, int(col)].abs() ** 2).to(torch.float32).contiguous()
    if _needs_grad(Ub) or q.requires_grad:
        return _raise_steps(kernel, q.contiguous(), k, [w])
    return kernel._plan.raise_layer(k + 1, w, q.contiguous())


def convolve_layers(kernel, qa: torch.Tensor, ka: int, qb: torch.Tensor, kb: int) -> torch.Tensor:
    """Distribution of the sum of two independent patterns (layers ka and kb, [B, S] each) on
    layer ka + kb: q(t) = sum_{ta + tb = t} qa(ta) qb(tb); a sum outside the (masked) basis drops
    its mass. Explicit S_a x S_b pairs: small layers only."""
    layers = kernel.layers
    occ_a, occ_b = layers.occupations(ka), layers.occupations(kb)
    occ_t = layers.occupations(ka + kb)
    index = {tuple(r): i for i, r in enumerate(occ_t.tolist())}
    ia, ib, tg = [], [], []
    for i, ta in enumerate(occ_a.tolist()):
        for j, tb in enumerate(occ_b.tolist()):
            t = index.get(tuple(x + y for x, y in zip(ta, tb)))
            if t is not None:
                ia.append(i); ib.append(j); tg.append(t)
    dev = qa.device
    ia_t, ib_t, tg_t = (torch.as_tensor(np.asarray(v, dtype=np.int64), device=dev) for v in (ia, ib, tg))
    out = torch.zeros((qa.shape[0], occ_t.shape[0]), dtype=qa.dtype, device=dev)
    return out.index_add(-1, tg_t, qa.index_select(-1, ia_t) * qb.index_select(-1, ib_t))


def species_probs(kernel, Ub: torch.Tensor, groups: Sequence[Sequence[int]], singles: Sequence[int]) -> torch.Tensor:
    """Distribution [B, S_n] for a species partition: `groups` the columns of each multi-photon
    species, `singles` the columns of the photons of a species of their own."""
    groups = [list(g) for g in groups]
    singles = list(singles) + [g[0] for g in groups if len(g) == 1]
    groups = sorted((g for g in groups if len(g) >= 2), key=len, reverse=True)
    q = layer_probs(kernel, Ub, groups[0] if groups else [])
    k = len(groups[0]) if groups else 0
    for g in groups[1:]:
        q, k = convolve_layers(kernel, q, k, layer_probs(kernel, Ub, g), len(g)), k + len(g)
    for c in singles:
        q, k = classical_raise(kernel, q, k, c, Ub), k + 1
    if k != kernel.n:
        raise ValueError(f"the partition covers {k} photons, the kernel has {kernel.n}")
    return q


def _subset_terms(kernel, Ub: torch.Tensor, cols: Sequence[int], bads: Sequence[Sequence[int]]) -> torch.Tensor:
    """sum over the subsets of species_probs(interfering = cols minus bad, singles = bad), for
    subsets that all leave out len(bad) photons: [B, S_n]. The subsets form the batch dimension:
    each row's U has the interfering columns first, so one forward with modes 0..g-1 covers all of
    them; the singles then enter by per-row raises (layer steps with the coefficient column
    |U[j, col]|^2 of the row)."""
    n, B, m = len(cols), Ub.shape[0], Ub.shape[-1]
    C, nb = len(bads), len(bads[0])
    dev = Ub.device
    goods = [[cols[j] for j in range(n) if j not in bad] for bad in bads]
    singles = [[cols[j] for j in bad] for bad in bads]
    g = n - nb
    grad = _needs_grad(Ub)
    if g >= 2:
        gi = torch.as_tensor(goods, dtype=torch.int64, device=dev)                  # [C, g]
        Ug = Ub[:, :, gi].permute(2, 0, 1, 3)                                        # [C, B, m, g]
        if g < m:
            Ug = torch.cat([Ug, torch.zeros((C, B, m, m - g), dtype=torch.complex64, device=dev)], dim=3)
        Up = Ug.reshape(C * B, m, m).contiguous()
        if grad:
            q = _raw_probs(kernel, Up, list(range(g)), g, 1.0)
        else:
            q = kernel._plan.forward(Up, modes=list(range(g)), layer=g, out_kind=nat.OUT_PROBS, scale=1.0,
                                     flags=nat.FLAG_COMPENSATED if kernel.compensated else 0)
        norms = []
        for good in goods:
            counts: dict = {}
            for c in good:
                counts[c] = counts.get(c, 0) + 1
            norms.append(1.0 / math.prod(math.factorial(v) for v in counts.values()))
        if any(v != 1.0 for v in norms):                                             # 1 / prod s_i!
            q = q * torch.as_tensor(norms, dtype=torch.float32, device=dev).repeat_interleave(B).unsqueeze(1)
        k = g
    else:
        q = torch.ones((C * B, 1), dtype=torch.float32, device=dev)
        singles = [sg + gd for sg, gd in zip(singles, goods)]
        k = 0
    if singles[0] and grad:
        Wcols = []
        for t in range(len(singles[0])):
            ct = torch.as_tensor([sg[t] for sg in singles], dtype=torch.int64, device=dev)  # [C]
            Wcols.append((Ub[:, :, ct].abs() ** 2).permute(2, 0, 1).reshape(C * B, m).to(torch.float32))
        q = _raise_steps(kernel, q.contiguous(), k, Wcols)
        k += len(singles[0])
    elif singles[0]:
        # The raises are the layer recurrence with a real coefficient column |U[j, col]|^2 on real
        # data: run them as complex layer steps (imaginary parts stay exactly zero, the real parts are
        # the same fused multiply-adds) so that the block kernel carries them.
        qc = torch.complex(q, torch.zeros_like(q))
        Uw = torch.zeros((C * B, m, m), dtype=torch.complex64, device=dev)
        for t in range(len(singles[0])):
            ct = torch.as_tensor([sg[t] for sg in singles], dtype=torch.int64, device=dev)  # [C]
            Uw[:, :, 0] = (Ub[:, :, ct].abs() ** 2).permute(2, 0, 1).reshape(C * B, m).to(torch.complex64)
            qc = kernel._plan.layer_step(Uw, k + 1, 0, qc.contiguous(), nat.OUT_RAW)
            k += 1
        q = qc.real
    if k != kernel.n:
        raise ValueError(f"the partition covers {k} photons, the kernel has {kernel.n}")
    return q.reshape(C, B, -1).sum(0)


def x_mixture(kernel, U: torch.Tensor, indist, order: int | None = None) -> torch.Tensor:
    """The x-model distribution on the final layer ([S] or [B, S], float32): the sum over the
    interfering subsets G of the input photons of x^|G| (1 - x)^(n - |G|) species_probs(G, rest).
    `indist` a float, a Distinguishability(x=...) or a tensor x (differentiable); `order=b` keeps
    the subsets with at most b distinguishable photons. Subsets with the same number of
    distinguishable photons run together (chunks of at most 2 GB of float32 terms)."""
    d = spec(indist)
    if d.x is None:
        raise NotImplementedError("strong simulation: the x-model only (random species need a partition sum)")
    Ub, batched = _batched_u(kernel, U)
    x = indist if isinstance(indist, torch.Tensor) else torch.tensor(d.x, dtype=torch.float32)
    x = x.to(dtype=torch.float32, device=Ub.device)
    cols = kernel.input_modes
    n = len(cols)
    b_max = n if order is None else max(0, min(int(order), n))
    per = max(1, (2 << 30) // (4 * kernel.layers.size(n) * Ub.shape[0]))
    out = None
    for b in range(b_max + 1):
        w = x ** (n - b) * (1 - x) ** b
        subsets = list(combinations(range(n), b))
        acc = None
        for i0 in range(0, len(subsets), per):
            term = _subset_terms(kernel, Ub, cols, subsets[i0:i0 + per])
            acc = term if acc is None else acc + term
        out = acc * w if out is None else out + acc * w
    return kernel._out(out, batched)
