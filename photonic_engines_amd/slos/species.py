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


def _complex_of(q: torch.Tensor) -> torch.dtype:
    """The complex dtype of the layer steps over real probabilities q (float64: complex128)."""
    return torch.complex128 if q.dtype == torch.float64 else torch.complex64


def _batched_u(kernel, U: torch.Tensor):
    """U [m,m] | [B,m,m] -> (contiguous [B,m,m] of the kernel's dtype on the compute GPU, was batched); the
    result stays attached to U's graph when U requires a gradient."""
    return kernel._prepare(U if _needs_grad(U) else U.detach())


def _layer_norm(kernel, g: int, device) -> torch.Tensor:
    """prod_j t_j! of the states of layer g as rdtype [S_g] (1 when no state can bunch)."""
    if kernel._output_norm_is_one:
        return torch.ones(1, dtype=kernel.rdtype, device=device)
    return torch.as_tensor(kernel.layers.prob_norms(g), dtype=kernel.rdtype, device=device)


def _raw_probs(kernel, Ub: torch.Tensor, modes: Sequence[int], g: int, scale: float) -> torch.Tensor:
    """|scale psi_g|^2 prod t! of layer g for the columns `modes` with U-gradients (SLOSRaw)."""
    from .autograd import SLOSRaw
    raw = SLOSRaw.apply(Ub, Ub.detach().contiguous(), kernel, list(modes), g)
    return (raw.real ** 2 + raw.imag ** 2) * (scale * scale) * _layer_norm(kernel, g, Ub.device)


def _raise_steps(kernel, q: torch.Tensor, k: int, Wcols: Sequence[torch.Tensor]) -> torch.Tensor:
    """Raises of q [B', S_k] (float32 or float64) by the coefficient columns Wcols[t] ([B', m], real, >= 0)
    as complex layer steps with their adjoint (LayerStep): [B', S_{k + len(Wcols)}]."""
    from .autograd import LayerStep
    Bp, m = q.shape[0], kernel.m
    cdt = _complex_of(q)
    qc = torch.complex(q, torch.zeros_like(q))
    for w in Wcols:
        W = torch.zeros((Bp, m, m), dtype=cdt, device=q.device)
        W = W.index_copy(2, torch.zeros(1, dtype=torch.int64, device=q.device), w.to(cdt).unsqueeze(2))
        qc = LayerStep.apply(W, qc, kernel._plan, k + 1, 0)
        k += 1
    return qc.real


def layer_probs(kernel, Ub: torch.Tensor, cols: Sequence[int]) -> torch.Tensor:
    """|<t| U |cols>|^2 on layer g = len(cols) of the kernel's basis, [B, S_g] rdtype: one forward
    over the columns `cols` with the input norm 1/sqrt(prod s_i!) and the layer's prod t_j!.
    g = 0 is the vacuum (probability 1)."""
    g = len(cols)
    B = Ub.shape[0]
    if g == 0:
        return torch.ones((B, 1), dtype=kernel.rdtype, device=Ub.device)
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
    w = (Ub[:, :, int(col)].abs() ** 2).to(q.dtype).contiguous()
    if _needs_grad(Ub) or q.requires_grad:
        return _raise_steps(kernel, q.contiguous(), k, [w])
    if q.dtype == torch.float64:  # FP64: the raise as a complex128 layer step
        return _raise_rows(kernel, q.contiguous(), k, w)
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


def _subset_forwards(kernel, Ub: torch.Tensor, cols: Sequence[int], goods: Sequence[Sequence[int]]) -> torch.Tensor:
    """The interfering part of species_probs for several subsets at once: |<t| U |good>|^2 on layer
    g = len(good) (g >= 2), normalised by 1 / prod s_i!, as [C, B, S_g]. The subsets form the batch
    dimension: each row's U has the interfering columns first, so one forward with modes 0..g-1
    covers all of them."""
    B, m = Ub.shape[0], Ub.shape[-1]
    C, g = len(goods), len(goods[0])
    dev = Ub.device
    gi = torch.as_tensor([list(gd) for gd in goods], dtype=torch.int64, device=dev)  # [C, g]
    Ug = Ub[:, :, gi].permute(2, 0, 1, 3)                                             # [C, B, m, g]
    if g < m:
        Ug = torch.cat([Ug, torch.zeros((C, B, m, m - g), dtype=Ub.dtype, device=dev)], dim=3)
    Up = Ug.reshape(C * B, m, m).contiguous()
    if _needs_grad(Ub):
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
    if any(v != 1.0 for v in norms):                                                  # 1 / prod s_i!
        q = q * torch.as_tensor(norms, dtype=q.dtype, device=dev).repeat_interleave(B).unsqueeze(1)
    return q.reshape(C, B, q.shape[-1])


def _raise_rows(kernel, q: torch.Tensor, k: int, W: torch.Tensor) -> torch.Tensor:
    """One classical raise of every row of q [R, S_k] (float32 or float64), row r by its own coefficient
    column W[r] ([m], real, >= 0): [R, S_{k+1}]. Without a gradient it runs as a complex layer step
    (imaginary parts stay exactly zero, the real parts are the same fused multiply-adds), so the
    block kernel carries it; with one, through LayerStep and its adjoint."""
    if _needs_grad(q) or _needs_grad(W):
        return _raise_steps(kernel, q.contiguous(), k, [W])
    R, m = q.shape[0], kernel.m
    Uw = torch.zeros((R, m, m), dtype=_complex_of(q), device=q.device)
    Uw[:, :, 0] = W.to(Uw.dtype)
    qc = torch.complex(q, torch.zeros_like(q)).contiguous()
    return kernel._plan.layer_step(Uw, k + 1, 0, qc, nat.OUT_RAW).real


def x_mixture(kernel, U: torch.Tensor, indist, order: int | None = None) -> torch.Tensor:
    """The x-model distribution on the final layer ([S] or [B, S], rdtype): the sum over the
    interfering subsets G of the input photons of x^|G| (1 - x)^(n - |G|) species_probs(G, rest).
    `indist` a float, a Distinguishability(x=...) or a tensor x (differentiable); `order=b` keeps
    the subsets with at most b distinguishable photons.

    The distinguishable photons D = rest enter species_probs by classical raises, which are linear
    and commute, so every term raises D in ascending photon order and the terms share their raises:
    A[T] collects, on layer n - |T|, everything still to be raised by the photons T. A subset's
    interfering forward (batched over the subsets of one size) starts A[D]; then, for |T| from the
    largest down, A[T] is raised by its smallest photon into A[T minus it], all sets of one size in
    one batched layer step, and the contributions are added in ascending order of T. A[()] is the
    result. The final layer then takes n raises instead of one per subset. An interfering set of
    fewer than two photons is itself classical: its photons join T (a single photon's
    distribution is its column |U[:, c]|^2)."""
    d = spec(indist)
    if d.x is None:
        raise NotImplementedError("strong simulation: the x-model only (random species need a partition sum)")
    Ub, batched = _batched_u(kernel, U)
    rdt = kernel.rdtype
    x = indist if isinstance(indist, torch.Tensor) else torch.tensor(d.x, dtype=rdt)
    x = x.to(dtype=rdt, device=Ub.device)
    cols = kernel.input_modes
    n = len(cols)
    B, m = Ub.shape[0], Ub.shape[-1]
    dev = Ub.device
    b_max = n if order is None else max(0, min(int(order), n))
    per = max(1, (2 << 30) // (rdt.itemsize * kernel.layers.size(n) * max(B, 1)))
    A: dict = {}

    def add(T, v):
        A[T] = v if T not in A else A[T] + v

    def start(L):
        # the terms whose distinguishable set has L photons (at least two interfering): weight
        # x^(n-L) (1-x)^L
        w = x ** (n - L) * (1 - x) ** L
        bads = list(combinations(range(n), L))
        for i0 in range(0, len(bads), per):
            chunk = bads[i0:i0 + per]
            goods = [[cols[j] for j in range(n) if j not in bad] for bad in chunk]
            F = _subset_forwards(kernel, Ub, cols, goods) * w
            for r, bad in enumerate(chunk):
                add(tuple(bad), F[r])

    # fewer than two interfering photons: the whole input is classical, every photon raised from the
    # vacuum (C(n, L) such terms for L = n and n - 1)
    wc = None
    for L in (n, n - 1):
        if 0 <= L <= b_max and n - L <= 1:
            t = math.comb(n, L) * (x ** (n - L) * (1 - x) ** L)
            wc = t if wc is None else wc + t
    if wc is not None and n >= 1:
        add(tuple(range(n)), wc * torch.ones((B, 1), dtype=rdt, device=dev))
    for L in range(n, 0, -1):
        if L <= b_max and n - L >= 2:
            start(L)
        Ts = sorted(T for T in A if len(T) == L)
        if not Ts:
            continue
        k = n - L
        # every raised row goes to A[T minus its smallest photon]. The rows are raised in an order
        # that makes each target's j-th contribution (its sources in ascending order of that
        # photon, Ts being sorted) part of one contiguous block: targets by descending source
        # count, block j = the j-th source of every target that has one. A target's contributions
        # are then summed source by source with contiguous prefix additions.
        srcs: dict = {}
        for T in Ts:
            srcs.setdefault(T[1:], []).append(T)
        order = sorted(srcs, key=lambda t: -len(srcs[t]))                  # stable: ties in set order
        cnt = [sum(1 for t in order if len(srcs[t]) > j) for j in range(len(srcs[order[0]]))]
        rows = [srcs[t][j] for j in range(len(cnt)) for t in order[:cnt[j]]]
        q = torch.cat([A[T] for T in rows], 0)                                        # [len(Ts) * B, S_k]
        ct = torch.as_tensor([cols[T[0]] for T in rows], dtype=torch.int64, device=dev)
        W = (Ub[:, :, ct].abs() ** 2).permute(2, 0, 1).reshape(len(rows) * B, m).to(rdt)
        r = _raise_rows(kernel, q, k, W)
        r = r.reshape(len(rows), B, r.shape[-1])
        for T in Ts:
            del A[T]
        acc = r[:cnt[0]].clone()
        o = cnt[0]
        for c in cnt[1:]:
            acc[:c] += r[o:o + c]
            o += c
        for ti, t in enumerate(order):
            add(t, acc[ti])
    if n >= 2:
        add((), _subset_forwards(kernel, Ub, cols, [list(cols)])[0] * (x ** n))
    if () not in A:  # no photon
        add((), torch.ones((B, 1), dtype=rdt, device=dev))
    out = A[()]
    return kernel._out(out, batched)
