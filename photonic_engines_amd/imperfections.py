"""Uniform output loss on the GPU: drop-in for the `loss=` path of the reference
(photonic_engines/imperfections.py:89-337: `_eta`, `binomial_pmf`, `LossyDistribution`,
`draw_indices`, `thin_distribution`) and `spec` (imperfections.py:508-526) for the `indist=`
argument check.

Removing one photon uniformly at random from a k-photon pattern t + e_j lands on t with
probability (t_j + 1) / k, so

    q_n = p,   q_{k-1}(t) = sum_j q_k(t + e_j) (t_j + 1) / k,   P_loss(layer k) = C(n,k) eta^k (1-eta)^(n-k) q_k.

Each descent step is one pull-form HIP kernel launch (every (k-1)-photon state gathers its m
children; no atomics, no edge tables); the binomial weights are torch ops so that eta may be a
tensor with a gradient, and the descent has its own adjoint kernel for gradients through p.
"""
from __future__ import annotations

import math

from dataclasses import dataclass

import numpy as np
import torch

from . import _native as nat

Eta = float | torch.Tensor


def _eta(eta) -> float | torch.Tensor | None:
    """None / >= 1 -> no loss; a float outside [0, 1] raises (imperfections.py:94-103)."""
    if eta is None:
        return None
    if isinstance(eta, torch.Tensor):
        return None if float(eta.detach()) >= 1.0 else eta
    e = float(eta)
    if not 0.0 <= e <= 1.0:
        raise ValueError(f"loss (transmission) must be in [0, 1], got {e}")
    return None if e >= 1.0 else e


def binomial_pmf(n: int, eta: Eta, dtype=torch.float64, device=None) -> torch.Tensor:
    """P(k of n photons survive), k = 0..n; differentiable in a tensor eta (imperfections.py:106-111)."""
    e = torch.as_tensor(eta, dtype=dtype, device=device)
    k = torch.arange(n + 1, dtype=dtype, device=e.device)
    logc = (torch.lgamma(torch.tensor(n + 1.0, dtype=dtype, device=e.device)) - torch.lgamma(k + 1)
            - torch.lgamma(n - k + 1))
    return torch.exp(logc) * e ** k * (1 - e) ** (n - k)


class Distinguishability:
    """Partial distinguishability (imperfections.py:465-506): the x-model (`x`, the pairwise
    overlap of the photons' internal states, HOM visibility x^2) or random species (`species`,
    accepted and validated; the strong simulation runs the x-model only)."""

    def __init__(self, x: float | None = None, species=None):
        if (x is None) == (species is None):
            raise ValueError("give exactly one of x (x-model) or species (random species probabilities)")
        self.x = None if x is None else float(x)
        self.species = None
        if self.x is not None and not 0.0 <= self.x <= 1.0:
            raise ValueError(f"x must be in [0, 1], got {x}")
        if species is not None:
            q = tuple(float(v) for v in species)
            if any(v < 0 for v in q) or sum(q) > 1.0 + 1e-12:
                raise ValueError("species probabilities must be non-negative and sum to at most 1")
            self.species = q

    @staticmethod
    def from_hom_visibility(V: float) -> "Distinguishability":
        """x-model with pairwise overlap x = sqrt(V), V the two-photon HOM dip visibility."""
        if not 0.0 <= float(V) <= 1.0:
            raise ValueError(f"HOM visibility must be in [0, 1], got {V}")
        return Distinguishability(x=math.sqrt(float(V)))

    @property
    def is_ideal(self) -> bool:
        if self.x is not None:
            return self.x >= 1.0
        return len(self.species) == 1 and self.species[0] >= 1.0


def spec(indist):
    """None / 1.0 / an ideal spec -> None (the lossless path); otherwise the spec (a float is the
    x-model). Overlaps outside [0, 1] raise ValueError (imperfections.py:508-526)."""
    if indist is None:
        return None
    if isinstance(indist, (list, tuple)) or (isinstance(indist, torch.Tensor) and indist.dim() == 1):
        xs = [float(v) for v in indist]
        if any(not 0.0 <= v <= 1.0 for v in xs):
            raise ValueError("per-photon overlaps must be in [0, 1]")
        return None if all(v >= 1.0 for v in xs) else Distinguishability(x=min(xs))
    d = indist if isinstance(indist, Distinguishability) else Distinguishability(
        x=float(indist.detach()) if isinstance(indist, torch.Tensor) else float(indist))
    return None if d.is_ideal else d


def truncation_bound(n: int, indist, order: int | None) -> float:
    """Mass dropped by the strong-side truncation `order=b` (imperfections.py:608-619): the
    partitions with more than b distinguishable photons, P(Bin(n, 1 - x) > b)."""
    d = spec(indist)
    if d is None or order is None or order >= n:
        return 0.0
    if d.x is None:
        raise NotImplementedError("truncation_bound: x-model only")
    p = 1.0 - d.x
    return float(sum(math.comb(n, b) * p ** b * (1 - p) ** (n - b) for b in range(order + 1, n + 1)))


MULTINOMIAL_MAX = 1 << 24


def draw_indices(p: torch.Tensor, shots: int, generator: torch.Generator | None = None) -> torch.Tensor:
    """`shots` indices from the unnormalised weights p [S], with replacement (imperfections.py:224-237):
    torch.multinomial up to 2^24 categories, the inverse CDF beyond."""
    if p.numel() <= MULTINOMIAL_MAX:
        return torch.multinomial(p, shots, replacement=True, generator=generator)
    cdf = torch.cumsum(p, dim=0)
    u = torch.rand(shots, dtype=cdf.dtype, device=cdf.device, generator=generator) * cdf[-1]
    return torch.searchsorted(cdf, u).clamp_(max=p.numel() - 1)


def thin_counts(counts: torch.Tensor, eta, generator: torch.Generator | None = None) -> torch.Tensor:
    """Output-side loss on samples (imperfections.py:116-126): every photon of counts [..., m] is kept
    independently with probability eta (binomial draws on the host). Same shape and dtype."""
    e = _eta(eta)
    if e is None:
        return counts
    p = torch.full(counts.shape, float(e), dtype=torch.float64, device="cpu")
    kept = torch.binomial(counts.to("cpu", torch.float64), p, generator=generator)
    return kept.to(counts.device, counts.dtype)


def thin_to_clicks(occupations, probs: torch.Tensor) -> tuple[np.ndarray, torch.Tensor]:
    """The distribution seen by threshold detectors (imperfections.py:429-438): patterns [C, m] of
    0/1 clicks and their probabilities [..., C] (the sum over the occupations with that click
    pattern). Patterns come in ascending order of sum_j click_j 2^j, the reference's order: the
    lexicographic order of the rows read from the last mode to the first."""
    cl = torch.as_tensor(np.asarray(occupations) > 0, dtype=torch.uint8)
    rev, inv = torch.unique(cl.flip(1), dim=0, return_inverse=True)
    out = torch.zeros((*probs.shape[:-1], rev.shape[0]), dtype=probs.dtype, device=probs.device)
    out.index_add_(-1, inv.to(probs.device), probs)
    return rev.flip(1).to(torch.int64).numpy(), out


@dataclass
class LossyDistribution:
    """A distribution over 0..n photons: probs[k] is [S_k] (or [B, S_k]) over occupations(k);
    the layers sum to the binomial weights (imperfections.py:164-218)."""
    layers: object
    probs: list
    _occ: list | None = None

    @property
    def n(self) -> int:
        return len(self.probs) - 1

    def occupations(self, k: int) -> np.ndarray:
        if self._occ is not None:
            return self._occ[k]
        return self.layers.occupations(k)

    def mass(self, k: int) -> torch.Tensor:
        """P(exactly k photons detected): scalar or [B]."""
        return self.probs[k].sum(dim=-1)

    def conditional(self, k: int) -> torch.Tensor:
        """The k-photon patterns given k detected (post-selection)."""
        return self.probs[k] / self.mass(k).unsqueeze(-1).clamp_min(1e-300)

    def flat(self) -> tuple[np.ndarray, torch.Tensor]:
        occ = np.concatenate([self.occupations(k) for k in range(self.n + 1)], axis=0)
        return occ, torch.cat(self.probs, dim=-1)

    def as_dict(self, tol: float = 0.0, batch: int | None = None) -> dict:
        occ, p = self.flat()
        if batch is not None:
            p = p[batch]
        p = p.detach().cpu().to(torch.float64).numpy()
        return {tuple(int(v) for v in o): float(q) for o, q in zip(occ, p) if q > tol}

    def clicks(self) -> tuple[np.ndarray, torch.Tensor]:
        occ, p = self.flat()
        return thin_to_clicks(occ, p)

    def sample(self, shots: int, generator: torch.Generator | None = None) -> torch.Tensor:
        """Samples [shots, m] (one distribution) (imperfections.py:213-218). With no generator or a
        CPU one: the reference's flat host draw, so the same generator gives the same samples.
        With a CUDA generator (which the reference's host draw cannot use): drawn on the GPU, the
        photon number first and then the pattern inside its layer (slos.sampling.sample_lossy_device)."""
        if any(q.dim() != 1 for q in self.probs):
            raise ValueError("one distribution at a time")
        if generator is not None and generator.device.type == "cuda":
            from .slos.sampling import sample_lossy_device
            return sample_lossy_device(self, shots, generator)
        occ, p = self.flat()
        return torch.as_tensor(occ)[draw_indices(p.detach().cpu().to(torch.float64), shots, generator)]


class _Descend(torch.autograd.Function):
    """q_{k-1} from q_k with the descent kernel; backward with its adjoint kernel."""

    @staticmethod
    def forward(ctx, q: torch.Tensor, plan, k: int) -> torch.Tensor:
        ctx.plan, ctx.k, ctx.dtype = plan, k, q.dtype
        return plan.loss_descend(k, q)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        return ctx.plan.loss_ascend(ctx.k, g.contiguous().to(ctx.dtype)), None, None


def carried_weights(n: int, eta: Eta, device, dtype=torch.float32) -> torch.Tensor | None:
    """The binomial weights of layers 0..n in the probabilities' dtype (float32, or float64) when
    the weighted blocks can be carried down the layers (a loss without gradient whose weights are
    all positive), else None."""
    e = _eta(eta)
    if e is None or (isinstance(e, torch.Tensor) and e.requires_grad):
        return None
    w = binomial_pmf(n, e, dtype=torch.float64, device=device).to(dtype)
    return w if bool((w > 0).all()) else None


def thin_distribution(layers, probs: torch.Tensor, eta: Eta, chunk: int | None = None,
                      _top_weighted: bool = False) -> LossyDistribution:
    """Uniform output loss on the n-photon distribution `probs` ([S] or [B, S] over
    layers.final_states), returned as a LossyDistribution over 0..n photons
    (imperfections.py:309-337). Differentiable in probs and eta. `chunk` is accepted for
    signature compatibility (the pull kernels need no edge chunks). `_top_weighted`: `probs` already
    carries the weight of layer n (the forward's readout applied it)."""
    n = layers.n
    e = _eta(eta)
    user_dev = probs.device
    batched = probs.dim() == 2
    dev = nat.require_gpu(probs.device if probs.device.type == "cuda" else None)
    rdt = torch.float64 if probs.dtype == torch.float64 else torch.float32   # FP64 kernels for float64 input
    q_top = (probs if batched else probs.unsqueeze(0)).to(device=dev, dtype=rdt).contiguous()
    plan = layers._plan
    grad = torch.is_grad_enabled() and (q_top.requires_grad or (isinstance(e, torch.Tensor) and e.requires_grad))
    w32 = None if grad else carried_weights(n, e, dev, dtype=rdt)
    if _top_weighted and w32 is None:
        raise ValueError("_top_weighted needs a loss whose weights can be carried down the layers")
    if w32 is not None:
        # carry the weights down the layers: block[k-1] = descent(block[k]) * w[k-1] / w[k], one pass
        # per layer and no unweighted copies
        ratio = (w32[:-1].double() / w32[1:].double()).cpu().tolist()
        blocks = [None] * (n + 1)
        blocks[n] = q_top if _top_weighted else q_top * w32[n]
        for k in range(n, 0, -1):
            blocks[k - 1] = plan.loss_descend(k, blocks[k], scale=ratio[k - 1])
        out = []
        for b in blocks:
            b = b if batched else b.squeeze(0)
            out.append(b.to(user_dev) if user_dev.type != "cuda" else b)
        return LossyDistribution(layers, out)
    q = [None] * (n + 1)
    q[n] = q_top
    grad = torch.is_grad_enabled() and q_top.requires_grad
    for k in range(n, 0, -1):
        q[k - 1] = _Descend.apply(q[k], plan, k) if grad else plan.loss_descend(k, q[k])
    if e is None:
        weights = torch.zeros(n + 1, dtype=rdt, device=dev)
        weights[n] = 1.0
    else:
        weights = binomial_pmf(n, e.to(dev) if isinstance(e, torch.Tensor) else e, dtype=torch.float64,
                               device=dev).to(rdt)
    blocks = []
    for k in range(n + 1):
        b = q[k] * weights[k]
        if not batched:
            b = b.squeeze(0)
        blocks.append(b.to(user_dev) if user_dev.type != "cuda" else b)
    return LossyDistribution(layers, blocks)
