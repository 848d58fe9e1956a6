"""Strong-then-sample on the device: `sample_from_probs` (photonic_engines/slos/emitter.py:215-230)
and the inverse-CDF branch of `draw_indices` (photonic_engines/imperfections.py:221-237).

Above 2^24 categories (torch.multinomial's cap) the reference takes a float64 cumulative sum of
the whole probability vector and a `searchsorted`; at n = 11, m = 32 that copy alone is 34 GB.
Here the float32 vector is read in place by slos_sample: FP64 sums of 64-entry chunks and
4096-entry blocks, a block CDF, and one wave per shot that finds the block, the chunk and the
entry, every FP64 addition in a fixed order. The drawn indices are unranked on the device into
occupation rows (slos_unrank_device). At or below 2^24 categories the draw is torch.multinomial,
as in the reference.
"""
from __future__ import annotations

import ctypes

import torch

from .. import _native as nat
from ..imperfections import MULTINOMIAL_MAX, thin_counts


def device_draw_indices(p: torch.Tensor, shots: int, generator: torch.Generator | None = None) -> torch.Tensor:
    """`shots` indices (int64, on p's GPU) drawn from the unnormalised float32 weights p [S] on the
    GPU, with replacement, by the two-level FP64 inverse CDF. Uniforms come from torch.rand
    (float64) on the generator's device."""
    if p.dim() != 1:
        raise ValueError("one distribution at a time: p must be [S]")
    dev = nat.require_gpu(p.device)
    q = p.detach().to(device=dev, dtype=torch.float32).contiguous()
    gdev = generator.device if generator is not None else dev
    u = torch.rand(int(shots), dtype=torch.float64, device=gdev, generator=generator).to(dev).contiguous()
    lib = nat.lib()
    need = ctypes.c_size_t()
    nat.check(lib.slos_sample_workspace(q.numel(), ctypes.byref(need)), "slos_sample_workspace")
    ws = torch.empty(max(1, need.value), dtype=torch.uint8, device=dev)
    idx = torch.empty(int(shots), dtype=torch.int64, device=dev)
    with torch.cuda.device(dev):
        nat.check(lib.slos_sample(nat.ptr(q), q.numel(), nat.ptr(u), int(shots), nat.ptr(idx), nat.ptr(ws), ws.numel(),
                                  nat.stream_handle(dev)), "slos_sample")
    return idx


def draw_indices(p: torch.Tensor, shots: int, generator: torch.Generator | None = None) -> torch.Tensor:
    """imperfections.py:221-237: torch.multinomial up to 2^24 categories; beyond, a float32 vector on
    the GPU goes to the device inverse CDF (no float64 copy), anything else to cumsum + searchsorted."""
    if p.numel() <= MULTINOMIAL_MAX:
        return torch.multinomial(p, shots, replacement=True, generator=generator)
    if p.device.type == "cuda" and p.dtype == torch.float32:
        return device_draw_indices(p, shots, generator)
    cdf = torch.cumsum(p.to(torch.float64), dim=0)
    u = torch.rand(shots, dtype=cdf.dtype, device=cdf.device, generator=generator) * cdf[-1]
    return torch.searchsorted(cdf, u).clamp_(max=p.numel() - 1)


def unrank_rows(layers, k: int, idx: torch.Tensor) -> torch.Tensor:
    """Occupation rows [len(idx), m] (int64, on idx's device) of the layer-k lex indices idx: on the
    GPU by slos_unrank_device, on the host by slos_unrank."""
    plan = layers._plan
    if idx.device.type == "cuda":
        ix = idx.to(torch.int64).contiguous()
        occ = torch.empty((ix.numel(), layers.m), dtype=torch.int32, device=ix.device)
        with torch.cuda.device(ix.device):
            nat.check(nat.lib().slos_unrank_device(plan._h, int(k), nat.ptr(ix), ix.numel(), nat.ptr(occ),
                                                   nat.stream_handle(ix.device)), "slos_unrank_device")
        return occ.to(torch.int64)
    pos = torch.as_tensor(layers.unrank(k, idx.cpu().numpy()).astype("int64"))
    out = torch.zeros((pos.shape[0], layers.m), dtype=torch.int64)
    if k:
        out.scatter_add_(1, pos, torch.ones_like(pos))
    return out


def sample_lossy_device(ld, shots: int, generator: torch.Generator) -> torch.Tensor:
    """Samples [shots, m] (int64, on the generator's GPU) of a LossyDistribution whose blocks are on
    the GPU, drawn there: the photon number k of every shot from the FP64 block masses
    (torch.multinomial on the generator), then the pattern inside layer k by draw_indices on that
    layer's float32 block (torch.multinomial up to 2^24 entries, else the device inverse CDF; no
    concatenated or float64 copy of the layers; float64 blocks, from complex128 kernels, draw by
    their FP64 cumsum) and the device unrank. The distribution is the
    reference's flat draw over the concatenated layers (imperfections.py:213-218); the draws
    differ (the reference draws on the host, from a CPU generator)."""
    if any(q.dim() != 1 for q in ld.probs):
        raise ValueError("one distribution at a time")
    dev = generator.device
    plan = ld.layers._plan
    m = ld.layers.m
    blocks = [q.detach().to(device=dev, dtype=torch.float64 if q.dtype == torch.float64 else torch.float32).contiguous()
              for q in ld.probs]
    for k, q in enumerate(blocks):
        if q.numel() != ld.layers.size(k):
            raise ValueError(f"block {k} has {q.numel()} entries, layer {k} has {ld.layers.size(k)}")
    masses = torch.cat([plan.total_mass(q.unsqueeze(0)) for q in blocks]).to(dev)
    ks = torch.multinomial(masses, int(shots), replacement=True, generator=generator)
    counts = torch.bincount(ks, minlength=len(blocks)).cpu().tolist()
    out = torch.zeros((int(shots), m), dtype=torch.int64, device=dev)
    for k, c in enumerate(counts):
        if c == 0 or k == 0:
            continue                                            # the vacuum row is all zeros
        sel = torch.nonzero(ks == k).squeeze(1)
        idx = draw_indices(blocks[k], int(c), generator)
        out[sel] = unrank_rows(ld.layers, k, idx)
    return out


def sample_from_probs(layers, probs, shots: int, generator: torch.Generator | None = None, loss=None) -> torch.Tensor:
    """Output occupations [shots, m] drawn from a probability vector over `layers.final_states`
    (emitter.py:215-230); probs need not be normalised. A LossyDistribution is sampled over its
    0..n photon layers; `loss=eta` on a lossless vector thins the samples (thin_counts)."""
    from ..imperfections import LossyDistribution
    if isinstance(probs, LossyDistribution):
        return thin_counts(probs.sample(shots, generator=generator), loss, generator=generator)
    if probs.dim() != 1:
        raise ValueError("one distribution at a time: probs must be [S]")
    if probs.numel() <= MULTINOMIAL_MAX:
        p = probs.detach().to(torch.float64)
        idx = torch.multinomial(p, shots, replacement=True, generator=generator)
    else:
        idx = draw_indices(probs.detach(), shots, generator)
    return thin_counts(unrank_rows(layers, layers.n, idx), loss, generator=generator)
