"""Hand-written adjoint of the SLOS recurrence behind torch.autograd.Function.

Replaces autograd through gather / mul / scatter_add_ (photonic_engines/slos/kernel.py:216-257)
and the opt-in recomputing layer (_CheckpointedLayer, kernel.py:40-80). A forward keeps only every
other layer (the even ones); the backward recomputes each odd layer once from the even layer below
it, then runs the recurrence backwards with the adjoint kernels:

  readout   g_n = 2 s^2 prod(t!) dL/dp * psi_n            (p = |s psi_n|^2 prod t!)
  layer k   g_{k-1}[t] = sum_j conj(U[j, i_k]) g_k[t + e_j]
            dL/dU[j, i_k] += sum_t conj(psi_{k-1}[t]) g_k[t + e_j]

Gradients follow torch's convention for complex leaves (U.grad = dL/dRe U + i dL/dIm U), as the
reference's autograd does; only the input columns i_k receive gradient.

  SLOSProbs  probabilities of the last layer (the fused readout), the drop-in probs(U);
  SLOSRaw    the raw (deferred-normalisation) amplitudes of layer `layer` for input columns
             `modes`: amplitudes(U) and the interfering subsets of the x-model;
  LayerStep  one layer step with a caller's coefficient matrix (the probability-level raise of the
             x-model runs as a complex layer step with the column |U[j, col]|^2).
"""
from __future__ import annotations

import torch

from .. import _native as nat


def _layer_flags(kernel) -> int:
    """The layer-step flags of the kernel's forward: compensated sums carry every layer as a hi/lo
    pair (as slos_forward does), so that the autograd forward equals the plain one bit for bit."""
    return (nat.FLAG_COMPENSATED | nat.FLAG_HILO) if kernel._flags() & nat.FLAG_COMPENSATED else 0


def _hi(psi: torch.Tensor) -> torch.Tensor:
    """The high parts [B, S] of a hi/lo layer [2, B, S] (a plain layer as is)."""
    return psi[0] if psi.dim() == 3 else psi


def _checkpointed_forward(plan, Ug: torch.Tensor, modes, layer: int, cf: int):
    """Layers 1..layer-1, keeping layer 0 and the even layers: (saved, psi_{layer-1})."""
    psi = torch.ones((Ug.shape[0], 1), dtype=Ug.dtype, device=Ug.device)
    if cf & nat.FLAG_HILO:
        psi = torch.stack([psi, torch.zeros_like(psi)])
    saved = {0: psi}
    for k in range(1, layer):
        psi = plan.layer_step(Ug, k, modes[k - 1], psi, nat.OUT_RAW, flags=cf)
        if k % 2 == 0:
            saved[k] = psi
    return saved, psi


def _adjoint_layers(plan, Ug: torch.Tensor, modes, layer: int, saved: dict, cf: int, g: torch.Tensor,
                    psi_top: torch.Tensor) -> torch.Tensor:
    """dL/dU (complex64 or complex128, like Ug) from g = dL/dpsi_layer, layers `layer` .. 1; psi_top = psi_{layer-1}."""

    def psi_of(k: int) -> torch.Tensor:
        if k in saved:
            return saved[k]
        return plan.layer_step(Ug, k, modes[k - 1], saved[k - 1], nat.OUT_RAW, flags=cf)

    gU = torch.zeros_like(Ug)
    psi_prev = psi_top
    for k in range(layer, 0, -1):
        if k < layer:
            psi_prev = psi_of(k - 1)
        g = plan.adjoint_layer(Ug, k, modes[k - 1], _hi(psi_prev), g, gU, want_prev=k > 1)
    return gU


# grad_p counts as sparse when at most one entry in SPARSE_READOUT of an evenly strided sample of
# about _SAMPLE entries (the whole of a smaller grad_p) is nonzero. The choice sets only the speed:
# both readout adjoints give the same gradient. (Counting all 4.3e9 entries at n = 11, m = 32 would
# cost about as much as the recompute it avoids.)
# Below SPARSE_MIN entries the recompute of layer n is cheap and the dense pass is taken without a
# look (the look costs a device synchronisation).
SPARSE_READOUT = 64
SPARSE_MIN = 1 << 24
_SAMPLE = 1 << 20


def _sparse(grad_p: torch.Tensor) -> bool:
    flat = grad_p.reshape(-1)
    if flat.numel() < SPARSE_MIN:
        return False
    sample = flat[::max(1, flat.numel() // _SAMPLE)]
    return int(torch.count_nonzero(sample)) * SPARSE_READOUT <= sample.numel()


def _user_grad(gU: torch.Tensor, ctx) -> torch.Tensor:
    gU = gU.to(device=ctx.user_device, dtype=ctx.user_dtype)
    return gU.squeeze(0) if len(ctx.user_shape) == 2 else gU


class SLOSProbs(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U_user: torch.Tensor, Ug: torch.Tensor, kernel) -> torch.Tensor:
        plan, modes, n = kernel._plan, kernel.input_modes, kernel.n
        cf = _layer_flags(kernel)                        # every layer as in the forward without grad
        saved, psi = _checkpointed_forward(plan, Ug, modes, n, cf)
        p = plan.layer_step(Ug, n, modes[n - 1], psi, nat.OUT_PROBS, scale=kernel._input_norm, flags=cf)
        ctx.kernel, ctx.saved, ctx.Ug, ctx.cf = kernel, saved, Ug, cf
        ctx.user_shape, ctx.user_dtype, ctx.user_device = U_user.shape, U_user.dtype, U_user.device
        return p

    @staticmethod
    def backward(ctx, grad_p: torch.Tensor):
        kernel, saved, Ug, cf = ctx.kernel, ctx.saved, ctx.Ug, ctx.cf
        plan, modes, n = kernel._plan, kernel.input_modes, kernel.n
        rdt = torch.float64 if Ug.dtype == torch.complex128 else torch.float32
        grad_p = grad_p.to(device=Ug.device, dtype=rdt).contiguous()
        psi_prev = saved[n - 1] if (n - 1) in saved else plan.layer_step(Ug, n - 1, modes[n - 2], saved[n - 2],
                                                                          nat.OUT_RAW, flags=cf)
        if Ug.dtype == torch.complex64 and not cf and _sparse(grad_p):
            # a loss over a few outputs: their amplitudes pulled from layer n-1, no recompute of layer n
            g = plan.readout_adjoint_pull(Ug, psi_prev, grad_p, kernel._input_norm)
        else:
            amps = _hi(plan.layer_step(Ug, n, modes[n - 1], psi_prev, nat.OUT_RAW, flags=cf))
            g = plan.readout_adjoint(amps, grad_p, kernel._input_norm)
            del amps
        gU = _adjoint_layers(plan, Ug, modes, n, saved, cf, g, psi_prev)
        ctx.saved = None
        return _user_grad(gU, ctx), None, None


class SLOSRaw(torch.autograd.Function):
    """Raw amplitudes [B, S_layer] (Ug's dtype) of layer `layer` for input columns `modes`."""

    @staticmethod
    def forward(ctx, U_user: torch.Tensor, Ug: torch.Tensor, kernel, modes, layer: int) -> torch.Tensor:
        plan = kernel._plan
        modes = [int(c) for c in modes]
        cf = _layer_flags(kernel)
        saved, psi = _checkpointed_forward(plan, Ug, modes, layer, cf)
        out = _hi(plan.layer_step(Ug, layer, modes[layer - 1], psi, nat.OUT_RAW, flags=cf))
        ctx.kernel, ctx.saved, ctx.Ug, ctx.cf, ctx.modes, ctx.layer = kernel, saved, Ug, cf, modes, layer
        ctx.user_shape, ctx.user_dtype, ctx.user_device = U_user.shape, U_user.dtype, U_user.device
        return out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        kernel, saved, Ug, cf, modes, layer = ctx.kernel, ctx.saved, ctx.Ug, ctx.cf, ctx.modes, ctx.layer
        plan = kernel._plan
        g = grad_out.to(device=Ug.device, dtype=Ug.dtype).contiguous()
        psi_prev = saved[layer - 1] if (layer - 1) in saved else plan.layer_step(
            Ug, layer - 1, modes[layer - 2], saved[layer - 2], nat.OUT_RAW, flags=cf)
        gU = _adjoint_layers(plan, Ug, modes, layer, saved, cf, g, psi_prev)
        ctx.saved = None
        return _user_grad(gU, ctx), None, None, None, None


class LayerStep(torch.autograd.Function):
    """psi_k = layer k of the recurrence from psi_{k-1} with coefficient column W[:, :, col]
    (complex64 or complex128 [B, m, m], differentiable) — one layer with its adjoint; psi_{k-1} is kept."""

    @staticmethod
    def forward(ctx, W: torch.Tensor, psi_prev: torch.Tensor, plan, k: int, col: int) -> torch.Tensor:
        W = W.contiguous()
        psi_prev = psi_prev.contiguous()
        ctx.save_for_backward(W, psi_prev)
        ctx.plan, ctx.k, ctx.col = plan, k, col
        return plan.layer_step(W, k, col, psi_prev, nat.OUT_RAW)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        W, psi_prev = ctx.saved_tensors
        gW = torch.zeros_like(W)
        gprev = ctx.plan.adjoint_layer(W, ctx.k, ctx.col, psi_prev, g.to(W.dtype).contiguous(), gW,
                                       want_prev=True)
        return gW, gprev, None, None, None
