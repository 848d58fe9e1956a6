"""Hand-written adjoint of FockSLOSKernel.probs behind torch.autograd.Function.

Replaces autograd through gather / mul / scatter_add_ (photonic_engines/slos/kernel.py:216-257)
and the opt-in recomputing layer (_CheckpointedLayer, kernel.py:40-80). Forward keeps only every
other layer (the even ones); backward recomputes each odd layer once from the even layer below it,
then runs the recurrence backwards with the adjoint kernels:

  readout   g_n = 2 s^2 prod(t!) dL/dp * psi_n            (p = |s psi_n|^2 prod t!)
  layer k   g_{k-1}[t] = sum_j conj(U[j, i_k]) g_k[t + e_j]
            dL/dU[j, i_k] += sum_t conj(psi_{k-1}[t]) g_k[t + e_j]

Gradients follow torch's convention for complex leaves (U.grad = dL/dRe U + i dL/dIm U), as the
reference's autograd does; only the input columns i_k receive gradient.
"""
from __future__ import annotations

import torch

from .. import _native as nat


class SLOSProbs(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U_user: torch.Tensor, Ug: torch.Tensor, kernel) -> torch.Tensor:
        plan, modes, n = kernel._plan, kernel.input_modes, kernel.n
        psi = torch.ones((Ug.shape[0], 1), dtype=torch.complex64, device=Ug.device)
        saved = {0: psi}
        for k in range(1, n):
            psi = plan.layer_step(Ug, k, modes[k - 1], psi, nat.OUT_RAW)
            if k % 2 == 0:
                saved[k] = psi
        p = plan.layer_step(Ug, n, modes[n - 1], psi, nat.OUT_PROBS, scale=kernel._input_norm,
                            flags=kernel._flags() & nat.FLAG_COMPENSATED)
        ctx.kernel, ctx.saved, ctx.Ug = kernel, saved, Ug
        ctx.user_shape, ctx.user_dtype, ctx.user_device = U_user.shape, U_user.dtype, U_user.device
        return p

    @staticmethod
    def backward(ctx, grad_p: torch.Tensor):
        kernel, saved, Ug = ctx.kernel, ctx.saved, ctx.Ug
        plan, modes, n = kernel._plan, kernel.input_modes, kernel.n
        grad_p = grad_p.to(device=Ug.device, dtype=torch.float32).contiguous()

        def layer(k: int) -> torch.Tensor:
            if k in saved:
                return saved[k]
            return plan.layer_step(Ug, k, modes[k - 1], saved[k - 1], nat.OUT_RAW)

        psi_prev = layer(n - 1)
        amps = plan.layer_step(Ug, n, modes[n - 1], psi_prev, nat.OUT_RAW)
        g = plan.readout_adjoint(amps, grad_p, kernel._input_norm)
        del amps
        gU = torch.zeros_like(Ug)
        for k in range(n, 0, -1):
            if k < n:
                psi_prev = layer(k - 1)
            g = plan.adjoint_layer(Ug, k, modes[k - 1], psi_prev, g, gU, want_prev=k > 1)
        ctx.saved = None
        gU = gU.to(device=ctx.user_device, dtype=ctx.user_dtype)
        if len(ctx.user_shape) == 2:
            gU = gU.squeeze(0)
        return gU, None, None
