"""FockSLOSKernel on MI355X: drop-in for photonic_engines/slos/kernel.py:83-293.

Same constructor arguments, the same output order (layers.final_states) and the same output
tensors. The recurrence (deferred normalisation, kernel.py:127-141)

    psi_{k+1}[c] = sum_{j distinct in c} U[b, j, i_{k+1}] * psi_k[c - e_j]

runs as one hand-written gfx950 HIP kernel launch per photon layer (pull form: every child sums
its parents, whose lex ranks are computed on the fly; nothing is stored but the two live layers),
and the last launch writes |amp * input_norm|^2 * prod_j t_j! directly as float32.

Device model: the engine always computes on an AMD GPU. `device` is where U is taken from and
where results are returned; a CPU `device` (the reference's default) therefore means host-resident
inputs and outputs around the same GPU computation. Without a visible GPU the engine raises.

Precision: complex64 end to end (float32 probabilities); `compensated=True` (keyword) accumulates
every child sum with fma-split products and two-sum (a compensated dot product), as an accuracy
option. `dtype=torch.complex128` computes every layer in complex128 and the probabilities in
float64 (kernel.py:117-118), on the thread-per-state FP64 kernels.
"""
from __future__ import annotations

import math
from collections.abc import Sequence

import numpy as np
import torch

from .. import _native as nat
from ..imperfections import LossyDistribution  # noqa: F401  (re-exported like the reference)
from ..masks import Mask
from ._plan import MAX_ROWS, Plan, _row_ranges
from .basis import RankedLayers, get_layers


def _maybe_thin(layers, p: torch.Tensor, loss):
    """loss None / >= 1: p itself; else its uniform-loss thinning (kernel.py:32-37)."""
    if loss is None or (not isinstance(loss, torch.Tensor) and float(loss) >= 1.0):
        return p
    from ..imperfections import thin_distribution
    return thin_distribution(layers, p, loss)


class FockSLOSKernel:
    """Precompiled SLOS evaluator for a fixed (n, m, mask, input); `probs(U)` is the timed op.

    U is [m, m] or [B, m, m]; probabilities come back as [S] or [B, S] rdtype (float32; float64 for complex128) over the final
    (masked) basis in layers.final_states order.
    """

    def __init__(self, n: int, m: int, input_occupation: Sequence[int],
                 mask: str = Mask.FOCK.value, device: str = "cpu",
                 dtype: torch.dtype = torch.complex64,
                 normalization: str = "deferred",
                 gradient_checkpointing: bool = False,
                 layout: str = "mode", layers=None, *, compensated: bool = False):
        # argument checks and their order follow kernel.py:96-126
        if layout not in ("mode", "edge"):
            raise ValueError("layout must be 'mode' or 'edge'")
        if layout == "mode" and (gradient_checkpointing or normalization != "deferred"):
            layout = "edge"
        self.layout = layout          # both layouts run the same pull kernel (no edge tables exist)
        if normalization not in ("deferred", "immediate"):
            raise ValueError(f"normalization must be 'deferred' or 'immediate', got {normalization!r}")
        if gradient_checkpointing and normalization != "deferred":
            raise ValueError("gradient_checkpointing requires deferred normalization")
        if dtype not in (torch.complex64, torch.complex128):
            raise ValueError(f"dtype must be torch.complex64 or torch.complex128, got {dtype}")
        if compensated and dtype == torch.complex128:
            raise ValueError("compensated=True applies to complex64; complex128 sums are FP64")
        self.gradient_checkpointing = gradient_checkpointing
        self.normalization = normalization
        self.compensated = bool(compensated)
        self.n, self.m, self.mask = int(n), int(m), Mask(mask).value
        self.device = torch.device(device)
        self.cdtype = dtype
        self.rdtype = torch.float32 if dtype == torch.complex64 else torch.float64
        self.layers: RankedLayers = layers if layers is not None else get_layers(n, m, self.mask, device=str(self.device))
        if self.layers.output_size() == 0:
            raise ValueError(f"empty output space for n={n} m={m} mask={mask}")
        self.input_modes: list[int] = [j for j, c in enumerate(input_occupation) for _ in range(int(c))]
        if len(self.input_modes) != n:
            raise ValueError(f"input has {len(self.input_modes)} photons, expected n={n}")
        norm = 1.0
        for c in input_occupation:
            norm *= math.factorial(int(c))
        self._input_norm = 1.0 / math.sqrt(norm)
        self._input_norm_is_one = norm == 1
        # prod t_j! == 1 for every state exactly when no state can hold two photons in a mode
        self._output_norm_is_one = self.mask in ("unbunched", "dual_rail", "sparse_1in4") or self.n < 2
        self._plan = Plan(self.n, self.m, self.mask, self.input_modes)
        self._output_prob_norm = None

    # ---- reference attributes ------------------------------------------------------------------
    @property
    def output_prob_norm(self) -> torch.Tensor:
        """prod_j t_j! per final state as an [S] rdtype tensor on `device` (kernel.py:139),
        built on first access; the engine itself never reads it (the norm is fused)."""
        if self._output_prob_norm is None:
            self._output_prob_norm = torch.as_tensor(self.layers.output_prob_norms()).to(
                dtype=self.rdtype, device=self.device)
        return self._output_prob_norm

    def output_size(self) -> int:
        """Size S of the output distribution (kernel.py:169-171)."""
        return self.layers.output_size()

    def memory_plan(self, batch: int = 1) -> dict:
        """Device bytes a probs(U) of `batch` unitaries needs (output, workspace, unitaries, total),
        by formula; probs raises torch.cuda.OutOfMemoryError before launching when they exceed
        the free device memory."""
        from ._plan import _dtype_flags
        return self._plan.memory_plan(batch, flags=_dtype_flags(self.cdtype, self._flags()))

    def _flags(self) -> int:
        f = nat.FLAG_COMPENSATED if self.compensated else 0
        if self.normalization == "immediate":
            f |= nat.FLAG_IMMEDIATE
        return f

    def _add_oom_hint(self, e: BaseException) -> None:
        """OOM guidance, a hint and never a retry (kernel.py:173-189)."""
        if not self.gradient_checkpointing and torch.is_grad_enabled():
            hint = ("hint: FockSLOSKernel(gradient_checkpointing=True) reduces peak GPU memory "
                    "in training, at the cost of a slower backward pass.")
            if hasattr(e, "add_note"):
                e.add_note(hint)
            elif len(e.args) == 1 and isinstance(e.args[0], str):
                e.args = (f"{e.args[0]}\n{hint}",)
            else:
                e.args = (*e.args, hint)

    def _prepare(self, U: torch.Tensor):
        """U [m,m] | [B,m,m] (any device) -> (contiguous [B,m,m] of the kernel's dtype on the GPU, batched)."""
        batched = U.dim() == 3
        if not batched:
            U = U.unsqueeze(0)
        if U.shape[-2:] != (self.m, self.m):
            raise ValueError(f"U must be [m, m] or [B, m, m] with m={self.m}, got {tuple(U.shape)}")
        dev = nat.require_gpu(U.device if U.device.type == "cuda" else self.device)
        return U.to(device=dev, dtype=self.cdtype).contiguous(), batched

    def _out(self, t: torch.Tensor, batched: bool) -> torch.Tensor:
        t = t if batched else t.squeeze(0)
        return t.to(self.device) if self.device.type != "cuda" else t

    # ---- amplitudes ------------------------------------------------------------------------------
    def amplitudes(self, U: torch.Tensor) -> torch.Tensor:
        """Output amplitudes over the final basis, [S] / [B, S] of the kernel's dtype (kernel.py:191-199):
        the raw (deferred-normalisation) amplitude times the input norm."""
        try:
            return self._amplitudes(U)
        except torch.cuda.OutOfMemoryError as e:
            self._add_oom_hint(e)
            raise

    def _amplitudes(self, U: torch.Tensor, input_modes: Sequence[int] | None = None) -> torch.Tensor:
        """With `input_modes` (g <= n columns) the recurrence stops on layer g of the same basis,
        raw, without the input norm (kernel.py:201-206, used by slos/species.py:94)."""
        if U.requires_grad and torch.is_grad_enabled() and (self.n if input_modes is None else len(input_modes)) > 0:
            return self._amplitudes_grad(U, input_modes)
        Ug, batched = self._prepare(U)
        if input_modes is None:
            out = self._plan.forward(Ug, out_kind=nat.OUT_AMPS, scale=self._input_norm, flags=self._flags())
        else:
            modes = [int(i) for i in input_modes]
            if not modes:
                out = torch.ones((Ug.shape[0], 1), dtype=self.cdtype, device=Ug.device)
            else:
                out = self._plan.forward(Ug, modes=modes, layer=len(modes), out_kind=nat.OUT_AMPS,
                                         scale=1.0, flags=self._flags() & ~nat.FLAG_IMMEDIATE)
        return self._out(out, batched)

    def _amplitudes_grad(self, U: torch.Tensor, input_modes: Sequence[int] | None) -> torch.Tensor:
        """_amplitudes with U requiring a gradient: the raw layer through the hand-written adjoint
        (SLOSRaw), then the input norm and, for immediate normalisation, sqrt(prod t_j!) as torch
        products (kernel.py:191-260 under autograd)."""
        from .autograd import SLOSRaw
        Ug, batched = self._prepare(U)
        modes = self.input_modes if input_modes is None else [int(i) for i in input_modes]
        if not modes:
            out = torch.ones((Ug.shape[0], 1), dtype=self.cdtype, device=Ug.device)
            return self._out(out, batched)
        raw = SLOSRaw.apply(U, Ug, self, modes, len(modes))
        if input_modes is None:
            if self._input_norm != 1.0:
                raw = raw * self._input_norm
            if self.normalization == "immediate" and not self._output_norm_is_one:
                raw = raw * torch.sqrt(self.output_prob_norm.to(raw.device))
        return self._out(raw, batched)

    # ---- probabilities ------------------------------------------------------------------------------
    def probs(self, U: torch.Tensor, loss=None, indist=None, order: int | None = None):
        """Output probabilities over the final basis (kernel.py:262-283). With `loss=eta` the
        uniform-loss LossyDistribution over 0..n photons is returned instead. With `indist=x`
        (partial distinguishability, x-model) the distribution is the mixture over the interfering
        subsets of the photons (slos/species.py:123-146, `species.x_mixture`); `order=b` keeps at
        most b distinguishable photons; a tensor x gets a gradient; loss is applied after."""
        if indist is not None:
            from ..imperfections import spec
            if spec(indist) is not None:
                from .species import x_mixture
                try:
                    p = x_mixture(self, U, indist, order=order)
                except torch.cuda.OutOfMemoryError as e:
                    self._add_oom_hint(e)
                    raise
                return _maybe_thin(self.layers, p, loss)
        try:
            if U.requires_grad and torch.is_grad_enabled() and self.n > 0:  # (n = 0: no dependence on U)
                from .autograd import SLOSProbs
                Ug, batched = self._prepare(U)
                p = SLOSProbs.apply(U, Ug, self)
                p = p if batched else p.squeeze(0)
                p = p.to(self.device) if self.device.type != "cuda" else p
            else:
                Ug, batched = self._prepare(U)
                # a loss without gradient: the readout applies the weight of layer n (|amp s|^2 w_n)
                w32 = None
                if loss is not None and not (isinstance(loss, torch.Tensor) and loss.requires_grad):
                    from ..imperfections import carried_weights
                    w32 = carried_weights(self.n, loss, Ug.device, dtype=self.rdtype)
                scale = self._input_norm * (math.sqrt(float(w32[self.n])) if w32 is not None else 1.0)
                p = self._out(self._plan.forward(Ug, out_kind=nat.OUT_PROBS, scale=scale, flags=self._flags()),
                              batched)
                if w32 is not None:
                    from ..imperfections import thin_distribution
                    return thin_distribution(self.layers, p, loss, _top_weighted=True)
        except torch.cuda.OutOfMemoryError as e:
            self._add_oom_hint(e)
            raise
        return _maybe_thin(self.layers, p, loss)

    def total_mass(self, p: torch.Tensor) -> torch.Tensor:
        """FP64 sum of a probability vector [S] / [B, S] on the GPU (deterministic, no atomics)."""
        q = p if p.dim() == 2 else p.unsqueeze(0)
        dev = nat.require_gpu(q.device if q.device.type == "cuda" else self.device)
        rdt = torch.float64 if q.dtype == torch.float64 else torch.float32
        out = self._plan.total_mass(q.to(device=dev, dtype=rdt).contiguous())
        return out if p.dim() == 2 else out.squeeze(0)

    def spot_probs(self, U: torch.Tensor, indices) -> torch.Tensor:
        """FP64 Ryser probabilities of the final states `indices` for one unitary U (complex128
        inputs kept exact): the accuracy oracle at sizes with no full FP64 run."""
        dev = nat.require_gpu(U.device if U.device.type == "cuda" else self.device)
        idx = torch.as_tensor(indices, dtype=torch.int64, device=dev)
        return self._plan.spot_probs(U.to(device=dev, dtype=torch.complex128), idx)

    # ---- grouping ----------------------------------------------------------------------------------
    def group_matrix(self, groups: torch.Tensor) -> "GroupedSLOSKernel":
        """Fold a group assignment (LongTensor [S] -> bin id) into the kernel: probabilities are
        summed into bins after the fused readout (kernel.py:286-294)."""
        return GroupedSLOSKernel(self, groups)


class GroupedSLOSKernel:
    """Grouped probabilities (kernel.py:584-620): the returned output is [B, n_groups] instead of
    [B, S]. The full recurrence and the fused |amp|^2 * prod t! readout run as for the wrapped
    kernel; the [B, S] probabilities are then summed into their groups by slos_group_sums, a keyed
    reduction without atomics (FP64 sums in a fixed order, reproducible run to run). The states
    are sorted by group once, here, into chunks of at most 4096 entries."""

    _CHUNK = 4096

    def __init__(self, kernel: FockSLOSKernel, groups: torch.Tensor):
        self.kernel = kernel
        g = torch.as_tensor(groups).to(torch.int64)
        if g.dim() != 1 or g.numel() != kernel.output_size():
            raise ValueError(f"groups must be [S] with S = {kernel.output_size()}")
        if g.numel() and int(g.min()) < 0:
            raise ValueError("group ids must be >= 0")
        self.groups = g.to(kernel.device)
        self.n_groups = int(self.groups.max().item()) + 1 if self.groups.numel() else 0
        self._sort = None

    def _sorted(self, dev: torch.device):
        """(order, chunks, first) on `dev`: the states in group order, the [begin, end) chunks of
        that list, and each group's first chunk (slos_group_sums)."""
        if self._sort is None or self._sort[0].device != dev:
            g = self.groups.to(dev)
            order = torch.argsort(g, stable=True)
            counts = torch.bincount(g, minlength=self.n_groups)
            nch = (counts + self._CHUNK - 1) // self._CHUNK
            first = torch.zeros(self.n_groups + 1, dtype=torch.int64, device=dev)
            first[1:] = torch.cumsum(nch, 0)
            seg = torch.cumsum(counts, 0) - counts
            cg = torch.repeat_interleave(torch.arange(self.n_groups, device=dev), nch)
            j = torch.arange(cg.numel(), device=dev) - first[cg]
            begin = seg[cg] + self._CHUNK * j
            end = torch.minimum(begin + self._CHUNK, seg[cg] + counts[cg])
            chunks = torch.stack([begin, end], 1).contiguous()
            self._sort = (order.contiguous(), chunks, first.to(torch.int32).contiguous())
        return self._sort

    def output_size(self) -> int:
        """Number of groups, the width of the returned distribution."""
        return self.n_groups

    def probs(self, U: torch.Tensor) -> torch.Tensor:
        """U [m, m] or [B, m, m] -> [n_groups] or [B, n_groups] (the kernel's rdtype), normalised like probs(U)."""
        p = self.kernel.probs(U)
        batched = p.dim() == 2
        q = (p if batched else p.unsqueeze(0)).contiguous()
        if q.requires_grad or q.device.type != "cuda":
            # autograd, or a kernel on the CPU device: the differentiable torch reduction
            idx = self.groups.to(q.device)
            out = torch.zeros((q.shape[0], self.n_groups), dtype=q.dtype, device=q.device)
            out = out.scatter_add(1, idx.unsqueeze(0).expand(q.shape[0], -1), q)
            return out if batched else out.squeeze(0)
        order, chunks, first = self._sorted(q.device)
        B = q.shape[0]
        n_chunks = chunks.shape[0]
        out = torch.empty((B, self.n_groups), dtype=q.dtype, device=q.device)
        ws = torch.empty(max(min(B, MAX_ROWS) * n_chunks, 1), dtype=torch.float64, device=q.device)
        fn = nat.lib().slos_group_sums_f64 if q.dtype == torch.float64 else nat.lib().slos_group_sums
        for i, j in _row_ranges(B):
            nat.check(fn(nat.ptr(q[i:j]), j - i, q.shape[1], nat.ptr(order), nat.ptr(chunks), n_chunks, nat.ptr(first),
                         self.n_groups, nat.ptr(out[i:j]), nat.ptr(ws), ws.numel() * 8, nat.stream_handle(q.device)),
                      "slos_group_sums")
        return out if batched else out.squeeze(0)


class SuperpositionSLOSKernel:
    """EBS kernel (kernel.py:296-431): K coherent input components of n photons each over one shared
    output basis. Every component is one forward of the same plan with its own input columns (the
    plan stores no edges, so nothing is rebuilt per component); the weighted amplitudes are summed
    on the GPU and squared with the deferred prod t_j!. `chunk_size` is accepted for signature
    compatibility: one component's layers are live at a time whatever its value."""

    def __init__(self, n: int, m: int, input_occupations, mask: str = Mask.FOCK.value, device: str = "cpu",
                 dtype: torch.dtype = torch.complex64, weights=None):
        if dtype not in (torch.complex64, torch.complex128):
            raise ValueError(f"dtype must be torch.complex64 or torch.complex128, got {dtype}")
        self.n, self.m, self.mask = int(n), int(m), Mask(mask).value
        self.device = torch.device(device)
        self.cdtype = dtype
        self.rdtype = torch.float32 if dtype == torch.complex64 else torch.float64
        self.layers: RankedLayers = get_layers(n, m, self.mask, device=str(self.device))
        if self.layers.output_size() == 0:
            raise ValueError(f"empty output space for n={n} m={m} mask={mask}")
        occ = [[int(c) for c in row] for row in input_occupations]
        self.K = len(occ)
        if self.K == 0:
            raise ValueError("need at least one input component")
        if any(len(r) != m or sum(r) != n for r in occ):
            raise ValueError(f"input_occupations must be [K, {m}] with {n} photons per component")
        self._modes = [[j for j, c in enumerate(r) for _ in range(c)] for r in occ]
        self._norms = [1.0 / math.sqrt(math.prod(math.factorial(c) for c in r)) for r in occ]
        self.default_weights = None if weights is None else torch.as_tensor(
            np.asarray(weights), dtype=dtype, device=self.device)
        self._plan = Plan(self.n, self.m, self.mask, self._modes[0])
        self._output_norm_is_one = self.mask in ("unbunched", "dual_rail", "sparse_1in4") or self.n < 2
        self._prob_norm = None

    @classmethod
    def from_state(cls, state, mask: str = Mask.FOCK.value, device: str = "cpu", dtype: torch.dtype = torch.complex64):
        """Kernel for a Fock superposition state (anything with n, m, occupations, amplitudes and
        norm()): its normalised amplitudes become the default weights."""
        return cls(state.n, state.m, state.occupations, mask=mask, device=device, dtype=dtype,
                   weights=np.asarray(state.amplitudes) / state.norm())

    def output_size(self) -> int:
        return self.layers.output_size()

    def _flags(self) -> int:
        return 0

    def amplitudes(self, U: torch.Tensor, weights=None, chunk_size: int | None = None) -> torch.Tensor:
        """Coherently summed output amplitudes [S] / [B, S] (the kernel's dtype): sum_k w_k * norm_k * raw_k.
        Default weights: the state's (from_state / weights=), else uniform 1/sqrt(K). Differentiable
        in U (the raw amplitudes through the hand-written adjoint, SLOSRaw) and in tensor weights."""
        grad_u = U.requires_grad and torch.is_grad_enabled()
        batched = U.dim() == 3
        Ub = (U if batched else U.unsqueeze(0))
        if Ub.shape[-2:] != (self.m, self.m):
            raise ValueError(f"U must be [m, m] or [B, m, m] with m={self.m}")
        dev = nat.require_gpu(U.device if U.device.type == "cuda" else self.device)
        cdt = self.cdtype
        Ub = Ub.to(device=dev, dtype=cdt).contiguous()
        if weights is None and self.default_weights is not None:
            w = self.default_weights
        elif weights is None:
            w = torch.full((self.K,), 1.0 / math.sqrt(self.K), dtype=cdt)
        elif isinstance(weights, torch.Tensor):
            w = weights.to(cdt)
        else:
            w = torch.as_tensor(np.asarray(weights), dtype=cdt)
        w = w.to(dev)
        if w.numel() != self.K:
            raise ValueError(f"weights must have K = {self.K} entries")
        # all components in one batched forward: row (k, b) holds U_b with component k's input
        # columns first, so modes 0..n-1 select them; the raw amplitudes then get the component's
        # input norm and weight and are summed over k
        K, B, m = self.K, Ub.shape[0], self.m
        cols = torch.as_tensor(self._modes, dtype=torch.int64, device=dev)            # [K, n]
        Ug = Ub[:, :, cols].permute(2, 0, 1, 3)                                     # [K, B, m, n]
        if self.n < m:
            Ug = torch.cat([Ug, torch.zeros((K, B, m, m - self.n), dtype=cdt, device=dev)], dim=3)
        Up = Ug.reshape(K * B, m, m).contiguous()
        if grad_u:
            from .autograd import SLOSRaw
            raw = SLOSRaw.apply(Up, Up.detach(), self, list(range(self.n)), self.n)
        else:
            raw = self._plan.forward(Up, modes=list(range(self.n)), layer=self.n, out_kind=nat.OUT_AMPS, scale=1.0)
        coef = (w * torch.as_tensor(self._norms, dtype=self.rdtype, device=dev)).to(cdt)
        amp = (raw.reshape(K, B, raw.shape[-1]) * coef.view(K, 1, 1)).sum(0)
        amp = amp if batched else amp.squeeze(0)
        return amp.to(self.device) if self.device.type != "cuda" else amp

    def probs(self, U: torch.Tensor, weights=None, chunk_size: int | None = None, loss=None):
        """Output distribution of the coherent superposition (`loss=eta`: thinned, as probs)."""
        amp = self.amplitudes(U, weights=weights, chunk_size=chunk_size)
        p = amp.real ** 2 + amp.imag ** 2
        if not self._output_norm_is_one:
            if self._prob_norm is None or self._prob_norm.device != p.device:  # prod t_j!, kept where p lives
                self._prob_norm = torch.as_tensor(self.layers.output_prob_norms()).to(device=p.device,
                                                                                      dtype=self.rdtype)
            p = p * self._prob_norm
        return _maybe_thin(self.layers, p, loss)
