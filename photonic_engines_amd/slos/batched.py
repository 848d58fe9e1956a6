"""Uneven batches: many SLOS instances of several shapes in one launch per photon layer.

The reference batches one shape at a time: `FockSLOSKernel.probs` takes U as [B, m, m] for a single
(n, m, input, mask) (photonic_engines/slos/kernel.py:208-215). `UnevenBatch` runs up to 8 such
kernels of different (n, m, input, mask) together. Layer K of every instance of every kernel with
n >= K is one persistent launch (slos_group_forward in include/slos_mi355x.h); a layer that only
one kernel reaches runs through that kernel's own single-plan launch (its 5-position-suffix and
group-order kernels). The forward can be captured once as a hipGraph and replayed (`capture` /
`replay`). A project-written autotuner (`tune`) picks each layer's launch: block or thread-per-state
group kernel, children per wave or states per thread, and persistent grid size, or one single-plan
launch per kernel, by timing the candidates on the caller's own buffers.

Each kernel's output is bit-for-bit the output of its own `FockSLOSKernel.probs(U)` (lossless,
float32 probabilities over `layers.final_states`).
"""
from __future__ import annotations

import ctypes
from collections.abc import Sequence

import torch

from .. import _native as nat
from .kernel import FockSLOSKernel

MAX_KERNELS = 8


class UnevenBatch:
    """Several FockSLOSKernel shapes run as one batch.

    `kernels` are FockSLOSKernel objects, at most 8 of them and all on the same device. They
    must stay alive while this batch is used. `probs(Us)` takes one [B_g, m_g, m_g] complex64
    tensor per kernel (B_g >= 1, different per kernel if wanted). It returns one [B_g, S_g]
    float32 tensor per kernel.
    """

    def __init__(self, kernels: Sequence[FockSLOSKernel], compensated: bool | None = None):
        kernels = list(kernels)
        if not 1 <= len(kernels) <= MAX_KERNELS:
            raise ValueError(f"an uneven batch holds 1 to {MAX_KERNELS} kernels, got {len(kernels)}")
        for k in kernels:
            if not isinstance(k, FockSLOSKernel):
                raise TypeError("UnevenBatch takes FockSLOSKernel objects")
            if k.normalization != "deferred":
                raise ValueError("UnevenBatch runs deferred normalisation only")
            if k.cdtype != torch.complex64:
                raise ValueError("UnevenBatch runs complex64 kernels (the group launch); a complex128 kernel "
                                 "computes through its own probs(U)")
        self.kernels = kernels
        comp = kernels[0].compensated if compensated is None else bool(compensated)
        self.flags = nat.FLAG_COMPENSATED if comp else 0
        self._lib = nat.lib()
        plans = (ctypes.c_void_p * len(kernels))(*[k._plan._h.value for k in kernels])
        h = ctypes.c_void_p()
        nat.check(self._lib.slos_group_create(len(kernels), plans, ctypes.byref(h)), "slos_group_create")
        self._h = h
        self._static = None

    @classmethod
    def from_specs(cls, specs, device: str = "cuda", compensated: bool = False) -> "UnevenBatch":
        """specs: (n, m, input_occupation) or (n, m, input_occupation, mask) per kernel."""
        ks = [FockSLOSKernel(s[0], s[1], list(s[2]), mask=(s[3] if len(s) > 3 else "fock"), device=device,
                             compensated=compensated) for s in specs]
        return cls(ks)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.slos_group_destroy(h)
            self._h = None

    # ---- argument marshalling ----------------------------------------------------------------
    def _prepare(self, Us: Sequence[torch.Tensor]):
        if len(Us) != len(self.kernels):
            raise ValueError(f"expected {len(self.kernels)} unitary batches, got {len(Us)}")
        out = []
        dev = None
        for k, U in zip(self.kernels, Us):
            if U.dim() == 2:
                U = U.unsqueeze(0)
            if U.dim() != 3 or U.shape[-2:] != (k.m, k.m):
                raise ValueError(f"U must be [B, {k.m}, {k.m}] for kernel (n={k.n}, m={k.m}), got {tuple(U.shape)}")
            d = nat.require_gpu(U.device if U.device.type == "cuda" else k.device)
            if dev is None:
                dev = d
            elif d != dev:
                raise ValueError("all unitary batches must be on one device")
            out.append(U.to(device=d, dtype=torch.complex64).contiguous())
        return out, dev

    def _args(self, Us, outs, ws):
        G = len(self.kernels)
        U_p = (ctypes.c_void_p * G)(*[u.data_ptr() for u in Us])
        B = (ctypes.c_int64 * G)(*[u.shape[0] for u in Us])
        O_p = (ctypes.c_void_p * G)(*[o.data_ptr() for o in outs])
        sc = (ctypes.c_double * G)(*[k._input_norm for k in self.kernels])
        return U_p, B, O_p, sc, ctypes.c_void_p(ws.data_ptr()), ws.numel()

    def workspace_bytes(self, batches: Sequence[int]) -> int:
        B = (ctypes.c_int64 * len(self.kernels))(*[int(b) for b in batches])
        out = ctypes.c_size_t()
        nat.check(self._lib.slos_group_workspace(self._h, B, self.flags, ctypes.byref(out)), "slos_group_workspace")
        return int(out.value)

    def _buffers(self, Us, dev):
        ws = torch.empty(max(1, self.workspace_bytes([u.shape[0] for u in Us])), dtype=torch.uint8, device=dev)
        outs = [torch.empty((u.shape[0], k.output_size()), dtype=torch.float32, device=dev)
                for k, u in zip(self.kernels, Us)]
        return ws, outs

    # ---- forward, tuning, graphs --------------------------------------------------------------
    def probs(self, Us: Sequence[torch.Tensor]) -> list[torch.Tensor]:
        """One eager forward of every kernel over its own unitary batch."""
        Us, dev = self._prepare(Us)
        ws, outs = self._buffers(Us, dev)
        with torch.cuda.device(dev):
            U_p, B, O_p, sc, wsp, wsn = self._args(Us, outs, ws)
            nat.check(self._lib.slos_group_forward(self._h, U_p, B, O_p, sc, self.flags, wsp, wsn,
                                                   nat.stream_handle(dev)), "slos_group_forward")
        return outs

    def tune(self, Us: Sequence[torch.Tensor], reps: int = 3) -> list[dict]:
        """Time the candidate launches of every layer on these inputs and keep the fastest; returns
        the chosen launch per layer (and leaves the last forward's output discarded)."""
        Us, dev = self._prepare(Us)
        ws, outs = self._buffers(Us, dev)
        with torch.cuda.device(dev):
            U_p, B, O_p, sc, wsp, wsn = self._args(Us, outs, ws)
            nat.check(self._lib.slos_group_tune(self._h, U_p, B, O_p, sc, self.flags, wsp, wsn, int(reps),
                                                nat.stream_handle(dev)), "slos_group_tune")
        self._static = None
        return self.config()

    def config(self) -> list[dict]:
        """The launch each layer uses: kernel ("block", "thread", "per_plan", "default"), per_wave, grid."""
        res = []
        for K in range(1, max(k.n for k in self.kernels) + 1):
            blk, pw, gr = ctypes.c_int(), ctypes.c_uint64(), ctypes.c_uint32()
            nat.check(self._lib.slos_group_config(self._h, K, ctypes.byref(blk), ctypes.byref(pw), ctypes.byref(gr)),
                      "slos_group_config")
            res.append({"layer": K, "kernel": {1: "block", 0: "thread", 2: "per_plan", -1: "default"}[blk.value],
                        "per_wave": int(pw.value), "grid": int(gr.value)})
        return res

    def capture(self, Us: Sequence[torch.Tensor]) -> None:
        """Capture one forward as a hipGraph over internal copies of these unitary batches; later
        `replay(Us)` copies new unitaries of the same shapes in and relaunches the graph."""
        Us, dev = self._prepare(Us)
        U_s = [u.clone() for u in Us]
        ws, outs = self._buffers(U_s, dev)
        with torch.cuda.device(dev):
            # one eager run first: device tables exist before the capture (nothing allocates inside it)
            U_p, B, O_p, sc, wsp, wsn = self._args(U_s, outs, ws)
            nat.check(self._lib.slos_group_forward(self._h, U_p, B, O_p, sc, self.flags, wsp, wsn,
                                                   nat.stream_handle(dev)), "slos_group_forward")
            torch.cuda.synchronize(dev)
            nat.check(self._lib.slos_group_capture(self._h, U_p, B, O_p, sc, self.flags, wsp, wsn),
                      "slos_group_capture")
        self._static = (dev, U_s, ws, outs)

    def replay(self, Us: Sequence[torch.Tensor] | None = None) -> list[torch.Tensor]:
        """Relaunch the captured forward on the current stream (after copying `Us` in, if given).
        Returns the captured output tensors, overwritten by this replay."""
        if self._static is None:
            raise RuntimeError("no captured forward: call capture(Us) first")
        dev, U_s, ws, outs = self._static
        if Us is not None:
            if len(Us) != len(U_s):
                raise ValueError(f"expected {len(U_s)} unitary batches, got {len(Us)}")
            for dst, src in zip(U_s, Us):
                dst.copy_(src.reshape(dst.shape))
        with torch.cuda.device(dev):
            nat.check(self._lib.slos_group_replay(self._h, nat.stream_handle(dev)), "slos_group_replay")
        return outs
