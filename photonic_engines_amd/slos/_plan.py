"""Python handle of one native plan (slos_plan_* in include/slos_mi355x.h).

A plan fixes (n, m, mask, input modes) and owns only the counting tables; every large buffer
(workspace, outputs) is a torch tensor from the caching allocator, passed by pointer, and every
launch goes to torch's current stream on the compute device.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _native as nat


def _dtype_flags(cdtype: torch.dtype, flags: int) -> int:
    """The forward flags for data of complex dtype `cdtype`: complex128 adds FLAG_F64 (FP64
    kernels; compensated sums are an FP32 option and are refused with it)."""
    if cdtype == torch.complex128:
        if flags & (nat.FLAG_COMPENSATED | nat.FLAG_HILO):
            raise ValueError("compensated sums apply to complex64; complex128 sums are FP64")
        return flags | nat.FLAG_F64
    return flags & ~nat.FLAG_F64


# Batch rows one launch takes: the batch is the launch grid's y dimension. Larger batches run as
# consecutive launches over row ranges (every kernel treats its rows independently, so each row's
# result is the same as in a single launch); an empty batch launches nothing.
MAX_ROWS = 65535


def _row_ranges(B: int) -> list[tuple[int, int]]:
    return [(i, min(B, i + MAX_ROWS)) for i in range(0, B, MAX_ROWS)]


def _out_dtypes(cdtype: torch.dtype) -> tuple[torch.dtype, torch.dtype]:
    """(probability dtype, amplitude dtype) of a forward over `cdtype` data."""
    return (torch.float64, torch.complex128) if cdtype == torch.complex128 else (torch.float32, torch.complex64)


class Plan:
    def __init__(self, n: int, m: int, mask: str, input_modes):
        if mask not in nat.MASK_IDS:
            raise ValueError(f"unknown mask {mask!r}")
        self.n, self.m, self.mask = int(n), int(m), mask
        self.modes = [int(i) for i in input_modes]
        if len(self.modes) != self.n:
            raise ValueError(f"input has {len(self.modes)} photons, expected n={n}")
        arr = (ctypes.c_int32 * self.n)(*self.modes)
        h = ctypes.c_void_p()
        self._lib = nat.lib()
        nat.check(self._lib.slos_plan_create(self.n, self.m, nat.MASK_IDS[mask], arr, ctypes.byref(h)),
                  "slos_plan_create")
        self._h = h
        self.sizes = [self._layer_size(k) for k in range(self.n + 1)]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.slos_plan_destroy(h)
            self._h = None

    # ---- host queries ---------------------------------------------------------------------
    def _layer_size(self, k: int) -> int:
        out = ctypes.c_uint64()
        nat.check(self._lib.slos_layer_size(self._h, k, ctypes.byref(out)), "slos_layer_size")
        return int(out.value)

    def rank(self, k: int, positions) -> np.ndarray:
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.int32).reshape(-1, k) if k else
                                   np.zeros((len(positions), 0), dtype=np.int32))
        out = np.zeros(pos.shape[0], dtype=np.uint64)
        nat.check(self._lib.slos_rank(self._h, k, pos.ctypes.data_as(nat._i32p), pos.shape[0],
                                      out.ctypes.data_as(nat._u64p)), "slos_rank")
        return out

    def unrank(self, k: int, idx) -> np.ndarray:
        idx = np.ascontiguousarray(np.asarray(idx, dtype=np.uint64).reshape(-1))
        out = np.zeros((idx.shape[0], k), dtype=np.int32)
        nat.check(self._lib.slos_unrank(self._h, k, idx.ctypes.data_as(nat._u64p), idx.shape[0],
                                        out.ctypes.data_as(nat._i32p)), "slos_unrank")
        return out

    def prob_norms(self, k: int, first: int = 0, count: int | None = None) -> np.ndarray:
        count = self.sizes[k] - first if count is None else count
        out = np.zeros(count, dtype=np.float64)
        nat.check(self._lib.slos_prob_norms(self._h, k, first, count, out.ctypes.data_as(nat._f64p)),
                  "slos_prob_norms")
        return out

    # ---- device work --------------------------------------------------------------------------
    def workspace_bytes(self, batch: int, layer: int, flags: int = 0) -> int:
        out = ctypes.c_size_t()
        nat.check(self._lib.slos_forward_workspace(self._h, batch, layer, int(flags), ctypes.byref(out)),
                  "slos_forward_workspace")
        return int(out.value)

    def memory_plan(self, batch: int, layer: int | None = None, out_kind: int = nat.OUT_PROBS, flags: int = 0) -> dict:
        """Device bytes of one forward by formula, before anything is allocated: the output
        [B, S_layer] (4 B a probability, 8 B a complex64 amplitude; twice that with FLAG_F64), the
        forward's workspace (the vacuum row and the two ping-pong layers, slos_forward_workspace)
        and the unitaries (8 m^2 B each, 16 with FLAG_F64)."""
        layer = self.n if layer is None else int(layer)
        w = 2 if flags & nat.FLAG_F64 else 1
        out_b = int(batch) * self.sizes[layer] * (4 if out_kind == nat.OUT_PROBS else 8) * w
        ws = self.workspace_bytes(batch, layer, flags)
        u_b = int(batch) * self.m * self.m * 8 * w
        return {"output": out_b, "workspace": ws, "unitaries": u_b, "total": out_b + ws + u_b}

    # below this many bytes a forward is not checked (the query costs more than such a forward)
    PREFLIGHT_MIN = 1 << 30

    @staticmethod
    def preflight(need: int, dev: torch.device, what: str, free_bytes=None) -> None:
        """Raise torch.cuda.OutOfMemoryError (the reference's exception for this condition), with
        the sizes, before anything is allocated or launched, if `need` bytes exceed the device's
        free memory plus what torch's cache holds unused. `free_bytes()` replaces that query.
        Needs below PREFLIGHT_MIN are not checked (torch's own allocation error still applies)."""
        if free_bytes is not None:
            free = int(free_bytes())
        elif need < Plan.PREFLIGHT_MIN:
            return
        else:
            # torch's unused cache first (no driver call; repeated forwards find their buffers there),
            # the device query (which can take milliseconds) only when that is not enough
            free = int(torch.cuda.memory_reserved(dev) - torch.cuda.memory_allocated(dev))
            if need <= free:
                return
            free += int(torch.cuda.mem_get_info(dev)[0])
        if need > free:
            raise torch.cuda.OutOfMemoryError(
                f"{what}: needs {need / 2**30:.2f} GiB more on {dev}, {free / 2**30:.2f} GiB are free")

    def forward(self, U: torch.Tensor, *, modes=None, layer: int | None = None, out_kind: int = nat.OUT_PROBS,
                scale: float = 1.0, flags: int = 0, out: torch.Tensor | None = None) -> torch.Tensor:
        """U: complex64 [B, m, m] on the compute device -> [B, S_layer] float32 / complex64;
        complex128 U -> float64 / complex128 (FLAG_F64, FP64 throughout)."""
        layer = self.n if layer is None else int(layer)
        dev = U.device
        B = U.shape[0]
        if U.dtype not in (torch.complex64, torch.complex128) or not U.is_contiguous() or U.shape[1:] != (self.m, self.m):
            raise ValueError("U must be a contiguous complex64 or complex128 [B, m, m] tensor")
        flags = _dtype_flags(U.dtype, flags)
        S = self.sizes[layer]
        ws_bytes = self.workspace_bytes(max(1, min(B, MAX_ROWS)), layer, flags)
        rdt, cdt = _out_dtypes(U.dtype)
        out_bytes = 0 if out is not None else B * S * (rdt if out_kind == nat.OUT_PROBS else cdt).itemsize
        self.preflight(out_bytes + ws_bytes, dev, f"SLOS forward (n={self.n}, m={self.m}, batch {B}: "
                       f"output {out_bytes:,} B, workspace {ws_bytes:,} B)")
        odt = rdt if out_kind == nat.OUT_PROBS else cdt
        if out is None:
            out = torch.empty((B, S), dtype=odt, device=dev)
        elif out.shape != (B, S) or out.dtype != odt or out.device != dev or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous {odt} tensor of shape {(B, S)} on {dev}")
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        mp = None
        if modes is not None:
            mp = (ctypes.c_int32 * layer)(*[int(i) for i in modes])
        for i, j in _row_ranges(B):
            nat.check(self._lib.slos_forward(self._h, nat.ptr(U[i:j]), j - i, mp, layer, out_kind, nat.ptr(out[i:j]),
                                             float(scale), int(flags), nat.ptr(ws), ws.numel(),
                                             nat.stream_handle(dev)), "slos_forward")
        return out

    def layer_step(self, U: torch.Tensor, k: int, col: int, psi_prev: torch.Tensor, out_kind: int = nat.OUT_RAW,
                   scale: float = 1.0, flags: int = 0, out: torch.Tensor | None = None) -> torch.Tensor:
        """Layer k from layer k-1 (input column col): raw complex64, probs or scaled amplitudes,
        into `out` if given (contiguous [B, S_k] of the output's dtype on U's device). With
        FLAG_HILO the raw layers in and out are hi/lo pairs [2, B, S] (compensated sums carry
        their low-order parts)."""
        B = U.shape[0]
        S = self.sizes[k]
        flags = _dtype_flags(U.dtype, flags)
        if psi_prev.dtype != U.dtype:
            raise ValueError(f"psi_prev must be {U.dtype} like U")
        rdt, cdt = _out_dtypes(U.dtype)
        dtype = rdt if out_kind == nat.OUT_PROBS else cdt
        shape = (2, B, S) if (flags & nat.FLAG_HILO and out_kind == nat.OUT_RAW) else (B, S)
        if flags & nat.FLAG_HILO and (psi_prev.dim() != 3 or psi_prev.shape[0] != 2):
            raise ValueError("FLAG_HILO: psi_prev must be a hi/lo pair [2, B, S_{k-1}]")
        if out is None:
            out = torch.empty(shape, dtype=dtype, device=U.device)
        elif out.shape != shape or out.dtype != dtype or out.device != U.device or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous {dtype} tensor of shape {shape} on {U.device}")
        hilo_in = bool(flags & nat.FLAG_HILO)
        hilo_out = len(shape) == 3
        ranges = _row_ranges(B)
        whole = len(ranges) == 1
        for i, j in ranges:
            if whole:
                pin, pout = psi_prev, out
            else:
                # a hi/lo pair [2, B, S] is passed as one contiguous block: its row ranges are copies
                pin = psi_prev[:, i:j].contiguous() if hilo_in else psi_prev[i:j]
                pout = torch.empty((2, j - i, S), dtype=dtype, device=U.device) if hilo_out else out[i:j]
            nat.check(self._lib.slos_layer_step(self._h, nat.ptr(U[i:j]), j - i, k, int(col), nat.ptr(pin), out_kind,
                                                nat.ptr(pout), float(scale), int(flags),
                                                nat.stream_handle(U.device)), "slos_layer_step")
            if not whole and hilo_out:
                out[:, i:j].copy_(pout)
        return out

    def readout_adjoint(self, amps: torch.Tensor, grad_p: torch.Tensor, scale: float) -> torch.Tensor:
        g = torch.empty_like(amps)
        f64 = amps.dtype == torch.complex128
        fn = self._lib.slos_readout_adjoint_f64 if f64 else self._lib.slos_readout_adjoint
        grad_p = grad_p.to(torch.float64 if f64 else torch.float32).contiguous()
        for i, j in _row_ranges(amps.shape[0]):
            nat.check(fn(self._h, j - i, nat.ptr(amps[i:j]), nat.ptr(grad_p[i:j]), float(scale), nat.ptr(g[i:j]),
                         nat.stream_handle(amps.device)), "slos_readout_adjoint")
        return g

    def readout_adjoint_pull(self, U: torch.Tensor, psi_prev: torch.Tensor, grad_p: torch.Tensor,
                             scale: float) -> torch.Tensor:
        """The readout adjoint from layer n-1 (psi_prev [B, S_{n-1}] complex64, of U): zero where
        grad_p is zero, elsewhere with the state's amplitude pulled from its parents."""
        B = U.shape[0]
        grad_p = grad_p.to(torch.float32).contiguous()
        g = torch.empty((B, self.sizes[self.n]), dtype=torch.complex64, device=U.device)
        for i, j in _row_ranges(B):
            nat.check(self._lib.slos_readout_adjoint_pull(self._h, nat.ptr(U[i:j]), j - i, nat.ptr(psi_prev[i:j]),
                                                          nat.ptr(grad_p[i:j]), float(scale), nat.ptr(g[i:j]),
                                                          nat.stream_handle(U.device)), "slos_readout_adjoint_pull")
        return g

    def adjoint_layer(self, U: torch.Tensor, k: int, col: int, psi_prev: torch.Tensor, grad_k: torch.Tensor,
                      grad_U: torch.Tensor, want_prev: bool = True) -> torch.Tensor | None:
        """grad of layer k-1 (if want_prev) and grad_U[:, :, col] += the layer's contribution."""
        B = U.shape[0]
        f64 = U.dtype == torch.complex128
        if psi_prev.dtype != U.dtype or grad_k.dtype != U.dtype or grad_U.dtype != U.dtype:
            raise ValueError(f"psi_prev, grad_k and grad_U must be {U.dtype} like U")
        wsb = ctypes.c_size_t()
        wsfn = self._lib.slos_adjoint_workspace_f64 if f64 else self._lib.slos_adjoint_workspace
        nat.check(wsfn(self._h, max(1, min(B, MAX_ROWS)), k, ctypes.byref(wsb)), "slos_adjoint_workspace")
        ws = torch.empty(max(int(wsb.value), 1), dtype=torch.uint8, device=U.device)
        gprev = torch.empty((B, self.sizes[k - 1]), dtype=U.dtype, device=U.device) if want_prev else None
        fn = self._lib.slos_adjoint_layer_f64 if f64 else self._lib.slos_adjoint_layer
        for i, j in _row_ranges(B):
            nat.check(fn(self._h, nat.ptr(U[i:j]), j - i, k, int(col), nat.ptr(psi_prev[i:j]), nat.ptr(grad_k[i:j]),
                         nat.ptr(gprev[i:j]) if want_prev else None, nat.ptr(grad_U[i:j]), 1, nat.ptr(ws), ws.numel(),
                         nat.stream_handle(U.device)), "slos_adjoint_layer")
        return gprev

    def loss_descend(self, k: int, q: torch.Tensor, flags: int = 0, scale: float = 1.0) -> torch.Tensor:
        """q_{k-1} = scale * descent of q_k ([B, S_k] float32, or float64 in FP64, contiguous on the GPU)."""
        out = torch.empty((q.shape[0], self.sizes[k - 1]), dtype=q.dtype, device=q.device)
        if q.dtype not in (torch.float32, torch.float64):
            raise ValueError(f"q must be float32 or float64, got {q.dtype}")
        for i, j in _row_ranges(q.shape[0]):
            if q.dtype == torch.float64:
                rc = self._lib.slos_loss_descend_f64(self._h, k, nat.ptr(q[i:j]), nat.ptr(out[i:j]), j - i,
                                                     float(scale), nat.stream_handle(q.device))
            else:
                rc = self._lib.slos_loss_descend(self._h, k, nat.ptr(q[i:j]), nat.ptr(out[i:j]), j - i, flags,
                                                 float(scale), nat.stream_handle(q.device))
            nat.check(rc, "slos_loss_descend")
        return out

    def loss_ascend(self, k: int, g: torch.Tensor) -> torch.Tensor:
        """Adjoint of loss_descend: [B, S_{k-1}] -> [B, S_k] (float32 or float64)."""
        out = torch.empty((g.shape[0], self.sizes[k]), dtype=g.dtype, device=g.device)
        fn = self._lib.slos_loss_ascend_f64 if g.dtype == torch.float64 else self._lib.slos_loss_ascend
        if g.dtype not in (torch.float32, torch.float64):
            raise ValueError(f"g must be float32 or float64, got {g.dtype}")
        for i, j in _row_ranges(g.shape[0]):
            nat.check(fn(self._h, k, nat.ptr(g[i:j]), nat.ptr(out[i:j]), j - i, nat.stream_handle(g.device)),
                      "slos_loss_ascend")
        return out

    def raise_layer(self, k: int, w: torch.Tensor, q: torch.Tensor) -> torch.Tensor:
        """q_k(c) = sum_{j distinct in c} w[b, j] q_{k-1}(c - e_j): w [B, m], q [B, S_{k-1}] float32 (GPU)."""
        out = torch.empty((q.shape[0], self.sizes[k]), dtype=torch.float32, device=q.device)
        for i, j in _row_ranges(q.shape[0]):
            nat.check(self._lib.slos_raise(self._h, k, nat.ptr(w[i:j]), nat.ptr(q[i:j]), nat.ptr(out[i:j]), j - i,
                                           nat.stream_handle(q.device)), "slos_raise")
        return out

    def total_mass(self, probs: torch.Tensor) -> torch.Tensor:
        """FP64 sum over the last axis of float32 probabilities [B, S] -> float64 [B] (float64
        probabilities: torch's FP64 sum on the same GPU)."""
        if probs.dtype == torch.float64:
            return probs.sum(-1)
        B, S = probs.shape
        out = torch.empty(B, dtype=torch.float64, device=probs.device)
        wsb = ctypes.c_size_t()
        nat.check(self._lib.slos_mass_workspace(max(1, min(B, MAX_ROWS)), S, ctypes.byref(wsb)), "slos_mass_workspace")
        ws = torch.empty(max(int(wsb.value), 1), dtype=torch.uint8, device=probs.device)
        for i, j in _row_ranges(B):
            nat.check(self._lib.slos_total_mass(nat.ptr(probs[i:j]), j - i, S, nat.ptr(out[i:j]), nat.ptr(ws),
                                                ws.numel(), nat.stream_handle(probs.device)), "slos_total_mass")
        return out

    def spot_probs(self, U128: torch.Tensor, idx: torch.Tensor) -> torch.Tensor:
        """FP64 Ryser probabilities of the layer-n states with lex indices idx (device uint64/int64)."""
        U128 = U128.to(torch.complex128).contiguous()
        idx = idx.to(device=U128.device, dtype=torch.int64).contiguous()
        out = torch.empty(idx.numel(), dtype=torch.float64, device=U128.device)
        nat.check(self._lib.slos_spot_probs(self._h, nat.ptr(U128), nat.ptr(idx), idx.numel(), nat.ptr(out),
                                            nat.stream_handle(U128.device)), "slos_spot_probs")
        return out
