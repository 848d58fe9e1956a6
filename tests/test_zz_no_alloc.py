"""slos_forward_noalloc (the forward without a workspace argument, for callers that cannot allocate)
through the C ABI: equal to slos_forward bit for bit on two calls in a row, the second smaller; a
forward whose layers exceed the fixed scratch and a plan with no tables on the GPU are refused."""
import ctypes

import pytest
import torch
from helpers import haar_unitary

from photonic_engines_amd import _native as nat

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _plan(n, m):
    h = ctypes.c_void_p()
    modes = (ctypes.c_int32 * n)(*range(n))
    nat.check(nat.lib().slos_plan_create(n, m, 0, modes, ctypes.byref(h)), "slos_plan_create")
    return h


def _size(h, k):
    s = ctypes.c_uint64()
    nat.check(nat.lib().slos_layer_size(h, k, ctypes.byref(s)), "slos_layer_size")
    return s.value


def _forward(h, n, U, out_kind, dtype, noalloc, flags=0):
    L = nat.lib()
    B = U.shape[0]
    out = torch.full((B, _size(h, n)), float("nan"), dtype=dtype, device=DEV)
    st = nat.stream_handle(torch.device(DEV))
    if noalloc:
        rc = L.slos_forward_noalloc(h, nat.ptr(U), B, None, n, out_kind, nat.ptr(out), 1.0, flags, st)
    else:
        need = ctypes.c_size_t()
        nat.check(L.slos_forward_workspace(h, B, n, flags, ctypes.byref(need)), "slos_forward_workspace")
        ws = torch.empty(need.value, dtype=torch.uint8, device=DEV)
        rc = L.slos_forward(h, nat.ptr(U), B, None, n, out_kind, nat.ptr(out), 1.0, flags, nat.ptr(ws), ws.numel(),
                            st)
    nat.check(rc, "forward")
    torch.cuda.synchronize()
    return out


def _unitaries(B, m, seed):
    return torch.stack([haar_unitary(m, seed + b) for b in range(B)]).to(torch.complex64).to(DEV).contiguous()


def test_noalloc_forward_equals_forward_twice_then_smaller():
    L = nat.lib()
    big, small = _plan(8, 20), _plan(6, 16)
    try:
        U1 = _unitaries(3, 20, 800)
        U2 = _unitaries(1, 16, 900)
        ref1 = _forward(big, 8, U1, nat.OUT_PROBS, torch.float32, noalloc=False)   # tables now resident
        ref2 = _forward(small, 6, U2, nat.OUT_AMPS, torch.complex64, noalloc=False)
        a1 = _forward(big, 8, U1, nat.OUT_PROBS, torch.float32, noalloc=True)
        a2 = _forward(small, 6, U2, nat.OUT_AMPS, torch.complex64, noalloc=True)   # smaller, after a larger one
        assert torch.equal(a1, ref1)
        assert torch.equal(a2, ref2)
        a1b = _forward(big, 8, U1, nat.OUT_PROBS, torch.float32, noalloc=True)     # the scratch keeps stale data
        assert torch.equal(a1b, ref1)
        assert abs(float(ref1.double().sum(1).max()) - 1.0) < 1e-5
    finally:
        L.slos_plan_destroy(big)
        L.slos_plan_destroy(small)


def test_noalloc_forward_complex128_equals_forward():
    """SLOS_FLAG_F64 (complex128 layers, twice the scratch per state): equal to slos_forward with
    the same flag bit for bit, on two calls in a row, the second smaller."""
    L = nat.lib()
    big, small = _plan(7, 18), _plan(5, 12)
    try:
        U1 = _unitaries(2, 18, 700).to(torch.complex128)
        U2 = _unitaries(1, 12, 710).to(torch.complex128)
        ref1 = _forward(big, 7, U1, nat.OUT_PROBS, torch.float64, noalloc=False, flags=nat.FLAG_F64)
        ref2 = _forward(small, 5, U2, nat.OUT_AMPS, torch.complex128, noalloc=False, flags=nat.FLAG_F64)
        a1 = _forward(big, 7, U1, nat.OUT_PROBS, torch.float64, noalloc=True, flags=nat.FLAG_F64)
        a2 = _forward(small, 5, U2, nat.OUT_AMPS, torch.complex128, noalloc=True, flags=nat.FLAG_F64)
        assert torch.equal(a1, ref1) and torch.equal(a2, ref2)
        assert abs(float(ref1.sum(1).max()) - 1.0) < 1e-5
    finally:
        L.slos_plan_destroy(big)
        L.slos_plan_destroy(small)


def test_noalloc_forward_refusals():
    L = nat.lib()
    st = nat.stream_handle(torch.device(DEV))
    fresh = _plan(4, 8)                      # no forward has run with it: no tables on the GPU
    huge = _plan(10, 24)                     # layer 9 alone is 225 MB of complex64
    try:
        U = _unitaries(1, 8, 1)
        out = torch.empty((1, _size(fresh, 4)), dtype=torch.float32, device=DEV)
        rc = L.slos_forward_noalloc(fresh, nat.ptr(U), 1, None, 4, nat.OUT_PROBS, nat.ptr(out), 1.0, 0, st)
        assert rc == nat.SLOS_EINVAL and b"no tables" in L.slos_last_error()
        Uh = _unitaries(1, 24, 2)
        outh = torch.empty((1, 4), dtype=torch.float32, device=DEV)   # never written: refused before any launch
        rc = L.slos_forward_noalloc(huge, nat.ptr(Uh), 1, None, 10, nat.OUT_PROBS, nat.ptr(outh), 1.0, 0, st)
        assert rc == nat.SLOS_ERANGE and b"scratch" in L.slos_last_error()
    finally:
        L.slos_plan_destroy(fresh)
        L.slos_plan_destroy(huge)
