"""The C ABI: the library loads, exports every symbol include/slos_mi355x.h declares, and its argument
checks map to the reference's exceptions — all without a GPU (no compute call is made)."""
import ctypes
import os
import re

import pytest

from photonic_engines_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "slos_mi355x.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(slos_\w+)\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = nat.lib()
    syms = header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
        assert s in nat.SIGNATURES, f"{s} has no ctypes signature"
    assert set(nat.SIGNATURES) == set(syms)
    assert nat.lib().slos_version().decode().startswith("slos_mi355x")


def _create(n, m, mask, modes):
    h = ctypes.c_void_p()
    arr = (ctypes.c_int32 * len(modes))(*modes)
    rc = nat.lib().slos_plan_create(n, m, mask, arr, ctypes.byref(h))
    return rc, h


def test_plan_errors():
    L = nat.lib()
    assert _create(3, 6, 9, [0, 1, 2])[0] == nat.SLOS_EINVAL
    assert b"unknown mask" in L.slos_last_error()
    assert _create(3, 6, 0, [0, 1, 7])[0] == nat.SLOS_EINVAL
    assert _create(-1, 6, 0, [])[0] == nat.SLOS_EINVAL
    rc0, h0 = _create(0, 6, 0, [])                                             # the vacuum: one state
    assert rc0 == nat.SLOS_OK
    one = ctypes.c_uint64()
    assert L.slos_layer_size(h0, 0, ctypes.byref(one)) == 0 and one.value == 1
    assert L.slos_plan_destroy(h0) == 0
    assert _create(40, 60, 0, [0] * 40)[0] == nat.SLOS_ERANGE
    assert _create(4, 3, 1, [0, 1, 2, 2])[0] == nat.SLOS_EINVAL               # unbunched C(3,4) = 0
    assert _create(30, 200, 0, [0] * 30)[0] == nat.SLOS_ERANGE               # C(229, 30) > 2^64
    rc, h = _create(3, 6, 0, [0, 2, 4])
    assert rc == nat.SLOS_OK
    out = ctypes.c_uint64()
    assert L.slos_layer_size(h, 3, ctypes.byref(out)) == 0 and out.value == 56
    assert L.slos_layer_size(h, 4, ctypes.byref(out)) == nat.SLOS_EINVAL
    # compute entry points validate before touching a device
    assert L.slos_forward(h, None, 1, None, 3, 0, None, 1.0, 0, None, 0, None) == nat.SLOS_EINVAL
    wsb = ctypes.c_size_t()
    assert L.slos_forward_workspace(h, 2, 3, 0, ctypes.byref(wsb)) == 0 and wsb.value >= 2 * 8 * (6 + 21)
    wsc = ctypes.c_size_t()
    assert L.slos_forward_workspace(h, 2, 3, nat.FLAG_COMPENSATED, ctypes.byref(wsc)) == 0
    assert wsc.value >= 2 * 2 * 8 * (6 + 21) and wsc.value > wsb.value          # hi/lo layers
    wsf = ctypes.c_size_t()                                                     # complex128 layers
    assert L.slos_forward_workspace(h, 2, 3, nat.FLAG_F64, ctypes.byref(wsf)) == 0
    assert wsf.value >= 2 * 16 * (6 + 21) and wsf.value > wsb.value
    assert L.slos_forward_workspace(h, 2, 3, nat.FLAG_F64 | nat.FLAG_COMPENSATED, ctypes.byref(wsf)) == nat.SLOS_EINVAL
    assert b"FP64" in L.slos_last_error()
    assert L.slos_forward(h, None, 1, None, 3, 0, None, 1.0, nat.FLAG_F64, None, 0, None) == nat.SLOS_EINVAL
    wa = ctypes.c_size_t()
    assert L.slos_adjoint_workspace_f64(h, 2, 3, ctypes.byref(wa)) == 0 and wa.value >= 2 * 6 * 16
    assert L.slos_adjoint_workspace_f64(h, 2, 4, ctypes.byref(wa)) == nat.SLOS_EINVAL
    assert L.slos_loss_descend_f64(h, 3, None, None, 1, 1.0, None) == nat.SLOS_EINVAL
    idx = (ctypes.c_uint64 * 1)(56)
    pos = (ctypes.c_int32 * 3)()
    assert L.slos_unrank(h, 3, idx, 1, pos) == nat.SLOS_ERANGE
    bad = (ctypes.c_int32 * 3)(2, 1, 0)                                         # not sorted
    assert L.slos_rank(h, 3, bad, 1, idx) == nat.SLOS_EINVAL
    assert L.slos_plan_destroy(h) == 0


def test_error_mapping():
    with pytest.raises(ValueError):
        nat.check(nat.SLOS_EINVAL, "x")
    with pytest.raises(OverflowError):
        nat.check(nat.SLOS_ERANGE, "x")
    import torch
    with pytest.raises(torch.cuda.OutOfMemoryError):
        nat.check(nat.SLOS_ENOMEM, "x")
    with pytest.raises(nat.NativeError):
        nat.check(nat.SLOS_EHIP, "x")


def test_group_host_checks():
    """Uneven-batch groups: creation limits, workspace arithmetic and replay-before-capture, host only."""
    L = nat.lib()
    plans = []
    for n, m in ((4, 12), (6, 16), (8, 20)):
        rc, h = _create(n, m, 0, list(range(n)))
        assert rc == nat.SLOS_OK
        plans.append(h)
    arr = (ctypes.c_void_p * 9)(*([plans[0].value] * 9))
    g = ctypes.c_void_p()
    assert L.slos_group_create(0, arr, ctypes.byref(g)) == nat.SLOS_EINVAL
    assert L.slos_group_create(9, arr, ctypes.byref(g)) == nat.SLOS_EINVAL
    nul = (ctypes.c_void_p * 2)(plans[0].value, None)
    assert L.slos_group_create(2, nul, ctypes.byref(g)) == nat.SLOS_EINVAL
    three = (ctypes.c_void_p * 3)(*[p.value for p in plans])
    assert L.slos_group_create(3, three, ctypes.byref(g)) == nat.SLOS_OK
    B = (ctypes.c_int64 * 3)(2, 3, 1)
    size = ctypes.c_size_t()
    assert L.slos_group_workspace(g, B, 0, ctypes.byref(size)) == nat.SLOS_OK
    # per plan: B ones + two ping-pong layers of B * max_{k<n} S_k complex64, each 256-byte aligned
    from math import comb

    def a256(x):
        return (x + 255) // 256 * 256
    exp = 0
    for (n, m), b in zip(((4, 12), (6, 16), (8, 20)), (2, 3, 1)):
        mx = max(comb(m + k - 1, k) for k in range(1, n))
        exp += a256(8 * b) + 2 * a256(8 * b * mx)
    assert size.value == exp
    size_c = ctypes.c_size_t()                                                  # compensated: hi/lo layers
    assert L.slos_group_workspace(g, B, nat.FLAG_COMPENSATED, ctypes.byref(size_c)) == nat.SLOS_OK
    assert size_c.value > size.value
    B0 = (ctypes.c_int64 * 3)(2, 0, 1)
    assert L.slos_group_workspace(g, B0, 0, ctypes.byref(size)) == nat.SLOS_EINVAL
    assert L.slos_group_replay(g, None) == nat.SLOS_EINVAL
    assert b"capture" in L.slos_last_error()
    blk, pw, gr = ctypes.c_int(), ctypes.c_uint64(), ctypes.c_uint32()
    assert L.slos_group_config(g, 8, ctypes.byref(blk), ctypes.byref(pw), ctypes.byref(gr)) == nat.SLOS_OK
    assert (blk.value, pw.value, gr.value) == (-1, 0, 0)
    assert L.slos_group_config(g, 9, ctypes.byref(blk), ctypes.byref(pw), ctypes.byref(gr)) == nat.SLOS_EINVAL
    assert L.slos_group_destroy(g) == nat.SLOS_OK
    for p in plans:
        assert L.slos_plan_destroy(p) == nat.SLOS_OK


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with status 2 and says so."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_memory_plan_and_preflight():
    """Memory pre-flight by formula, host only: the plan's byte counts, and a need above the free
    memory raises torch.cuda.OutOfMemoryError (the reference's exception) before any allocation."""
    import math

    import torch

    from photonic_engines_amd.slos._plan import Plan
    p = Plan(6, 10, "fock", [0, 1, 2, 3, 4, 5])
    mp = p.memory_plan(2)
    assert mp["output"] == 2 * 4 * math.comb(15, 6)
    assert mp["workspace"] == p.workspace_bytes(2, 6)
    assert mp["unitaries"] == 2 * 10 * 10 * 8
    assert mp["total"] == mp["output"] + mp["workspace"] + mp["unitaries"]
    assert p.memory_plan(1, layer=3, out_kind=nat.OUT_RAW)["output"] == 8 * math.comb(12, 3)
    dev = torch.device("cuda", 0)
    Plan.preflight(mp["total"], dev, "fits", free_bytes=lambda: mp["total"])
    with pytest.raises(torch.cuda.OutOfMemoryError, match="needs"):
        Plan.preflight(mp["total"], dev, "too big", free_bytes=lambda: mp["total"] - 1)
    mf = p.memory_plan(2, flags=nat.FLAG_F64)                                   # complex128: twice the bytes
    assert mf["output"] == 2 * mp["output"] and mf["unitaries"] == 2 * mp["unitaries"]
    assert mf["workspace"] == p.workspace_bytes(2, 6, nat.FLAG_F64) > mp["workspace"]
    from photonic_engines_amd.slos import FockSLOSKernel
    with pytest.raises(ValueError, match="complex64"):
        FockSLOSKernel(6, 10, [1] * 6 + [0] * 4, dtype=torch.complex128, compensated=True)
    with pytest.raises(ValueError, match="dtype"):
        FockSLOSKernel(6, 10, [1] * 6 + [0] * 4, dtype=torch.complex32)
    k128 = FockSLOSKernel(6, 10, [1] * 6 + [0] * 4, dtype=torch.complex128)
    assert k128.rdtype == torch.float64 and k128.memory_plan(2)["output"] == mf["output"]
    from photonic_engines_amd.slos import UnevenBatch
    with pytest.raises(ValueError, match="complex64"):
        UnevenBatch([k128])
    big = Plan(12, 64, "fock", list(range(12)))
    assert big.memory_plan(1)["output"] == 4 * math.comb(75, 12)              # ~105 TB of probabilities
