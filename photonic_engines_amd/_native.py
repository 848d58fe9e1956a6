"""ctypes binding of the HIP library behind include/slos_mi355x.h.

The library is built in-tree (`make -C photonic_engines_amd/csrc`, or `__graft_entry__.build()`)
into photonic_engines_amd/_build/libslos_mi355x.so. There is no fallback: if the library is
missing or no GPU is visible, the engine raises.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (binds the library to torch's HIP runtime: one runtime per process)

_HERE = os.path.dirname(os.path.abspath(__file__))
# SLOS_MI355X_LIB selects another build of the same library (e.g. an instrumented one)
LIB_PATH = os.environ.get("SLOS_MI355X_LIB") or os.path.join(_HERE, "_build", "libslos_mi355x.so")

SLOS_OK, SLOS_EINVAL, SLOS_ENOMEM, SLOS_EHIP, SLOS_ERANGE, SLOS_ENODEV = range(6)
MASK_IDS = {"fock": 0, "unbunched": 1, "dual_rail": 2, "sparse_1in4": 3, "tail_empty": 4}
OUT_PROBS, OUT_AMPS, OUT_RAW = 0, 1, 2
FLAG_COMPENSATED, FLAG_IMMEDIATE, FLAG_SIMPLE, FLAG_HILO, FLAG_F64 = 1, 2, 4, 8, 16

# every symbol include/slos_mi355x.h declares: (name, restype, argtypes)
_c = ctypes
_u64p, _i32p, _f64p, _vp = _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_int32), _c.POINTER(_c.c_double), _c.c_void_p
SIGNATURES = {
    "slos_version": (_c.c_char_p, []),
    "slos_last_error": (_c.c_char_p, []),
    "slos_plan_create": (_c.c_int, [_c.c_int, _c.c_int, _c.c_int, _i32p, _c.POINTER(_vp)]),
    "slos_plan_destroy": (_c.c_int, [_vp]),
    "slos_layer_size": (_c.c_int, [_vp, _c.c_int, _u64p]),
    "slos_rank": (_c.c_int, [_vp, _c.c_int, _i32p, _c.c_int64, _u64p]),
    "slos_unrank": (_c.c_int, [_vp, _c.c_int, _u64p, _c.c_int64, _i32p]),
    "slos_prob_norms": (_c.c_int, [_vp, _c.c_int, _c.c_uint64, _c.c_int64, _f64p]),
    "slos_forward_workspace": (_c.c_int, [_vp, _c.c_int64, _c.c_int, _c.c_uint, _c.POINTER(_c.c_size_t)]),
    "slos_forward": (_c.c_int, [_vp, _vp, _c.c_int64, _i32p, _c.c_int, _c.c_int, _vp, _c.c_double, _c.c_uint,
                                _vp, _c.c_size_t, _vp]),
    "slos_forward_noalloc": (_c.c_int, [_vp, _vp, _c.c_int64, _i32p, _c.c_int, _c.c_int, _vp, _c.c_double,
                                        _c.c_uint, _vp]),
    "slos_layer_step": (_c.c_int, [_vp, _vp, _c.c_int64, _c.c_int, _c.c_int, _vp, _c.c_int, _vp, _c.c_double,
                                   _c.c_uint, _vp]),
    "slos_readout_adjoint": (_c.c_int, [_vp, _c.c_int64, _vp, _vp, _c.c_double, _vp, _vp]),
    "slos_readout_adjoint_pull": (_c.c_int, [_vp, _vp, _c.c_int64, _vp, _vp, _c.c_double, _vp, _vp]),
    "slos_adjoint_workspace": (_c.c_int, [_vp, _c.c_int64, _c.c_int, _c.POINTER(_c.c_size_t)]),
    "slos_adjoint_layer": (_c.c_int, [_vp, _vp, _c.c_int64, _c.c_int, _c.c_int, _vp, _vp, _vp, _vp, _c.c_int,
                                      _vp, _c.c_size_t, _vp]),
    "slos_loss_descend": (_c.c_int, [_vp, _c.c_int, _vp, _vp, _c.c_int64, _c.c_uint, _c.c_double, _vp]),
    "slos_loss_ascend": (_c.c_int, [_vp, _c.c_int, _vp, _vp, _c.c_int64, _vp]),
    "slos_raise": (_c.c_int, [_vp, _c.c_int, _vp, _vp, _vp, _c.c_int64, _vp]),
    "slos_loss_descend_f64": (_c.c_int, [_vp, _c.c_int, _vp, _vp, _c.c_int64, _c.c_double, _vp]),
    "slos_loss_ascend_f64": (_c.c_int, [_vp, _c.c_int, _vp, _vp, _c.c_int64, _vp]),
    "slos_readout_adjoint_f64": (_c.c_int, [_vp, _c.c_int64, _vp, _vp, _c.c_double, _vp, _vp]),
    "slos_adjoint_workspace_f64": (_c.c_int, [_vp, _c.c_int64, _c.c_int, _c.POINTER(_c.c_size_t)]),
    "slos_group_sums_f64": (_c.c_int, [_vp, _c.c_int64, _c.c_uint64, _vp, _vp, _c.c_uint32, _vp, _c.c_int, _vp,
                                       _vp, _c.c_size_t, _vp]),
    "slos_adjoint_layer_f64": (_c.c_int, [_vp, _vp, _c.c_int64, _c.c_int, _c.c_int, _vp, _vp, _vp, _vp, _c.c_int,
                                          _vp, _c.c_size_t, _vp]),
    "slos_mass_workspace": (_c.c_int, [_c.c_int64, _c.c_uint64, _c.POINTER(_c.c_size_t)]),
    "slos_total_mass": (_c.c_int, [_vp, _c.c_int64, _c.c_uint64, _vp, _vp, _c.c_size_t, _vp]),
    "slos_group_sums": (_c.c_int, [_vp, _c.c_int64, _c.c_uint64, _vp, _vp, _c.c_uint32, _vp, _c.c_int, _vp, _vp,
                                   _c.c_size_t, _vp]),
    "slos_spot_probs": (_c.c_int, [_vp, _vp, _vp, _c.c_int64, _vp, _vp]),
    "slos_sample_workspace": (_c.c_int, [_c.c_uint64, _c.POINTER(_c.c_size_t)]),
    "slos_sample": (_c.c_int, [_vp, _c.c_uint64, _vp, _c.c_int64, _vp, _vp, _c.c_size_t, _vp]),
    "slos_unrank_device": (_c.c_int, [_vp, _c.c_int, _vp, _c.c_int64, _vp, _vp]),
    "slos_group_create": (_c.c_int, [_c.c_int, _c.POINTER(_vp), _c.POINTER(_vp)]),
    "slos_group_destroy": (_c.c_int, [_vp]),
    "slos_group_workspace": (_c.c_int, [_vp, _c.POINTER(_c.c_int64), _c.c_uint, _c.POINTER(_c.c_size_t)]),
    "slos_group_forward": (_c.c_int, [_vp, _c.POINTER(_vp), _c.POINTER(_c.c_int64), _c.POINTER(_vp), _f64p,
                                      _c.c_uint, _vp, _c.c_size_t, _vp]),
    "slos_group_tune": (_c.c_int, [_vp, _c.POINTER(_vp), _c.POINTER(_c.c_int64), _c.POINTER(_vp), _f64p,
                                   _c.c_uint, _vp, _c.c_size_t, _c.c_int, _vp]),
    "slos_group_config": (_c.c_int, [_vp, _c.c_int, _c.POINTER(_c.c_int), _u64p, _c.POINTER(_c.c_uint32)]),
    "slos_group_capture": (_c.c_int, [_vp, _c.POINTER(_vp), _c.POINTER(_c.c_int64), _c.POINTER(_vp), _f64p,
                                      _c.c_uint, _vp, _c.c_size_t]),
    "slos_group_replay": (_c.c_int, [_vp, _vp]),
}

_lib = None


class NativeError(RuntimeError):
    """A failed call into the HIP library (the message is the library's own)."""


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"HIP library not built: {LIB_PATH} is missing "
                              "(run `make -C photonic_engines_amd/csrc` or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    """Map a library status to the exception the reference raises for the same condition."""
    if rc == SLOS_OK:
        return
    msg = f"{what}: {lib().slos_last_error().decode()}"
    if rc == SLOS_EINVAL:
        raise ValueError(msg)
    if rc == SLOS_ENOMEM:
        raise torch.cuda.OutOfMemoryError(msg)
    if rc == SLOS_ERANGE:
        raise OverflowError(msg)
    raise NativeError(msg)


def stream_handle(device: torch.device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def require_gpu(device: torch.device | None = None) -> torch.device:
    """The compute device: `device` if it is a GPU, else the current GPU. Raises without one."""
    if not torch.cuda.is_available():
        raise RuntimeError("photonic_engines_amd computes on an AMD GPU (MI355X / gfx950); none is visible")
    if device is not None and device.type == "cuda":
        return device if device.index is not None else torch.device("cuda", torch.cuda.current_device())
    return torch.device("cuda", torch.cuda.current_device())
