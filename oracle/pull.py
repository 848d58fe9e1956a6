"""ctypes face of the C++ oracle (oracle/pull_oracle.cpp) — test infrastructure only."""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

MASK_IDS = {"fock": 0, "unbunched": 1, "dual_rail": 2, "sparse_1in4": 3, "tail_empty": 4}
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build() -> str:
    """Compile the oracle library (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_layer_size.argtypes = [ctypes.c_int] * 4 + [u64p]
        L.oracle_rank.argtypes = [ctypes.c_int] * 4 + [ip, ctypes.c_int64, u64p]
        L.oracle_unrank.argtypes = [ctypes.c_int] * 4 + [u64p, ctypes.c_int64, ip]
        L.oracle_forward_probs.argtypes = [ctypes.c_int] * 3 + [ip, dp, ctypes.c_double, ctypes.c_int, dp, ctypes.c_int]
        L.oracle_forward_amps.argtypes = [ctypes.c_int] * 3 + [ip, dp, ctypes.c_double, dp, ctypes.c_int]
        L.oracle_forward_sample.argtypes = [ctypes.c_int] * 3 + [ip, dp, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, dp, dp]
        L.oracle_permanent.argtypes = [ctypes.c_int, dp, dp]
        L.oracle_spot_probs.argtypes = [ctypes.c_int] * 3 + [ip, dp, ctypes.c_double, u64p, ctypes.c_int64, dp, ctypes.c_int]
        _lib = L
    return _lib


def _modes(occ, n=None, m=None, U=None):
    """Input modes of the photons, one per photon in ascending mode order. With n and m given, the
    occupation must have m entries summing to n and U (if given) must be [m, m]: the C routines
    index U by these modes and would read outside it otherwise."""
    if m is not None and len(occ) != m:
        raise ValueError(f"occupation has {len(occ)} modes, expected m = {m}")
    if n is not None and sum(int(c) for c in occ) != n:
        raise ValueError(f"occupation holds {sum(int(c) for c in occ)} photons, expected n = {n}")
    if U is not None and np.shape(U) != (m, m):
        raise ValueError(f"U must be [{m}, {m}], got {np.shape(U)}")
    return [j for j, c in enumerate(occ) for _ in range(int(c))]


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def layer_size(n, m, mask, k) -> int:
    out = ctypes.c_uint64()
    if lib().oracle_layer_size(n, m, MASK_IDS[mask], k, ctypes.byref(out)):
        raise ValueError("invalid space")
    return int(out.value)


def rank(n, m, mask, pos: np.ndarray) -> np.ndarray:
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    cnt, k = pos.shape
    out = np.zeros(cnt, dtype=np.uint64)
    lib().oracle_rank(n, m, MASK_IDS[mask], k, _ptr(pos, ctypes.c_int), cnt, _ptr(out, ctypes.c_uint64))
    return out


def unrank(n, m, mask, k, idx) -> np.ndarray:
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    out = np.zeros((idx.shape[0], k), dtype=np.int32)
    lib().oracle_unrank(n, m, MASK_IDS[mask], k, _ptr(idx, ctypes.c_uint64), idx.shape[0], _ptr(out, ctypes.c_int))
    return out


def forward_probs(n, m, occ, U, mask="fock", prec=128, threads=0) -> np.ndarray:
    """Lossless probabilities of the pull recurrence in complex128 (prec=128) or complex64 (64)."""
    modes = np.asarray(_modes(occ, n, m, U), dtype=np.int32)
    Ud = np.ascontiguousarray(np.asarray(U, dtype=np.complex128)).view(np.float64)
    in_norm = 1.0 / math.sqrt(math.prod(math.factorial(int(c)) for c in occ))
    out = np.zeros(layer_size(n, m, mask, n), dtype=np.float64)
    rc = lib().oracle_forward_probs(n, m, MASK_IDS[mask], _ptr(modes, ctypes.c_int), _ptr(Ud, ctypes.c_double),
                                    in_norm, prec, _ptr(out, ctypes.c_double), threads)
    if rc:
        raise RuntimeError("oracle_forward_probs failed")
    return out


def forward_amps(n, m, occ, U, mask="fock", threads=0) -> np.ndarray:
    """complex128 raw (deferred-normalisation) amplitudes times the input norm."""
    modes = np.asarray(_modes(occ, n, m, U), dtype=np.int32)
    Ud = np.ascontiguousarray(np.asarray(U, dtype=np.complex128)).view(np.float64)
    in_norm = 1.0 / math.sqrt(math.prod(math.factorial(int(c)) for c in occ))
    out = np.zeros(layer_size(n, m, mask, n), dtype=np.complex128)
    rc = lib().oracle_forward_amps(n, m, MASK_IDS[mask], _ptr(modes, ctypes.c_int), _ptr(Ud, ctypes.c_double),
                                   in_norm, _ptr(out.view(np.float64), ctypes.c_double), threads)
    if rc:
        raise RuntimeError("oracle_forward_amps failed")
    return out


def permanent(A) -> complex:
    A = np.ascontiguousarray(np.asarray(A, dtype=np.complex128))
    out = np.zeros(2)
    lib().oracle_permanent(A.shape[0], _ptr(A.view(np.float64), ctypes.c_double), _ptr(out, ctypes.c_double))
    return complex(out[0], out[1])


def spot_probs(n, m, occ, U, idx, mask="fock", threads=0) -> np.ndarray:
    """Ryser (complex128) probabilities of the outputs with lex indices idx."""
    modes = np.asarray(_modes(occ, n, m, U), dtype=np.int32)
    Ud = np.ascontiguousarray(np.asarray(U, dtype=np.complex128)).view(np.float64)
    s_fact = float(math.prod(math.factorial(int(c)) for c in occ))
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    out = np.zeros(idx.shape[0])
    rc = lib().oracle_spot_probs(n, m, MASK_IDS[mask], _ptr(modes, ctypes.c_int), _ptr(Ud, ctypes.c_double), s_fact,
                                 _ptr(idx, ctypes.c_uint64), idx.shape[0], _ptr(out, ctypes.c_double), threads)
    if rc:
        raise RuntimeError("oracle_spot_probs failed")
    return out


def forward_sample(n, m, occ, U, last_count, mask="fock", prec=64, threads=0):
    """Timed bounded sample of the pull recurrence: layers 1..n-1 complete plus the first
    `last_count` states of layer n. Returns (prefix_seconds, last_seconds, states_done, checksum)."""
    modes = np.asarray(_modes(occ, n, m, U), dtype=np.int32)
    Ud = np.ascontiguousarray(np.asarray(U, dtype=np.complex128)).view(np.float64)
    t = np.zeros(2)
    cs = np.zeros(1)
    rc = lib().oracle_forward_sample(n, m, MASK_IDS[mask], _ptr(modes, ctypes.c_int), _ptr(Ud, ctypes.c_double), prec,
                                     int(last_count), threads, _ptr(t, ctypes.c_double), _ptr(cs, ctypes.c_double))
    if rc:
        raise RuntimeError("oracle_forward_sample failed")
    return float(t[0]), float(t[1]), min(int(last_count), layer_size(n, m, mask, n)), float(cs[0])
