"""The GPU kernels' index code executed on the CPU: tests/native/block_emu.cpp compiles the product
headers photonic_engines_amd/csrc/{block_index,pull_index,space}.hpp with g++ and runs the block
kernel's wave / block / 64-lane loop, the descent and its adjoint, against the oracle."""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest
from helpers import haar_unitary, occupation_list

from oracle import pull
from oracle import slos_oracle as so

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    out = tmp_path_factory.mktemp("emu") / "libemu.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wno-unknown-pragmas", "-o", str(out),
                    os.path.join(ROOT, "tests", "native", "block_emu.cpp")], check=True)
    L = ctypes.CDLL(str(out))
    L.emu_child_edges.restype = ctypes.c_int64
    L.emu_check_tuple_words.restype = ctypes.c_int64
    L.emu_check_tuple_words5.restype = ctypes.c_int64
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


FWD = [(4, 6, "fock", None), (5, 7, "fock", None), (6, 6, "fock", [3, 0, 2, 0, 1, 0]), (5, 8, "unbunched", None),
       (6, 12, "dual_rail", None), (4, 16, "sparse_1in4", None), (5, 9, "tail_empty", [0, 2, 1, 1, 1, 0, 0, 0, 0]),
       (7, 9, "fock", [1, 0, 2, 1, 0, 3, 0, 0, 0]), (8, 10, "fock", None), (6, 8, "fock", [6, 0, 0, 0, 0, 0, 0, 0]),
       (7, 12, "fock", None), (5, 20, "sparse_1in4", None), (9, 5, "fock", [2, 2, 2, 2, 1]), (10, 5, "fock", [3, 0, 4, 1, 2]),
       (11, 4, "fock", [5, 0, 3, 3]), (12, 4, "fock", [3, 3, 3, 3]), (13, 3, "fock", [5, 4, 4]), (9, 12, "unbunched", None),
       (10, 20, "dual_rail", None), (5, 40, "unbunched", None), (7, 20, "fock", None), (13, 9, "fock", [2, 2, 2, 2, 1, 1, 1, 1, 1])]


@pytest.mark.parametrize("suf", [4, 5, 6])
@pytest.mark.parametrize("n,m,mask,occ", FWD)
def test_block_kernel_decomposition(emu, n, m, mask, occ, suf):
    """The block kernel's wave loop with 4-position suffixes on every block layer, with 5-position
    suffixes on the layers K = 6..13 (the device's choice where its table exists), and (suf 6) with
    the layers K = 7..11 of a multiset space in group order (the group kernel's work items and sharing)."""
    if suf == 5 and n < 6:
        pytest.skip("no layer with a 5-position suffix block")
    if suf == 6 and not (n >= 7 and mask in ("fock", "tail_empty")):
        pytest.skip("no group-order last layer")
    occ = occ or occupation_list(n, m, mask)
    assert len(occ) == m and sum(occ) == n                 # the emulator indexes U by these modes
    U = haar_unitary(m, n + 7 * m).numpy()
    ref = pull.forward_probs(n, m, occ, U, mask)
    modes = np.array([j for j, c in enumerate(occ) for _ in range(c)], dtype=np.int32)
    Ud = np.ascontiguousarray(U).view(np.float64)
    innorm = 1 / math.sqrt(math.prod(math.factorial(c) for c in occ))
    for per_wave in (64, 100, 256, 1000, 4096):
        out = np.zeros(ref.shape[0])
        rc = emu.emu_forward_suf(n, m, pull.MASK_IDS[mask], _p(modes), _p(Ud), ctypes.c_double(innorm),
                                 ctypes.c_uint64(per_wave), suf, _p(out))
        assert rc == 0, "some child was not written exactly once" if rc == 2 else rc
        # the kernel forms prod t_j! in float32: exact below 2^24, one rounding (2^-24 relative) above
        assert np.all(np.abs(out - ref) <= 1e-14 + 2.4e-7 * np.abs(ref))


@pytest.mark.parametrize("n,m,mask", [(4, 6, "fock"), (5, 5, "fock"), (4, 8, "unbunched"), (4, 8, "dual_rail"),
                                      (3, 12, "sparse_1in4"), (4, 7, "tail_empty"), (6, 4, "fock")])
def test_descent_adjoint_and_child_edges(emu, n, m, mask):
    rng = np.random.default_rng(n * m)
    lay = so.enumerate_layers(n, m, mask)
    for k in range(1, n + 1):
        q = rng.random(len(lay[k]))
        out = np.zeros(len(lay[k - 1]))
        assert emu.emu_descend(n, m, pull.MASK_IDS[mask], k, _p(q), _p(out)) == 0
        assert np.abs(out - so.descend(lay, q, k)[0]).max() < 1e-12
        g = rng.random(len(lay[k - 1]))
        gk = np.zeros(len(lay[k]))
        assert emu.emu_ascend(n, m, pull.MASK_IDS[mask], k, _p(g), _p(gk)) == 0
        assert abs(gk @ q - g @ out) < 1e-10 * abs(g @ out)
        cap = len(lay[k]) * m
        buf = np.zeros(4 * cap, dtype=np.int64)
        E = emu.emu_child_edges(n, m, pull.MASK_IDS[mask], k, cap, _p(buf))
        got = sorted(map(tuple, buf[:4 * E].reshape(-1, 4).tolist()))
        par, chi, mode, fac = so.edges(lay[k - 1], lay[k])
        exp = sorted(zip(par.tolist(), chi.tolist(), mode.tolist(), np.round(fac ** 2).astype(int).tolist()))
        assert got == exp


@pytest.mark.parametrize("m,mask", [(4, "fock"), (9, "fock"), (32, "fock"), (64, "fock"), (10, "tail_empty"),
                                    (64, "tail_empty"), (33, "unbunched"), (64, "unbunched"), (20, "dual_rail"),
                                    (24, "sparse_1in4")])
def test_tuple_words_and_local_parent_ranks(emu, m, mask):
    """Every suffix tuple: the 4-byte word's positions and marks, the four local-parent ranks from the
    per-position terms cg[3][x+1], cg[2][x+1], cg[1][x], and the run data from the mark code equal the
    direct unrank / rank / run computation. The same for every 5-tuple (8-byte words, terms
    cg[4][x+1], cg[3][x+1], cg[2][x+1], cg[1][x] packed into the suffix table's two words)."""
    assert emu.emu_check_tuple_words(m, pull.MASK_IDS[mask]) == 0
    assert emu.emu_check_tuple_words5(m, pull.MASK_IDS[mask]) == 0


@pytest.mark.parametrize("n,m,mask,k,first", [(6, 5, "fock", 6, 0), (7, 6, "fock", 7, 0), (7, 6, "fock", 6, 0),
                                              (8, 9, "fock", 8, 0), (9, 7, "tail_empty", 9, 0), (10, 12, "fock", 10, 0),
                                              (12, 6, "fock", 12, 0), (13, 5, "fock", 13, 0),
                                              (11, 32, "fock", 11, 0), (11, 32, "fock", 11, 560549704),
                                              (11, 32, "fock", 11, 1121099408 - 200000), (11, 32, "fock", 10, 136719440)])
def test_descend_block_child_indices(emu, n, m, mask, k, first):
    """The block descent's child of each parent and mode (segment bases, the stream of block (P, a),
    the tail offsets in block (P, s0)) and its copies equal slos::for_each_child, in the same order,
    for 200,000 parents from `first` (n=11, m=32: the start, middle and end of layers 10 and 9)."""
    emu.emu_check_descend_block.restype = ctypes.c_int64
    emu.emu_check_descend_block.argtypes = [ctypes.c_int] * 4 + [ctypes.c_uint64] * 2
    assert emu.emu_check_descend_block(n, m, pull.MASK_IDS[mask], k, first, 200_000) == 0


@pytest.mark.parametrize("n,m,mask,occ", [(6, 6, "fock", None), (7, 5, "fock", [3, 0, 2, 1, 1]),
                                          (8, 7, "tail_empty", [2, 2, 1, 1, 1, 1, 0]), (9, 4, "fock", [9, 0, 0, 0])])
def test_descend_block_emulation(emu, n, m, mask, occ):
    """The block descent emulated in float32 (the device's order of operations) against the
    double-precision pull descent on the oracle's loss of a real distribution."""
    occ = occ or occupation_list(n, m, mask)
    U = haar_unitary(m, 5 * n + m).numpy()
    q = pull.forward_probs(n, m, occ, U, mask)
    for k in range(n, slos_min_k(), -1):
        q32 = np.ascontiguousarray(q, dtype=np.float32)
        out32 = np.zeros(len(so.enumerate_layers(k - 1, m, mask)[k - 1]), dtype=np.float32)
        assert emu.emu_descend_block(n, m, pull.MASK_IDS[mask], k, _p(q32), _p(out32)) == 0
        ref = np.zeros(len(out32))
        q64 = np.ascontiguousarray(q, dtype=np.float64)
        assert emu.emu_descend(n, m, pull.MASK_IDS[mask], k, _p(q64), _p(ref)) == 0
        assert np.abs(out32 - ref).max() <= 1e-6 * max(ref.max(), 1e-30)
        assert abs(out32.astype(np.float64).sum() - q64.sum()) < 1e-5     # the descent keeps the mass
        q = ref


def slos_min_k():
    """The block descent covers k - 1 >= 5 (a parent prefix of at least one position)."""
    return 5
