"""The product's lex-ranked layers (photonic_engines_amd.slos.basis, host side of the C ABI) against
the oracle's reference-order enumeration, plus size-independent properties at the frontier size."""
import math

import numpy as np
import pytest
import torch
from helpers import occupation_list

from oracle import slos_oracle as so
from photonic_engines_amd.masks import Mask, graded_state_space_size, is_valid, state_space_size
from photonic_engines_amd.slos import FockSLOSKernel, RankedLayers, clear_layers_cache, get_layers

CASES = [(4, 6, "fock"), (5, 7, "fock"), (3, 9, "fock"), (4, 8, "unbunched"), (4, 8, "dual_rail"),
         (3, 12, "sparse_1in4"), (4, 7, "tail_empty"), (6, 6, "fock"), (2, 5, "tail_empty"), (3, 10, "sparse_1in4")]


@pytest.mark.parametrize("n,m,mask", CASES)
def test_states_equal_reference_order(n, m, mask):
    L = RankedLayers(n, m, mask)
    E = so.enumerate_layers(n, m, mask)
    for k in range(n + 1):
        assert L.size(k) == E[k].shape[0]
        assert np.array_equal(L.states[k].astype(np.int64), E[k].astype(np.int64)), (k,)
        assert [tuple(t) for t in L.layers[k]] == [tuple(r) for r in E[k].tolist()]


@pytest.mark.parametrize("n,m,mask", CASES)
def test_edges_and_norms_equal_oracle(n, m, mask):
    L = RankedLayers(n, m, mask)
    E = so.enumerate_layers(n, m, mask)
    for k in range(n):
        pa, ca, ja, fa = L.edges[k]
        got = sorted(zip(pa.tolist(), ca.tolist(), ja.tolist(), np.round(fa.astype(np.float64), 6).tolist()))
        par, chi, mode, fac = so.edges(E[k], E[k + 1])
        exp = sorted(zip(par.tolist(), chi.tolist(), mode.tolist(), np.round(fac, 6).tolist()))
        assert got == exp
        p, c, off = L.edges_by_mode(k)
        assert off[0] == 0 and off[-1] == len(p) and (np.diff(off) >= 0).all()
    assert np.array_equal(L.output_prob_norms(), so.prob_norms(E[n]))
    assert L.final_states == [tuple(r) for r in so.to_occupation(E[n], m).tolist()]


def test_index_of_and_rank_roundtrip():
    L = get_layers(4, 9, "fock")
    for i, occ in enumerate(L.final_states):
        assert L.index_of(occ) == i


def test_frontier_rank_properties():
    """n=11, m=32 (4.28e9 states, never materialised): unrank/rank round trip, strict lex order of
    consecutive states, first and last state, prod t! from the run structure."""
    L = RankedLayers(11, 32, "fock")
    S = L.output_size()
    assert S == math.comb(42, 11) == 4_280_561_376
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.integers(0, S, 2000, dtype=np.uint64), np.array([0, 1, S - 2, S - 1], dtype=np.uint64)])
    pos = L.unrank(11, idx)
    assert (np.diff(pos.astype(np.int64), axis=1) >= 0).all()
    assert np.array_equal(L.rank(11, pos), idx)
    nxt = L.unrank(11, idx[idx < S - 1] + 1).astype(np.int64)
    cur = pos[idx < S - 1].astype(np.int64)
    diff = nxt - cur
    first_nz = np.argmax(diff != 0, axis=1)
    assert (diff[np.arange(len(diff)), first_nz] > 0).all()          # lexicographically increasing
    assert pos[-4].tolist() == [0] * 11 and pos[-1].tolist() == [31] * 11
    norms = L.prob_norms_range(11, 0, 5)
    assert norms.tolist() == [math.factorial(11), math.factorial(10), math.factorial(10), math.factorial(10),
                              math.factorial(10)]
    with pytest.raises(MemoryError):
        L.states[11]


def test_masks_module():
    assert state_space_size(6, 12, Mask.FOCK) == 12376
    assert state_space_size(6, 12, "unbunched") == 924
    assert state_space_size(6, 12, "dual_rail") == 64
    assert state_space_size(2, 8, "sparse_1in4") == math.comb(2, 2) * 4 ** 2
    assert state_space_size(3, 7, "tail_empty") == 35
    assert graded_state_space_size(3, 6) == sum(math.comb(6 + k - 1, k) for k in range(4))
    assert not is_valid(3, 6, "sparse_1in4") and is_valid(1, 8, "sparse_1in4")
    with pytest.raises(ValueError):
        state_space_size(4, 6, "dual_rail")


def test_kernel_constructor_checks():                                            # kernel.py:96-126
    occ = occupation_list(3, 6)
    with pytest.raises(ValueError):
        FockSLOSKernel(3, 6, occ, layout="bad")
    with pytest.raises(ValueError):
        FockSLOSKernel(3, 6, occ, normalization="bad")
    with pytest.raises(ValueError):
        FockSLOSKernel(3, 6, occ, normalization="immediate", gradient_checkpointing=True)
    with pytest.raises(ValueError):
        FockSLOSKernel(3, 6, [1, 1, 0, 0, 0, 0])
    with pytest.raises(ValueError):
        FockSLOSKernel(3, 6, occ, dtype=torch.float32)
    with pytest.raises(ValueError):
        FockSLOSKernel(3, 6, occ, dtype=torch.complex128, compensated=True)
    k128 = FockSLOSKernel(3, 6, occ, dtype=torch.complex128)                    # kernel.py:117-118
    assert (k128.cdtype, k128.rdtype) == (torch.complex128, torch.float64)
    with pytest.raises(ValueError):
        FockSLOSKernel(4, 3, [1, 1, 1], mask="unbunched")        # C(3, 4) = 0: empty output space
    k = FockSLOSKernel(3, 6, occ, normalization="immediate")
    assert k.layout == "edge" and k.output_size() == 56
    assert FockSLOSKernel(3, 6, [2, 1, 0, 0, 0, 0]).input_modes == [0, 0, 1]


def test_layers_cache():
    clear_layers_cache()
    a = get_layers(3, 6, "fock", device="cuda")
    assert get_layers(n=3, m=6, mask="fock") is a
    clear_layers_cache()
    assert get_layers(3, 6) is not a
