"""GPU parity of the block-structured loss descent (slos_loss_descend on multiset spaces with
k - 1 > 4; imperfections.py:254-276): bit-identical to the thread-per-state kernel, equal to the
oracle's thinning, and mass-preserving at n=11, m=32."""
import numpy as np
import pytest
import torch
from helpers import haar_unitary, occupation_list

from oracle import pull
from oracle import slos_oracle as so
from photonic_engines_amd import _native as nat
from photonic_engines_amd.slos import FockSLOSKernel

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("n,m,mask,occ", [(6, 6, "fock", None), (7, 9, "fock", [1, 0, 2, 1, 0, 3, 0, 0, 0]),
                                          (8, 10, "fock", None), (9, 14, "tail_empty", None),
                                          (10, 5, "fock", [3, 0, 4, 1, 2]), (12, 4, "fock", [3, 3, 3, 3]),
                                          (6, 64, "fock", None)])
def test_block_descent_bitwise_equals_thread_kernel(n, m, mask, occ):
    """Every descent step k -> k-1 (k - 1 >= 5) on random weights, one and three rows: the block
    kernel and the thread-per-state kernel sum the same children in the same order."""
    occ = occ or occupation_list(n, m, mask)
    k = FockSLOSKernel(n, m, occ, mask=mask, device=DEV)
    plan = k._plan
    g = torch.Generator(device=DEV).manual_seed(n * 100 + m)
    for kk in range(n, 5, -1):
        for B in (1, 3):
            q = torch.rand((B, plan.sizes[kk]), generator=g, device=DEV, dtype=torch.float32)
            a = plan.loss_descend(kk, q)
            b = plan.loss_descend(kk, q, flags=nat.FLAG_SIMPLE)
            assert torch.equal(a, b), (kk, B)


@pytest.mark.parametrize("occ,eta", [((1, 0, 2, 1, 0, 1, 1, 0, 0), 0.6), ((1,) * 7 + (0,) * 3, 0.85)])
def test_block_descent_thinning_matches_oracle(occ, eta):                       # test_loss.py:80-96
    m, n = len(occ), sum(occ)
    U = haar_unitary(m, 40 + n)
    ld = FockSLOSKernel(n, m, list(occ), device=DEV).probs(U.to(torch.complex64), loss=eta)
    ref = so.thin_distribution(so.enumerate_layers(n, m), pull.forward_probs(n, m, occ, U.numpy()), eta)
    for j in range(n + 1):
        assert np.abs(ld.probs[j].cpu().numpy() - ref[j][0]).max() < 1e-6


def test_block_descent_frontier_n11_m32():
    """n=11, m=32: the descent of the 4.3e9 float32 probabilities to layer 10 equals the thread
    kernel's bit for bit, and keeps the mass (FP64 sums)."""
    n, m = 11, 32
    k = FockSLOSKernel(n, m, [1] * n + [0] * (m - n), device=DEV)
    p = k.probs(haar_unitary(m, 11032).to(torch.complex64)).unsqueeze(0)
    a = k._plan.loss_descend(n, p)
    b = k._plan.loss_descend(n, p, flags=nat.FLAG_SIMPLE)
    assert torch.equal(a, b)
    del b
    s11 = float(k.total_mass(p.squeeze(0)))
    s10 = float(k.total_mass(a.squeeze(0)))
    assert abs(s10 - s11) < 1e-6
