"""Batches that one launch cannot take. The reference accepts U [B, m, m] for any B its memory holds
(kernel.py:208-215); one launch here takes at most 65,535 rows (the launch grid's y dimension), so
larger row counts run as consecutive launches over row ranges, and an empty batch launches nothing.
Every kernel treats its rows independently: the rows of a split batch must equal, bit for bit, the
same rows run in batches of one launch, and a few of them the oracle."""
import numpy as np
import pytest
import torch

from oracle import pull
from photonic_engines_amd.slos import FockSLOSKernel, SuperpositionSLOSKernel
from photonic_engines_amd.slos._plan import MAX_ROWS

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _haar_batch(B: int, m: int, seed: int) -> torch.Tensor:
    """B seeded Haar unitaries [B, m, m] complex128 (QR of complex Gaussians with the phase fix)."""
    g = torch.Generator().manual_seed(seed)
    z = torch.complex(torch.randn(B, m, m, generator=g, dtype=torch.float64),
                      torch.randn(B, m, m, generator=g, dtype=torch.float64))
    q, r = torch.linalg.qr(z)
    d = torch.diagonal(r, dim1=-2, dim2=-1)
    return q * (d / d.abs()).unsqueeze(-2)


def _in_launches(fn, U):
    """fn over U in row ranges that each fit one launch, concatenated."""
    return torch.cat([fn(U[i:i + MAX_ROWS]) for i in range(0, U.shape[0], MAX_ROWS)], 0)


def test_forward_beyond_one_launch():
    n, m, B = 2, 3, MAX_ROWS + 5
    occ = [1, 1, 0]
    U = _haar_batch(B, m, 1).to(torch.complex64)
    k = FockSLOSKernel(n, m, occ, device=DEV)
    p = k.probs(U)
    assert p.shape == (B, 6)
    assert torch.equal(p, _in_launches(k.probs, U))
    for b in (0, MAX_ROWS - 1, MAX_ROWS, B - 1):
        ref = pull.forward_probs(n, m, occ, U[b].to(torch.complex128).numpy())
        assert np.abs(p[b].double().cpu().numpy() - ref).max() < 1e-6, b
    assert float((k.total_mass(p) - 1).abs().max()) < 1e-5
    a = k.amplitudes(U)
    assert torch.equal(a, _in_launches(k.amplitudes, U))


def test_loss_and_gradient_beyond_one_launch():
    n, m, B = 2, 3, MAX_ROWS + 3
    occ = [1, 0, 1]
    U = _haar_batch(B, m, 2).to(torch.complex64)
    k = FockSLOSKernel(n, m, occ, device=DEV)
    ld = k.probs(U, loss=0.6)
    parts = [k.probs(U[i:i + MAX_ROWS], loss=0.6) for i in range(0, B, MAX_ROWS)]
    for j in range(n + 1):
        assert ld.probs[j].shape[0] == B
        assert torch.equal(ld.probs[j], torch.cat([q.probs[j] for q in parts], 0)), j
    Ug = U.to(DEV).requires_grad_(True)
    k.probs(Ug)[:, 1].sum().backward()
    g_all = Ug.grad.clone()
    Ut = U[MAX_ROWS:].to(DEV).requires_grad_(True)
    k.probs(Ut)[:, 1].sum().backward()
    assert torch.equal(g_all[MAX_ROWS:], Ut.grad)
    assert torch.count_nonzero(g_all[:, :, [0, 2]]).item() > 0


def test_compensated_gradient_beyond_one_launch():
    """compensated=True under autograd steps through hi/lo layer pairs [2, B, S], whose row ranges are
    passed as copies: the probabilities equal the plain compensated forward's, bit for bit, and the
    gradient rows of the second range equal those of the same rows run alone."""
    n, m, B = 3, 4, MAX_ROWS + 2
    occ = [1, 1, 0, 1]
    U = _haar_batch(B, m, 5).to(torch.complex64).to(DEV)
    k = FockSLOSKernel(n, m, occ, device=DEV, compensated=True)
    a = k.probs(U)
    Ug = U.clone().requires_grad_(True)
    b = k.probs(Ug)
    assert torch.equal(a, b.detach())
    b[:, 3].sum().backward()
    Ut = U[MAX_ROWS:].clone().requires_grad_(True)
    k.probs(Ut)[:, 3].sum().backward()
    assert torch.equal(Ug.grad[MAX_ROWS:], Ut.grad)


def test_xmodel_rows_beyond_one_launch():
    """n=3 at x=0.7: the forwards of the one-photon-distinguishable subsets run as 3 B rows and the
    raises as more; with B = 30,000 both exceed one launch."""
    n, m, B = 3, 4, 30_000
    occ = [1, 1, 1, 0]
    U = _haar_batch(B, m, 3).to(torch.complex64)
    k = FockSLOSKernel(n, m, occ, device=DEV)
    p = k.probs(U, indist=0.7)
    half = B // 2
    assert torch.equal(p, torch.cat([k.probs(U[:half], indist=0.7), k.probs(U[half:], indist=0.7)], 0))
    assert float((p.double().sum(1) - 1).abs().max()) < 1e-5


def test_superposition_rows_beyond_one_launch():
    """K = 3 components: the forward runs 3 B rows."""
    n, m, B = 2, 4, 30_000
    ks = SuperpositionSLOSKernel(n, m, [[1, 1, 0, 0], [0, 1, 0, 1], [2, 0, 0, 0]], device=DEV)
    U = _haar_batch(B, m, 4).to(torch.complex64)
    p = ks.probs(U)
    half = B // 2
    assert torch.equal(p, torch.cat([ks.probs(U[:half]), ks.probs(U[half:])], 0))


def test_empty_batch():
    n, m = 3, 5
    occ = [1, 1, 1, 0, 0]
    k = FockSLOSKernel(n, m, occ, device=DEV)
    S = k.output_size()
    U0 = torch.zeros((0, m, m), dtype=torch.complex64)
    p = k.probs(U0)
    assert p.shape == (0, S) and p.dtype == torch.float32
    assert k.amplitudes(U0).shape == (0, S)
    ld = k.probs(U0, loss=0.5)
    assert [q.shape for q in ld.probs] == [(0, k.layers.size(j)) for j in range(n + 1)]
    assert k.probs(U0, indist=0.5).shape == (0, S)
    assert k.total_mass(p).shape == (0,)
    Ug = U0.to(DEV).requires_grad_(True)
    k.probs(Ug).sum().backward()
    assert Ug.grad is not None and Ug.grad.shape == (0, m, m)
