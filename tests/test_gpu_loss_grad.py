"""GPU parity of the loss descent (probs(U, loss=eta)) and of the hand-written adjoint (gradients of
probs), against the oracle and the golden finite-difference gradient."""
import math

import numpy as np
import pytest
import torch
from helpers import first_modes, golden, haar_unitary, occupation_list

from oracle import pull
from oracle import slos_oracle as so
from photonic_engines_amd.imperfections import LossyDistribution, binomial_pmf
from photonic_engines_amd.slos import FockSLOSKernel

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("occ,eta", [((1, 1, 1, 0), 0.7), ((2, 0, 1, 0, 0), 0.9), ((1, 1, 0, 0), 0.5),
                                     ((1, 0, 1, 0, 1, 0, 1, 0), 0.3)])
def test_thinned_distribution_matches_oracle(occ, eta):                         # test_loss.py:80-96
    m, n = len(occ), sum(occ)
    U = haar_unitary(m, 11 + n)
    k = FockSLOSKernel(n, m, list(occ), device=DEV)
    ld = k.probs(U.to(torch.complex64), loss=eta)
    assert isinstance(ld, LossyDistribution) and ld.n == n
    L = so.enumerate_layers(n, m)
    ref = so.thin_distribution(L, pull.forward_probs(n, m, occ, U.numpy()), eta)
    for j in range(n + 1):
        assert np.abs(ld.probs[j].cpu().numpy() - ref[j][0]).max() < 1e-6
    d = ld.as_dict()
    assert abs(sum(d.values()) - 1) < 1e-5
    pk = binomial_pmf(n, eta)
    assert all(abs(float(ld.mass(j)) - float(pk[j])) < 1e-5 for j in range(n + 1))
    assert k.probs(U.to(torch.complex64), loss=1.0).shape == (k.output_size(),)


def test_lossy_n11_m22_block_descent_against_fp64():
    """n=11, m=22 (129M outputs; the descents k = 11..6 run on the block structure) at eta = 0.9:
    every block of the complex64 engine against the complex128 engine's FP64 thread-per-state
    descents of the same U, and the FP64 block masses against the binomial law."""
    n, m, eta = 11, 22, 0.9
    U = haar_unitary(m, 1000 * n + m)
    occ = first_modes(n, m)
    ld = FockSLOSKernel(n, m, occ, device=DEV).probs(U.to(torch.complex64).to(DEV), loss=eta)
    ld64 = FockSLOSKernel(n, m, occ, device=DEV, dtype=torch.complex128).probs(U.to(DEV), loss=eta)
    pk = binomial_pmf(n, eta)
    for j in range(n + 1):
        d = (ld.probs[j].double() - ld64.probs[j]).abs()
        assert float(d.max()) <= 1e-7, j
        assert float(d.sum()) <= 1e-5, j
        assert abs(float(ld.probs[j].double().sum()) - float(pk[j])) < 1e-5, j


@pytest.mark.parametrize("mask,n,m", [("unbunched", 3, 7), ("dual_rail", 3, 6), ("tail_empty", 3, 7),
                                      ("sparse_1in4", 2, 12)])
def test_thinning_masks(mask, n, m):
    occ = occupation_list(n, m, mask)
    U = haar_unitary(m, 3)
    ld = FockSLOSKernel(n, m, occ, mask=mask, device=DEV).probs(U.to(torch.complex64), loss=0.6)
    L = so.enumerate_layers(n, m, mask)
    ref = so.thin_distribution(L, pull.forward_probs(n, m, occ, U.numpy(), mask), 0.6)
    for j in range(n + 1):
        assert np.abs(ld.probs[j].cpu().numpy() - ref[j][0]).max() < 1e-6


def test_thinning_batched_and_differentiable():                                 # test_loss.py:99-110
    m, occ = 4, [1, 1, 0, 0]
    U = torch.stack([haar_unitary(m, 1), haar_unitary(m, 2)]).to(torch.complex64)
    k = FockSLOSKernel(2, m, occ, device="cpu")
    eta = torch.tensor(0.8, dtype=torch.float64, requires_grad=True)
    ld = k.probs(U, loss=eta)
    assert ld.probs[2].shape == (2, k.output_size())
    ld.mass(2).sum().backward()
    assert abs(float(eta.grad) - 2 * 2 * 0.8) < 1e-5
    single = k.probs(U[1], loss=0.8).as_dict()
    d1 = ld.as_dict(batch=1)
    assert max(abs(d1[t] - single[t]) for t in single) < 1e-7


def test_tutorial_losses():                                                      # test_tutorial.py:191-204
    U6 = haar_unitary(6, 3)
    k = FockSLOSKernel(3, 6, [1, 1, 1, 0, 0, 0], device=DEV)
    ld = k.probs(U6.to(torch.complex64), loss=0.8)
    assert abs(float(ld.mass(2)) - 3 * 0.8 ** 2 * 0.2) < 1e-6
    assert ld.probs[2].shape[0] == ld.occupations(2).shape[0] == 21
    assert torch.allclose(ld.conditional(3), k.probs(U6.to(torch.complex64)), atol=1e-6)


def test_gradient_through_loss_descent():
    """d/dp of a weighted lossy mass through the descent's adjoint kernel = finite differences."""
    n, m = 3, 5
    U = haar_unitary(m, 9).to(torch.complex64)
    k = FockSLOSKernel(n, m, first_modes(n, m), device=DEV)
    p = k.probs(U).to(DEV).detach().requires_grad_(True)
    w = torch.linspace(0.5, 1.5, 15, device=DEV)
    from photonic_engines_amd.imperfections import thin_distribution
    ld = thin_distribution(k.layers, p, 0.7)
    (ld.probs[2] * w).sum().backward()
    # reference: the same linear map on the oracle, applied to unit vectors
    L = so.enumerate_layers(n, m)
    S = p.shape[0]
    J = np.stack([so.thin_distribution(L, np.eye(S)[i], 0.7)[2][0] for i in range(S)])   # [S, S_2]
    ref = J @ w.cpu().numpy()
    assert np.abs(p.grad.cpu().numpy() - ref).max() < 1e-5


def test_gradient_flows():                                                       # test_kernel.py:74-80
    U = haar_unitary(4, 5).to(torch.complex64).requires_grad_(True)
    p = FockSLOSKernel(2, 4, occupation_list(2, 4), device="cpu").probs(U)
    (p * p).sum().backward()
    assert U.grad is not None and torch.isfinite(U.grad).all() and U.grad.abs().sum() > 0


def test_c4_gradient_matches_finite_differences():
    """n=6, m=16: dL/dU of L = -mean log p[idx] (1,024 golden indices) from the adjoint kernels
    against complex128 central finite differences (h = 1e-6)."""
    U = torch.as_tensor(golden("c4_u.npy")).to(torch.complex64).to(DEV).requires_grad_(True)
    idx = torch.as_tensor(golden("c4_idx.npy"), device=DEV)
    k = FockSLOSKernel(6, 16, first_modes(6, 16), device=DEV)
    p = k.probs(U)
    loss = -torch.log(p[idx]).mean()
    assert abs(float(loss) - golden("golden.json")["c4_loss"]) < 1e-4
    loss.backward()
    g = U.grad.cpu().numpy()
    g_fd = golden("c4_grad_fd.npy")
    rel = np.abs(g - g_fd).max() / np.abs(g_fd).max()
    assert rel < 1e-3, rel
    assert np.abs(g[:, 6:]).max() == 0.0


@pytest.mark.parametrize("n,m,mask,occ", [(3, 6, "fock", [1, 0, 1, 0, 1, 0]), (3, 4, "fock", [2, 0, 1, 0]),
                                          (4, 8, "dual_rail", None), (3, 7, "tail_empty", None),
                                          (5, 7, "fock", [1, 1, 1, 1, 1, 0, 0]), (4, 9, "unbunched", None)])
@pytest.mark.parametrize("sparse", [False, True])
def test_gradient_matches_restatement(n, m, mask, occ, sparse):
    """Weighted-sum loss over every state, batched: adjoint kernels vs complex128 autograd of the
    oracle's stored-edge recurrence written in torch (test_kernel.py:131-146 shape). `sparse`: the
    weights of six states in seven are zero (the readout adjoint's zero-gradient entries)."""
    occ = occ or occupation_list(n, m, mask)
    Ub = torch.stack([haar_unitary(m, s) for s in (6, 7)])
    k = FockSLOSKernel(n, m, occ, mask=mask, device=DEV)
    U = Ub.to(torch.complex64).to(DEV).requires_grad_(True)
    p = k.probs(U)
    w = torch.linspace(0.5, 1.5, p.shape[-1], device=DEV)
    if sparse:
        w[torch.arange(p.shape[-1], device=DEV) % 7 != 3] = 0.0
    (p * w).sum().backward()
    L = so.enumerate_layers(n, m, mask)
    modes = [j for j, c in enumerate(occ) for _ in range(c)]
    s_norm = float(math.prod(math.factorial(c) for c in occ))
    Ur = Ub.clone().requires_grad_(True)
    tot = 0
    for b in range(2):
        psi = torch.ones(1, dtype=torch.complex128)
        for kk, i in enumerate(modes):
            par, chi, mode, _ = so.edges(L[kk], L[kk + 1])
            contrib = Ur[b][torch.as_tensor(mode), i] * psi[torch.as_tensor(par)]
            psi = torch.zeros(L[kk + 1].shape[0], dtype=torch.complex128).index_add(0, torch.as_tensor(chi), contrib)
        pr = psi.abs() ** 2 / s_norm * torch.as_tensor(so.prob_norms(L[n]), dtype=torch.float64)
        tot = tot + (pr * w.cpu().double()).sum()
    tot.backward()
    g, gr = U.grad.cpu().numpy(), Ur.grad.numpy()
    assert np.abs(g - gr).max() < 1e-5 * max(1.0, np.abs(gr).max())


def test_sparse_readout_adjoint_equals_dense(monkeypatch):
    """A loss over a few outputs takes the readout adjoint from layer n-1 (each nonzero entry's
    amplitude pulled from its parents, slos_readout_adjoint_pull); the dense pass over the recomputed
    last layer gives the same gradient."""
    from photonic_engines_amd.slos import autograd as ag
    n, m = 6, 12
    k = FockSLOSKernel(n, m, first_modes(n, m), device=DEV)
    Ub = torch.stack([haar_unitary(m, s) for s in (21, 22)]).to(torch.complex64).to(DEV)
    S = k.output_size()
    idx = torch.as_tensor(np.random.default_rng(3).choice(S, 40, replace=False), device=DEV)

    def grad():
        U = Ub.clone().requires_grad_(True)
        p = k.probs(U)
        (-torch.log(p[0, idx]).mean() - torch.log(p[1, idx[:20]]).mean()).backward()
        return U.grad

    g_dense = grad()                                        # small: the dense pass without a look
    monkeypatch.setattr(ag, "SPARSE_MIN", 0)
    assert 60 * ag.SPARSE_READOUT <= 2 * S                  # this loss now takes the sparse path
    g_sparse = grad()
    scale = float(g_dense.abs().max())
    assert float((g_sparse - g_dense).abs().max()) <= 1e-5 * scale
    assert torch.count_nonzero(g_sparse[:, :, n:]) == 0


def test_gradient_n11_m22_against_fp64():
    """n=11, m=22 (129M outputs): L = -mean log p over 1,024 outputs, the complex64 gradient (the
    readout adjoint pulled from layer 10, block-structured adjoint layers) against the complex128
    engine's gradient of the same loss (FP64 thread-per-state adjoint kernels)."""
    n, m = 11, 22
    U = haar_unitary(m, 1000 * n + m)
    occ = first_modes(n, m)
    k = FockSLOSKernel(n, m, occ, device=DEV)
    S = k.output_size()
    idx = torch.as_tensor(np.random.default_rng(1122).choice(S, 1024, replace=False), device=DEV)

    def grad(kern, Ud):
        Ug = Ud.clone().requires_grad_(True)
        loss = -torch.log(kern.probs(Ug)[idx]).mean()
        loss.backward()
        return Ug.grad

    g32 = grad(k, U.to(torch.complex64).to(DEV)).to(torch.complex128)
    g64 = grad(FockSLOSKernel(n, m, occ, device=DEV, dtype=torch.complex128), U.to(DEV))
    scale = float(g64.abs().max())
    assert float((g32 - g64).abs().max()) <= 1e-3 * scale
    assert torch.count_nonzero(g64[:, n:]) == 0 and torch.count_nonzero(g64[:, :n]) > 0


def test_compensated_probs_equal_with_and_without_grad():
    """FockSLOSKernel(compensated=True): every layer is compensated whether or not U requires a
    gradient (the autograd forward runs the same launches as the plain forward)."""
    n, m = 7, 12
    k = FockSLOSKernel(n, m, [1] * n + [0] * (m - n), device=DEV, compensated=True)
    U = haar_unitary(m, 712).to(torch.complex64).to(DEV)
    a = k.probs(U)
    b = k.probs(U.clone().requires_grad_(True))
    assert torch.equal(a, b.detach())


@pytest.mark.parametrize("pull", [False, True])
def test_gradient_n8_m24_block_adjoint(pull, monkeypatch):
    """n=8, m=24 (7.9M outputs; layers 6..8 run the block-structured adjoint): dL/dU of
    L = -mean log p[idx] over 1,024 golden indices against complex128 torch autograd of the
    stored-edge restatement (tests/golden/g8_*.npy); `pull`: the readout adjoint from layer 7
    (the sparse path of large layers) instead of the recomputed layer 8."""
    if pull:
        from photonic_engines_amd.slos import autograd as ag
        monkeypatch.setattr(ag, "SPARSE_MIN", 0)
    U = torch.as_tensor(golden("g8_u.npy")).to(torch.complex64).to(DEV).requires_grad_(True)
    idx = torch.as_tensor(golden("g8_idx.npy"), device=DEV)
    k = FockSLOSKernel(8, 24, first_modes(8, 24), device=DEV)
    loss = -torch.log(k.probs(U)[idx]).mean()
    assert abs(float(loss.detach()) - golden("golden.json")["g8_loss"]) < 1e-3
    loss.backward()
    g, gr = U.grad.cpu().numpy(), golden("g8_grad.npy")
    assert np.abs(g - gr).max() / np.abs(gr).max() < 1e-3
    assert np.abs(g[:, 8:]).max() == 0.0


def _restated_raw(Ur, n, m, occ, mask="fock"):
    """Raw deferred amplitudes of the oracle's stored-edge recurrence in complex128 torch (autograd)."""
    L = so.enumerate_layers(n, m, mask)
    modes = [j for j, c in enumerate(occ) for _ in range(c)]
    psi = torch.ones(1, dtype=torch.complex128)
    for kk, i in enumerate(modes):
        par, chi, mode, _ = so.edges(L[kk], L[kk + 1])
        contrib = Ur[torch.as_tensor(mode), i] * psi[torch.as_tensor(par)]
        psi = torch.zeros(L[kk + 1].shape[0], dtype=torch.complex128).index_add(0, torch.as_tensor(chi), contrib)
    return psi, L


@pytest.mark.parametrize("n,m,occ,norm", [(4, 7, [1, 1, 0, 1, 1, 0, 0], "deferred"), (3, 6, [2, 0, 1, 0, 0, 0], "deferred"),
                                          (3, 5, [1, 0, 2, 0, 0], "immediate"), (7, 9, None, "deferred")])
def test_amplitudes_gradient_matches_restatement(n, m, occ, norm):
    """amplitudes(U) under autograd (kernel.py:191-260): the raw layer through the adjoint kernels,
    the input norm and (immediate) sqrt(prod t!) as torch products, against complex128 autograd of
    the oracle's recurrence, for a loss that uses the complex amplitudes themselves."""
    occ = occ or [1] * n + [0] * (m - n)
    U0 = haar_unitary(m, 30 + n + m)
    k = FockSLOSKernel(n, m, occ, device=DEV, normalization=norm)
    U = U0.to(torch.complex64).to(DEV).requires_grad_(True)
    a = k.amplitudes(U)
    S = a.shape[0]
    c = torch.exp(1j * torch.linspace(0.0, 3.0, S, dtype=torch.float64)).to(torch.complex64)
    w = torch.linspace(0.5, 1.5, S, dtype=torch.float32)
    loss = (a * c.to(DEV)).real.sum() + (a.abs() ** 2 * w.to(DEV)).sum()
    loss.backward()
    Ur = U0.clone().requires_grad_(True)
    psi, L = _restated_raw(Ur, n, m, occ)
    ar = psi / math.sqrt(math.prod(math.factorial(v) for v in occ))
    if norm == "immediate":
        ar = ar * torch.sqrt(torch.as_tensor(so.prob_norms(L[n]), dtype=torch.float64))
    assert np.abs(a.detach().cpu().numpy() - ar.detach().numpy()).max() < 1e-6
    lr = (ar * c.to(torch.complex128)).real.sum() + (ar.abs() ** 2 * w.double()).sum()
    lr.backward()
    g, gr = U.grad.cpu().numpy(), Ur.grad.numpy()
    assert np.abs(g - gr).max() < 2e-5 * max(1.0, np.abs(gr).max())


@pytest.mark.parametrize("n,m,x,order", [(3, 6, 0.8, None), (4, 6, 0.6, None), (4, 7, 0.9, 1)])
def test_xmodel_gradient_in_u_matches_finite_differences(n, m, x, order):
    """probs(U, indist=x) is differentiable in U (species.py:37-41): the interfering subsets through
    SLOSRaw, the raises as complex layer steps with their adjoint; dL/dU of L = sum w p against
    complex128 central finite differences (h = 1e-6) of the oracle's species mixture."""
    occ = [1] * n + [0] * (m - n)
    U0 = haar_unitary(m, 70 + n + m)
    k = FockSLOSKernel(n, m, occ, device=DEV)
    U = U0.to(torch.complex64).to(DEV).requires_grad_(True)
    p = k.probs(U, indist=x, order=order)
    w = torch.linspace(0.5, 1.5, p.shape[-1], device=DEV)
    (p * w).sum().backward()
    wn = w.cpu().double().numpy()
    Un = U0.numpy()

    def L(Uv):
        return float(so.species_mixture(Uv, n, m, occ, x, order=order) @ wn)

    assert np.abs(p.detach().cpu().numpy() - so.species_mixture(Un, n, m, occ, x, order=order)).max() < 1e-6
    h = 1e-6
    g_fd = np.zeros((m, m), dtype=np.complex128)
    for j in range(m):
        for i in range(n):                               # only the input columns carry gradient
            for unit in (1.0, 1j):
                Up, Um = Un.copy(), Un.copy()
                Up[j, i] += unit * h
                Um[j, i] -= unit * h
                g_fd[j, i] += (L(Up) - L(Um)) / (2 * h) * (1.0 if unit == 1.0 else 1j)
    g = U.grad.cpu().numpy()
    assert np.abs(g - g_fd).max() < 2e-5 * max(1.0, np.abs(g_fd).max())
    assert np.abs(g[:, n:]).max() == 0.0


def test_gradient_wide_modes_adjoint_lds_above_64k():
    """n=2, m=200: the thread adjoint's per-lane mode sums take 102 KB of LDS (more than the 64 KB of
    earlier CDNA parts). dL/dU of L = sum_s w_s p_s against complex128 torch autograd of the closed
    form p(i, j) = |U[i,0] U[j,1] + U[j,0] U[i,1]|^2 (i < j), 2 |U[i,0] U[i,1]|^2 (i = j)."""
    m = 200
    U0 = haar_unitary(m, 200)
    k = FockSLOSKernel(2, m, [1, 1] + [0] * (m - 2), device=DEV)
    U = U0.to(torch.complex64).to(DEV).requires_grad_(True)
    p = k.probs(U)
    w = torch.linspace(0.5, 1.5, p.shape[-1], device=DEV)
    (p * w).sum().backward()
    i, j = np.triu_indices(m)
    i, j = torch.as_tensor(i), torch.as_tensor(j)
    Ur = U0.clone().requires_grad_(True)
    amp = Ur[i, 0] * Ur[j, 1] + Ur[j, 0] * Ur[i, 1]
    pr = amp.abs() ** 2 / torch.where(i == j, 2.0, 1.0).double()
    assert np.abs(p.detach().double().cpu().numpy() - pr.detach().numpy()).max() < 1e-6
    (pr * w.cpu().double()).sum().backward()
    g, gr = U.grad.cpu().numpy(), Ur.grad.numpy()
    assert np.abs(g - gr).max() < 1e-5 * max(1.0, np.abs(gr).max())
    assert np.abs(g[:, 2:]).max() == 0.0
