"""GPU parity of coherent superposition inputs (SuperpositionSLOSKernel, kernel.py:296-431) and grouped
probabilities (GroupedSLOSKernel / group_matrix, kernel.py:286-294, 584-620) against the oracle.
Tolerance: complex64 engine, max|dp| <= 1e-6."""
import numpy as np
import pytest
import torch
from helpers import haar_unitary

from oracle import slos_oracle as so
from photonic_engines_amd.slos import FockSLOSKernel, SuperpositionSLOSKernel

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _err(p, ref):
    return float(np.abs(p.detach().double().cpu().numpy() - ref).max())


@pytest.mark.parametrize("n,m,occs,mask", [
    (3, 6, [[1, 1, 1, 0, 0, 0], [2, 0, 0, 1, 0, 0], [0, 0, 1, 1, 1, 0]], "fock"),
    (2, 4, [[1, 0, 1, 0], [0, 1, 0, 1]], "fock"),
    (3, 8, [[1, 0, 1, 0, 1, 0, 0, 0], [0, 1, 0, 1, 0, 1, 0, 0]], "unbunched"),
    (6, 9, [[1] * 6 + [0] * 3, [0] * 3 + [1] * 6, [2, 0, 2, 0, 2, 0, 0, 0, 0]], "fock")])
def test_superposition_matches_oracle(n, m, occs, mask):
    U = haar_unitary(m, 3 * n + m).to(torch.complex64)
    rng = np.random.default_rng(n + m)
    w = rng.normal(size=len(occs)) + 1j * rng.normal(size=len(occs))
    w = w / np.linalg.norm(w)
    k = SuperpositionSLOSKernel(n, m, occs, mask=mask, device=DEV)
    U64 = U.to(torch.complex128).numpy()
    p = k.probs(U, weights=w)
    assert _err(p, so.superposition_probs(U64, n, m, occs, w, mask)) < 1e-6
    # default weights: uniform 1/sqrt(K); batched U; from a state object
    pu = k.probs(torch.stack([U, U]))
    ref_u = so.superposition_probs(U64, n, m, occs, [len(occs) ** -0.5] * len(occs), mask)
    assert _err(pu[1], ref_u) < 1e-6 and torch.equal(pu[0], pu[1])

    class State:
        pass
    st = State()
    st.n, st.m, st.occupations, st.amplitudes = n, m, occs, 2 * w
    st.norm = lambda: 2.0
    assert _err(SuperpositionSLOSKernel.from_state(st, mask=mask, device=DEV).probs(U), p.cpu().double().numpy()) < 1e-7
    # loss keeps the mass of the (masked) lossless distribution, which is 1 for fock only
    ld = k.probs(U, weights=w, loss=0.8)
    assert abs(float(sum(b.double().sum() for b in ld.probs)) - float(p.double().sum())) < 1e-5


def test_grouped_probabilities():
    n, m = 4, 8
    k = FockSLOSKernel(n, m, [1, 0, 1, 0, 1, 0, 1, 0], device=DEV)
    S = k.output_size()
    groups = torch.as_tensor(np.random.default_rng(1).integers(0, 17, S))
    g = k.group_matrix(groups)
    assert g.output_size() == int(groups.max()) + 1
    U = haar_unitary(m, 5)
    pg = g.probs(U.to(torch.complex64))
    ref = so.reference_probs(U.numpy(), n, m, [1, 0, 1, 0, 1, 0, 1, 0])
    refg = np.bincount(groups.numpy(), weights=ref, minlength=g.output_size())
    assert _err(pg, refg) < 1e-6
    pb = g.probs(torch.stack([U, U]).to(torch.complex64))
    assert pb.shape == (2, g.output_size()) and _err(pb[1], refg) < 1e-6
    ident = k.group_matrix(torch.arange(S)).probs(U.to(torch.complex64))
    assert torch.equal(ident, k.probs(U.to(torch.complex64)))
    with pytest.raises(ValueError):
        k.group_matrix(torch.zeros(S + 1, dtype=torch.int64))


def test_superposition_gradient_matches_finite_differences():
    """SuperpositionSLOSKernel under autograd (the reference's kernel is a torch graph,
    kernel.py:369-420): dL/dU and dL/dw of L = sum v p against complex128 central finite
    differences (h = 1e-6) of the oracle's superposition_probs."""
    n, m = 3, 6
    occs = [[1, 1, 1, 0, 0, 0], [2, 0, 0, 1, 0, 0], [0, 0, 1, 1, 1, 0]]
    U0 = haar_unitary(m, 41)
    rng = np.random.default_rng(3)
    w0 = rng.normal(size=3) + 1j * rng.normal(size=3)
    k = SuperpositionSLOSKernel(n, m, occs, device=DEV)
    U = U0.to(torch.complex64).to(DEV).requires_grad_(True)
    w = torch.as_tensor(w0, dtype=torch.complex64, device=DEV).requires_grad_(True)
    p = k.probs(U, weights=w)
    v = torch.linspace(0.5, 1.5, p.shape[0], device=DEV)
    (p * v).sum().backward()
    vn = v.cpu().double().numpy()

    def L(Uv, wv):
        return float(so.superposition_probs(Uv, n, m, occs, wv) @ vn)

    h = 1e-6
    Un = U0.numpy()
    gU = np.zeros((m, m), dtype=np.complex128)
    for j in range(m):
        for i in range(m):
            for unit in (1.0, 1j):
                Up, Um = Un.copy(), Un.copy()
                Up[j, i] += unit * h
                Um[j, i] -= unit * h
                gU[j, i] += (L(Up, w0) - L(Um, w0)) / (2 * h) * (1.0 if unit == 1.0 else 1j)
    gw = np.zeros(3, dtype=np.complex128)
    for c in range(3):
        for unit in (1.0, 1j):
            wp, wm = w0.copy(), w0.copy()
            wp[c] += unit * h
            wm[c] -= unit * h
            gw[c] += (L(Un, wp) - L(Un, wm)) / (2 * h) * (1.0 if unit == 1.0 else 1j)
    assert np.abs(U.grad.cpu().numpy() - gU).max() < 2e-5 * max(1.0, np.abs(gU).max())
    assert np.abs(w.grad.cpu().numpy() - gw).max() < 2e-5 * max(1.0, np.abs(gw).max())


def test_grouped_sums_large_groups_empty_groups_and_reproducible():
    """n=8, m=24 (7.9M states) summed into 9 interleaved groups (state index mod 9), with empty
    groups between them (ids 2 * (index mod 9)): every group of ~0.9M states over many
    4096-entry chunks, against the FP64 bincount of the same probabilities; two calls give
    identical sums (no atomics); under autograd the torch reduction carries the gradient."""
    n, m = 8, 24
    k = FockSLOSKernel(n, m, [1] * n + [0] * (m - n), device=DEV)
    groups = 2 * (torch.arange(k.output_size(), dtype=torch.int64) % 9)
    g = k.group_matrix(groups)
    U = haar_unitary(m, 824).to(torch.complex64).to(DEV)
    a = g.probs(U)
    b = g.probs(U)
    assert torch.equal(a, b)
    p = k.probs(U).double().cpu().numpy()
    ref = np.bincount(groups.numpy(), weights=p, minlength=g.output_size())
    assert a.shape == (2 * n + 1,)
    assert float(a[1::2].abs().max()) == 0.0
    assert np.abs(a.double().cpu().numpy() - ref).max() < 1e-6
    Ug = U.clone().requires_grad_(True)
    pg = g.probs(Ug)
    pg[0].backward()
    assert Ug.grad is not None and float(Ug.grad.abs().sum()) > 0


def test_superposition_n9_m18_against_fp64():
    """n=9, m=18 (3,124,550 outputs), K = 3 components with complex weights: the complex64 kernel
    against the complex128 kernel's FP64 forward of the same U and weights, and the FP64 mass
    (sum |w_k|^2 for these mutually orthogonal Fock inputs)."""
    n, m = 9, 18
    occs = [[1] * 9 + [0] * 9, [0] * 9 + [1] * 9, [2, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 0, 0]]
    w = torch.tensor([0.6, 0.48j, 0.64], dtype=torch.complex128)
    U = haar_unitary(m, 918)
    p = SuperpositionSLOSKernel(n, m, occs, device=DEV).probs(U.to(torch.complex64).to(DEV), weights=w)
    p64 = SuperpositionSLOSKernel(n, m, occs, device=DEV, dtype=torch.complex128).probs(U.to(DEV), weights=w)
    assert p.shape == (3_124_550,)
    assert float((p.double() - p64).abs().max()) <= 1e-6
    assert abs(float(p64.sum()) - float((w.abs() ** 2).sum())) < 1e-9
