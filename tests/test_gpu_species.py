"""GPU parity of partial distinguishability (FockSLOSKernel.probs(indist=x), the x-model) against
the oracle's Gram-matrix distribution and its species-mixture restatement
(photonic_engines/slos/species.py:37-146; the reference's own checks are
tests/test_imperfections.py:432-460). Tolerance: complex64 engine, max|dp| <= 1e-6."""
import numpy as np
import pytest
import torch
from helpers import haar_unitary

from oracle import slos_oracle as so
from photonic_engines_amd.imperfections import Distinguishability, truncation_bound
from photonic_engines_amd.slos import FockSLOSKernel

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL = 1e-6


def _err(p, ref):
    return float(np.abs(p.detach().double().cpu().numpy() - ref).max())


def test_x_model_equals_gram_oracle():                                       # test_imperfections.py:432-441
    n, m, x = 4, 6, 0.7
    U = haar_unitary(m, 3)
    occ = [1, 1, 0, 1, 1, 0]
    k = FockSLOSKernel(n, m, occ, device=DEV)
    S = x * np.ones((n, n)) + (1 - x) * np.eye(n)
    p = k.probs(U.to(torch.complex64), indist=x)
    U64 = U.to(torch.complex64).to(torch.complex128).numpy()
    assert _err(p, so.gram_probs(U64, n, m, occ, S)) < TOL
    assert abs(float(p.double().sum()) - 1) < 1e-5
    # the ideal spec is the lossless path itself, bit for bit
    assert torch.equal(k.probs(U.to(torch.complex64), indist=1.0), k.probs(U.to(torch.complex64)))
    assert torch.equal(k.probs(U.to(torch.complex64), indist=Distinguishability(x=1.0)), k.probs(U.to(torch.complex64)))


@pytest.mark.parametrize("n,m,occ,mask,x", [(3, 6, [2, 0, 1, 0, 0, 0], "fock", 0.0), (3, 6, [2, 0, 1, 0, 0, 0], "fock", 0.45),
                                            (4, 8, [1, 0, 1, 0, 1, 0, 1, 0], "unbunched", 0.6),
                                            (3, 6, [1, 0, 1, 0, 1, 0], "dual_rail", 0.3),
                                            (3, 7, [1, 1, 0, 1, 0, 0, 0], "tail_empty", 0.8),
                                            (5, 7, [1, 1, 1, 1, 1, 0, 0], "fock", 0.9)])
def test_x_model_masks_and_bunched_inputs(n, m, occ, mask, x):
    U = haar_unitary(m, 10 * n + m)
    k = FockSLOSKernel(n, m, occ, mask=mask, device=DEV)
    p = k.probs(U.to(torch.complex64), indist=x)
    U64 = U.to(torch.complex64).to(torch.complex128).numpy()
    assert _err(p, so.species_mixture(U64, n, m, occ, x, mask)) < TOL


def test_order_truncation_batch_gradient_and_loss():
    n, m, x = 4, 6, 0.7
    occ = [1, 1, 0, 1, 1, 0]
    k = FockSLOSKernel(n, m, occ, device=DEV)
    Us = torch.stack([haar_unitary(m, 3), haar_unitary(m, 4)]).to(torch.complex64)
    for b in (0, 1, 2):
        pb = k.probs(Us[0], indist=x, order=b)
        assert abs((1 - float(pb.double().sum())) - truncation_bound(n, x, b)) < 1e-5
        U64 = Us[0].to(torch.complex128).numpy()
        assert _err(pb, so.species_mixture(U64, n, m, occ, x, order=b)) < TOL
    xt = torch.tensor(x, requires_grad=True)
    pb = k.probs(Us, indist=xt)
    assert pb.shape == (2, k.output_size())
    assert _err(pb[0], k.probs(Us[0], indist=x).double().cpu().numpy()) == 0.0
    pb[1].sum().backward()
    assert abs(float(xt.grad)) < 1e-4                                       # the mass is 1 at every x
    # d/dx of one probability against a central difference of the oracle
    xt.grad = None
    q = k.probs(Us[0], indist=xt)
    q[5].backward()
    U64 = Us[0].to(torch.complex128).numpy()
    h = 1e-4
    fd = (so.species_mixture(U64, n, m, occ, x + h)[5] - so.species_mixture(U64, n, m, occ, x - h)[5]) / (2 * h)
    assert abs(float(xt.grad) - fd) < 1e-4
    ld = k.probs(Us[0], indist=x, loss=0.8)
    assert abs(float(sum(blk.double().sum() for blk in ld.probs)) - 1) < 1e-5


def test_u_gradient_flows_on_this_path():
    """U requiring a gradient: the same probabilities as without (to float32 rounding; the
    differentiable path squares the raw amplitudes in torch), and a finite gradient confined to the
    input columns (its values are checked against finite differences in test_gpu_loss_grad.py)."""
    k = FockSLOSKernel(3, 5, [1, 1, 1, 0, 0], device=DEV)
    U0 = haar_unitary(5, 1).to(torch.complex64)
    U = U0.clone().requires_grad_(True)
    p = k.probs(U, indist=0.5)
    assert torch.allclose(p.detach().cpu(), k.probs(U0, indist=0.5).cpu(), atol=1e-7)
    (p * torch.linspace(0.5, 1.5, p.shape[0], device=p.device)).sum().backward()
    assert torch.isfinite(U.grad).all() and U.grad.abs().max() > 1e-3
    assert U.grad[:, 3:].abs().max() == 0


@pytest.mark.parametrize("n,m,mask,occ", [(6, 9, "fock", None), (7, 6, "fock", [2, 0, 1, 3, 1, 0]),
                                          (6, 12, "unbunched", None), (6, 8, "tail_empty", None)])
def test_raise_as_complex_layer_step_equals_raise_kernel(n, m, mask, occ):
    """The x-model's raise (a layer step whose U column is |U[j, col]|^2, on real data) equals the
    probability-level raise kernel slos_raise on every layer, per batch row."""
    from helpers import occupation_list
    from photonic_engines_amd import _native as nat
    occ = occ or occupation_list(n, m, mask)
    k = FockSLOSKernel(n, m, occ, mask=mask, device=DEV)
    plan = k._plan
    g = torch.Generator(device=DEV).manual_seed(n + m)
    for kk in range(1, n + 1):
        q = torch.rand((3, plan.sizes[kk - 1]), generator=g, device=DEV)
        w = torch.rand((3, m), generator=g, device=DEV)
        a = plan.raise_layer(kk, w, q)
        Uw = torch.zeros((3, m, m), dtype=torch.complex64, device=DEV)
        Uw[:, :, 0] = w.to(torch.complex64)
        b = plan.layer_step(Uw, kk, 0, torch.complex(q, torch.zeros_like(q)), nat.OUT_RAW)
        assert torch.all(b.imag == 0)
        assert torch.allclose(b.real, a, rtol=2e-6, atol=0), kk


@pytest.mark.parametrize("order", [None, 2])
def test_x_model_n10_m20_against_fp64(order):
    """n=10, m=20 (20,030,010 outputs, every subset of the 10 photons, or the subsets with at most
    2 distinguishable photons): the complex64 x-model against the complex128 engine's FP64 x-model
    of the same U, and the FP64 mass."""
    n, m, x = 10, 20, 0.8
    U = haar_unitary(m, 1020)
    occ = [1] * n + [0] * (m - n)
    p = FockSLOSKernel(n, m, occ, device=DEV).probs(U.to(torch.complex64).to(DEV), indist=x, order=order)
    p64 = FockSLOSKernel(n, m, occ, device=DEV, dtype=torch.complex128).probs(U.to(DEV), indist=x, order=order)
    assert p.shape == (20_030_010,)
    assert float((p.double() - p64).abs().max()) <= TOL
    mass = float(p64.sum())
    if order is None:
        assert abs(mass - 1) < 1e-9
    else:
        assert abs(mass - (1 - truncation_bound(n, x, order))) < 1e-9
