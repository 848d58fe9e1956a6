"""GPU parity of FockSLOSKernel(dtype=torch.complex128) (kernel.py:92,117-118: every layer in
complex128, probabilities in float64) against the complex128 oracle. The engine's FP64 kernels sum
the same parents as the oracle, in its own order, so the bar is FP64 rounding: every output within
1e-10 of its own size plus 1e-15, where the complex64 engine is held to 1e-6 absolute."""
import math

import numpy as np
import pytest
import torch
from helpers import first_modes, golden, haar_unitary, occupation_list

from oracle import pull
from oracle import slos_oracle as so
from photonic_engines_amd import _native as nat
from photonic_engines_amd.slos import FockSLOSKernel, SuperpositionSLOSKernel

pytestmark = pytest.mark.gpu
DEV = "cuda"
C128 = torch.complex128


def _close(got, ref, rel=1e-10, floor=1e-15, what=""):
    g = got.detach().cpu().numpy() if isinstance(got, torch.Tensor) else np.asarray(got)
    err = np.abs(g - ref)
    bad = err > rel * np.abs(ref) + floor
    assert not bad.any(), f"{what}: {int(bad.sum())} outputs off, max|d| = {err.max():.3e}"
    return err.max()


CASES = [(3, 6, "fock", None), (3, 6, "unbunched", None), (3, 6, "dual_rail", None), (2, 8, "sparse_1in4", None),
         (3, 7, "tail_empty", None), (3, 4, "fock", [2, 0, 1, 0]), (6, 6, "fock", [3, 0, 2, 0, 1, 0]),
         (7, 9, "fock", [1, 0, 2, 1, 0, 3, 0, 0, 0]), (8, 10, "fock", None), (9, 12, "unbunched", None),
         (4, 130, "fock", None), (13, 3, "fock", [5, 4, 4]), (1, 5, "fock", None)]


@pytest.mark.parametrize("n,m,mask,occ", CASES)
def test_complex128_matches_oracle(n, m, mask, occ):
    """probs and amplitudes (deferred and immediate) in complex128 against the complex128 pull
    oracle on every mask, bunched inputs and m = 130; the output dtypes are the reference's."""
    occ = occ or occupation_list(n, m, mask)
    if len(occ) < m:
        occ = occ + [0] * (m - len(occ))
    U = haar_unitary(m, n * 31 + m).numpy()
    k = FockSLOSKernel(n, m, occ, mask=mask, device=DEV, dtype=C128)
    assert k.rdtype == torch.float64 and k.output_prob_norm.dtype == torch.float64
    p = k.probs(torch.as_tensor(U))
    assert p.dtype == torch.float64 and p.device.type == "cuda"
    _close(p, pull.forward_probs(n, m, occ, U, mask), what=f"probs {n},{m},{mask}")
    a = k.amplitudes(torch.as_tensor(U))
    assert a.dtype == torch.complex128
    ref_a = pull.forward_amps(n, m, occ, U, mask)
    assert np.abs(a.cpu().numpy() - ref_a).max() <= 1e-13 * max(1.0, np.abs(ref_a).max())
    if k.output_size() <= 100_000:   # (the oracle enumerates every layer in Python)
        ki = FockSLOSKernel(n, m, occ, mask=mask, device=DEV, dtype=C128, normalization="immediate")
        ai = ki.amplitudes(torch.as_tensor(U)).cpu().numpy()
        t = np.sqrt(so.prob_norms(so.enumerate_layers(n, m, mask)[n]))
        assert np.abs(ai - ref_a * t).max() <= 1e-13 * max(1.0, np.abs(ref_a * t).max())


def test_complex128_ryser_and_c1_golden():
    """C1 (n=3, m=12) against its golden Ryser probabilities, now to FP64 rounding."""
    U = torch.as_tensor(golden("c1_u.npy"))
    occ = [1, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0]
    k = FockSLOSKernel(3, 12, occ, device=DEV, dtype=C128)
    p = k.probs(U)
    assert p.shape == (364,) and p.dtype == torch.float64
    _close(p, golden("c1_probs.npy"), rel=1e-9, what="C1")
    assert abs(float(k.total_mass(p)) - 1) < 1e-13


def test_complex128_c2_full():
    """n=8, m=24: all 7,888,725 outputs against the complex128 pull oracle, the 10,000 golden Ryser
    spots, and the FP64 mass."""
    U = golden("c2_u.npy")
    occ = first_modes(8, 24)
    k = FockSLOSKernel(8, 24, occ, device=DEV, dtype=C128)
    p = k.probs(torch.as_tensor(U).to(DEV))
    ref = pull.forward_probs(8, 24, occ, U)
    pc = p.cpu().numpy()
    _close(pc, ref, what="C2")
    assert 0.5 * np.abs(pc - ref).sum() < 1e-12
    idx = golden("c2_spot_idx.npy").astype(np.int64)
    _close(pc[idx], golden("c2_spot_probs.npy"), rel=1e-9, what="C2 Ryser spots")
    assert abs(float(k.total_mass(p)) - 1) < 1e-12
    # the complex64 engine on the same U agrees to its own bar
    p64 = FockSLOSKernel(8, 24, occ, device=DEV).probs(torch.as_tensor(U).to(torch.complex64).to(DEV))
    assert (p64.double() - p).abs().max().item() <= 1e-6


def test_complex128_n11_m22_full():
    """n=11, m=22 (129,024,480 outputs, the shape the reference publishes its probs(U) time for,
    results/keep/sbs_20260919_product_slos_r2_n11.csv:2): every output of the complex128 engine
    against the complex128 pull oracle, the TVD and the FP64 mass."""
    n, m = 11, 22
    U = haar_unitary(m, 1000 * n + m)
    occ = first_modes(n, m)
    p = FockSLOSKernel(n, m, occ, device=DEV, dtype=C128).probs(U.to(DEV))
    ref = pull.forward_probs(n, m, occ, U.numpy())
    pc = p.cpu().numpy()
    _close(pc, ref, what="n=11, m=22")
    assert 0.5 * np.abs(pc - ref).sum() < 1e-12
    assert abs(float(pc.sum()) - 1) < 1e-12


def test_complex128_batched_vacuum_and_errors():
    n, m = 5, 10
    occ = occupation_list(n, m)
    Ub = torch.stack([haar_unitary(m, s) for s in (1, 2, 3)])
    k = FockSLOSKernel(n, m, occ, device=DEV, dtype=C128)
    pb = k.probs(Ub)
    for b in range(3):
        assert torch.equal(pb[b], k.probs(Ub[b]))
        _close(pb[b], pull.forward_probs(n, m, occ, Ub[b].numpy()), what=f"batch row {b}")
    # host-resident U and outputs (the reference's CPU device)
    kc = FockSLOSKernel(n, m, occ, device="cpu", dtype=C128)
    pcpu = kc.probs(Ub[0])
    assert pcpu.device.type == "cpu" and pcpu.dtype == torch.float64 and torch.equal(pcpu, pb[0].cpu())
    # the vacuum: one state of probability 1
    k0 = FockSLOSKernel(0, 4, [0, 0, 0, 0], device=DEV, dtype=C128)
    p0 = k0.probs(haar_unitary(4, 1))
    assert p0.dtype == torch.float64 and p0.tolist() == [1.0]
    with pytest.raises(ValueError):
        FockSLOSKernel(n, m, occ, device=DEV, dtype=C128, compensated=True)
    with pytest.raises(ValueError):
        FockSLOSKernel(n, m, occ, device=DEV, dtype=torch.float64)


@pytest.mark.parametrize("occ,eta", [([1, 0, 1, 0, 1, 0, 0, 0], 0.7), ([2, 0, 1, 0, 1, 1, 0, 0], 0.35)])
def test_complex128_loss_matches_oracle(occ, eta):
    """probs(U, loss=eta) in float64 against the oracle's thinning of the complex128 distribution;
    d/deta (the FP64 descent under the differentiable binomial weights) against central finite
    differences of the oracle; device sampling of the float64 blocks."""
    n, m = sum(occ), len(occ)
    U = haar_unitary(m, 11)
    k = FockSLOSKernel(n, m, occ, device=DEV, dtype=C128)
    ld = k.probs(U, loss=eta)
    L = so.enumerate_layers(n, m)
    ref = [r[0] for r in so.thin_distribution(L, pull.forward_probs(n, m, occ, U.numpy()), eta)]   # [1, S_k] rows
    for kk in range(n + 1):
        assert ld.probs[kk].dtype == torch.float64
        _close(ld.probs[kk], ref[kk], rel=1e-9, what=f"layer {kk}")
    assert abs(sum(float(ld.mass(kk)) for kk in range(n + 1)) - 1) < 1e-12
    # gradient in eta: L(eta) = sum_k c_k . q_k
    c = [torch.linspace(0.1, 1.0, int(ref[kk].shape[0]), dtype=torch.float64) for kk in range(n + 1)]
    e = torch.tensor(eta, dtype=torch.float64, requires_grad=True)
    ldg = k.probs(U, loss=e)
    tot = sum((ldg.probs[kk].cpu() * c[kk]).sum() for kk in range(n + 1))
    tot.backward()
    h = 1e-6
    base = pull.forward_probs(n, m, occ, U.numpy())

    def f(v):
        r = so.thin_distribution(L, base, v)
        return sum(float(r[kk][0] @ c[kk].numpy()) for kk in range(n + 1))
    fd = (f(eta + h) - f(eta - h)) / (2 * h)
    assert abs(float(e.grad) - fd) < 1e-6 * max(1.0, abs(fd)), (float(e.grad), fd)
    # device sampling of the float64 blocks (the FP64 cumsum path above 2^24, multinomial below)
    gen = torch.Generator(device=DEV).manual_seed(5)
    s = ld.sample(2000, generator=gen)
    assert s.shape == (2000, m) and int(s.sum(1).max()) <= n and int(s.min()) >= 0


def test_complex128_loss_gradient_in_u():
    """dL/dU through the FP64 forward, the FP64 descent and its adjoint (loss=0.6 with U requiring
    a gradient) against a central finite difference of the FP64 engine along a random direction."""
    n, m = 4, 8
    occ = [1, 0, 1, 0, 1, 0, 1, 0]
    U0 = haar_unitary(m, 13)
    k = FockSLOSKernel(n, m, occ, device=DEV, dtype=C128)
    c = [torch.linspace(0.2, 1.0, k.layers.size(kk), dtype=torch.float64, device=DEV) for kk in range(n + 1)]

    def f(V):
        ld = k.probs(V, loss=0.6)
        return sum((ld.probs[kk] * c[kk]).sum() for kk in range(n + 1))
    U = U0.to(DEV).requires_grad_(True)
    f(U).backward()
    rng = np.random.default_rng(17)
    D = torch.as_tensor(rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m)))
    h = 1e-6
    with torch.no_grad():
        fd = (float(f((U0 + h * D).to(DEV))) - float(f((U0 - h * D).to(DEV)))) / (2 * h)
    dirg = float((U.grad.cpu().conj() * D).real.sum())
    assert abs(dirg - fd) < 1e-7 * max(1.0, abs(fd)), (dirg, fd)


@pytest.mark.parametrize("n,m,mask,occ", [(3, 6, "fock", [1, 0, 1, 0, 1, 0]), (3, 4, "fock", [2, 0, 1, 0]),
                                          (4, 8, "dual_rail", None), (5, 7, "fock", [1, 1, 1, 1, 1, 0, 0])])
def test_complex128_gradient_matches_restatement(n, m, mask, occ):
    """dL/dU in complex128 through the FP64 adjoint kernels against complex128 autograd of the
    oracle's stored-edge recurrence written in torch (test_kernel.py:131-146 shape), to FP64
    rounding."""
    occ = occ or occupation_list(n, m, mask)
    Ub = torch.stack([haar_unitary(m, s) for s in (6, 7)])
    k = FockSLOSKernel(n, m, occ, mask=mask, device=DEV, dtype=C128)
    U = Ub.to(DEV).requires_grad_(True)
    p = k.probs(U)
    assert p.dtype == torch.float64
    w = torch.linspace(0.5, 1.5, p.shape[-1], device=DEV, dtype=torch.float64)
    (p * w).sum().backward()
    assert U.grad.dtype == C128
    L = so.enumerate_layers(n, m, mask)
    modes = [j for j, c in enumerate(occ) for _ in range(c)]
    s_norm = float(math.prod(math.factorial(c) for c in occ))
    Ur = Ub.clone().requires_grad_(True)
    tot = 0
    for b in range(2):
        psi = torch.ones(1, dtype=C128)
        for kk, i in enumerate(modes):
            par, chi, mode, _ = so.edges(L[kk], L[kk + 1])
            contrib = Ur[b][torch.as_tensor(mode), i] * psi[torch.as_tensor(par)]
            psi = torch.zeros(L[kk + 1].shape[0], dtype=C128).index_add(0, torch.as_tensor(chi), contrib)
        pr = psi.abs() ** 2 / s_norm * torch.as_tensor(so.prob_norms(L[n]), dtype=torch.float64)
        tot = tot + (pr * w.cpu()).sum()
    tot.backward()
    g, gr = U.grad.cpu().numpy(), Ur.grad.numpy()
    assert np.abs(g - gr).max() < 1e-11 * max(1.0, np.abs(gr).max())
    # amplitudes(U) under autograd, immediate normalisation: the same restatement's amplitudes
    Ua = Ub[0].to(DEV).requires_grad_(True)
    ka = FockSLOSKernel(n, m, occ, mask=mask, device=DEV, dtype=C128, normalization="immediate")
    a = ka.amplitudes(Ua)
    va = torch.linspace(-1.0, 1.0, a.shape[-1], device=DEV, dtype=torch.float64)
    (a.real * va).sum().backward()
    Ur2 = Ub[0].clone().requires_grad_(True)
    psi = torch.ones(1, dtype=C128)
    for kk, i in enumerate(modes):
        par, chi, mode, _ = so.edges(L[kk], L[kk + 1])
        psi = torch.zeros(L[kk + 1].shape[0], dtype=C128).index_add(
            0, torch.as_tensor(chi), Ur2[torch.as_tensor(mode), i] * psi[torch.as_tensor(par)])
    ar = psi / math.sqrt(s_norm) * torch.sqrt(torch.as_tensor(so.prob_norms(L[n]), dtype=torch.float64))
    (ar.real * va.cpu()).sum().backward()
    assert np.abs(a.detach().cpu().numpy() - ar.detach().numpy()).max() < 1e-13 * max(1.0, float(ar.detach().abs().max()))
    assert np.abs(Ua.grad.cpu().numpy() - Ur2.grad.numpy()).max() < 1e-11 * max(1.0, float(Ur2.grad.abs().max()))


def test_complex128_gradient_n8_m24():
    """n=8, m=24: the FP64 adjoint at the C2 shape against the complex64 engine's gradient (its
    own bar) and finite differences of the FP64 forward along one random direction."""
    n, m = 8, 24
    occ = first_modes(n, m)
    U0 = haar_unitary(m, 824)
    idx = torch.as_tensor(np.random.default_rng(3).integers(0, 7888725, 512), device=DEV)
    k = FockSLOSKernel(n, m, occ, device=DEV, dtype=C128)
    U = U0.to(DEV).requires_grad_(True)
    loss = -torch.log(k.probs(U)[idx]).mean()
    loss.backward()
    g = U.grad
    D = torch.as_tensor(np.random.default_rng(4).normal(size=(m, m)) + 1j * np.random.default_rng(5).normal(size=(m, m)))
    h = 1e-6

    def f(V):
        with torch.no_grad():
            return float(-torch.log(k.probs(V.to(DEV))[idx]).mean())
    fd = (f(U0 + h * D) - f(U0 - h * D)) / (2 * h)
    dirg = float((g.cpu().conj() * D).real.sum())                     # torch's convention: dL/dRe + i dL/dIm
    assert abs(dirg - fd) < 1e-5 * max(1.0, abs(fd)), (dirg, fd)
    k32 = FockSLOSKernel(n, m, occ, device=DEV)
    U32 = U0.to(torch.complex64).to(DEV).requires_grad_(True)
    (-torch.log(k32.probs(U32)[idx]).mean()).backward()
    rel = (U32.grad.to(C128) - g).abs().max().item() / g.abs().max().item()
    assert rel < 1e-3, rel


def test_complex128_xmodel_superposition_and_grouped():
    """The x-model, the coherent superposition and grouped probabilities in complex128 / float64
    against their oracles, to FP64 rounding."""
    n, m = 4, 7
    occ = [1, 1, 0, 1, 1, 0, 0]
    U = haar_unitary(m, 47)
    U64 = U.numpy()
    k = FockSLOSKernel(n, m, occ, device=DEV, dtype=C128)
    for x, order in ((0.8, None), (0.5, 2)):
        p = k.probs(U, indist=x, order=order)
        assert p.dtype == torch.float64
        _close(p, so.species_mixture(U64, n, m, occ, x, order=order), rel=1e-9, floor=1e-14, what=f"x={x}")
    Ug = U.clone().requires_grad_(True)
    k.probs(Ug, indist=0.8).sum().backward()
    assert Ug.grad.dtype == C128 and torch.isfinite(Ug.grad).all()
    occs = [[1, 1, 1, 0, 0, 0, 0], [2, 0, 0, 1, 0, 0, 0], [0, 0, 1, 1, 1, 0, 0]]
    w = np.array([0.6, 0.3 + 0.4j, -0.5j])
    w = w / np.linalg.norm(w)
    ks = SuperpositionSLOSKernel(3, m, occs, device=DEV, dtype=C128)
    ps = ks.probs(U, weights=w)
    assert ps.dtype == torch.float64
    _close(ps, so.superposition_probs(U64, 3, m, occs, w, "fock"), rel=1e-9, floor=1e-14, what="superposition")
    S = k.output_size()
    groups = torch.as_tensor(np.random.default_rng(1).integers(0, 13, S))
    pg = k.group_matrix(groups).probs(U)
    assert pg.dtype == torch.float64
    ref = pull.forward_probs(n, m, occ, U64)
    _close(pg, np.bincount(groups.numpy(), weights=ref, minlength=13), rel=1e-12, what="grouped")
    assert torch.equal(pg, k.group_matrix(groups).probs(U))


def test_complex128_layer_step_and_dtype_checks():
    """The F64 layer step equals the F64 forward's layers bit for bit (same kernel); mixed dtypes
    and compensated F64 are refused."""
    n, m = 5, 9
    occ = occupation_list(n, m)
    k = FockSLOSKernel(n, m, occ, device=DEV, dtype=C128)
    plan = k._plan
    U = haar_unitary(m, 9).to(DEV).unsqueeze(0).contiguous()
    psi = torch.ones((1, 1), dtype=C128, device=DEV)
    for kk in range(1, n):
        psi = plan.layer_step(U, kk, k.input_modes[kk - 1], psi, nat.OUT_RAW)
    p = plan.layer_step(U, n, k.input_modes[n - 1], psi, nat.OUT_PROBS, scale=k._input_norm)
    assert torch.equal(p, plan.forward(U, scale=k._input_norm))
    with pytest.raises(ValueError):
        plan.layer_step(U, 1, 0, torch.ones((1, 1), dtype=torch.complex64, device=DEV), nat.OUT_RAW)
    with pytest.raises(ValueError):
        plan.forward(U, flags=nat.FLAG_COMPENSATED)
    mp = k.memory_plan(2)
    assert mp["output"] == 2 * 8 * k.output_size() and mp["unitaries"] == 2 * 16 * m * m


def _full_compare(p32: torch.Tensor, p64: torch.Tensor, chunk: int = 1 << 28):
    """max|dp|, TVD and the max relative error (over outputs above 1e-12) of a float32
    distribution against a float64 one, on the GPU in chunks."""
    mx = tvd = rel = 0.0
    for i in range(0, p64.numel(), chunk):
        a, b = p32[i:i + chunk].double(), p64[i:i + chunk]
        d = (a - b).abs()
        mx = max(mx, float(d.max()))
        tvd += float(d.sum())
        big = b > 1e-12
        if bool(big.any()):
            rel = max(rel, float((d[big] / b[big]).max()))
    return mx, 0.5 * tvd, rel


def test_complex128_c5_ryser_spots_and_every_complex64_output():
    """n=11, m=32 (4,280,561,376 outputs) in complex128: the 10,000 golden Ryser spots to FP64
    rounding and the FP64 mass; then every output of the complex64 engine on the same U against
    it: max|dp| <= 1e-6 (the frontier workload's bar, held by every output, not only the spots)."""
    props = torch.cuda.get_device_properties(0)
    if props.total_memory < 120 * 2 ** 30:
        pytest.skip("needs ~80 GB of device memory")
    U = torch.as_tensor(golden("c5_u.npy"))
    occ = first_modes(11, 32)
    k = FockSLOSKernel(11, 32, occ, device=DEV, dtype=C128)
    p64 = k.probs(U.to(DEV))
    assert p64.shape == (4_280_561_376,) and p64.dtype == torch.float64
    assert abs(float(k.total_mass(p64)) - 1) < 1e-12
    idx = torch.as_tensor(golden("c5_spot_idx.npy").astype(np.int64), device=DEV)
    _close(p64[idx], golden("c5_spot_probs.npy"), rel=1e-8, floor=1e-22, what="C5 Ryser spots")
    torch.cuda.empty_cache()
    p32 = FockSLOSKernel(11, 32, occ, device=DEV).probs(U.to(torch.complex64).to(DEV))
    mx, tvd, rel = _full_compare(p32, p64)
    assert mx <= 1e-6, mx
    assert tvd < 1e-4, tvd
    assert rel < 1e-2, rel

