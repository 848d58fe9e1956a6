"""GPU parity of the forward path (FockSLOSKernel.probs / amplitudes through the C ABI) against the
oracle. Tolerances: the engine computes in complex64; the bar is max|dp| <= 1e-6 against complex128
(the accuracy target of the frontier workload), with the max relative error reported by assertion
messages."""
import math

import numpy as np
import pytest
import torch
from helpers import first_modes, golden, haar_unitary, input_norm, occupation_list

from oracle import pull
from oracle import slos_oracle as so
from photonic_engines_amd import _native as nat
from photonic_engines_amd.slos import FockSLOSKernel, get_layers

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL = 1e-6


def _check(p_gpu, p_ref, tol=TOL, what=""):
    p = p_gpu.detach().double().cpu().numpy()
    err = np.abs(p - p_ref).max()
    assert err <= tol, f"{what}: max|dp| = {err:.3e}"
    return err


@pytest.mark.parametrize("mask", ["fock", "unbunched", "dual_rail", "tail_empty"])
def test_matches_permanents_all_masks(mask):                                  # test_kernel.py:38-46
    n, m = 3, 6
    U = haar_unitary(m, 7)
    occ = occupation_list(n, m, mask)
    p = FockSLOSKernel(n, m, occ, mask=mask, device=DEV).probs(U.to(torch.complex64))
    _check(p, so.reference_probs(U.numpy(), n, m, occ, mask), what=mask)


CASES = [(3, 6, "fock", None), (3, 6, "unbunched", None), (3, 6, "dual_rail", None), (2, 8, "sparse_1in4", None),
         (3, 7, "tail_empty", None), (3, 4, "fock", [2, 0, 1, 0]), (5, 7, "fock", None), (4, 8, "dual_rail", None),
         (3, 12, "sparse_1in4", None), (4, 7, "tail_empty", [0, 2, 1, 1, 0, 0, 0]), (6, 6, "fock", [3, 0, 2, 0, 1, 0]),
         (7, 9, "fock", [1, 0, 2, 1, 0, 3, 0, 0, 0]), (8, 10, "fock", None), (6, 12, "dual_rail", None),
         (5, 20, "sparse_1in4", None), (6, 8, "fock", [6, 0, 0, 0, 0, 0, 0, 0]), (9, 11, "fock", None),
         (2, 129, "fock", None), (1, 5, "fock", None), (4, 130, "fock", None), (5, 60, "unbunched", None),
         # prefix widths d = n - 4 of 5..9 (every stream-slot count the block kernel is built for)
         (9, 5, "fock", [2, 2, 2, 2, 1]), (10, 5, "fock", [3, 0, 4, 1, 2]), (11, 4, "fock", [5, 0, 3, 3]),
         (12, 4, "fock", [3, 3, 3, 3]), (13, 3, "fock", [5, 4, 4]), (9, 12, "unbunched", None),
         (10, 20, "dual_rail", None),
         # 5-position suffixes with 7 and 8 distinct prefix modes (layers 12 and 13)
         (13, 9, "fock", [2, 2, 2, 2, 1, 1, 1, 1, 1]),
         # the block kernel's widest mode count (6-bit positions in the tuple words)
         (5, 64, "fock", None), (6, 64, "unbunched", None)]


@pytest.mark.parametrize("n,m,mask,occ", CASES)
def test_matches_oracle(n, m, mask, occ):
    occ = occ or occupation_list(n, m, mask)
    if len(occ) < m:
        occ = occ + [0] * (m - len(occ))
    U = haar_unitary(m, n * 31 + m).numpy()
    ref = pull.forward_probs(n, m, occ, U, mask)
    k = FockSLOSKernel(n, m, occ, mask=mask, device=DEV)
    p = k.probs(torch.as_tensor(U).to(torch.complex64))
    _check(p, ref, what=f"{n},{m},{mask}")
    # the thread-per-state kernel (ablation arm) agrees too
    ps = k._plan.forward(torch.as_tensor(U).to(torch.complex64).to(DEV).unsqueeze(0).contiguous(), scale=input_norm(occ),
                         flags=nat.FLAG_SIMPLE)
    _check(ps[0], ref, what=f"simple {n},{m},{mask}")
    amp = k.amplitudes(torch.as_tensor(U).to(torch.complex64))
    ref_amp = pull.forward_amps(n, m, occ, U, mask)
    # raw (deferred) amplitudes of bunched states grow like sqrt(prod t!): relative to their scale
    assert np.abs(amp.cpu().numpy() - ref_amp).max() < 1e-5 * max(1.0, np.abs(ref_amp).max())


def test_c1_golden():
    U = torch.as_tensor(golden("c1_u.npy"))
    occ = [1, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0]
    k = FockSLOSKernel(3, 12, occ, device=DEV)
    p = k.probs(U.to(torch.complex64))
    assert p.shape == (364,) and p.dtype == torch.float32
    _check(p, golden("c1_probs.npy"), what="C1")
    assert abs(float(k.total_mass(p)) - 1) < 1e-6


def test_c2_full_and_spot():
    """n=8, m=24: all 7,888,725 outputs against the complex128 pull oracle, 10,000 against Ryser."""
    U = golden("c2_u.npy")
    occ = first_modes(8, 24)
    k = FockSLOSKernel(8, 24, occ, device=DEV)
    p = k.probs(torch.as_tensor(U).to(torch.complex64).to(DEV))
    ref = pull.forward_probs(8, 24, occ, U)
    err = _check(p, ref, what="C2")
    pc = p.double().cpu().numpy()
    tvd = 0.5 * np.abs(pc - ref).sum()
    assert tvd < 1e-4, tvd
    idx = golden("c2_spot_idx.npy").astype(np.int64)
    assert np.abs(pc[idx] - golden("c2_spot_probs.npy")).max() <= TOL
    assert abs(float(k.total_mass(p)) - 1) < 1e-5
    # compensated sums carry every layer's rounding error (hi/lo layers): against the exact result
    # for the complex64 U the engine sees (complex128 arithmetic on the rounded U), only the
    # final float32 rounding is left; against the complex128 U, the TVD drops below the plain one
    U32 = torch.as_tensor(U).to(torch.complex64)
    kc = FockSLOSKernel(8, 24, occ, device=DEV, compensated=True)
    pcmp = kc.probs(U32.to(DEV)).double().cpu().numpy()
    exact = pull.forward_probs(8, 24, occ, U32.to(torch.complex128).numpy())
    big = exact > 1e-12
    rel_c = (np.abs(pcmp - exact) / exact)[big]
    rel_p = (np.abs(pc - exact) / exact)[big]
    assert rel_c.max() < 1e-6, rel_c.max()                        # float32 output rounding, ~6e-8
    assert np.median(rel_c) < 0.25 * np.median(rel_p), (np.median(rel_c), np.median(rel_p))
    assert 0.5 * np.abs(pcmp - ref).sum() < tvd
    _check(torch.as_tensor(pcmp), ref, what="C2 compensated")


def test_n9_m22_full_against_complex128():
    """n=9, m=22 (14,307,150 outputs; layers 7..9 in group order at a realistic m): every output
    against the complex128 pull oracle, relative to its own size, plus the TVD and the FP64 mass."""
    n, m = 9, 22
    U = haar_unitary(m, 9022)
    occ = first_modes(n, m)
    k = FockSLOSKernel(n, m, occ, device=DEV)
    p = k.probs(U.to(torch.complex64).to(DEV))
    ref = pull.forward_probs(n, m, occ, U.numpy())
    pc = p.double().cpu().numpy()
    assert np.abs(pc - ref).max() <= 1e-9
    assert np.all(np.abs(pc - ref) <= 1e-3 * ref + 1e-12)
    assert 0.5 * np.abs(pc - ref).sum() < 1e-5
    assert abs(float(k.total_mass(p)) - 1) < 1e-5


@pytest.mark.parametrize("n,m,S,alternate", [(11, 22, 129_024_480, False), (12, 24, 834_451_800, False),
                                             (10, 20, 20_030_010, True)])
def test_published_shapes_full_against_complex128(n, m, S, alternate):
    """The shapes the reference publishes its own times for: n=11, m=22 (its probs(U),
    results/keep/sbs_20260919_product_slos_r2_n11.csv:2), n=12, m=24 (its C engine,
    results/keep/cpu48_scaling_v2.csv:14; layer 12 runs the 5-position-suffix block kernel with 7
    stream slots) and n=10, m=20 with the photons in alternate modes
    (results/keep/bench_emitter_slos_h100_20260909_modeslab.csv:4): every output against the
    complex128 pull oracle, relative to its own size, plus the TVD and the FP64 mass."""
    U = haar_unitary(m, 1000 * n + m)
    occ = [1, 0] * n if alternate else first_modes(n, m)
    k = FockSLOSKernel(n, m, occ, device=DEV)
    p = k.probs(U.to(torch.complex64).to(DEV))
    assert p.shape == (S,)
    ref = pull.forward_probs(n, m, occ, U.numpy())
    pc = p.double().cpu().numpy()
    d = np.abs(pc - ref)
    assert d.max() <= 1e-9
    assert np.all(d <= 1e-3 * ref + 1e-12)
    assert 0.5 * d.sum() < 1e-5
    assert abs(float(k.total_mass(p)) - 1) < 1e-5


def test_batched_equals_looped_and_deterministic():                            # test_kernel.py:64-71
    n, m = 5, 10
    occ = occupation_list(n, m)
    k = FockSLOSKernel(n, m, occ, device=DEV)
    Us = torch.stack([haar_unitary(m, s).to(torch.complex64) for s in (1, 2, 3)]).to(DEV)
    pb = k.probs(Us)
    assert pb.shape == (3, k.output_size())
    for b in range(3):
        assert torch.equal(pb[b], k.probs(Us[b]))
    assert torch.equal(pb, k.probs(Us))


@pytest.mark.parametrize("n,m,mask,occ", [(9, 14, "fock", None), (8, 16, "fock", [2, 0, 1, 0, 3, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0]),
                                          (7, 18, "tail_empty", None), (6, 24, "unbunched", None),
                                          (10, 9, "fock", [2, 2, 1, 1, 1, 1, 1, 1, 0])])
def test_block_kernel_bitwise_equals_thread_kernel(n, m, mask, occ):
    """Every launch form sums a child's parents in ascending photon position: the block kernel (64-child
    steps, 128-child adjacent pairs, staged or global local parents) and the thread-per-state kernel
    give bit-identical layers, probabilities and scaled amplitudes; compensated too."""
    occ = occ or occupation_list(n, m, mask)
    k = FockSLOSKernel(n, m, occ, mask=mask, device=DEV)
    U = torch.stack([haar_unitary(m, 77 + s).to(torch.complex64) for s in range(2)]).to(DEV).contiguous()
    for flags in (0, nat.FLAG_COMPENSATED):
        for kind in (nat.OUT_PROBS, nat.OUT_AMPS):
            a = k._plan.forward(U, out_kind=kind, scale=input_norm(occ), flags=flags)
            b = k._plan.forward(U, out_kind=kind, scale=input_norm(occ), flags=flags | nat.FLAG_SIMPLE)
            assert torch.equal(a, b), (flags, kind)


def test_host_device_roundtrip_and_amplitude_norms():                          # test_tutorial.py:45-59
    U = haar_unitary(8, 0)
    k = FockSLOSKernel(n=3, m=8, input_occupation=[1, 0, 1, 0, 1, 0, 0, 0], mask="fock", device="cpu")
    p = k.probs(U.to(torch.complex64))
    assert p.device.type == "cpu" and p.shape == (120,) and abs(float(p.sum()) - 1) < 1e-6
    amp = k.amplitudes(U.to(torch.complex64)) * torch.sqrt(torch.as_tensor(k.layers.output_prob_norms(), dtype=torch.float32))
    assert torch.allclose(amp.abs() ** 2, p, atol=1e-6)
    ki = FockSLOSKernel(3, 8, [1, 0, 1, 0, 1, 0, 0, 0], normalization="immediate", device="cpu")
    assert torch.allclose(ki.amplitudes(U.to(torch.complex64)), amp, atol=1e-6)
    assert torch.allclose(ki.probs(U.to(torch.complex64)), p, atol=1e-7)


def test_immediate_equals_deferred_bunched():                                  # test_kernel.py:96-113
    for n, m, mask, occ in [(2, 4, "fock", [1, 0, 1, 0]), (3, 6, "fock", [1, 0, 1, 0, 1, 0]),
                            (2, 4, "unbunched", [1, 0, 1, 0]), (4, 8, "dual_rail", [1, 0, 1, 0, 1, 0, 1, 0]),
                            (3, 6, "tail_empty", [1, 1, 1, 0, 0, 0]), (3, 4, "fock", [2, 0, 1, 0])]:
        U = haar_unitary(m, 13).to(torch.complex64)
        pd = FockSLOSKernel(n, m, occ, mask=mask, device=DEV).probs(U)
        pi = FockSLOSKernel(n, m, occ, mask=mask, normalization="immediate", device=DEV).probs(U)
        assert torch.allclose(pd, pi, atol=1e-6)
        ref = so.reference_probs(U.to(torch.complex128).numpy(), n, m, occ, mask)
        _check(pd, ref, what=f"{n},{m},{mask},{occ}")


def test_subset_amplitudes():
    """_amplitudes(U, input_modes=...) stops at layer g of the same basis, raw (kernel.py:201-206)."""
    n, m = 4, 7
    occ = first_modes(n, m)
    k = FockSLOSKernel(n, m, occ, device=DEV)
    U = haar_unitary(m, 5)
    sub = k._amplitudes(U.to(torch.complex64), input_modes=[1, 3])
    ref = pull.forward_amps(2, m, [0, 1, 0, 1, 0, 0, 0], U.numpy())
    assert sub.shape == (math.comb(m + 1, 2),)
    assert np.abs(sub.cpu().numpy() - ref).max() < 1e-6


def test_frontier_shape_mass_and_spot_checks():
    """n=11, m=32 (4.28e9 outputs): FP64 total mass of the float32 output and 10,000 seeded outputs
    against complex128 Ryser permanents (computed on the CPU when the fixtures were made)."""
    props = torch.cuda.get_device_properties(0)
    if props.total_memory < 40 * 2 ** 30:
        pytest.skip("needs ~26 GB of device memory")
    U = torch.as_tensor(golden("c5_u.npy"))
    k = FockSLOSKernel(11, 32, first_modes(11, 32), device=DEV)
    p = k.probs(U.to(torch.complex64).to(DEV))
    assert p.shape == (4_280_561_376,)
    mass = float(k.total_mass(p))
    assert abs(mass - 1) < 1e-4, mass
    idx = torch.as_tensor(golden("c5_spot_idx.npy").astype(np.int64), device=DEV)
    got = p[idx].double().cpu().numpy()
    ref = golden("c5_spot_probs.npy")
    assert np.abs(got - ref).max() <= TOL
    rel = np.abs(got - ref) / np.maximum(ref, 1e-30)
    assert np.median(rel) < 1e-5, np.median(rel)
    assert np.all(np.abs(got - ref) <= 2e-4 * ref + 1e-16), rel.max()     # every spot, relative
    # the device FP64 Ryser oracle agrees with the host one
    gs = k.spot_probs(U, idx[:256]).cpu().numpy()
    assert np.abs(gs - ref[:256]).max() < 1e-12 * max(1.0, ref[:256].max())
    del p
    torch.cuda.empty_cache()
    # compensated sums (hi/lo layers) are more accurate on the same 10,000 spots
    kc = FockSLOSKernel(11, 32, first_modes(11, 32), device=DEV, compensated=True)
    pc = kc.probs(U.to(torch.complex64).to(DEV))
    gc = pc[idx].double().cpu().numpy()
    rel_c = np.abs(gc - ref) / np.maximum(ref, 1e-30)
    assert rel_c.max() < rel.max(), (rel_c.max(), rel.max())
    assert np.median(rel_c) < np.median(rel), (np.median(rel_c), np.median(rel))
    assert abs(float(kc.total_mass(pc)) - 1) < 1e-4


def _phase_dft(m: int, seed: int) -> torch.Tensor:
    """A cheap m x m unitary for wide-mode tests: diag(phases) . DFT / sqrt(m) . diag(phases)."""
    g = np.random.default_rng(seed)
    d1 = np.exp(2j * np.pi * g.random(m))
    d2 = np.exp(2j * np.pi * g.random(m))
    k = np.arange(m)
    F = np.exp(-2j * np.pi * np.outer(k, k) / m) / math.sqrt(m)
    return torch.as_tensor((d1[:, None] * F) * d2[None, :])


def test_two_photons_wide_modes_counting_table_above_64k():
    """n=2, m=3000 (4.5M outputs): the counting table, U column and mode list need 108 KB of LDS,
    more than the 64 KB of earlier CDNA parts; every probability against the closed form
    |U[i,0] U[j,1] + U[j,0] U[i,1]|^2 (i < j), 2 |U[i,0] U[i,1]|^2 (i = j), in lex order of (i, j)."""
    m = 3000
    U = _phase_dft(m, 3000)
    p = FockSLOSKernel(2, m, [1, 1] + [0] * (m - 2), device=DEV).probs(U.to(torch.complex64)).double().cpu().numpy()
    i, j = np.triu_indices(m)
    u0, u1 = U[:, 0].numpy(), U[:, 1].numpy()
    ref = np.abs(u0[i] * u1[j] + u0[j] * u1[i]) ** 2
    ref[i == j] /= 2
    assert p.shape == ref.shape
    _check(torch.as_tensor(p), ref, tol=1e-9, what="n=2 m=3000")
    assert abs(p.sum() - 1) < 1e-4


def test_impossible_size_raises_before_launch():
    """n=12, m=64 needs ~105 TB of output: probs(U) raises torch.cuda.OutOfMemoryError from the
    pre-flight, with the sizes, before it allocates or launches anything."""
    k = FockSLOSKernel(12, 64, [1] * 12 + [0] * 52, device=DEV)
    assert k.memory_plan(1)["total"] > 10 ** 14
    U = torch.eye(64, dtype=torch.complex64, device=DEV)
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated()
    with pytest.raises(torch.cuda.OutOfMemoryError, match="needs"):
        k.probs(U)
    assert torch.cuda.memory_allocated() == before


def test_vacuum_input_and_single_mode():
    """No photon (n = 0): one output state, probability 1, amplitude 1, batched too; and every
    photon in one mode of a 1-mode space: probability |U|^(2n) = 1 (kernel.py:90-126 accept both)."""
    k0 = FockSLOSKernel(0, 5, [0] * 5, device=DEV)
    U5 = torch.stack([haar_unitary(5, s) for s in (1, 2)]).to(torch.complex64).to(DEV)
    assert k0.output_size() == 1
    assert torch.equal(k0.probs(U5[0]), torch.ones(1, device=DEV))
    assert torch.equal(k0.probs(U5), torch.ones(2, 1, device=DEV))
    assert torch.equal(k0.amplitudes(U5[0]), torch.ones(1, dtype=torch.complex64, device=DEV))
    k1 = FockSLOSKernel(4, 1, [4], device=DEV)
    U1 = torch.tensor([[complex(math.cos(0.3), math.sin(0.3))]], dtype=torch.complex64, device=DEV)
    assert abs(float(k1.probs(U1)[0]) - 1.0) < 1e-6
    Ug = U5[0].clone().requires_grad_(True)
    assert torch.equal(k0.probs(Ug).detach(), torch.ones(1, device=DEV))
