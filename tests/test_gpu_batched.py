"""GPU parity of uneven batches (UnevenBatch / slos_group_*): every kernel's output is bit-for-bit
its own FockSLOSKernel.probs(U) (the same summation order in every launch form), tuned launches
and hipGraph replays included; the batched uneven workload (512 unitaries, 1/3 each of (4,12),
(6,16), (8,20)) is checked instance by instance against complex128 / FP64 oracles."""
import numpy as np
import pytest
import torch
from helpers import haar_unitary

from oracle import pull
from photonic_engines_amd.slos import FockSLOSKernel, UnevenBatch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _us(m, count, seed):
    return torch.stack([haar_unitary(m, seed + i) for i in range(count)]).to(torch.complex64).to(DEV)


def _singles(batch, Us):
    return [k.probs(U) for k, U in zip(batch.kernels, Us)]


SPECS = [(4, 12, [1] * 4 + [0] * 8), (6, 16, [1] * 6 + [0] * 10), (8, 20, [1] * 8 + [0] * 12)]


def test_group_equals_single_kernels_bitwise():
    b = UnevenBatch.from_specs(SPECS)
    Us = [_us(12, 3, 100), _us(16, 2, 200), _us(20, 2, 300)]
    got = b.probs(Us)
    for g, r in zip(got, _singles(b, Us)):
        assert g.shape == r.shape and torch.equal(g, r)


def test_group_masks_and_bunched_inputs():
    specs = [(3, 8, [2, 0, 1, 0, 0, 0, 0, 0], "fock"), (4, 8, [1, 0, 1, 0, 1, 0, 1, 0], "unbunched"),
             (3, 6, [1, 0, 1, 0, 1, 0], "dual_rail"), (1, 5, [0, 1, 0, 0, 0], "fock"),
             (6, 9, [0, 3, 0, 1, 1, 1, 0, 0, 0], "tail_empty"), (9, 5, [2, 2, 2, 2, 1], "fock")]
    b = UnevenBatch.from_specs(specs)
    Us = [_us(s[1], 2, 17 * i) for i, s in enumerate(specs)]
    got = b.probs(Us)
    for s, g, r, U in zip(specs, got, _singles(b, Us), Us):
        assert torch.equal(g, r), s
        ref = pull.forward_probs(s[0], s[1], s[2], U[0].cpu().to(torch.complex128).numpy(), s[3])
        assert np.abs(g[0].double().cpu().numpy() - ref).max() < 1e-6, s


def test_compensated_group_equals_compensated_single():
    b = UnevenBatch.from_specs(SPECS, compensated=True)
    Us = [_us(12, 2, 5), _us(16, 2, 6), _us(20, 1, 7)]
    for g, r in zip(b.probs(Us), _singles(b, Us)):
        assert torch.equal(g, r)


def test_tuned_launches_and_graph_replays():
    b = UnevenBatch.from_specs(SPECS)
    Us = [_us(12, 4, 1), _us(16, 3, 2), _us(20, 2, 3)]
    cfg = b.tune(Us, reps=1)
    assert len(cfg) == 8 and all(c["kernel"] in ("block", "thread", "per_plan") for c in cfg)
    assert all(c["kernel"] in ("thread", "per_plan") for c in cfg[:4])  # layers <= 4 have no block form
    ref = _singles(b, Us)
    for g, r in zip(b.probs(Us), ref):
        assert torch.equal(g, r)
    b.capture(Us)
    for rep in range(3):
        Us2 = [_us(12, 4, 50 + rep), _us(16, 3, 60 + rep), _us(20, 2, 70 + rep)]
        out = b.replay(Us2)
        torch.cuda.synchronize()
        ref2 = _singles(b, Us2)  # eager work between the replays
        torch.cuda.synchronize()
        for g, r in zip(out, ref2):
            assert torch.equal(g, r), f"replay {rep}"


def test_batched_uneven_512_every_instance():
    """512 seeded Haar unitaries, 1/3 each of (n, m) = (4,12), (6,16), (8,20), one graph launch;
    every instance against the complex128 pull oracle ((4,12), (6,16), and the first (8,20)
    instances in full) or FP64 Ryser permanents (64 seeded outputs) plus its FP64 total mass."""
    counts = [171, 171, 170]
    b = UnevenBatch.from_specs(SPECS)
    Us = [_us(s[1], c, 10_000 * (i + 1)) for i, (s, c) in enumerate(zip(SPECS, counts))]
    b.tune(Us, reps=1)
    b.capture(Us)
    out = b.replay()
    torch.cuda.synchronize()
    rng = np.random.default_rng(512)
    worst = 0.0
    for (n, m, occ), k, U, P in zip(SPECS, b.kernels, Us, out):
        mass = k.total_mass(P).cpu().numpy()
        assert np.abs(mass - 1).max() < 1e-5, (n, m)
        U128 = U.cpu().to(torch.complex128)
        for i in range(U.shape[0]):
            if m < 20 or i < 3:
                ref = pull.forward_probs(n, m, occ, U128[i].numpy(), "fock")
                err = np.abs(P[i].double().cpu().numpy() - ref).max()
            else:
                idx = torch.as_tensor(np.sort(rng.choice(P.shape[1], 64, replace=False)).astype(np.int64), device=DEV)
                ref = k.spot_probs(U128[i], idx).cpu().numpy()
                err = np.abs(P[i, idx].double().cpu().numpy() - ref).max()
            worst = max(worst, err)
            assert err <= 1e-6, (n, m, i, err)
    assert worst > 0.0
