"""GPU parity of the strong-simulation backend face (photonic_engines/adapters/perceval/backends.py:
54-244, here `adapters.perceval.StrongBackend`; the Perceval plugin class itself needs the perceval
package, which is absent, so its Perceval-type plumbing is parity unpinned). States are looked up by
rank, never through a dict of the whole basis."""
import math

import numpy as np
import pytest
import torch
from helpers import haar_unitary

from oracle import slos_oracle as so
from photonic_engines_amd.adapters.perceval import StrongBackend

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _amp(U, occ_in, occ_out):
    ins, outs = so.input_modes(occ_in), so.input_modes(occ_out)
    norm = math.sqrt(math.prod(math.factorial(c) for c in occ_in) * math.prod(math.factorial(c) for c in occ_out))
    return so.permanent_ryser(np.asarray(U)[np.ix_(outs, ins)]) / norm


def test_fock_input_amplitudes_probabilities_and_samples():
    m, occ = 6, [1, 0, 1, 0, 1, 0]
    U = haar_unitary(m, 4)
    Uc = U.to(torch.complex64).to(torch.complex128).numpy()
    be = StrongBackend(device=DEV)
    be.set_circuit(U.numpy())
    be.set_input_state(occ)
    final = so.to_occupation(so.enumerate_layers(3, m)[3], m).tolist()
    ref = so.reference_probs(Uc, 3, m, occ)
    for i in (0, 7, 55, len(final) - 1):
        t = final[i]
        assert abs(be.probability(t) - ref[i]) < 1e-6
        assert abs(be.prob_amplitude(t) - _amp(Uc, occ, t)) < 1e-6
    assert be.probability([1, 1, 0, 0, 0, 0]) == 0.0 and be.prob_amplitude([3, 0, 0, 0, 0, 1]) == 0j
    ap = be.all_prob()
    assert len(ap) == len(final) and abs(sum(ap) - 1) < 1e-5
    ev = be.evolve()
    assert len(ev) == len(final) and abs(ev[tuple(final[7])] - _amp(Uc, occ, final[7])) < 1e-6
    s = be.samples(500, generator=torch.Generator(device=DEV).manual_seed(1))
    assert len(s) == 500 and all(sum(t) == 3 for t in s)


def test_complex128_backend():
    """StrongBackend(dtype=torch.complex128) builds complex128 kernels (backends.py:59-77,219 pass the
    dtype on): probabilities and amplitudes to FP64 rounding against Ryser permanents of the exact
    U, a lossy block in float64, and samples."""
    m, occ = 7, [1, 0, 1, 1, 0, 0, 0]
    U = haar_unitary(m, 41)
    Un = U.numpy()
    be = StrongBackend(device=DEV, dtype=torch.complex128)
    be.set_circuit(Un)
    be.set_input_state(occ)
    assert be._kernels[3].cdtype == torch.complex128
    final = so.to_occupation(so.enumerate_layers(3, m)[3], m).tolist()
    ref = so.reference_probs(Un, 3, m, occ)
    for i in (0, 9, 40, len(final) - 1):
        t = final[i]
        assert abs(be.probability(t) - ref[i]) < 1e-13
        assert abs(be.prob_amplitude(t) - _amp(Un, occ, t)) < 1e-13
    assert abs(sum(be.all_prob()) - 1) < 1e-12
    lossy = StrongBackend(device=DEV, dtype=torch.complex128, loss=0.6)
    lossy.set_circuit(Un)
    lossy.set_input_state(occ)
    lossy._ensure()
    ld = lossy._lossy[3]
    assert all(q.dtype == torch.float64 for q in ld.probs)
    assert abs(sum(float(ld.mass(k)) for k in range(4)) - 1) < 1e-12
    s = be.samples(300, generator=torch.Generator(device=DEV).manual_seed(2))
    assert len(s) == 300 and all(sum(t) == 3 for t in s)


def test_rank_lookup_at_the_frontier():
    n, m = 11, 32
    be = StrongBackend(device=DEV)
    U = haar_unitary(m, 11032)
    be.set_circuit(U.numpy())
    be.set_input_state([1] * n + [0] * (m - n))
    k = be._kernels[n]
    idx = [0, 123_456_789, 4_280_561_375]
    rows = k.layers.unrank(n, idx)
    spot = k.spot_probs(U.to(torch.complex64).to(torch.complex128), idx).cpu().numpy()
    for r, p in zip(rows, spot):
        occ = [0] * m
        for j in r:
            occ[int(j)] += 1
        assert abs(be.probability(occ) - p) < 1e-6


def test_superposition_indist_species_and_loss():
    m = 6
    U = haar_unitary(m, 9)
    Uc = U.to(torch.complex64).to(torch.complex128).numpy()

    class Sup:
        pass
    sup = Sup()
    sup.occupations = [[1, 1, 1, 0, 0, 0], [0, 0, 1, 1, 1, 0]]
    sup.amplitudes = [0.8, 0.6j]
    be = StrongBackend(device=DEV)
    be.set_circuit(U.numpy())
    be.set_input_state(sup)
    final = so.to_occupation(so.enumerate_layers(3, m)[3], m).tolist()
    ref = so.superposition_probs(Uc, 3, m, sup.occupations, sup.amplitudes)
    assert max(abs(be.probability(t) - ref[i]) for i, t in enumerate(final)) < 1e-6

    occ = [1, 1, 0, 1, 0, 0]
    bi = StrongBackend(device=DEV, indist=0.6)
    bi.set_circuit(U.numpy())
    bi.set_input_state(occ)
    mix = so.species_mixture(Uc, 3, m, occ, 0.6)
    assert max(abs(bi.probability(t) - mix[i]) for i, t in enumerate(final)) < 1e-6
    with pytest.raises(NotImplementedError):
        bi.prob_amplitude(final[0])

    bs = StrongBackend(device=DEV)
    bs.set_circuit(U.numpy())
    bs.set_input_state(occ, species=([], [0, 1, 3]))             # every photon a species of its own
    cl = so.species_mixture(Uc, 3, m, occ, 0.0)
    assert max(abs(bs.probability(t) - cl[i]) for i, t in enumerate(final)) < 1e-6

    bl = StrongBackend(device=DEV, loss=0.7)
    bl.set_circuit(U.numpy())
    bl.set_input_state(occ)
    layers = so.enumerate_layers(3, m)
    thin = so.thin_distribution(layers, so.reference_probs(Uc, 3, m, occ), 0.7)
    for k in range(4):
        occs = so.to_occupation(layers[k], m).tolist()
        for i in range(0, len(occs), max(1, len(occs) // 7)):
            assert abs(bl.probability(occs[i]) - thin[k][0, i]) < 1e-6
    assert abs(sum(bl.prob_distribution().values()) - 1) < 1e-5
    with pytest.raises(NotImplementedError):
        bl.all_prob()
