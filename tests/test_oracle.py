"""Pins of the oracle (oracle/): every function is checked against something that does not share its
code — brute force, closed forms, the permutation-sum permanent, known answers of the reference's
own tests, itertools' order, or an independent route — and planted mistakes in the oracle are shown
to fail those checks."""
import hashlib
import math
from itertools import permutations

import numpy as np
import pytest
import torch
from helpers import golden, haar_unitary, occupation_list

from oracle import pull
from oracle import slos_oracle as so

MASK_CASES = [(2, 4, "fock"), (3, 4, "fock"), (2, 5, "unbunched"), (2, 6, "dual_rail"), (2, 8, "sparse_1in4"),
              (2, 5, "tail_empty"), (3, 5, "fock")]                                    # test_basis.py:39-40
ORDER_CASES = [(4, 6, "fock"), (5, 7, "fock"), (3, 9, "fock"), (4, 8, "unbunched"), (4, 8, "dual_rail"),
               (3, 12, "sparse_1in4"), (4, 7, "tail_empty"), (6, 6, "fock")]          # test_basis_vectorized.py:55-56


# ---- permanents --------------------------------------------------------------------------------

def test_permanent_known_answers():
    assert abs(so.permanent_ryser(np.array([[1, 2], [3, 4]])) - 10) < 1e-12       # test_kernel.py:25-27
    for n in range(1, 8):
        assert abs(so.permanent_ryser(np.ones((n, n))) - math.factorial(n)) < 1e-9  # test_reference.py:20
        assert abs(pull.permanent(np.ones((n, n))) - math.factorial(n)) < 1e-9


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_ryser_equals_permutation_sum(n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    brute = sum(np.prod([A[i, p[i]] for i in range(n)]) for p in permutations(range(n)))
    assert abs(so.permanent_ryser(A) - brute) < 1e-12 * max(1, abs(brute))
    assert abs(pull.permanent(A) - brute) < 1e-12 * max(1, abs(brute))


def test_planted_sign_error_in_ryser_is_caught():
    """A Ryser with the classic extra (-1)^n (reference test_reference.py:10-11) fails the pin."""
    A = np.random.default_rng(3).normal(size=(3, 3)) + 0j
    brute = sum(np.prod([A[i, p[i]] for i in range(3)]) for p in permutations(range(3)))
    wrong = (-1) ** 3 * so.permanent_ryser(A)
    assert abs(wrong - brute) > 1e-3


# ---- basis order, sizes, edges --------------------------------------------------------------------

@pytest.mark.parametrize("n,m,mask", MASK_CASES)
def test_states_match_brute_force(n, m, mask):
    L = so.enumerate_layers(n, m, mask)
    for k in range(n + 1):
        occ = [tuple(o) for o in so.to_occupation(L[k], m).tolist()]
        assert len(set(occ)) == len(occ)
        assert set(occ) == so.brute_states(k, m, mask)


@pytest.mark.parametrize("n,m,mask", ORDER_CASES)
def test_order_equals_recursive_builder(n, m, mask):
    L = so.enumerate_layers(n, m, mask)
    R = so.enumerate_layers_recursive(n, m, mask)
    for k in range(n + 1):
        assert [tuple(r) for r in L[k].tolist()] == R[k]


@pytest.mark.parametrize("n,m,mask", [(4, 7, "fock"), (3, 9, "unbunched"), (6, 5, "fock")])
def test_order_equals_itertools(n, m, mask):
    assert [tuple(r) for r in so.enumerate_layers(n, m, mask)[n].tolist()] == so.itertools_order(n, m, mask)


@pytest.mark.parametrize("mask,n,m", [("fock", 3, 6), ("unbunched", 3, 6), ("dual_rail", 3, 6),
                                      ("sparse_1in4", 2, 8), ("tail_empty", 3, 7)])
def test_sizes_closed_form(mask, n, m):                                             # test_kernel.py:30-35
    assert so.enumerate_layers(n, m, mask)[n].shape[0] == so.state_space_size(n, m, mask)


def test_size_known_answers():
    assert so.state_space_size(6, 12, "fock") == 12376                          # test_tutorial.py:33-35
    assert so.state_space_size(6, 12, "unbunched") == 924
    assert so.state_space_size(6, 12, "dual_rail") == 64
    assert so.enumerate_layers(3, 8)[3].shape[0] == 120                          # test_tutorial.py:50
    assert so.enumerate_layers(3, 6)[2].shape[0] == 21                           # test_tutorial.py:196


@pytest.mark.parametrize("n,m,mask", MASK_CASES)
def test_edges_match_creation_operator(n, m, mask):
    L = so.enumerate_layers(n, m, mask)
    occ = [[tuple(o) for o in so.to_occupation(L[k], m).tolist()] for k in range(n + 1)]
    for k in range(n):
        par, chi, mode, fac = so.edges(L[k], L[k + 1])
        got = {(p, c, j, round(f, 9)) for p, c, j, f in zip(par.tolist(), chi.tolist(), mode.tolist(), fac.tolist())}
        assert got == so.brute_edges(occ[k], occ[k + 1])


def test_golden_order_hashes():
    for key, v in golden("golden.json")["order"].items():
        n, m, mask = key.split(",")
        L = so.enumerate_layers(int(n), int(m), mask)[int(n)].astype(np.int16)
        assert L.shape[0] == v["size"]
        assert hashlib.sha256(L.tobytes()).hexdigest() == v["sha256"], key


def test_c_oracle_rank_matches_enumeration():
    for n, m, mask in ORDER_CASES + [(5, 9, "tail_empty"), (3, 10, "sparse_1in4")]:
        L = so.enumerate_layers(n, m, mask)
        for k in range(1, n + 1):
            assert (pull.rank(n, m, mask, L[k].astype(np.int32)) == np.arange(L[k].shape[0])).all()
            assert (pull.unrank(n, m, mask, k, np.arange(L[k].shape[0])) == L[k]).all()


# ---- amplitudes and probabilities --------------------------------------------------------------------

@pytest.mark.parametrize("mask", ["fock", "unbunched", "dual_rail", "tail_empty"])
def test_oracle_matches_permanents(mask):                                         # test_kernel.py:38-46
    n, m = 3, 6
    U = haar_unitary(m, 7).numpy()
    occ = occupation_list(n, m, mask)
    p = so.OracleSLOS(n, m, occ, mask).probs(U)
    assert np.abs(p - so.reference_probs(U, n, m, occ, mask)).max() < 1e-12


def test_oracle_bunched_input_and_immediate_normalisation():                        # test_kernel.py:96-113
    U = haar_unitary(4, 13).numpy()
    for occ in ([2, 0, 1, 0], [0, 3, 0, 0], [1, 1, 1, 0]):
        d = so.OracleSLOS(3, 4, occ).probs(U)
        i = so.OracleSLOS(3, 4, occ, normalization="immediate").probs(U)
        assert np.abs(d - so.reference_probs(U, 3, 4, occ)).max() < 1e-12
        assert np.abs(d - i).max() < 1e-12


def test_invariants():
    n, m = 3, 6
    occ = occupation_list(n, m)
    assert abs(so.OracleSLOS(n, m, occ).probs(haar_unitary(m, 11).numpy()).sum() - 1) < 1e-12   # test_kernel.py:49-54
    # U = I: the output is the input with certainty
    o = so.OracleSLOS(n, m, occ)
    p = o.probs(np.eye(m))
    idx = [tuple(r) for r in so.to_occupation(o.layers[n], m).tolist()].index(tuple(occ))
    assert p[idx] == 1.0 and p.sum() == 1.0
    # a permutation U moves the photons deterministically
    perm = np.eye(m)[[1, 2, 3, 4, 5, 0]]
    q = o.probs(perm)
    out = tuple(int(v) for v in perm @ np.asarray(occ))
    assert q[[tuple(r) for r in so.to_occupation(o.layers[n], m).tolist()].index(out)] == pytest.approx(1.0)
    # Hong-Ou-Mandel: two photons on a 50:50 beam splitter never leave in different modes
    bs = np.array([[1, 1], [1, -1]]) / math.sqrt(2)
    h = so.OracleSLOS(2, 2, [1, 1]).probs(bs)                                     # order (2,0), (1,1), (0,2)
    assert abs(h[1]) < 1e-15 and h[0] == pytest.approx(0.5) and h[2] == pytest.approx(0.5)


def test_planted_transpose_is_caught():
    """The oracle with U transposed (U[i, j] instead of U[j, i]) disagrees with the permanents."""
    U = haar_unitary(6, 7).numpy()
    occ = occupation_list(3, 6)
    good = so.reference_probs(U, 3, 6, occ)
    assert np.abs(so.OracleSLOS(3, 6, occ).probs(U.T) - good).max() > 1e-3


@pytest.mark.parametrize("n,m,mask,occ", [(3, 6, "fock", None), (3, 6, "unbunched", None), (3, 6, "dual_rail", None),
                                          (2, 8, "sparse_1in4", None), (3, 7, "tail_empty", None),
                                          (3, 4, "fock", [2, 0, 1, 0]), (5, 7, "fock", None),
                                          (6, 6, "fock", [3, 0, 2, 0, 1, 0]), (4, 8, "dual_rail", None)])
def test_c_pull_oracle_matches_numpy_oracle(n, m, mask, occ):
    occ = occ or occupation_list(n, m, mask)
    U = haar_unitary(m, n * 31 + m).numpy()
    o = so.OracleSLOS(n, m, occ, mask)
    assert np.abs(o.amplitudes(U) - pull.forward_amps(n, m, occ, U, mask)).max() < 1e-14
    assert np.abs(o.probs(U) - pull.forward_probs(n, m, occ, U, mask)).max() < 1e-14


# ---- loss --------------------------------------------------------------------------------------------

def test_binomial_known_answers():                                                 # test_loss.py:57-59
    p = so.binomial_pmf(5, 0.3)
    assert abs(p.sum() - 1) < 1e-12 and abs(p[5] - 0.3 ** 5) < 1e-12


@pytest.mark.parametrize("occ,eta", [((1, 1, 1, 0), 0.7), ((2, 0, 1, 0, 0), 0.9), ((1, 1, 0, 0), 0.5)])
def test_thinning_matches_pattern_route(occ, eta):                                  # test_loss.py:80-96
    m, n = len(occ), sum(occ)
    U = haar_unitary(m, 11 + n).numpy()
    L = so.enumerate_layers(n, m)
    p = so.OracleSLOS(n, m, list(occ)).probs(U)
    blocks = so.thin_distribution(L, p, eta)
    d = {}
    for k in range(n + 1):
        for o, v in zip(so.to_occupation(L[k], m).tolist(), blocks[k][0]):
            d[tuple(o)] = v
    ref = so.thin_patterns_dict(so.to_occupation(L[n], m), p, eta)
    assert max(abs(d[t] - ref.get(t, 0.0)) for t in d) < 1e-12
    w = so.binomial_pmf(n, eta)
    assert all(abs(blocks[k].sum() - w[k]) < 1e-12 for k in range(n + 1))
    one = so.thin_distribution(L, p, 1.0)
    assert np.array_equal(one[n][0], p) and all(one[k].sum() == 0 for k in range(n))


def test_tutorial_loss_mass():                                                      # test_tutorial.py:191-204
    U = haar_unitary(6, 3).numpy()
    L = so.enumerate_layers(3, 6)
    p = so.OracleSLOS(3, 6, [1, 1, 1, 0, 0, 0]).probs(U)
    b = so.thin_distribution(L, p, 0.8)
    assert abs(b[2].sum() - 3 * 0.8 ** 2 * 0.2) < 1e-12 and b[2].shape[1] == 21


# ---- golden fixtures ---------------------------------------------------------------------------------

def test_golden_c1_probs():
    U = golden("c1_u.npy")
    occ = [1, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0]
    p = golden("c1_probs.npy")
    assert np.abs(so.reference_probs(U, 3, 12, occ) - p).max() < 1e-15
    assert np.abs(pull.forward_probs(3, 12, occ, U) - p).max() < 1e-15
    assert np.array_equal(U, haar_unitary(12, 3012).numpy())


@pytest.mark.parametrize("tag,n,m", [("c2", 8, 24), ("c5", 11, 32)])
def test_golden_spot_probs_against_numpy_ryser(tag, n, m):
    U, idx, ps = golden(f"{tag}_u.npy"), golden(f"{tag}_spot_idx.npy"), golden(f"{tag}_spot_probs.npy")
    occ = [1] * n + [0] * (m - n)
    pos = pull.unrank(n, m, "fock", n, idx[:12])
    for q in range(12):
        t = np.bincount(pos[q], minlength=m)
        assert abs(so.prob_of_occupation(U, occ, t) - ps[q]) < 1e-13 * max(1.0, ps[q])


def test_golden_c3_generator_hash():
    shapes = [(4, 12), (6, 16), (8, 20)]
    parts = [haar_unitary(shapes[i % 3][1], 500000 + i).numpy().astype(np.complex64).ravel() for i in range(512)]
    assert hashlib.sha256(np.concatenate(parts).tobytes()).hexdigest() == golden("golden.json")["c3_sha256"]


def _torch_restatement_loss(U, idx, n=6, m=16):
    """complex128 autograd through our own torch restatement of the stored-edge recurrence."""
    L = so.enumerate_layers(n, m)
    psi = torch.ones(1, dtype=torch.complex128)
    for k, i in enumerate(range(n)):
        par, chi, mode, _ = so.edges(L[k], L[k + 1])
        contrib = U[torch.as_tensor(mode), i] * psi[torch.as_tensor(par)]
        psi = torch.zeros(L[k + 1].shape[0], dtype=torch.complex128).index_add(0, torch.as_tensor(chi), contrib)
    p = (psi.abs() ** 2) * torch.as_tensor(so.prob_norms(L[n]), dtype=torch.float64)
    return -torch.log(p[torch.as_tensor(idx)]).mean()


def test_golden_c4_fd_gradient_matches_autograd_restatement():
    U = torch.as_tensor(golden("c4_u.npy")).clone().requires_grad_(True)
    idx = golden("c4_idx.npy")
    loss = _torch_restatement_loss(U, idx)
    assert abs(float(loss) - golden("golden.json")["c4_loss"]) < 1e-12
    loss.backward()
    g_fd = golden("c4_grad_fd.npy")
    g = U.grad.numpy()
    assert np.abs(g - g_fd).max() < 1e-6 * np.abs(g_fd).max()
    assert np.abs(g[:, 6:]).max() == 0.0                 # only the input columns receive gradient


# ---- partial distinguishability (x-model) ---------------------------------------------------------

def _classical(U, occ_in, occ_out):
    """Distinguishable photons: per(|U|^2[out rows][:, in cols]) / prod t! (multinomial transport)."""
    ins, outs = so.input_modes(occ_in), so.input_modes(occ_out)
    W = np.abs(np.asarray(U)) ** 2
    return so.permanent_brute(W[np.ix_(outs, ins)]).real / math.prod(math.factorial(int(c)) for c in occ_out)


def test_gram_limits_are_permanents():
    """S = all ones: |per M|^2 / prod t!; S = identity: per(|M|^2) / prod t! (two independent routes)."""
    n, m = 3, 5
    U = so.haar_unitary(m, 21)
    occ = [1, 0, 1, 1, 0]
    final = so.to_occupation(so.enumerate_layers(n, m)[n], m)
    ones = so.gram_probs(U, n, m, occ, np.ones((n, n)))
    assert np.abs(ones - so.reference_probs(U, n, m, occ)).max() < 1e-13
    eye = so.gram_probs(U, n, m, occ, np.eye(n))
    assert np.abs(eye - np.array([_classical(U, occ, t) for t in final.tolist()])).max() < 1e-13
    assert abs(ones.sum() - 1) < 1e-12 and abs(eye.sum() - 1) < 1e-12


@pytest.mark.parametrize("n,m,seed,x", [(3, 5, 1, 0.7), (4, 5, 2, 0.35), (3, 6, 3, 0.0), (3, 6, 4, 1.0)])
def test_species_mixture_equals_gram(n, m, seed, x):
    """The x-model as a mixture over interfering subsets equals the Gram-matrix distribution with
    S = x J + (1 - x) I (the identity the reference's strong path rests on)."""
    U = so.haar_unitary(m, seed)
    occ = [1] * n + [0] * (m - n)
    S = x * np.ones((n, n)) + (1 - x) * np.eye(n)
    mix = so.species_mixture(U, n, m, occ, x)
    assert np.abs(mix - so.gram_probs(U, n, m, occ, S)).max() < 1e-13


def test_species_mixture_limits_bunched_and_truncation():
    n, m = 3, 6
    U = so.haar_unitary(m, 5)
    occ = [2, 0, 1, 0, 0, 0]
    final = so.to_occupation(so.enumerate_layers(n, m)[n], m)
    assert np.abs(so.species_mixture(U, n, m, occ, 1.0) - so.reference_probs(U, n, m, occ)).max() < 1e-13
    cl = np.array([_classical(U, occ, t) for t in final.tolist()])
    assert np.abs(so.species_mixture(U, n, m, occ, 0.0) - cl).max() < 1e-13
    x = 0.6
    for b in range(n + 1):
        dropped = sum(math.comb(n, c) * (1 - x) ** c * x ** (n - c) for c in range(b + 1, n + 1))
        assert abs(so.species_mixture(U, n, m, occ, x, order=b).sum() - (1 - dropped)) < 1e-12


def test_species_mixture_catches_a_planted_mistake():
    """A raise with |U[col, j]|^2 (transposed U) instead of |U[j, col]|^2 breaks the Gram identity."""
    n, m, x = 3, 5, 0.5
    U = so.haar_unitary(m, 9)
    occ = [1, 1, 1, 0, 0]
    S = x * np.ones((n, n)) + (1 - x) * np.eye(n)
    bad = so.species_mixture(np.asarray(U).T, n, m, [1, 1, 1, 0, 0], x)
    assert np.abs(bad - so.gram_probs(U, n, m, occ, S)).max() > 1e-3


# ---- strong-then-sample: the two-level inverse CDF -----------------------------------------------

def test_two_level_inverse_cdf_matches_the_cumsum_draw():
    """The device sampler's operation order draws the same indices as the reference's float64 cumsum +
    searchsorted (imperfections.py:234-237) except within rounding of a CDF step; zero-probability
    entries and whole zero blocks are never drawn; the edge uniforms land on the first / last support."""
    rng = np.random.default_rng(3)
    S = 3 * 4096 * 1024 + 777                      # several prefix segments, ragged tail
    p = rng.random(S).astype(np.float32) ** 4
    p[rng.random(S) < 0.2] = 0
    p[5 * 4096:9 * 4096] = 0                       # empty blocks
    u = np.concatenate([[0.0, 1.0 - 2 ** -53], rng.random(4000)])
    a = so.draw_indices_two_level(p, u)
    b = so.draw_indices_cumsum(p, u)
    cdf = np.cumsum(p.astype(np.float64))
    near = np.abs(cdf[np.minimum(b, S - 1)] - u * cdf[-1]) < 1e-9 * cdf[-1]
    assert np.all((a == b) | near) and (a != b).sum() <= 2
    assert np.all(p[a] > 0)
    assert a[0] == np.nonzero(p)[0][0] and a[1] == np.nonzero(p)[0][-1]


def test_two_level_inverse_cdf_frequencies():
    """Empirical frequencies of 400k draws from a 10-category law match it (chi-square, 9 dof)."""
    rng = np.random.default_rng(4)
    w = np.array([0.0, 3, 1, 0, 7, 2, 2, 5, 0.5, 4.5], dtype=np.float32)
    u = rng.random(400_000)
    idx = so.draw_indices_two_level(w, u)
    f = np.bincount(idx, minlength=10)
    e = w / w.sum() * u.size
    keep = e > 0
    chi2 = (((f - e) ** 2)[keep] / e[keep]).sum()
    assert f[~keep].sum() == 0 and chi2 < 27.9      # p = 0.001 at 9 dof


# ---- coherent superpositions ---------------------------------------------------------------------

def test_superposition_oracle_pins():
    """One component with weight 1 is the single-input distribution; the coherent sum equals the
    recurrence amplitudes (an independent route) combined by the weights; orthogonal Fock inputs keep
    the total mass sum |w_k|^2."""
    n, m = 3, 6
    U = so.haar_unitary(m, 31)
    occs = [[1, 1, 1, 0, 0, 0], [2, 0, 0, 1, 0, 0], [0, 0, 1, 1, 1, 0]]
    w = np.array([0.6, 0.3 - 0.5j, -0.2 + 0.1j])
    w = w / np.linalg.norm(w)
    one = so.superposition_probs(U, n, m, occs[:1], [1.0])
    assert np.abs(one - so.reference_probs(U, n, m, occs[0])).max() < 1e-13
    sup = so.superposition_probs(U, n, m, occs, w)
    amps = [so.OracleSLOS(n, m, o).amplitudes(U) for o in occs]
    norms = so.OracleSLOS(n, m, occs[0]).norms
    rec = np.abs(sum(wk * a for wk, a in zip(w, amps))) ** 2 * norms
    assert np.abs(sup - rec).max() < 1e-13
    assert abs(sup.sum() - 1) < 1e-12


def test_stored_edge_restatement_autograd_is_the_fd_gradient():
    """The torch stored-edge restatement (oracle/stored_edge.py), whose complex128 autograd made the
    n=8, m=24 gradient fixture: its probabilities equal the oracle's at C1, and its autograd
    gradient at C4 equals the complex128 central finite differences of the golden C4 fixture."""
    from oracle.stored_edge import StoredEdgeSLOS
    U1 = golden("c1_u.npy")
    p1 = StoredEdgeSLOS(3, 12, [1, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0], dtype=torch.complex128).probs(torch.as_tensor(U1))
    assert np.abs(p1.numpy() - golden("c1_probs.npy")).max() < 1e-15
    U4 = torch.as_tensor(golden("c4_u.npy")).requires_grad_(True)
    idx = torch.as_tensor(golden("c4_idx.npy"))
    p4 = StoredEdgeSLOS(6, 16, [1] * 6 + [0] * 10, dtype=torch.complex128).probs(U4)
    (-torch.log(p4[idx]).mean()).backward()
    g_fd = golden("c4_grad_fd.npy")
    assert np.abs(U4.grad.numpy() - g_fd).max() < 1e-6 * np.abs(g_fd).max()
