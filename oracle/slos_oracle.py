"""numpy restatement of the reference SLOS algorithm (the oracle; test infrastructure only).

Each function restates, in code written for this repository, the reference function cited in its
docstring (paths relative to the reference repository root). Everything here is float64 /
complex128 and deliberately simple; it is meant for sizes that finish in seconds.
"""
from __future__ import annotations

import math
from itertools import combinations, permutations, product

import numpy as np

MASKS = ("fock", "unbunched", "dual_rail", "sparse_1in4", "tail_empty")


# ---------------------------------------------------------------------------------------------
# basis: caps, block constraint, layer enumeration in the reference order
# ---------------------------------------------------------------------------------------------

def per_mode_max(mask: str, n: int, m: int) -> list[int]:
    """Per-mode occupation cap. Restates photonic_engines/slos/basis.py:40-48."""
    if mask == "fock":
        return [n] * m
    if mask in ("unbunched", "dual_rail", "sparse_1in4"):
        return [1] * m
    if mask == "tail_empty":
        return [n] * (m - 2) + [0, 0]
    raise ValueError(f"unknown mask {mask!r}")


def block_shift(mask: str) -> int:
    """Block width exponent (at most one photon per 2**shift modes). Restates basis.py:51-57."""
    return {"dual_rail": 1, "sparse_1in4": 2}.get(mask, 0)


def enumerate_layers(n: int, m: int, mask: str = "fock") -> list[np.ndarray]:
    """Layers 0..n as int16 arrays [S_k, k] of sorted photon positions, in the reference order.

    Restates FockLayers._children (photonic_engines/slos/basis.py:120-148): the children of
    layer k are every parent (in order) followed by every admissible mode a >= last(parent)
    (in increasing a); a == last is admissible while the trailing run is below the mode cap;
    with blocks, a must not share its block with last(parent).
    """
    cap = np.asarray(per_mode_max(mask, n, m), dtype=np.int64)
    shift = block_shift(mask)
    modes = np.arange(m, dtype=np.int64)
    layers = [np.zeros((1, 0), dtype=np.int16)]
    run = np.zeros(1, dtype=np.int64)             # length of the trailing run of equal modes
    for k in range(n):
        par = layers[-1]
        S = par.shape[0]
        if k == 0:
            adm = (cap >= 1)[None, :].copy()
        else:
            last = par[:, -1].astype(np.int64)
            adm = (modes[None, :] > last[:, None]) & (cap >= 1)[None, :]
            adm[np.arange(S), last] = run < cap[last]
            if shift:
                adm &= (modes[None, :] >> shift) != (last[:, None] >> shift)
        rows, a = np.nonzero(adm)
        child = np.empty((rows.shape[0], k + 1), dtype=np.int16)
        child[:, :k] = par[rows]
        child[:, k] = a
        if k == 0:
            run = np.ones(rows.shape[0], dtype=np.int64)
        else:
            same = a == par[rows, -1]
            run = np.where(same, run[rows] + 1, 1)
        layers.append(child)
    return layers


def enumerate_layers_recursive(n: int, m: int, mask: str = "fock") -> list[list[tuple]]:
    """Independent recursive builder (depth-first over sorted tuples, pure Python) used to pin
    `enumerate_layers`. Same ordering rule as the reference's own order oracle
    (tests/slos/test_basis_vectorized.py:11-36): slots filled left to right, modes ascending."""
    cap = per_mode_max(mask, n, m)
    shift = block_shift(mask)
    out = []
    for k in range(n + 1):
        acc, cur = [], []

        def rec(slot, start):
            if slot == k:
                acc.append(tuple(cur))
                return
            for a in range(start, m):
                if cur and shift and (a >> shift) == (cur[-1] >> shift):
                    continue
                if cur.count(a) + 1 > cap[a]:
                    continue
                cur.append(a)
                rec(slot + 1, a)
                cur.pop()
        rec(0, 0)
        out.append(acc)
    return out


def mask_ok(occ, mask: str, m: int) -> bool:
    """Occupation-vector admissibility. Restates tests/slos/test_basis.py:15-26."""
    if mask == "fock":
        return True
    if mask == "unbunched":
        return all(x <= 1 for x in occ)
    if mask == "dual_rail":
        return all(occ[2 * i] + occ[2 * i + 1] <= 1 for i in range(m // 2))
    if mask == "sparse_1in4":
        return all(sum(occ[b:b + 4]) <= 1 for b in range(0, m, 4))
    if mask == "tail_empty":
        return occ[-1] == 0 and occ[-2] == 0
    raise ValueError(mask)


def brute_states(k: int, m: int, mask: str) -> set:
    """All admissible occupation vectors with k photons (brute force over the box)."""
    return {occ for occ in product(range(k + 1), repeat=m) if sum(occ) == k and mask_ok(occ, mask, m)}


def state_space_size(n: int, m: int, mask: str = "fock") -> int:
    """Closed-form sizes. Restates photonic_engines/masks.py:60-74 (validity: masks.py:31-45)."""
    if mask == "fock":
        return math.comb(m + n - 1, n)
    if mask == "unbunched":
        return math.comb(m, n)
    if mask == "dual_rail":
        return 2 ** (m // 2)
    if mask == "sparse_1in4":
        return math.comb(m // 4, n) * 4 ** n
    if mask == "tail_empty":
        return math.comb(m - 2 + n - 1, n)
    raise ValueError(mask)


def to_occupation(states: np.ndarray, m: int) -> np.ndarray:
    """[S, k] positions -> [S, m] occupations. Restates basis.py:244-252."""
    S, k = states.shape
    occ = np.zeros((S, m), dtype=np.int64)
    ar = np.arange(S)
    for j in range(k):
        np.add.at(occ, (ar, states[:, j].astype(np.int64)), 1)
    return occ


def prob_norms(states: np.ndarray) -> np.ndarray:
    """prod_j t_j! per state, from run lengths. Restates basis.py:260-272."""
    S, k = states.shape
    out = np.ones(S, dtype=np.int64)
    pos = np.ones(S, dtype=np.int64)
    for j in range(1, k):
        pos = np.where(states[:, j] == states[:, j - 1], pos + 1, 1)
        out *= pos
    return out


def edges(parents: np.ndarray, children: np.ndarray):
    """Raising edges parent -> child: (parent_pos, child_pos, mode, factor) with one edge per
    (child, distinct mode), parent = child minus one photon of that mode, factor =
    sqrt(count of the mode in the child). Restates FockLayers._edges (basis.py:160-193) with a
    dict lookup instead of packed-key searchsorted."""
    index = {tuple(r): i for i, r in enumerate(parents.tolist())}
    pp, cp, md, fc = [], [], [], []
    for c, row in enumerate(children.tolist()):
        L = len(row)
        i = 0
        while i < L:
            a = row[i]
            j = i + 1
            while j < L and row[j] == a:
                j += 1
            pp.append(index[tuple(row[:i] + row[i + 1:])])
            cp.append(c)
            md.append(a)
            fc.append(math.sqrt(j - i))
            i = j
    return (np.asarray(pp, dtype=np.int64), np.asarray(cp, dtype=np.int64),
            np.asarray(md, dtype=np.int64), np.asarray(fc, dtype=np.float64))


def brute_edges(layer_occ_k, layer_occ_k1):
    """Edges from the creation-operator rule on occupation vectors (tests/slos/test_basis.py:52-70)."""
    child_idx = {o: i for i, o in enumerate(layer_occ_k1)}
    out = set()
    m = len(layer_occ_k[0]) if layer_occ_k else 0
    for p_pos, p in enumerate(layer_occ_k):
        for j in range(m):
            child = (*p[:j], p[j] + 1, *p[j + 1:])
            if child in child_idx:
                out.add((p_pos, child_idx[child], j, round(math.sqrt(p[j] + 1), 9)))
    return out


# ---------------------------------------------------------------------------------------------
# the stored-edge recurrence and readout
# ---------------------------------------------------------------------------------------------

def input_modes(occ) -> list[int]:
    """Photon k uses column i_k, modes expanded in ascending order. Restates kernel.py:124."""
    return [j for j, c in enumerate(occ) for _ in range(int(c))]


class OracleSLOS:
    """complex128 restatement of FockSLOSKernel (photonic_engines/slos/kernel.py:83-283), edge
    layout: per layer one gather of parents, one multiply by U[:, mode, i_k], one scatter-add
    into the children (kernel.py:231-257), deferred normalisation (kernel.py:127-141)."""

    def __init__(self, n: int, m: int, occ, mask: str = "fock", normalization: str = "deferred"):
        self.n, self.m, self.mask = n, m, mask
        self.normalization = normalization
        self.layers = enumerate_layers(n, m, mask)
        if self.layers[n].shape[0] == 0:
            raise ValueError("empty output space")
        self.modes = input_modes(occ)
        if len(self.modes) != n:
            raise ValueError("photon-count mismatch")
        self.input_norm = 1.0 / math.sqrt(math.prod(math.factorial(int(c)) for c in occ))
        self.edges = [edges(self.layers[k], self.layers[k + 1]) for k in range(n)]
        self.norms = prob_norms(self.layers[n]).astype(np.float64)

    def output_size(self) -> int:
        return self.layers[self.n].shape[0]

    def amplitudes(self, U, modes=None) -> np.ndarray:
        """Raw amplitudes (kernel.py:201-260): U [m,m] or [B,m,m] -> [S] or [B,S]."""
        U = np.asarray(U, dtype=np.complex128)
        batched = U.ndim == 3
        if not batched:
            U = U[None]
        B = U.shape[0]
        modes = self.modes if modes is None else list(modes)
        psi = np.ones((B, 1), dtype=np.complex128)
        for k, i in enumerate(modes):
            par, chi, mode, fac = self.edges[k]
            S1 = self.layers[k + 1].shape[0]
            nxt = np.zeros((B, S1), dtype=np.complex128)
            for b in range(B):
                contrib = U[b, mode, i] * psi[b, par]
                if self.normalization == "immediate":
                    contrib = contrib * fac
                nxt[b] = (np.bincount(chi, weights=contrib.real, minlength=S1)
                          + 1j * np.bincount(chi, weights=contrib.imag, minlength=S1))
            psi = nxt
        psi = psi * self.input_norm
        return psi if batched else psi[0]

    def probs(self, U) -> np.ndarray:
        """|amp|^2 * prod t_j! (kernel.py:262-283, lossless, no indist)."""
        a = self.amplitudes(U)
        p = a.real ** 2 + a.imag ** 2
        if self.normalization != "immediate":
            p = p * self.norms
        return p


# ---------------------------------------------------------------------------------------------
# permanents
# ---------------------------------------------------------------------------------------------

def permanent_brute(A) -> complex:
    """Permutation sum (tests/slos/test_reference.py:8-20 uses it to pin Ryser)."""
    A = np.asarray(A, dtype=np.complex128)
    n = A.shape[0]
    if n == 0:
        return 1.0 + 0j
    return complex(sum(np.prod([A[i, p[i]] for i in range(n)]) for p in permutations(range(n))))


def permanent_ryser(A) -> complex:
    """Ryser's formula, per(A) = sum_S (-1)^(n-|S|) prod_i sum_{j in S} a_ij, with all subset
    row sums built incrementally (subset = subset-without-lowest-bit + one column).
    Restates photonic_engines/slos/reference.py:24-36."""
    A = np.asarray(A, dtype=np.complex128)
    n = A.shape[0]
    if n == 0:
        return 1.0 + 0j
    N = 1 << n
    sums = np.zeros((N, n), dtype=np.complex128)
    pop = np.zeros(N, dtype=np.int64)
    for s in range(1, N):
        low = (s & -s).bit_length() - 1
        sums[s] = sums[s & (s - 1)] + A[:, low]
        pop[s] = pop[s & (s - 1)] + 1
    sign = np.where((n - pop) % 2 == 0, 1.0, -1.0)
    return complex((sign[1:] * np.prod(sums[1:], axis=1)).sum())


def reference_probs(U, n: int, m: int, occ, mask: str = "fock") -> np.ndarray:
    """Permanent-based probabilities in the kernel's order. Restates reference.py:88-105."""
    U = np.asarray(U, dtype=np.complex128)
    ins = input_modes(occ)
    s_norm = math.prod(math.factorial(int(c)) for c in occ)
    final = to_occupation(enumerate_layers(n, m, mask)[n], m)
    out = np.zeros(final.shape[0])
    for idx, t in enumerate(final.tolist()):
        outs = input_modes(t)
        if not outs:
            out[idx] = 1.0
            continue
        t_norm = math.prod(math.factorial(int(c)) for c in t)
        amp = permanent_ryser(U[np.ix_(outs, ins)]) / math.sqrt(s_norm * t_norm)
        out[idx] = abs(amp) ** 2
    return out


def prob_of_occupation(U, occ_in, occ_out) -> float:
    """|per(U[out rows][:, in cols])|^2 / (prod s! prod t!) for one output pattern."""
    U = np.asarray(U, dtype=np.complex128)
    ins, outs = input_modes(occ_in), input_modes(occ_out)
    s_norm = math.prod(math.factorial(int(c)) for c in occ_in)
    t_norm = math.prod(math.factorial(int(c)) for c in occ_out)
    return abs(permanent_ryser(U[np.ix_(outs, ins)])) ** 2 / (s_norm * t_norm)


# ---------------------------------------------------------------------------------------------
# uniform loss
# ---------------------------------------------------------------------------------------------

def gram_probs(U, n: int, m: int, occ, S, mask: str = "fock") -> np.ndarray:
    """Partial distinguishability from the internal-state Gram matrix S (n x n, S[a, b] = <phi_a|phi_b>),
    in the kernel's order: P(t) = (1 / prod t!) sum_{sigma, rho} prod_k M[k, sigma_k] conj(M[k, rho_k])
    S[rho_k, sigma_k], M = U[output rows][:, input columns] (distinct input columns). S = all ones
    gives |per M|^2 / prod t!, S = identity gives per(|M|^2) / prod t!. This is the closed form the
    reference's tests check the x-model against (tests/test_imperfections.py:159-171); brute force
    over permutation pairs, small n only."""
    from itertools import permutations
    U = np.asarray(U, dtype=np.complex128)
    S = np.asarray(S, dtype=np.complex128)
    ins = input_modes(occ)
    assert len(set(ins)) == len(ins), "gram_probs: distinct input columns"
    final = to_occupation(enumerate_layers(n, m, mask)[n], m)
    perms = list(permutations(range(n)))
    out = np.zeros(final.shape[0])
    for idx, t in enumerate(final.tolist()):
        rows = input_modes(t)
        M = U[np.ix_(rows, ins)]
        acc = 0j
        for sg in perms:
            a = np.prod([M[k, sg[k]] for k in range(n)])
            for rh in perms:
                acc += a * np.prod([np.conj(M[k, rh[k]]) * S[rh[k], sg[k]] for k in range(n)])
        out[idx] = acc.real / math.prod(math.factorial(int(c)) for c in t)
    return out


def species_mixture(U, n: int, m: int, occ, x: float, mask: str = "fock", order=None) -> np.ndarray:
    """The x-model as a mixture over interfering subsets, in the kernel's order. Restates
    photonic_engines/slos/species.py:104-146 with permanents: the interfering photons G give
    |per(U[t rows][:, G cols])|^2 / (prod s! prod t!) on the g-photon patterns of layer g; each other
    photon is raised classically, q'(t + e_j) += q(t) |U[j, col]|^2; patterns outside the masked
    layers drop their mass; weights x^|G| (1 - x)^(n - |G|); `order` keeps at most `order`
    distinguishable photons. Dict-based, small n only."""
    from itertools import combinations
    U = np.asarray(U, dtype=np.complex128)
    cols = input_modes(occ)
    layers = [set(map(tuple, to_occupation(L, m).tolist())) for L in enumerate_layers(n, m, mask)]
    b_max = n if order is None else max(0, min(int(order), n))
    total: dict = {}
    for b in range(b_max + 1):
        w = x ** (n - b) * (1 - x) ** b
        for bad in combinations(range(n), b):
            good = [cols[j] for j in range(n) if j not in bad]
            g = len(good)
            if g >= 2:
                s_norm = 1
                for c in set(good):
                    s_norm *= math.factorial(good.count(c))
                dist = {}
                for t in layers[g]:
                    rows = input_modes(t)
                    amp = permanent_ryser(U[np.ix_(rows, good)])
                    dist[t] = abs(amp) ** 2 / (s_norm * math.prod(math.factorial(int(c)) for c in t))
                k, singles = g, [cols[j] for j in bad]
            else:
                dist, k = {tuple([0] * m): 1.0}, 0
                singles = [cols[j] for j in bad] + good
            for c in singles:
                new: dict = {}
                for t, p in dist.items():
                    for j in range(m):
                        t2 = list(t)
                        t2[j] += 1
                        t2 = tuple(t2)
                        if t2 in layers[k + 1]:
                            new[t2] = new.get(t2, 0.0) + p * abs(U[j, c]) ** 2
                dist, k = new, k + 1
            for t, p in dist.items():
                total[t] = total.get(t, 0.0) + w * p
    final = to_occupation(enumerate_layers(n, m, mask)[n], m)
    return np.array([total.get(tuple(t), 0.0) for t in final.tolist()])


def superposition_probs(U, n: int, m: int, occs, weights, mask: str = "fock") -> np.ndarray:
    """Coherent superposition of K Fock inputs (photonic_engines/slos/kernel.py:296-431), in the
    kernel's order: |sum_k w_k per(U[t rows][:, cols_k]) / sqrt(prod s_k! prod t!)|^2, with the
    permanents of reference.py:24-36."""
    U = np.asarray(U, dtype=np.complex128)
    final = to_occupation(enumerate_layers(n, m, mask)[n], m)
    out = np.zeros(final.shape[0])
    for idx, t in enumerate(final.tolist()):
        rows = input_modes(t)
        t_norm = math.prod(math.factorial(int(c)) for c in t)
        amp = 0j
        for occ, w in zip(occs, weights):
            s_norm = math.prod(math.factorial(int(c)) for c in occ)
            amp += w * permanent_ryser(U[np.ix_(rows, input_modes(occ))]) / math.sqrt(s_norm * t_norm)
        out[idx] = abs(amp) ** 2
    return out


def draw_indices_cumsum(p, uniforms) -> np.ndarray:
    """The inverse-CDF branch of draw_indices (photonic_engines/imperfections.py:234-237): a float64
    cumulative sum, u = uniform * total, the first index whose CDF reaches u, clamped to S - 1."""
    cdf = np.cumsum(np.asarray(p, dtype=np.float64))
    u = np.asarray(uniforms, dtype=np.float64) * cdf[-1]
    return np.minimum(np.searchsorted(cdf, u, side="left"), cdf.size - 1)


def draw_indices_two_level(p, uniforms, chunk: int = 64, chunks: int = 64, seg: int = 1024) -> np.ndarray:
    """The device sampler's inverse CDF with the same FP64 operations in the same order: chunk sums
    (in index order), block sums (in chunk order), block CDF (in-order within segments of `seg`
    blocks, plus the in-order exclusive prefix of the segment totals), then per draw the block, the
    chunk and the entry whose in-order prefix first reaches the remainder. Indices equal the plain
    cumulative-sum ones except where u lies within rounding of a CDF step."""
    p = np.asarray(p, dtype=np.float32).astype(np.float64)
    S = p.size
    blk = chunk * chunks
    nb = -(-S // blk)
    q = np.zeros(nb * blk)
    q[:S] = p
    cum_e = np.cumsum(q.reshape(nb, chunks, chunk), axis=2)          # in-order within a chunk
    csum = cum_e[:, :, -1]
    cum_c = np.cumsum(csum, axis=1)                                  # in-order over the chunks
    bsum = cum_c[:, -1]
    nseg = -(-nb // seg)
    bpad = np.zeros(nseg * seg)
    bpad[:nb] = bsum
    local = np.cumsum(bpad.reshape(nseg, seg), axis=1)
    tot = local[:, -1]
    off = np.concatenate([[0.0], np.cumsum(tot)[:-1]])               # run starts at 0.0, in order
    cdf = (off[:, None] + local).reshape(-1)[:nb]
    u = np.asarray(uniforms, dtype=np.float64) * cdf[-1]
    b = np.minimum(np.searchsorted(cdf, u, side="left"), nb - 1)
    r = u - np.where(b > 0, cdf[np.maximum(b - 1, 0)], 0.0)
    out = np.empty(u.size, dtype=np.int64)
    for s in range(u.size):
        cc = cum_c[b[s]]
        hit = np.nonzero(cc >= r[s])[0]
        c = int(hit[0]) if hit.size else chunks - 1
        r2 = r[s] - (cc[c - 1] if c > 0 else 0.0)
        ce = cum_e[b[s], c]
        hit = np.nonzero(ce >= r2)[0]
        k = int(hit[0]) if hit.size else chunk - 1
        out[s] = min(int(b[s]) * blk + c * chunk + k, S - 1)
    return out


def binomial_pmf(n: int, eta: float) -> np.ndarray:
    """P(k of n survive), k = 0..n, through lgamma. Restates imperfections.py:106-111."""
    k = np.arange(n + 1, dtype=np.float64)
    lg = np.vectorize(math.lgamma)
    logc = math.lgamma(n + 1.0) - lg(k + 1) - lg(n - k + 1)
    return np.exp(logc) * eta ** k * (1 - eta) ** (n - k)


def descend(layers, q_k: np.ndarray, k: int) -> np.ndarray:
    """q_{k-1}(t) = sum_j q_k(t + e_j) (t_j + 1) / k over the k-1 -> k edges (edge weight =
    occupation of the raised mode in the child). Restates imperfections.py:254-276."""
    par, chi, mode, fac = edges(layers[k - 1], layers[k])
    w = (fac ** 2) / k
    q_k = np.atleast_2d(q_k)
    out = np.zeros((q_k.shape[0], layers[k - 1].shape[0]))
    for b in range(q_k.shape[0]):
        out[b] = np.bincount(par, weights=q_k[b, chi] * w, minlength=layers[k - 1].shape[0])
    return out


def thin_distribution(layers, probs: np.ndarray, eta) -> list[np.ndarray]:
    """Uniform output loss: blocks[k] = C(n,k) eta^k (1-eta)^(n-k) q_k, q_n = probs, q_{k-1} =
    descend(q_k). eta None or >= 1 keeps only the n-photon block. Restates
    imperfections.py:309-337 (and _eta :94-103)."""
    n = len(layers) - 1
    probs = np.atleast_2d(np.asarray(probs, dtype=np.float64))
    q = [None] * (n + 1)
    q[n] = probs
    for k in range(n, 0, -1):
        q[k - 1] = descend(layers, q[k], k)
    if eta is None or float(eta) >= 1.0:
        w = np.zeros(n + 1)
        w[n] = 1.0
    else:
        if not 0.0 <= float(eta) <= 1.0:
            raise ValueError("eta outside [0, 1]")
        w = binomial_pmf(n, float(eta))
    return [q[k] * w[k] for k in range(n + 1)]


def thin_patterns_dict(occ_final: np.ndarray, probs: np.ndarray, eta: float) -> dict:
    """Independent route (restates imperfections.py:340-409 for one photon number): delete one
    photon at a time from explicit patterns, weight t_j / k, merge equal patterns in a dict."""
    n = int(occ_final[0].sum())
    layer = {tuple(int(x) for x in o): float(p) for o, p in zip(occ_final, probs)}
    per_k = {n: dict(layer)}
    for k in range(n, 0, -1):
        nxt = {}
        for occ, p in per_k[k].items():
            for j, c in enumerate(occ):
                if c:
                    t = list(occ)
                    t[j] -= 1
                    t = tuple(t)
                    nxt[t] = nxt.get(t, 0.0) + p * c / k
        per_k[k - 1] = nxt
    w = binomial_pmf(n, eta)
    out = {}
    for k, d in per_k.items():
        for t, p in d.items():
            out[t] = out.get(t, 0.0) + p * w[k]
    return out


# ---------------------------------------------------------------------------------------------
# inputs
# ---------------------------------------------------------------------------------------------

def haar_unitary(m: int, seed: int):
    """Haar-random complex128 U: QR of a seeded float64 Gaussian pair with the d/|d| phase fix.
    Restates tests/helpers.py:10-17 (torch generator, so the same seed gives the same U)."""
    import torch
    g = torch.Generator().manual_seed(seed)
    z = torch.randn(m, m, generator=g, dtype=torch.float64) + 1j * torch.randn(
        m, m, generator=g, dtype=torch.float64)
    q, r = torch.linalg.qr(z)
    d = torch.diagonal(r)
    return (q * (d / d.abs()).unsqueeze(0)).numpy()


def occupation_list(n: int, m: int, mask: str = "fock") -> list[int]:
    """Input placement of the reference tests (tests/helpers.py:26-33)."""
    if mask == "dual_rail":
        return [1, 0] * (m // 2)
    if m >= 2 * n:
        return [1, 0] * n + [0] * (m - 2 * n)
    return [1] * n + [0] * (m - n)


def itertools_order(n: int, m: int, mask: str):
    """For fock / unbunched the reference order equals itertools' (combinations_with_replacement
    / combinations) order; an independent order check."""
    from itertools import combinations_with_replacement
    if mask == "fock":
        return list(combinations_with_replacement(range(m), n))
    if mask == "unbunched":
        return list(combinations(range(m), n))
    raise ValueError(mask)
