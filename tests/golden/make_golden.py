"""Writes the golden fixtures of tests/golden/ from the oracle (oracle/), never from the reference.

    python tests/golden/make_golden.py

Workloads (seeds follow seed = 1000 n + m; every U is the phase-fixed Haar QR of a seeded float64
Gaussian pair, the generator of the reference's tests/helpers.py:10-17):
  C1  n=3,  m=12, input [1,0,1,0,1,0,0,0,0,0,0,0], fock       -> all 364 probabilities (complex128)
  C2  n=8,  m=24, input [1]*8 + [0]*16                         -> 10,000 seeded outputs, Ryser complex128
  C3  512 instances (4,12) / (6,16) / (8,20) by i mod 3, photons in the first n modes,
      seed 500000 + i                                          -> sha256 of the complex64 batch
  C4  n=6,  m=16, first n modes                                -> 1,024 indices drawn from the exact
      distribution (numpy PCG64 seed 4), and dL/dU of L = -mean log p[idx] by complex128 central
      finite differences (h = 1e-6) on Re and Im of every entry
  C5  n=11, m=32, input [1]*11 + [0]*21                        -> 10,000 seeded outputs, Ryser complex128
  order  sha256 of the [S, n] int16 position arrays of the reference-order enumeration
  G8  n=8, m=24, first n modes, U of seed 8024                 -> 1,024 indices drawn from the exact
      distribution (numpy PCG64 seed 8), and dL/dU of L = -mean log p[idx] by complex128 torch
      autograd of the stored-edge restatement (oracle/stored_edge.py)

    python tests/golden/make_golden.py g8      (only G8)
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pull  # noqa: E402
from oracle.slos_oracle import OracleSLOS, enumerate_layers, haar_unitary, reference_probs  # noqa: E402

C1_OCC = [1, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0]
ORDER_CASES = [(3, 12, "fock"), (4, 12, "fock"), (6, 16, "fock"), (4, 8, "unbunched"), (4, 8, "dual_rail"),
               (3, 12, "sparse_1in4"), (4, 9, "tail_empty"), (8, 12, "fock")]


def seed_of(n, m):
    return 1000 * n + m


def first_modes(n, m):
    return [1] * n + [0] * (m - n)


def c3_batch() -> np.ndarray:
    """The 512 complex64 unitaries of C3, concatenated in instance order (variable m)."""
    shapes = [(4, 12), (6, 16), (8, 20)]
    parts = []
    for i in range(512):
        n, m = shapes[i % 3]
        parts.append(haar_unitary(m, 500000 + i).astype(np.complex64).ravel())
    return np.concatenate(parts)


def loss_c4(U, idx):
    p = pull.forward_probs(6, 16, first_modes(6, 16), U, prec=128)
    return -np.mean(np.log(p[idx]))


def grad_n8m24() -> float:
    """G8: the indices, the gradient and the loss value (written to g8_*.npy; returns the loss)."""
    import torch
    from oracle.stored_edge import StoredEdgeSLOS
    n, m = 8, 24
    U = haar_unitary(m, 8024)
    ref = StoredEdgeSLOS(n, m, first_modes(n, m), dtype=torch.complex128)
    Ut = torch.as_tensor(U).requires_grad_(True)
    p = ref.probs(Ut)
    pn = p.detach().numpy()
    idx = np.random.default_rng(8).choice(pn.shape[0], size=1024, replace=True, p=pn / pn.sum()).astype(np.int64)
    loss = -torch.log(p[torch.as_tensor(idx)]).mean()
    loss.backward()
    np.save(os.path.join(HERE, "g8_u.npy"), U)
    np.save(os.path.join(HERE, "g8_idx.npy"), idx)
    np.save(os.path.join(HERE, "g8_grad.npy"), Ut.grad.numpy())
    return float(loss)


def main():
    out = {}
    # C1
    U1 = haar_unitary(12, seed_of(3, 12))
    p1 = OracleSLOS(3, 12, C1_OCC).probs(U1)
    r1 = reference_probs(U1, 3, 12, C1_OCC)
    assert np.abs(p1 - r1).max() < 1e-15
    np.save(os.path.join(HERE, "c1_u.npy"), U1)
    np.save(os.path.join(HERE, "c1_probs.npy"), p1)
    # C2, C5 spot checks
    for tag, n, m in (("c2", 8, 24), ("c5", 11, 32)):
        U = haar_unitary(m, seed_of(n, m))
        S = math.comb(m + n - 1, n)
        idx = np.sort(np.random.default_rng(seed_of(n, m)).choice(S, size=10_000, replace=False)).astype(np.uint64)
        ps = pull.spot_probs(n, m, first_modes(n, m), U, idx)
        np.save(os.path.join(HERE, f"{tag}_u.npy"), U)
        np.save(os.path.join(HERE, f"{tag}_spot_idx.npy"), idx)
        np.save(os.path.join(HERE, f"{tag}_spot_probs.npy"), ps)
    # C3
    out["c3_sha256"] = hashlib.sha256(c3_batch().tobytes()).hexdigest()
    # C4
    U4 = haar_unitary(16, seed_of(6, 16))
    p4 = pull.forward_probs(6, 16, first_modes(6, 16), U4, prec=128)
    idx4 = np.random.default_rng(4).choice(p4.shape[0], size=1024, replace=True, p=p4 / p4.sum()).astype(np.int64)
    h = 1e-6
    g = np.zeros((16, 16), dtype=np.complex128)
    for j in range(16):
        for i in range(16):
            for part, unit in ((0, 1.0), (1, 1j)):
                Up, Um = U4.copy(), U4.copy()
                Up[j, i] += unit * h
                Um[j, i] -= unit * h
                d = (loss_c4(Up, idx4) - loss_c4(Um, idx4)) / (2 * h)
                g[j, i] += d * (1.0 if part == 0 else 1j)
    np.save(os.path.join(HERE, "c4_u.npy"), U4)
    np.save(os.path.join(HERE, "c4_idx.npy"), idx4)
    np.save(os.path.join(HERE, "c4_grad_fd.npy"), g)
    out["c4_loss"] = float(loss_c4(U4, idx4))
    # order hashes
    order = {}
    for n, m, mask in ORDER_CASES:
        L = enumerate_layers(n, m, mask)[n].astype(np.int16)
        order[f"{n},{m},{mask}"] = {"size": int(L.shape[0]), "sha256": hashlib.sha256(L.tobytes()).hexdigest()}
    out["order"] = order
    out["g8_loss"] = grad_n8m24()
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    if sys.argv[1:] == ["g8"]:
        with open(os.path.join(HERE, "golden.json")) as fh:
            rec = json.load(fh)
        rec["g8_loss"] = grad_n8m24()
        with open(os.path.join(HERE, "golden.json"), "w") as fh:
            json.dump(rec, fh, indent=1, sort_keys=True)
    else:
        main()
