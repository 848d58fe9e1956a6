"""Shared test helpers: seeded inputs, golden fixtures, tolerances."""
from __future__ import annotations

import json
import math
import os

import numpy as np
import torch

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def haar_unitary(m: int, seed: int) -> torch.Tensor:
    """Haar-random complex128 U, the generator of the reference tests (tests/helpers.py:10-17)."""
    g = torch.Generator().manual_seed(seed)
    z = torch.randn(m, m, generator=g, dtype=torch.float64) + 1j * torch.randn(m, m, generator=g, dtype=torch.float64)
    q, r = torch.linalg.qr(z)
    d = torch.diagonal(r)
    return q * (d / d.abs()).unsqueeze(0)


def occupation_list(n: int, m: int, mask: str = "fock") -> list[int]:
    """Input placement of the reference tests (tests/helpers.py:26-33)."""
    if mask == "dual_rail":
        return [1, 0] * (m // 2)
    if m >= 2 * n:
        return [1, 0] * n + [0] * (m - 2 * n)
    return [1] * n + [0] * (m - n)


def first_modes(n: int, m: int) -> list[int]:
    return [1] * n + [0] * (m - n)


def golden(name: str):
    if name.endswith(".json"):
        with open(os.path.join(GOLDEN, name)) as fh:
            return json.load(fh)
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def input_norm(occ) -> float:
    return 1.0 / math.sqrt(math.prod(math.factorial(int(c)) for c in occ))


def gpu_ok() -> bool:
    return torch.cuda.is_available()
