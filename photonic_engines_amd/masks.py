"""Output computation-space restrictions (masks) and their state-space sizes.

Same names, validity rules and closed-form sizes as photonic_engines/masks.py:18-74:

  FOCK         full basis: C(m + n - 1, n)
  UNBUNCHED    at most 1 photon per mode: C(m, n), valid only when n <= m
  DUAL_RAIL    at most 1 photon per consecutive mode pair; valid for m even, n == m // 2: 2**(m/2)
  SPARSE_1IN4  at most 1 photon per consecutive 4-mode block: C(m/4, n) 4**n
  TAIL_EMPTY   last 2 modes empty: C(m - 2 + n - 1, n)
"""
from __future__ import annotations

from enum import Enum
from math import comb


class Mask(str, Enum):
    """Names of the supported output spaces (a str enum: members equal their value)."""

    FOCK = "fock"
    UNBUNCHED = "unbunched"
    DUAL_RAIL = "dual_rail"
    SPARSE_1IN4 = "sparse_1in4"
    TAIL_EMPTY = "tail_empty"


def is_valid(n: int, m: int, mask: str) -> bool:
    """Whether (n, m) is a valid configuration for the mask (masks.py:31-45)."""
    mask = Mask(mask).value
    if n < 0 or m <= 0:
        return False
    if mask == "fock":
        return True
    if mask == "unbunched":
        return n <= m
    if mask == "dual_rail":
        return m % 2 == 0 and n == m // 2
    if mask == "sparse_1in4":
        return m % 4 == 0 and 1 <= n <= m // 4
    return m > 2 and n <= m - 2


def state_space_size(n: int, m: int, mask: str = Mask.FOCK.value) -> int:
    """Number of output states for (n, m) under `mask`; raises ValueError if invalid (masks.py:60-74)."""
    if not is_valid(n, m, mask):
        raise ValueError(f"invalid config n={n} m={m} for mask {mask!r}")
    mask = Mask(mask).value
    if mask == "fock":
        return comb(m + n - 1, n)
    if mask == "unbunched":
        return comb(m, n)
    if mask == "dual_rail":
        return 2 ** (m // 2)
    if mask == "sparse_1in4":
        return comb(m // 4, n) * 4 ** n
    return comb(m - 2 + n - 1, n)


def graded_state_space_size(n: int, m: int, mask: str = Mask.FOCK.value) -> int:
    """States over all photon numbers 0..n (masks.py:48-57)."""
    for k in range(n + 1):
        if not is_valid(k, m, mask):
            raise ValueError(f"mask {mask!r} has no {k}-photon layer: no graded space for n={n}, m={m}")
    return sum(state_space_size(k, m, mask) for k in range(n + 1))
