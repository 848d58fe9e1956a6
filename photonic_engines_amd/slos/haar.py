"""Seeded Haar-random unitaries: QR of a float64 complex Gaussian with the d/|d| phase fix, drawn
from a torch.Generator so that a seed gives the same U as the reference's test helper
(tests/helpers.py:10-17 and bench/bench_product_slos.py:33-37)."""
from __future__ import annotations

import torch


def haar_unitary(m: int, seed: int, dtype: torch.dtype = torch.complex128, device="cpu") -> torch.Tensor:
    """Haar-random m x m unitary, deterministic in `seed` (computed in complex128 on the host)."""
    g = torch.Generator().manual_seed(int(seed))
    z = torch.randn(m, m, generator=g, dtype=torch.float64) + 1j * torch.randn(m, m, generator=g, dtype=torch.float64)
    q, r = torch.linalg.qr(z)
    d = torch.diagonal(r)
    return (q * (d / d.abs()).unsqueeze(0)).to(device=device, dtype=dtype)
