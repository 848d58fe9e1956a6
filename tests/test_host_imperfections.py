"""Host-side pieces of the loss path that need no GPU: thin_to_clicks (imperfections.py:429-438)
against an independent accumulation by click-pattern key, and the binomial pmf's known answers."""
import math

import numpy as np
import torch

from photonic_engines_amd.imperfections import binomial_pmf, thin_to_clicks


def test_thin_to_clicks_against_key_accumulation():
    rng = np.random.default_rng(0)
    occ = rng.integers(0, 3, size=(300, 7))
    p = torch.rand(2, 300, dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    pat, q = thin_to_clicks(occ, p)
    keys = ((occ > 0) * (2 ** np.arange(7))).sum(1)
    uk = np.unique(keys)
    ref = np.zeros((2, uk.size))
    for i, key in enumerate(keys):
        ref[:, np.searchsorted(uk, key)] += p[:, i].numpy()
    assert np.allclose(q.numpy(), ref, rtol=1e-14, atol=0)
    assert np.array_equal((pat * (2 ** np.arange(7))).sum(1), uk)        # the reference's pattern order
    assert set(np.unique(pat)) <= {0, 1}
    assert abs(float(q.sum()) - float(p.sum())) < 1e-12


def test_binomial_pmf_known_answers():                                  # test_loss.py:57-59
    assert abs(float(binomial_pmf(5, 0.3)[5]) - 0.3 ** 5) < 1e-15
    assert abs(float(binomial_pmf(3, 0.8)[2]) - 3 * 0.8 ** 2 * 0.2) < 1e-15
    pmf = binomial_pmf(9, 0.7)
    assert abs(float(pmf.sum()) - 1) < 1e-14
    assert all(abs(float(pmf[k]) - math.comb(9, k) * 0.7 ** k * 0.3 ** (9 - k)) < 1e-15 for k in range(10))
