"""Synthetic inputs of one site update with chosen charge ranges (test
infrastructure).

The three bonds carry every charge of a square range, with `chi` slots
spread as evenly as the blocks allow, so that sectors tau = (t, t) in the
middle of the ranges have L and R of order chi / 2 or more.  Gamma entries
exist only where the local occupation (left charge minus right charge) is
non-negative and below d, as in an MPO built by the update itself; values are
complex Gaussians.  LR is a positive, decaying profile.
"""
import numpy as np


def square_charges(chi, lo, hi, rng=None):
    """chi slots over the charges lo..hi (both components), sorted."""
    qs = [(a, b) for a in range(lo, hi + 1) for b in range(lo, hi + 1)]
    n = np.full(len(qs), chi // len(qs))
    n[: chi - n.sum()] += 1
    return np.array([q for q, k in zip(qs, n) for _ in range(k)], np.int32)


def synthetic_update(d, chi, left, centre, right, seed):
    """(CL, CC, CR, Glc, Gcr, LR) for charge ranges left, centre, right =
    (lo, hi) each."""
    rng = np.random.default_rng(seed)
    CL = square_charges(chi, *left)
    CC = square_charges(chi, *centre)
    CR = square_charges(chi, *right)

    def gamma(Ca, Cb):
        occ0 = Ca[:, None, 0] - Cb[None, :, 0]
        occ1 = Ca[:, None, 1] - Cb[None, :, 1]
        mask = (occ0 >= 0) & (occ1 >= 0) & (occ0 < d) & (occ1 < d)
        g = rng.standard_normal((chi, chi)) + 1j * rng.standard_normal((chi, chi))
        return np.where(mask, g, 0.0) / np.sqrt(chi)

    Glc = gamma(CL, CC)
    Gcr = gamma(CC, CR)
    LR = np.exp(-np.arange(chi) / (chi / 4.0)) * (0.5 + rng.random(chi))
    return CL, CC, CR, Glc, Gcr, LR
