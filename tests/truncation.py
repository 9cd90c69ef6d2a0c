"""Comparison of two truncated states whose tridiagonal-route eigensolves
rounded differently (shared by the GPU tests)."""
import numpy as np


def same_truncation(la, ca, lb, cb, rel=1e-10, max_flips=4):
    """Lambda [M-1, chi] and charges [M+1, chi, 2] of two runs of the
    tridiagonal route whose Householder row sums were added in different
    orders (LDS atomics, block shapes): every bond's singular
    values, sorted, agree to `rel` of the bond's largest; where both runs keep
    the same number of values the bond's charges are equal; elsewhere the
    counts differ by at most `max_flips`: values that sit at a sector's drop
    threshold (lambda <= drop_rel x the sector's largest), where the Gram
    route's absolute rounding (~eps x the sector's largest eigenvalue) decides
    whether they are kept.  Returns the bonds whose kept counts differ."""
    flips = []
    for i in range(la.shape[0]):
        sa, sb = np.sort(la[i])[::-1], np.sort(lb[i])[::-1]
        assert np.abs(sa - sb).max() <= rel * max(sa.max(), 1e-300), i
        ka, kb = int((sa > 0).sum()), int((sb > 0).sum())
        if ka == kb:
            assert np.array_equal(ca[i + 1], cb[i + 1]), i
        else:
            assert abs(ka - kb) <= max_flips, (i, ka, kb)
            flips.append(i)
    assert np.array_equal(ca[0], cb[0]) and np.array_equal(ca[-1], cb[-1])
    return flips
