"""The kept-set comparison the GPU tests use between runs that round
differently (tests/truncation.py), on hand-made states."""
import numpy as np
import pytest

from truncation import same_truncation


def _state(vals_per_bond, chi=6, M=3, seed=0):
    rng = np.random.default_rng(seed)
    lam = np.zeros((M - 1, chi))
    ch = np.full((M + 1, chi, 2), 9, np.int32)
    for i, v in enumerate(vals_per_bond):
        lam[i, :len(v)] = v
        ch[i + 1, :len(v)] = rng.integers(0, 4, (len(v), 2))
    return lam, ch


def test_identical_states_have_no_flips():
    la, ca = _state([[1.0, 0.5, 1e-3], [0.9, 0.1]])
    assert same_truncation(la, ca, la.copy(), ca.copy()) == []


def test_extra_value_at_the_noise_floor_is_a_flip():
    la, ca = _state([[1.0, 0.5, 1e-3], [0.9, 0.1]])
    lb, cb = la.copy(), ca.copy()
    lb[1, 2] = 1e-14  # bond 1 keeps one more value, far below 1e-10 of its largest
    cb[2, 2] = [1, 2]
    assert same_truncation(la, ca, lb, cb) == [1]


def test_value_change_above_tolerance_fails():
    la, ca = _state([[1.0, 0.5, 1e-3], [0.9, 0.1]])
    lb = la.copy()
    lb[0, 1] = 0.5 + 1e-8
    with pytest.raises(AssertionError):
        same_truncation(la, ca, lb, ca)


def test_charges_differing_with_equal_counts_fail():
    la, ca = _state([[1.0, 0.5, 1e-3], [0.9, 0.1]])
    cb = ca.copy()
    cb[1, 0] = cb[1, 0] + 1  # bond 0 (charge row 1) reordered, counts equal
    with pytest.raises(AssertionError):
        same_truncation(la, ca, la, cb)


def test_large_extra_value_fails():
    la, ca = _state([[1.0, 0.5, 1e-3], [0.9, 0.1]])
    lb = la.copy()
    lb[1, 2] = 1e-3  # an extra kept value far above the noise floor
    with pytest.raises(AssertionError):
        same_truncation(la, ca, lb, ca)
