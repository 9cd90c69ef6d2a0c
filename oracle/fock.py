"""Fock-space beam-splitter matrix elements (FP64), three independent forms.

Convention (fixed by how the reference indexes its table, see
src/update_MPO_kernel.cu:491-497 and src/MPO.py:254):

    U3[n1, n2, p] = < n1, n2 | U(B) | p, n1 + n2 - p >

i.e. the first two indices are the photon numbers of (mode, mode+1) AFTER the
beam splitter and the third is the photon number of ``mode`` BEFORE it.  ``B``
is the single-particle transfer matrix, new = B @ old:  <j|U(B)|i> = B[j, i].
For the reference parametrisation (src/cuda_kernels.py:37-43,
src/MPO.py:8-10)  B = [[t, r], [-conj(r), conj(t)]].
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def _fact(n: int) -> float:
    return float(math.factorial(n))


def _comb(n: int, k: int) -> float:
    if k < 0 or k > n:
        return 0.0
    return float(math.comb(n, k))


def bs_table_reference(d: int, t: complex, r: complex) -> np.ndarray:
    """Reference binomial formula, restated in complex128.

    Follows src/MPO.py:11-17 (``bs_coeff`` and the n, m, l, k ranges) and
    src/u_kernel.cu:37-70 (same sum, log-space in the CUDA kernel).  Returns
    U3[n, m, l] (the reference stores it as Unitary[n, m, l, n+m-l]).
    Entries outside l in [max(0, n+m+1-d), min(d, n+m+1)) stay 0.
    """
    ct = np.conj(t)
    cr = np.conj(r)
    U = np.zeros((d, d, d), dtype=np.complex128)
    for n in range(d):
        for m in range(d):
            for l in range(max(0, n + m + 1 - d), min(d, n + m + 1)):
                acc = 0.0 + 0.0j
                for k in range(max(0, l - m), min(l + 1, n + 1, d)):
                    coef = math.sqrt(_fact(l) * _fact(n + m - l) / _fact(n) / _fact(m))
                    coef *= _comb(n, k) * _comb(m, l - k)
                    acc += coef * (t ** k) * (ct ** (m - l + k)) * (r ** (n - k)) * ((-cr) ** (l - k))
                U[n, m, l] = acc
    return U


def reference_bs_matrix(t: complex, r: complex) -> np.ndarray:
    """Single-particle matrix B (new = B @ old) of the reference's (t, r)."""
    return np.array([[t, r], [-np.conj(r), np.conj(t)]], dtype=np.complex128)


def bs_table(d: int, B: np.ndarray) -> np.ndarray:
    """U3 for a general 2x2 matrix B (any U(2) element), by the binomial sum.

    <n1, n2 | U(B) | p, q> = sqrt(n1! n2! / (p! q!)) *
        sum_k C(p, k) C(q, n1-k) B00^k B10^(p-k) B01^(n1-k) B11^(q-n1+k)
    from a0^+ -> B00 a0^+ + B10 a1^+,  a1^+ -> B01 a0^+ + B11 a1^+.
    Entries with n1 + n2 - p outside [0, d) are left 0 (as the reference).
    """
    B = np.asarray(B, dtype=np.complex128)
    b00, b01, b10, b11 = B[0, 0], B[0, 1], B[1, 0], B[1, 1]
    U = np.zeros((d, d, d), dtype=np.complex128)
    for n1 in range(d):
        for n2 in range(d):
            s = n1 + n2
            for p in range(d):
                q = s - p
                if q < 0 or q >= d:
                    continue
                acc = 0.0 + 0.0j
                for k in range(max(0, n1 - q), min(p, n1) + 1):
                    acc += (_comb(p, k) * _comb(q, n1 - k) * b00 ** k * b10 ** (p - k)
                            * b01 ** (n1 - k) * b11 ** (q - n1 + k))
                U[n1, n2, p] = acc * math.sqrt(_fact(n1) * _fact(n2) / (_fact(p) * _fact(q)))
    return U


def permanent(A: np.ndarray) -> complex:
    """Ryser's formula (O(2^n n)); small n only."""
    A = np.asarray(A, dtype=np.complex128)
    n = A.shape[0]
    if n == 0:
        return 1.0 + 0.0j
    total = 0.0 + 0.0j
    for subset in range(1, 1 << n):
        cols = [j for j in range(n) if subset >> j & 1]
        rowsums = A[:, cols].sum(axis=1)
        total += (-1) ** len(cols) * np.prod(rowsums)
    return (-1) ** n * total


def fock_amplitude_permanent(B: np.ndarray, out_occ, in_occ) -> complex:
    """<out|U(B)|in> = Perm(B[rows(out), cols(in)]) / sqrt(prod out! prod in!)."""
    rows = list(itertools.chain.from_iterable([i] * n for i, n in enumerate(out_occ)))
    cols = list(itertools.chain.from_iterable([i] * n for i, n in enumerate(in_occ)))
    if len(rows) != len(cols):
        return 0.0 + 0.0j
    norm = math.sqrt(np.prod([_fact(n) for n in out_occ]) * np.prod([_fact(n) for n in in_occ]))
    return permanent(np.asarray(B)[np.ix_(rows, cols)]) / norm


def bs_table_permanent(d: int, B: np.ndarray) -> np.ndarray:
    """U3 from the permanent formula (pins the two binomial forms)."""
    U = np.zeros((d, d, d), dtype=np.complex128)
    for n1 in range(d):
        for n2 in range(d):
            for p in range(d):
                q = n1 + n2 - p
                if 0 <= q < d:
                    U[n1, n2, p] = fock_amplitude_permanent(B, (n1, n2), (p, q))
    return U


def canonicalisation_bs(d: int) -> np.ndarray:
    """U3 used by the reference canonicalisation sweep: Rand_BS_MPS(0, d, 0).

    src/MPO.py:6-10,151: RandomState(0); t = exp(2 pi i u0) (|t| = 1), r = 0.
    """
    rs = np.random.RandomState(0)
    t = np.exp(1j * rs.rand() * 2 * np.pi)
    r = 0.0 * np.exp(1j * rs.rand() * 2 * np.pi)
    return bs_table_reference(d, t, r)
