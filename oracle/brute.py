"""Dense Fock-space density-matrix simulator (no truncation), M <= 5 modes.

Pins the whole MPO chain (init + updates + observables) on small cases: with
chi large enough nothing is truncated and the MPO must reproduce this exactly.
It shares only the BS table convention U3[n1, n2, p] = <n1, n2|U|p, q>.
"""
from __future__ import annotations

import numpy as np

from .fock import bs_table
from .mpo import input_density, layer_modes


def two_mode_matrix(U3: np.ndarray) -> np.ndarray:
    d = U3.shape[0]
    G = np.zeros((d, d, d, d), np.complex128)   # [n1, n2, p, q]
    for n1 in range(d):
        for n2 in range(d):
            for p in range(d):
                q = n1 + n2 - p
                if 0 <= q < d:
                    G[n1, n2, p, q] = U3[n1, n2, p]
    return G.reshape(d * d, d * d)


def initial_rho(M, d, input_modes, eta):
    rho = np.ones((1, 1), np.complex128)
    vac = np.zeros((d, d))
    vac[0, 0] = 1
    for j in range(M):
        r = input_density(d, eta) if j in set(input_modes) else vac
        rho = np.kron(rho, r)
    return rho


def apply_bs(rho, M, d, mode, U3):
    G = two_mode_matrix(U3)
    shape = [d] * M
    T = rho.reshape(shape + shape)
    # ket side
    T = np.moveaxis(T, [mode, mode + 1], [0, 1])
    s = T.shape
    T = (G @ T.reshape(d * d, -1)).reshape(s)
    T = np.moveaxis(T, [0, 1], [mode, mode + 1])
    # bra side
    T = np.moveaxis(T, [M + mode, M + mode + 1], [0, 1])
    s = T.shape
    T = (G.conj() @ T.reshape(d * d, -1)).reshape(s)
    T = np.moveaxis(T, [0, 1], [M + mode, M + mode + 1])
    return T.reshape(d ** M, d ** M)


def run(M, d, input_modes, eta, mesh):
    rho = initial_rho(M, d, input_modes, eta)
    max_input = max(input_modes)
    for layer in range(M):
        for i, mode in layer_modes(M, layer, max_input):
            rho = apply_bs(rho, M, d, mode, bs_table(d, mesh[layer, i]))
    return rho


def marginals(rho, M, d):
    p = np.real(np.diag(rho)).reshape([d] * M)
    P = np.zeros((M, d))
    for j in range(M):
        axes = tuple(a for a in range(M) if a != j)
        P[j] = p.sum(axis=axes)
    return P


def pattern_probability(rho, M, d, pattern):
    p = np.real(np.diag(rho)).reshape([d] * M)
    return float(p[tuple(pattern)])
