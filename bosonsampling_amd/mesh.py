"""Synthetic interferometer meshes: Haar-random unitaries, Clements rectangular
decomposition, and the brick-wall schedule of the reference.

The reference has no Haar/Clements code (it samples reflectivities,
src/interferometer.py:9-54); its brick-wall layer order is
src/beam_splitter_parallel/Full.py:152 -- layer l acts on (m, m+1) for
m = l % 2, l % 2 + 2, ...  This module produces, for an M x M unitary, an
array ``mesh[l, i]`` of 2x2 single-particle matrices (new = B @ old on modes
(m, m+1), m = l % 2 + 2 i) such that, applied layer by layer, they realise the
unitary up to an output phase screen (which photon-number statistics ignore).
"""
from __future__ import annotations

import math

import numpy as np



def clements_bs_matrix(theta: float, phi: float) -> np.ndarray:
    """Clements et al. (Optica 3, 1460 (2016)) element T(theta, phi), 2x2 block:
    [[e^{i phi} cos(theta), -sin(theta)], [e^{i phi} sin(theta), cos(theta)]]."""
    c, s = math.cos(theta), math.sin(theta)
    e = np.exp(1j * phi)
    return np.array([[e * c, -s], [e * s, c]], dtype=np.complex128)


def haar_unitary(M: int, seed: int) -> np.ndarray:
    """Haar unitary: QR of a complex Ginibre matrix with Mezzadri's phase fix."""
    rng = np.random.default_rng(seed)
    Z = (rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))) / math.sqrt(2.0)
    Q, R = np.linalg.qr(Z)
    ph = np.diag(R) / np.abs(np.diag(R))
    return Q * ph[None, :]


def _embed(M: int, m: int, B: np.ndarray) -> np.ndarray:
    G = np.eye(M, dtype=np.complex128)
    G[m:m + 2, m:m + 2] = B
    return G


def clements_decompose(U: np.ndarray):
    """Clements et al., Optica 3, 1460 (2016), rectangular decomposition.

    Returns (seq, D): ``seq`` is the list of (m, theta, phi) in APPLICATION
    order (first applied first) and D the output phase diagonal, with
    U = diag(D) @ G_K @ ... @ G_1, G = T_m(theta, phi) embedded.
    """
    U = np.array(U, dtype=np.complex128)
    N = U.shape[0]
    right = []   # T's applied as U <- U @ T^dagger (mixing columns m, m+1)
    left = []    # T's applied as U <- T @ U (mixing rows m, m+1)
    for i in range(N - 1):
        if i % 2 == 0:
            for j in range(i + 1):
                x, m = N - 1 - j, i - j
                a, b = U[x, m], U[x, m + 1]
                theta = math.atan2(abs(a), abs(b))
                phi = float(np.angle(a) - np.angle(b)) if abs(a) > 0 and abs(b) > 0 else 0.0
                T = clements_bs_matrix(theta, phi)
                U = U @ _embed(N, m, T).conj().T
                right.append((m, theta, phi))
        else:
            for j in range(1, i + 2):
                x, y = N + j - i - 2, j - 1
                m = x - 1
                a, b = U[m, y], U[x, y]
                theta = math.atan2(abs(b), abs(a))
                phi = float(np.pi + np.angle(b) - np.angle(a)) if abs(a) > 0 and abs(b) > 0 else 0.0
                T = clements_bs_matrix(theta, phi)
                U = _embed(N, m, T) @ U
                left.append((m, theta, phi))
    D = np.diag(U).copy()
    # U_orig = L_1^+ ... L_k^+ D R_p ... R_1 ; push each L^+ through D:
    # T(th, ph)^+ diag(d1, d2) = diag(-e^{-i ph} d2, d2) T(th, ph'),  e^{i ph'} = -d1/d2
    moved = []
    for (m, theta, phi) in reversed(left):
        d1, d2 = D[m], D[m + 1]
        phi_new = float(np.angle(-d1 / d2))
        D[m] = -np.exp(-1j * phi) * d2
        D[m + 1] = d2
        moved.append((m, theta, phi_new))
    # U = D T'_1 ... T'_k R_p ... R_1 with moved = [T'_k, ..., T'_1]
    seq = list(right) + list(moved)
    return seq, D


def schedule_brickwall(M: int, seq):
    """Assign the ordered BS list to the reference brick wall
    (src/beam_splitter_parallel/Full.py:152).  Returns mesh[M, M//2, 2, 2]
    (identity where a slot is unused) and the number of layers used."""
    n_pos = M // 2
    mesh = np.zeros((M, n_pos, 2, 2), dtype=np.complex128)
    mesh[:, :] = np.eye(2)
    last = [-1] * M
    used = np.zeros((M, n_pos), dtype=bool)
    n_layers = 0
    for (m, theta, phi) in seq:
        lay = max(last[m], last[m + 1]) + 1
        if lay % 2 != m % 2:
            lay += 1
        if lay >= M:
            raise ValueError("sequence does not fit the M-layer brick wall")
        pos = (m - lay % 2) // 2
        if used[lay, pos]:
            raise ValueError("slot collision in brick wall")
        mesh[lay, pos] = clements_bs_matrix(theta, phi)
        used[lay, pos] = True
        last[m] = last[m + 1] = lay
        n_layers = max(n_layers, lay + 1)
    return mesh, used, n_layers


def haar_clements_mesh(M: int, seed: int):
    """mesh[l, i] (2x2, new = B @ old) for a Haar unitary of the given seed,
    plus the unitary W realised by the mesh (W = D^-1 U, columns = inputs)."""
    U = haar_unitary(M, seed)
    seq, D = clements_decompose(U)
    mesh, used, _ = schedule_brickwall(M, seq)
    return mesh, mesh_unitary(mesh)


def mesh_unitary(mesh: np.ndarray) -> np.ndarray:
    """Single-particle unitary realised by applying mesh layer by layer."""
    M = mesh.shape[0]
    W = np.eye(M, dtype=np.complex128)
    for lay in range(M):
        for i, m in enumerate(range(lay % 2, M - 1, 2)):
            W = _embed(M, m, mesh[lay, i]) @ W
    return W


def reference_phases(seed: int, r: float):
    """(t, r) complex for a reference BS: np.random.seed(seed) then two
    np.random.rand() draws, alpha for t then beta for r.

    Follows src/beam_splitter_parallel/Rank.py:56-57 and
    src/cuda_kernels.py:18-43 (r == 0: t = 1 and only one draw, for r's phase).
    """
    rs = np.random.RandomState(int(seed))
    if r == 0:
        return 1.0 + 0.0j, complex(0.0 * np.exp(1j * rs.rand() * 2 * np.pi))
    t = math.sqrt(1 - r ** 2) * np.exp(1j * rs.rand() * 2 * np.pi)
    rr = r * np.exp(1j * rs.rand() * 2 * np.pi)
    return complex(t), complex(rr)


def reference_bs_matrix(t: complex, r: complex) -> np.ndarray:
    """Single-particle matrix B (new = B @ old) of the reference's (t, r):
    [[t, r], [-r*, t*]] (src/MPO.py:8-11 as indexed by
    src/update_MPO_kernel.cu:491-497)."""
    return np.array([[t, r], [-np.conj(r), np.conj(t)]], dtype=np.complex128)


class _HaarReflectivity:
    """pdf (k - idx)(1 - x)^(k - idx - 1) on [0, 1] (src/interferometer.py:4-7),
    sampled with scipy's generic rv_continuous machinery as the reference does."""

    _dist = None

    @classmethod
    def rvs(cls, k, idx):
        if cls._dist is None:
            from scipy.stats import rv_continuous

            class _pdf(rv_continuous):
                def _pdf(self, x, k, idx):
                    return (k - idx) * (1 - x) ** (k - idx - 1)

            cls._dist = _pdf(a=0, b=1, name="my_pdf")
        return cls._dist.rvs(k, idx)


def reflectivity_and_seeds(n_modes: int, seed=None):
    """Reference mesh parameters: per-BS reflectivity |r| from the
    Haar-induced law and an integer phase seed (src/interferometer.py:9-54).
    Uses numpy's global RNG exactly as the reference does.  Slots no beam
    splitter uses (the last one of odd layers) are 0 here, where the reference
    leaves np.empty's contents; odd n_modes raise IndexError, as in the
    reference."""
    if seed is not None:
        np.random.seed(seed)
    reflectivity = np.zeros([n_modes, n_modes // 2])
    seeds = np.zeros([n_modes, n_modes // 2], dtype=int)
    for k in range(n_modes - 1):
        if k < n_modes / 2:
            temp1, temp2 = 2 * k + 1, 2
            l, i = 2 * k, 0
            while l >= 0:
                if temp1 > 0:
                    T = _HaarReflectivity.rvs(2 * k + 2, temp1)
                    temp1 -= 2
                else:
                    T = _HaarReflectivity.rvs(2 * k + 2, temp2)
                    temp2 += 2
                reflectivity[i, k - (i + 1) // 2] = np.sqrt(1 - T)
                seeds[i, k - (i + 1) // 2] = np.random.randint(0, 13579)
                l -= 1
                i += 1
        else:
            temp1, temp2 = 2 * n_modes - (2 * k + 3), 2
            first_layer = 2 * k - n_modes + 2
            for i in range(2 * n_modes - 2 * k - 2):
                if temp1 >= 0:
                    T = _HaarReflectivity.rvs(2 * n_modes - (2 * k + 1), temp1)
                    temp1 -= 2
                else:
                    T = _HaarReflectivity.rvs(2 * n_modes - (2 * k + 1), temp2)
                    temp2 += 2
                reflectivity[first_layer + i, n_modes // 2 - 1 - (i + 1) // 2] = np.sqrt(1 - T)
                seeds[first_layer + i, n_modes // 2 - 1 - (i + 1) // 2] = np.random.randint(0, 13579)
    return reflectivity, seeds


def reference_mesh(M: int, reflectivity: np.ndarray, seeds: np.ndarray) -> np.ndarray:
    """mesh[l, i] for reference-style parameters (per-BS |r| and integer seed),
    as src/beam_splitter_parallel/Full.py:152-161 + Rank.py:56-57 consume
    them: B = [[t, r], [-r*, t*]] with the phases drawn from the seed."""
    n_pos = M // 2
    mesh = np.zeros((M, n_pos, 2, 2), dtype=np.complex128)
    mesh[:, :] = np.eye(2)
    for lay in range(M):
        for i, m in enumerate(range(lay % 2, M - 1, 2)):
            t, r = reference_phases(int(seeds[lay, i]), float(reflectivity[lay, i]))
            mesh[lay, i] = reference_bs_matrix(t, r)
    return mesh
