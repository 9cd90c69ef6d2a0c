"""Pins the FP64 oracle: closed forms, library routines and brute force.

The reference ships no tests, fixtures or golden vectors (its CPU path
src/cpu_algorithm is an empty file), so every oracle function is pinned
here by mathematics chosen so that a plausible slip (dropped term, wrong
sign or index, transposed operand) fails at least one check.
"""
import math
import os

import numpy as np
import pytest
import scipy.linalg

from oracle import brute, fock, mpo
from bosonsampling_amd import mesh


# ------------------------------------------------------------------ fock --

@pytest.mark.parametrize("d", [2, 4, 6])
def test_bs_three_forms_agree(d):
    rng = np.random.default_rng(d)
    for _ in range(3):
        B = mesh.haar_unitary(2, int(rng.integers(1 << 30)))
        a = fock.bs_table(d, B)
        b = fock.bs_table_permanent(d, B)
        assert np.abs(a - b).max() < 1e-12
    t, r = mesh.reference_phases(3, 0.42)
    ref = fock.bs_table_reference(d, t, r)
    assert np.abs(ref - fock.bs_table(d, fock.reference_bs_matrix(t, r))).max() < 1e-13


def test_bs_single_photon_is_B():
    B = mesh.haar_unitary(2, 5)
    U = fock.bs_table(3, B)
    # <1,0|U|1,0> = B00, <1,0|U|0,1> = B01, <0,1|U|1,0> = B10, <0,1|U|0,1> = B11
    assert np.isclose(U[1, 0, 1], B[0, 0]) and np.isclose(U[1, 0, 0], B[0, 1])
    assert np.isclose(U[0, 1, 1], B[1, 0]) and np.isclose(U[0, 1, 0], B[1, 1])


def test_bs_reference_single_photon_entries():
    """U3[1,0,1] = t, U3[1,0,0] = r, U3[0,1,0] = t*, U3[0,1,1] = -r*
    (how src/update_MPO_kernel.cu:491-497 reads src/MPO.py:11)."""
    t, r = mesh.reference_phases(17, 0.6)
    U = fock.bs_table_reference(3, t, r)
    assert np.isclose(U[1, 0, 1], t) and np.isclose(U[1, 0, 0], r)
    assert np.isclose(U[0, 1, 0], np.conj(t)) and np.isclose(U[0, 1, 1], -np.conj(r))


@pytest.mark.parametrize("d", [5, 9])
def test_bs_unitary_per_photon_sector(d):
    B = mesh.haar_unitary(2, 9)
    U = fock.bs_table(d, B)
    for s in range(d):          # total photons s < d: no cutoff truncation
        idx = [(n1, s - n1) for n1 in range(s + 1)]
        Ms = np.array([[U[n1, n2, p] for p in range(s + 1)] for (n1, n2) in idx])
        assert np.abs(Ms @ Ms.conj().T - np.eye(s + 1)).max() < 1e-12


def test_bs_matches_matrix_exponential():
    """U(B) for B = exp(i H) equals exp of the second-quantised generator
    restricted to each photon-number sector (scipy.linalg.expm)."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    H = (A + A.conj().T) / 2
    B = scipy.linalg.expm(1j * H)
    d = 6
    U = fock.bs_table(d, B)
    for s in range(1, d):
        # a_i^+ a_j in the basis |n1, s-n1>, generator sum_ij H_ij a_i^+ a_j
        n = s + 1
        G = np.zeros((n, n), complex)
        for k in range(n):
            n1, n2 = k, s - k
            G[k, k] += H[0, 0] * n1 + H[1, 1] * n2
            if n2 > 0:   # a_0^+ a_1 |n1, n2> = sqrt((n1+1) n2) |n1+1, n2-1>
                G[k + 1, k] += H[0, 1] * math.sqrt((n1 + 1) * n2)
            if n1 > 0:
                G[k - 1, k] += H[1, 0] * math.sqrt(n1 * (n2 + 1))
        E = scipy.linalg.expm(1j * G)
        Us = np.array([[U[n1, s - n1, p] for p in range(n)] for n1 in range(n)])
        assert np.abs(E - Us).max() < 1e-11


def test_hong_ou_mandel_dip():
    B = mesh.clements_bs_matrix(math.pi / 4, 0.3)
    U = fock.bs_table(3, B)
    assert abs(U[1, 1, 1]) < 1e-15            # |t|^2 - |r|^2 = 0 at 50:50


def test_identity_phase_bs_is_diagonal():
    U = fock.bs_table_reference(5, np.exp(0.7j), 0.0)
    for n in range(5):
        for m in range(5):
            for l in range(5):
                if l != n and n + m - l < 5 and n + m - l >= 0:
                    assert U[n, m, l] == 0


# --------------------------------------------------------------- mesh -----

@pytest.mark.parametrize("M,seed", [(4, 0), (8, 0), (16, 2), (32, 1), (64, 100)])
def test_clements_reconstructs_unitary(M, seed):
    U = mesh.haar_unitary(M, seed)
    seq, D = mesh.clements_decompose(U)
    mm, used, nl = mesh.schedule_brickwall(M, seq)
    slots = np.zeros_like(used)
    for lay in range(M):
        slots[lay, :len(range(lay % 2, M - 1, 2))] = True
    assert len(seq) == M * (M - 1) // 2 and np.array_equal(used, slots) and nl == M
    W = mesh.mesh_unitary(mm)
    assert np.abs(np.diag(D) @ W - U).max() < 1e-12


def test_haar_unitary_is_unitary_and_seeded():
    U = mesh.haar_unitary(12, 4)
    assert np.abs(U @ U.conj().T - np.eye(12)).max() < 1e-13
    assert np.array_equal(U, mesh.haar_unitary(12, 4))
    assert not np.array_equal(U, mesh.haar_unitary(12, 5))


def test_reference_phase_rng_order():
    """Rank.py:56-57 + cuda_kernels.py:37-43: seed, then alpha (t) then beta (r)."""
    rs = np.random.RandomState(123)
    a, b = rs.rand(), rs.rand()
    t, r = mesh.reference_phases(123, 0.3)
    assert np.isclose(t, math.sqrt(1 - 0.09) * np.exp(2j * np.pi * a))
    assert np.isclose(r, 0.3 * np.exp(2j * np.pi * b))
    t0, r0 = mesh.reference_phases(123, 0.0)
    assert t0 == 1 and r0 == 0


def test_reference_reflectivities_shape_and_range():
    refl, seeds = mesh.reflectivity_and_seeds(8, seed=3)
    assert refl.shape == (8, 4) and seeds.shape == (8, 4)
    valid = np.zeros((8, 4), bool)
    for lay in range(8):
        for i, m in enumerate(range(lay % 2, 7, 2)):
            valid[lay, i] = True
    assert np.all((refl[valid] >= 0) & (refl[valid] <= 1))
    assert np.all((seeds[valid] >= 0) & (seeds[valid] < 13579))


# ---------------------------------------------------------------- mpo -----

def _run(M, d, chi, inputs, eta, seed):
    mm, W = mesh.haar_clements_mesh(M, seed)
    st = mpo.initialize(M, d, chi, inputs, eta)
    prob, EE = mpo.simulate(st, mm, inputs)
    return st, mm, W, prob, EE


@pytest.mark.parametrize("M,d,chi,inputs,eta,seed", [(4, 4, 64, [0, 1, 2], 0.7, 3),
                                                      (5, 3, 64, [1, 3], 0.8, 4),
                                                      (4, 5, 256, [0, 1, 2, 3], 0.5, 0)])
def test_mpo_matches_brute_force(M, d, chi, inputs, eta, seed):
    st, mm, W, prob, EE = _run(M, d, chi, inputs, eta, seed)
    rho = brute.run(M, d, inputs, eta, mm)
    assert np.allclose(prob, 1.0, atol=1e-12)
    assert np.abs(brute.marginals(rho, M, d) - mpo.marginals(st)).max() < 1e-12
    for pattern in ([1] * len(inputs) + [0] * (M - len(inputs)), [0] * M):
        assert abs(mpo.pattern_probability(st, pattern)
                   - brute.pattern_probability(rho, M, d, pattern)) < 1e-12


def test_mean_photon_number_closed_form():
    M, d, inputs, eta = 8, 5, [0, 2, 4, 6], 0.6
    st, mm, W, prob, EE = _run(M, d, 256, inputs, eta, 11)
    nbar = eta * (np.abs(W[:, inputs]) ** 2).sum(axis=1)
    P = mpo.marginals(st)
    assert np.abs((P * np.arange(d)).sum(axis=1) - nbar).max() < 1e-11
    assert np.allclose(P.sum(axis=1), 1.0, atol=1e-12)


def test_pattern_probability_permanent():
    M, d, inputs, eta = 6, 4, [0, 1, 2], 0.8
    st, mm, W, prob, EE = _run(M, d, 128, inputs, eta, 5)
    out = [1, 3, 4]
    pattern = [1 if j in out else 0 for j in range(M)]
    perm = fock.permanent(W[np.ix_(out, inputs)])
    assert abs(mpo.pattern_probability(st, pattern) - eta ** 3 * abs(perm) ** 2) < 1e-12


def test_literal_contraction_matches_block_form():
    M, d, chi, inputs = 6, 4, 16, [0, 1, 2]
    mm, _ = mesh.haar_clements_mesh(M, 2)
    st = mpo.initialize(M, d, chi, inputs, 0.7)
    mpo.simulate(st, mm, inputs, layers=range(3))
    mode = 2
    U3 = fock.bs_table(d, mm[3, 0])
    sec = mpo.contract_sectors(d, U3, st.charge[mode], st.charge[mode + 1], st.charge[mode + 2],
                               st.Gamma[mode], st.Gamma[mode + 1])
    assert sec
    for s in sec:
        rows, cols, C = mpo.contract_sector_literal(d, U3, st.charge[mode], st.charge[mode + 1],
                                                    st.charge[mode + 2], st.Gamma[mode],
                                                    st.Gamma[mode + 1], s.tau)
        assert np.array_equal(rows, s.rows) and np.array_equal(cols, s.cols)
        assert np.abs(C - s.C).max() < 1e-14


def test_initial_state_is_product():
    M, d, chi, inputs, eta = 7, 5, 32, [1, 3, 4, 6], 0.55
    st = mpo.initialize(M, d, chi, inputs, eta)
    assert abs(mpo.trace(st) - 1.0) < 1e-13
    P = mpo.marginals(st)
    for j in range(M):
        exp = [1 - eta, eta] if j in inputs else [1.0, 0.0]
        assert np.allclose(P[j, :2], exp, atol=1e-13) and np.allclose(P[j, 2:], 0, atol=1e-13)
    for b in range(M + 1):
        key = st.charge[b][:, 0] * (d + 1) + st.charge[b][:, 1]
        assert np.all(np.diff(key) >= 0)


def test_truncation_keeps_top_chi_without_renormalisation():
    st, mm, W, prob, EE = _run(8, 5, 16, [0, 1, 2, 3], 0.5, 0)
    assert prob[-1] < prob[0]                  # trace drifts (Rank.py has no renorm)
    assert (st.Lambda > 0).sum(axis=1).max() <= 16


def test_select_top_ties_and_order():
    kept = mpo.select_top([np.array([3.0, 1.0]), np.array([2.0, 1.0, 0.5])], 3)
    assert kept == [(0, 0), (0, 1), (1, 0)]   # tie at 1.0: lower concat index wins
    # twins split by rounding noise are still ties: the lower index wins
    for eps in (1e-15, -1e-15, 3e-13, 5e-7):
        kept = mpo.select_top([np.array([3.0]), np.array([1.0]), np.array([1.0 * (1 + eps)])], 2)
        assert kept == [(0, 0), (1, 0)]
    # genuinely different values are not ties
    kept = mpo.select_top([np.array([3.0]), np.array([1.0]), np.array([1.0 + 1e-5])], 2)
    assert kept == [(0, 0), (2, 0)]
    kept = mpo.select_top([np.array([3.0]), np.array([5.0])], 10)
    assert kept == [(0, 0), (1, 0)]


# --------------------------------------------------------------- entropy --
# src/summary.py:3-8,28-30.  Closed forms pin the base (log2), the squaring of
# Lambda and the normalisation: a restated formula with ln, with Lambda in
# place of Lambda^2 or without the column normalisation fails one of them.

def test_entropy_equal_values_is_log2_k():
    for k in (1, 2, 3, 7, 64):
        lam = np.zeros((1, 80))
        lam[0, :k] = 0.37
        assert abs(mpo.entanglement_entropy(lam)[0] - np.log2(k)) < 1e-13


def test_entropy_two_values_is_binary_entropy():
    a, b = 0.9, 0.2
    p = a * a / (a * a + b * b)        # weights are Lambda^2, normalised
    h = -(p * np.log2(p) + (1 - p) * np.log2(1 - p))
    lam = np.array([[a, 0.0, b, 0.0]])
    assert abs(mpo.entanglement_entropy(lam)[0] - h) < 1e-13
    # Lambda (not squared) would give a different value
    q = a / (a + b)
    assert abs(h + (q * np.log2(q) + (1 - q) * np.log2(1 - q))) > 1e-2


def test_entropy_single_value_scale_and_zero_bond():
    lam = np.array([[0.0, 3.0, 0.0], [0.0, 0.0, 0.0]])
    S = mpo.entanglement_entropy(lam)
    assert S[0] == 0.0
    assert S[1] == 0.0                  # nansum over 0/0 in the reference
    rng = np.random.default_rng(4)
    lam = rng.random((3, 50))
    assert np.allclose(mpo.entanglement_entropy(7.5 * lam), mpo.entanglement_entropy(lam),
                       rtol=0, atol=1e-13)


# ------------------------------------------------ reference mesh sampler --
# src/interferometer.py:9-54 draws reflectivities through scipy's generic
# rv_continuous sampler (numeric inverse CDF) and integer phase seeds from
# numpy's global RNG; tests/golden/reflectivity.npz holds its own output
# (tests/golden/make_reflectivity.py).  The mirror must give the same values
# bit for bit on every slot that holds a beam splitter.

def test_reflectivity_and_seeds_match_reference_vectors():
    from golden.make_reflectivity import used_slots
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "reflectivity.npz"))
    for M, seed in z["cases"]:
        r, s = mesh.reflectivity_and_seeds(int(M), int(seed))
        mask = used_slots(int(M))
        assert np.array_equal(np.where(mask, r, 0.0), z[f"r_{M}_{seed}"])
        assert np.array_equal(np.where(mask, s, 0), z[f"s_{M}_{seed}"])


def test_reflectivity_and_seeds_odd_modes_raise_like_the_reference():
    with pytest.raises(IndexError):
        mesh.reflectivity_and_seeds(5, 1)
