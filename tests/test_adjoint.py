"""Pins the oracle's reverse-mode functions (oracle/adjoint.py) by central
finite differences of the oracle's forward functions."""
import math

import numpy as np
import pytest

from oracle import adjoint, fock, mpo
from bosonsampling_amd import mesh as meshmod

H = 1e-6


def _fd_complex(f, x, idx_list):
    """Real-packed FD gradient dL/dRe + i dL/dIm at the given indices."""
    out = []
    for idx in idx_list:
        g = 0j
        for unit in (1.0, 1j):
            xp = x.copy()
            xm = x.copy()
            xp[idx] += unit * H
            xm[idx] -= unit * H
            g += unit * (f(xp) - f(xm)) / (2 * H)
        out.append(g)
    return np.array(out)


@pytest.mark.parametrize("shape", [(7, 11), (11, 7), (9, 9)])
def test_sector_vjp_gauge_invariant_functional(shape):
    """f = Re<Z, A diag(q) B> + p.s + Re<Q, A A^H> depends on the kept
    subspace only (A = V_P, B = A^H C, s = S_P), so it has a gradient."""
    rng = np.random.default_rng(sum(shape))
    L, R = shape
    C = rng.standard_normal((L, R)) + 1j * rng.standard_normal((L, R))
    LRc = rng.uniform(0.3, 1.0, R)
    kP = 4
    Z = rng.standard_normal((L, R)) + 1j * rng.standard_normal((L, R))
    q = rng.uniform(0.5, 1.5, kP)
    p = rng.standard_normal(kP)
    Q = rng.standard_normal((L, L)) + 1j * rng.standard_normal((L, L))

    def parts(C, LRc):
        V, S = mpo._svd_left(C * LRc[None, :])
        A = V[:, :kP]
        return V, S, A, A.conj().T @ C, S[:kP]

    def f_of(C, LRc):
        _, _, A, B, s = parts(C, LRc)
        return (np.real(np.sum(np.conj(Z) * (A @ np.diag(q) @ B))) + p @ s
                + np.real(np.sum(np.conj(Q) * (A @ A.conj().T))))

    V, S, A, B, s = parts(C, LRc)
    gA = Z @ B.conj().T @ np.diag(q) + (Q + Q.conj().T) @ A
    gB = np.diag(q) @ A.conj().T @ Z
    gC, gLR = adjoint.sector_vjp(C, LRc, V, S, np.arange(kP), gA, gB, p)
    idx = [(i, j) for i in range(L) for j in range(R)][::5]
    fd = _fd_complex(lambda x: f_of(x, LRc), C, idx)
    got = np.array([gC[i] for i in idx])
    assert np.abs(got - fd).max() <= 1e-6 * np.abs(fd).max()
    fd_lr = np.array([(f_of(C, LRc + H * e) - f_of(C, LRc - H * e)) / (2 * H)
                      for e in np.eye(R)])
    assert np.abs(gLR - fd_lr).max() <= 1e-6 * np.abs(fd_lr).max()


def _state(M=6, d=4, chi=12, inputs=(0, 1, 2), eta=0.8, seed=3, layers=2):
    mm, _ = meshmod.haar_clements_mesh(M, seed)
    st = mpo.initialize(M, d, chi, list(inputs), eta)
    mpo.simulate(st, mm, list(inputs), layers=range(layers))
    return st, mm


def test_contract_vjp_linear_functional():
    st, mm = _state()
    d, mode = st.d, 1
    U3 = fock.bs_table(d, mm[2, 0])
    CL, CC, CR = st.charge[mode], st.charge[mode + 1], st.charge[mode + 2]
    sec = mpo.contract_sectors(d, U3, CL, CC, CR, st.Gamma[mode], st.Gamma[mode + 1])
    rng = np.random.default_rng(0)
    Zs = [rng.standard_normal(s.C.shape) + 1j * rng.standard_normal(s.C.shape) for s in sec]

    def f(Glc, Gcr, U3x):
        ss = mpo.contract_sectors(d, U3x, CL, CC, CR, Glc, Gcr)
        return sum(np.real(np.sum(np.conj(z) * s.C)) for z, s in zip(Zs, ss))

    gGlc, gGcr, gU3 = adjoint.contract_vjp(d, U3, CL, CC, CR, st.Gamma[mode], st.Gamma[mode + 1],
                                           sec, Zs)
    nz = np.argwhere(np.abs(st.Gamma[mode]) > 0)[::7]
    fd = _fd_complex(lambda x: f(x, st.Gamma[mode + 1], U3), st.Gamma[mode], [tuple(i) for i in nz])
    assert np.abs(np.array([gGlc[tuple(i)] for i in nz]) - fd).max() <= 1e-7 * np.abs(fd).max()
    nz = np.argwhere(np.abs(st.Gamma[mode + 1]) > 0)[::7]
    fd = _fd_complex(lambda x: f(st.Gamma[mode], x, U3), st.Gamma[mode + 1], [tuple(i) for i in nz])
    assert np.abs(np.array([gGcr[tuple(i)] for i in nz]) - fd).max() <= 1e-7 * np.abs(fd).max()
    nz = np.argwhere(np.abs(U3) > 0)[::3]
    fd = _fd_complex(lambda x: f(st.Gamma[mode], st.Gamma[mode + 1], x), U3, [tuple(i) for i in nz])
    assert np.abs(np.array([gU3[tuple(i)] for i in nz]) - fd).max() <= 1e-7 * np.abs(fd).max()


def test_bs_table_vjp():
    d = 5
    rng = np.random.default_rng(1)
    B = meshmod.haar_unitary(2, 4)
    G = rng.standard_normal((d, d, d)) + 1j * rng.standard_normal((d, d, d))
    f = lambda Bx: np.real(np.sum(np.conj(G) * fock.bs_table(d, Bx)))
    gB = adjoint.bs_table_vjp(d, B, G)
    fd = _fd_complex(f, B, [(0, 0), (0, 1), (1, 0), (1, 1)])
    assert np.abs(gB.reshape(-1) - fd).max() <= 1e-7 * np.abs(fd).max()


def test_clements_vjp():
    th, ph = 0.7, -1.3
    gB = np.array([[0.3 + 0.2j, -1.1j], [0.5, 0.9 - 0.4j]])
    f = lambda t, p: np.real(np.sum(np.conj(gB) * meshmod.clements_bs_matrix(t, p)))
    gt, gp = adjoint.clements_vjp(th, ph, gB)
    assert abs(gt - (f(th + H, ph) - f(th - H, ph)) / (2 * H)) < 1e-8
    assert abs(gp - (f(th, ph + H) - f(th, ph - H)) / (2 * H)) < 1e-8


def test_observable_loss_gradient():
    st, mm = _state(layers=6)
    M = st.M
    pattern = [1, 0, 1, 0, 1, 0]
    w = np.linspace(0.1, 0.6, M)
    loss, P, nw, g = adjoint.observable_loss(st, pattern, w)
    assert abs(P - mpo.pattern_probability(st, pattern)) < 1e-14
    assert abs(nw - float(mpo.marginals(st) @ np.arange(st.d) @ w)) < 1e-12
    for b in (0, 2, 5):
        nz = np.argwhere(np.abs(st.Gamma[b]) > 0)[::3]

        def f(x, b=b):
            s2 = st.copy()
            s2.Gamma[b] = x
            return adjoint.observable_loss(s2, pattern, w)[0]

        fd = _fd_complex(f, st.Gamma[b], [tuple(i) for i in nz])
        assert np.abs(np.array([g[b][tuple(i)] for i in nz]) - fd).max() <= 1e-6 * np.abs(fd).max()


def _clements_params(M, seed):
    """(theta, phi) per brick-wall slot of the Haar/Clements mesh."""
    U = meshmod.haar_unitary(M, seed)
    seq, _ = meshmod.clements_decompose(U)
    params = np.zeros((M, M // 2, 2))
    last = [-1] * M
    for (m, th, ph) in seq:
        lay = max(last[m], last[m + 1]) + 1
        if lay % 2 != m % 2:
            lay += 1
        params[lay, (m - lay % 2) // 2] = (th, ph)
        last[m] = last[m + 1] = lay
    return params


@pytest.mark.parametrize("chi,ckpt", [(64, None), (16, 3), (8, 2)])
def test_whole_run_gradient_matches_finite_differences(chi, ckpt):
    """-log P(pattern) + sum w_j <n_j> through every update, with and without
    truncation, against central differences (step 1e-6) on (theta, phi)."""
    M, d, inputs, eta = 6, 4, [0, 1, 2], 0.8
    params = _clements_params(M, 5)
    st0 = mpo.initialize(M, d, chi, inputs, eta)
    pattern = [1, 0, 1, 0, 1, 0]
    w = np.linspace(0.2, -0.3, M)
    loss, grad = adjoint.loss_and_grad(st0, params, inputs, pattern, w, checkpoint_every=ckpt)

    def L(p):
        st = st0.copy()
        mesh = adjoint.mesh_from_params(M, p)
        for layer in range(M):
            for i, mode in mpo.layer_modes(M, layer, max(inputs)):
                mpo.apply_update(st, mode, fock.bs_table(d, mesh[layer, i]))
        return adjoint.observable_loss(st, pattern, w)[0]

    assert abs(loss - L(params)) < 1e-12
    used = [(lay, i) for lay in range(M) for i, _ in mpo.layer_modes(M, lay, max(inputs))]
    fd = np.zeros_like(params)
    for lay, i in used:
        for c in range(2):
            pp, pm = params.copy(), params.copy()
            pp[lay, i, c] += H
            pm[lay, i, c] -= H
            fd[lay, i, c] = (L(pp) - L(pm)) / (2 * H)
    scale = np.abs(fd).max()
    print(f"chi={chi}: max |grad - fd| / max|fd| = {np.abs(grad - fd).max() / scale:.2e}")
    assert np.abs(grad - fd).max() <= 1e-5 * scale
