"""FP64 reverse-mode restatement of the site update, the truncated SVD, the
BS table and the output observables (test infrastructure, like the rest of
oracle/).

The reference has no derivative code; these are the adjoints of the oracle's
own forward functions (oracle/mpo.py, oracle/fock.py), written from the
matrix-calculus identities named at each function, and pinned in
tests/test_adjoint.py by central finite differences of the forward.

Convention: for a real loss L and a complex array z, the gradient is
g = dL/dRe(z) + i dL/dIm(z), so that dL = Re(sum(conj(g) * dz)).  Linear
maps y = A x then give g_x = A^H g_y and g_A = g_y x^H.
"""
from __future__ import annotations

import math

import numpy as np

from . import mpo
from .fock import _comb, _fact

# eigenvalue pairs closer than this (relative to the largest) get no
# eigenvector coupling term (their 1/(lambda_j - lambda_i) is dropped)
DEGEN_REL = 1e-13


# --------------------------------------------------------------- forward --

def site_update_tape(d, chi, U3, LR, CL, CC, CR, Glc, Gcr, drop_rel=mpo.DROP_REL):
    """mpo.site_update's forward, keeping the per-sector factors and the
    kept (sector, position) list for the adjoint."""
    sec = mpo.contract_sectors(d, U3, CL, CC, CR, Glc, Gcr)
    for s in sec:
        T = s.C * np.asarray(LR, np.float64)[s.cols][None, :]
        V, S = mpo._svd_left(T)
        keep = S > drop_rel * S[0] if (S.size and S[0] > 0) else np.zeros(S.size, bool)
        s.V, s.S = V[:, keep], S[keep]
        s.W = s.V.conj().T @ s.C
    kept = mpo.select_top([s.S for s in sec], chi)
    return sec, kept


# -------------------------------------------------------- eigen adjoint --

def _eig_mid(lam, J, g_lam):
    """H = sym(J o F) + diag(g_lam), F_ij = 1 / (lam_j - lam_i) (i != j):
    the adjoint of a Hermitian eigendecomposition G = Y diag(lam) Y^H is
    G_bar = Y H Y^H for eigenvector gradient Y_bar = Y J and eigenvalue
    gradient g_lam (Giles, "An extended collection of matrix derivative
    results", 2008, sec. 3.1, complex Hermitian case)."""
    diff = lam[None, :] - lam[:, None]
    scale = max(float(np.abs(lam).max()), 1e-300)
    with np.errstate(divide="ignore", invalid="ignore"):
        F = np.where(np.abs(diff) > DEGEN_REL * scale, 1.0 / diff, 0.0)
    K = J * F
    return 0.5 * (K + K.conj().T) + np.diag(g_lam)


def sector_vjp(C, LRc, V, S, P, gA, gB, gs):
    """Adjoint of one sector: T = C diag(LRc), T = V_all S_all W^H (thin),
    outputs A = V[:, P], B = A^H C, s = S[P].  Given gA (L x |P|), gB
    (|P| x R), gs (|P|), returns (gC, gLRc).

    Uses the Gram matrix of the smaller side, as the device path does:
      L <= R: G = T T^H, Y = left singular vectors, A = Y_P;
      L >  R: G = T^H T, Y = right singular vectors, A = T Y_P / s_P.
    """
    L, R = C.shape
    T = C * LRc[None, :]
    A = V[:, P]
    sP = S[P]
    gC = A @ gB
    gA = gA + C @ gB.conj().T
    if L <= R:
        lam, Y = np.linalg.eigh(T @ T.conj().T)
        lam, Y = lam[::-1], Y[:, ::-1].copy()
        Y[:, P] = A                               # the forward's gauge on kept columns
        gY = np.zeros((L, L), np.complex128)
        gY[:, P] = gA
        g_lam = np.zeros(L)
        g_lam[P] = gs / (2 * sP)
        H = _eig_mid(lam, Y.conj().T @ gY, g_lam)
        gT = 2 * (Y @ H @ Y.conj().T) @ T
    else:
        lam, Y = np.linalg.eigh(T.conj().T @ T)
        lam, Y = lam[::-1], Y[:, ::-1].copy()
        YP = (T.conj().T @ A) / sP[None, :]
        Y[:, P] = YP
        gT = (gA / sP[None, :]) @ YP.conj().T
        gYP = T.conj().T @ gA / sP[None, :]
        gsP = gs - np.real(np.einsum("ij,ij->j", A.conj(), gA)) / sP
        gY = np.zeros((R, R), np.complex128)
        gY[:, P] = gYP
        g_lam = np.zeros(R)
        g_lam[P] = gsP / (2 * sP)
        H = _eig_mid(lam, Y.conj().T @ gY, g_lam)
        gT = gT + 2 * T @ (Y @ H @ Y.conj().T)
    gC = gC + gT * LRc[None, :]
    gLRc = np.real(np.einsum("ij,ij->j", gT.conj(), C))
    return gC, gLRc


# --------------------------------------------------- contraction adjoint --

def contract_vjp(d, U3, CL, CC, CR, Glc, Gcr, sec, gCs):
    """Adjoint of mpo.contract_sectors for the sectors `sec` (same order)
    and their C gradients gCs.  Returns (gGlc, gGcr, gU3)."""
    bl, bc, br = (mpo.charge_blocks(x, d) for x in (CL, CC, CR))
    ic = {q: np.arange(*bc[q]) for q in bc}
    gGlc = np.zeros_like(Glc, dtype=np.complex128)
    gGcr = np.zeros_like(Gcr, dtype=np.complex128)
    gU3 = np.zeros_like(U3, dtype=np.complex128)
    for s, gC in zip(sec, gCs):
        t = s.tau
        ro = 0
        for a in sorted(bl):
            if not (a[0] >= t[0] and a[1] >= t[1]):
                continue
            A = np.arange(*bl[a])
            co = 0
            for b in sorted(br):
                if not (b[0] <= t[0] and b[1] <= t[1]):
                    continue
                Bx = np.arange(*br[b])
                blk = gC[ro:ro + A.size, co:co + Bx.size]
                for c in ic:
                    if not (a[0] >= c[0] >= b[0] and a[1] >= c[1] >= b[1]):
                        continue
                    i0 = (a[0] - t[0], t[0] - b[0], a[0] - c[0])
                    i1 = (a[1] - t[1], t[1] - b[1], a[1] - c[1])
                    x, y = U3[i0], U3[i1]
                    u = x * np.conj(y)
                    Gl = Glc[np.ix_(A, ic[c])]
                    Gr = Gcr[np.ix_(ic[c], Bx)]
                    gGlc[np.ix_(A, ic[c])] += np.conj(u) * (blk @ Gr.conj().T)
                    gGcr[np.ix_(ic[c], Bx)] += np.conj(u) * (Gl.conj().T @ blk)
                    gu = np.sum(np.conj(Gl @ Gr) * blk)
                    gU3[i0] += gu * y
                    gU3[i1] += np.conj(gu) * x
                co += Bx.size
            ro += A.size
    return gGlc, gGcr, gU3


def site_update_vjp(d, chi, U3, LR, CL, CC, CR, Glc, Gcr, gG0, gG1, gLam):
    """Adjoint of mpo.site_update (no edge cut, no Lambda filter): gradients
    of the new Gamma[mode] (gG0), Gamma[mode+1] (gG1) and Lambda[mode]
    (gLam) to (gGlc, gGcr, gLR, gU3)."""
    sec, kept = site_update_tape(d, chi, U3, LR, CL, CC, CR, Glc, Gcr)
    LR = np.asarray(LR, np.float64)
    by_sec = {}
    for j, (si, p) in enumerate(kept):
        by_sec.setdefault(si, []).append((p, j))
    gCs = []
    gLR = np.zeros_like(LR)
    for si, s in enumerate(sec):
        L, R = s.C.shape
        if si not in by_sec:
            gCs.append(np.zeros((L, R), np.complex128))
            continue
        P = np.array([p for p, _ in by_sec[si]])
        slots = np.array([j for _, j in by_sec[si]])
        gA = gG0[np.ix_(s.rows, slots)]
        gB = gG1[np.ix_(slots, s.cols)]
        gs = np.asarray(gLam)[slots]
        gC, gLRc = sector_vjp(s.C, LR[s.cols], s.V, s.S, P, gA, gB, gs)
        gCs.append(gC)
        gLR[s.cols] += gLRc
    gGlc, gGcr, gU3 = contract_vjp(d, U3, CL, CC, CR, Glc, Gcr, sec, gCs)
    return gGlc, gGcr, gLR, gU3


# ------------------------------------------------------ BS table adjoint --

def bs_table_vjp(d, B, gU3):
    """Adjoint of fock.bs_table: U3 is a holomorphic polynomial in the four
    entries of B, so g_B = sum gU3 * conj(dU3 / dB)."""
    B = np.asarray(B, np.complex128)
    b = [B[0, 0], B[0, 1], B[1, 0], B[1, 1]]
    gB = np.zeros(4, np.complex128)
    for n1 in range(d):
        for n2 in range(d):
            s = n1 + n2
            for p in range(d):
                q = s - p
                if q < 0 or q >= d:
                    continue
                nrm = math.sqrt(_fact(n1) * _fact(n2) / (_fact(p) * _fact(q)))
                g = gU3[n1, n2, p]
                for k in range(max(0, n1 - q), min(p, n1) + 1):
                    c = _comb(p, k) * _comb(q, n1 - k) * nrm
                    e = [k, n1 - k, p - k, q - n1 + k]          # powers of b00, b01, b10, b11
                    for v in range(4):
                        if e[v] == 0:
                            continue
                        term = c * e[v] * b[v] ** (e[v] - 1)
                        for w in range(4):
                            if w != v:
                                term *= b[w] ** e[w]
                        gB[v] += g * np.conj(term)
    return gB.reshape(2, 2)


def clements_vjp(theta, phi, gB):
    """(dL/dtheta, dL/dphi) for B = [[e^{i phi} c, -s], [e^{i phi} s, c]]."""
    c, s, e = math.cos(theta), math.sin(theta), np.exp(1j * phi)
    dth = np.array([[-e * s, -c], [e * c, -s]])
    dph = np.array([[1j * e * c, 0], [1j * e * s, 0]])
    return (float(np.real(np.sum(np.conj(gB) * dth))),
            float(np.real(np.sum(np.conj(gB) * dph))))


# ---------------------------------------------------- observable adjoint --

def observable_loss(state, pattern, weights):
    """L = -log P(pattern) + sum_j weights[j] * <n_j>, with the gradient of
    L with respect to every Gamma[b] (dense [M, chi, chi]).  P and <n_j> are
    Re of chain products over the ket = bra diagonal (mpo.pattern_probability,
    mpo.marginals); each is linear in every Gamma[b], so the gradient is the
    outer product of the left and right environments around site b."""
    M, chi, d = state.M, state.chi, state.d
    masks = [mpo._diag_mask(state, b) for b in range(M + 1)]
    occ = [state.charge[b][:, 0][:, None] - state.charge[b + 1][:, 0][None, :] for b in range(M)]
    keep = [(occ[b] == int(pattern[b])) for b in range(M)]
    w = np.asarray(weights, np.float64)
    # pattern chain
    lp = [None] * (M + 1)
    lp[0] = masks[0].astype(np.complex128)
    for b in range(M):
        lp[b + 1] = (lp[b] @ (state.Gamma[b] * keep[b])) * masks[b + 1]
    rp = [None] * (M + 1)
    rp[M] = np.zeros(chi, np.complex128)
    rp[M][0] = 1.0
    rp[M] = rp[M] * masks[M]
    for b in range(M - 1, -1, -1):
        rp[b] = ((state.Gamma[b] * keep[b]) @ rp[b + 1]) * masks[b]
    zP = np.sum(lp[M] * rp[M])
    P = float(np.real(zP))
    # plain chain and weighted-occupation chain
    l0 = [None] * (M + 1)
    lw = [None] * (M + 1)
    l0[0] = masks[0].astype(np.complex128)
    lw[0] = np.zeros(chi, np.complex128)
    for b in range(M):
        l0[b + 1] = (l0[b] @ state.Gamma[b]) * masks[b + 1]
        lw[b + 1] = (lw[b] @ state.Gamma[b] + w[b] * (l0[b] @ (state.Gamma[b] * occ[b]))) * masks[b + 1]
    r0 = [None] * (M + 1)
    rw = [None] * (M + 1)
    r0[M] = rp[M].copy()
    rw[M] = np.zeros(chi, np.complex128)
    for b in range(M - 1, -1, -1):
        r0[b] = (state.Gamma[b] @ r0[b + 1]) * masks[b]
        rw[b] = (state.Gamma[b] @ rw[b + 1] + w[b] * ((state.Gamma[b] * occ[b]) @ r0[b + 1])) * masks[b]
    nw = float(np.real(np.sum(lw[M] * r0[M])))
    loss = -math.log(P) + nw
    g = np.zeros((M, chi, chi), np.complex128)
    for b in range(M):
        dzP = np.outer(lp[b], rp[b + 1]) * keep[b]
        dzw = (np.outer(lw[b], r0[b + 1]) + np.outer(l0[b], rw[b + 1])
               + w[b] * np.outer(l0[b], r0[b + 1]) * occ[b])
        g[b] = np.conj(-dzP / P + dzw)
    return loss, P, nw, g


# ------------------------------------------------------------ whole run --

def mesh_from_params(M, params):
    """mesh[l, i] = clements_bs_matrix(theta, phi) from params[l, i] =
    (theta, phi); slots beyond a layer's last pair stay the identity."""
    from bosonsampling_amd.mesh import clements_bs_matrix
    mesh = np.zeros((M, M // 2, 2, 2), np.complex128)
    mesh[:, :] = np.eye(2)
    for lay in range(M):
        for i, _ in enumerate(range(lay % 2, M - 1, 2)):
            mesh[lay, i] = clements_bs_matrix(*params[lay, i])
    return mesh


def loss_and_grad(state0, params, inputs, pattern, weights, checkpoint_every=None):
    """Loss of the run from state0 through the mesh given by params
    [M, M//2, 2] = (theta, phi), and its gradient with respect to params.
    checkpoint_every: keep a state copy every that many layers and recompute
    the layers in between during the backward sweep (None: keep every
    update's input)."""
    from oracle import fock
    M, d, chi = state0.M, state0.d, state0.chi
    mesh = mesh_from_params(M, params)
    max_input = max(inputs)
    k = checkpoint_every or 1
    ckpt = {}
    st = state0.copy()
    for layer in range(M):
        if layer % k == 0:
            ckpt[layer] = st.copy()
        for i, mode in mpo.layer_modes(M, layer, max_input):
            mpo.apply_update(st, mode, fock.bs_table(d, mesh[layer, i]))
    loss, P, nw, gG = observable_loss(st, pattern, weights)
    gL = np.zeros_like(st.Lambda)
    grad = np.zeros_like(params)
    for seg in sorted(ckpt, reverse=True):
        # recompute the segment, keeping each update's input
        st = ckpt[seg].copy()
        tape = []
        for layer in range(seg, min(seg + k, M)):
            for i, mode in mpo.layer_modes(M, layer, max_input):
                U3 = fock.bs_table(d, mesh[layer, i])
                LR = st.edge if mode == M - 2 else st.Lambda[mode + 1]
                tape.append((layer, i, mode, U3, LR.copy(), st.charge[mode:mode + 3].copy(),
                             st.Gamma[mode].copy(), st.Gamma[mode + 1].copy()))
                mpo.apply_update(st, mode, U3)
        for layer, i, mode, U3, LR, ch, G0, G1 in reversed(tape):
            gGlc, gGcr, gLR, gU3 = site_update_vjp(d, chi, U3, LR, ch[0], ch[1], ch[2], G0, G1,
                                                   gG[mode], gG[mode + 1], gL[mode])
            gG[mode], gG[mode + 1] = gGlc, gGcr
            gL[mode] = 0.0
            if mode < M - 2:
                gL[mode + 1] += gLR
            gB = bs_table_vjp(d, mesh[layer, i], gU3)
            grad[layer, i] = clements_vjp(*params[layer, i], gB)
    return loss, grad
