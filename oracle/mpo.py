"""FP64 restatement of the reference's MPO state, site update, layer schedule
and observables.

State layout (reference: src/beam_splitter_parallel/Full.py:56-63):
  Gamma  [M, chi, chi]   complex128   Gamma[i] = Lambda_{i-1} . Gamma_i
  Lambda [M-1, chi]      float64      Lambda[i] lives on bond i+1
  charge [M+1, chi, 2]   int32        bond charges (ket, bra photons to the
                                      right of the bond), lexicographically
                                      sorted, unused slots padded with d.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import fock


# --------------------------------------------------------------------------
# charge-block bookkeeping (the role of src/mpo_sort.py:215-259)
# --------------------------------------------------------------------------

def charge_blocks(charges: np.ndarray, d: int):
    """Contiguous blocks of a lexicographically sorted charge array.

    Returns a dict (c0, c1) -> (start, stop) for valid charges (< d).
    Raises if the array is not sorted (src/mpo_sort.py relies on it too).
    """
    ch = np.asarray(charges)
    key = ch[:, 0].astype(np.int64) * (d + 1) + ch[:, 1]
    if np.any(np.diff(key) < 0):
        raise ValueError("charges are not lexicographically sorted")
    blocks = {}
    n = ch.shape[0]
    i = 0
    while i < n:
        c0, c1 = int(ch[i, 0]), int(ch[i, 1])
        j = i
        while j < n and ch[j, 0] == c0 and ch[j, 1] == c1:
            j += 1
        if c0 < d and c1 < d:
            blocks[(c0, c1)] = (i, j)
        i = j
    return blocks


def sector_range(d: int, CL: np.ndarray, CR: np.ndarray):
    """tau0, tau1 loop bounds, src/beam_splitter_parallel/Rank.py:78-91."""
    v0 = CL[:, 0] != d
    v1 = CL[:, 1] != d
    lc0 = int(CL[v0, 0].max()) if v0.any() else 0
    lc1 = int(CL[v1, 1].max()) if v1.any() else 0
    sr0 = int(CR[:, 0].min())
    sr1 = int(CR[:, 1].min())
    return range(sr0, lc0 + 1), range(sr1, lc1 + 1)


def bs_weight(U3: np.ndarray, cl, ch, cr, tau) -> complex:
    """u(cl, ch, cr; tau), src/update_MPO_kernel.cu:441-456,491-497 and
    src/MPO.py:254: U3[cl0-t0, t0-cr0, cl0-ch0] * conj(U3[cl1-t1, t1-cr1, cl1-ch1])
    if cl >= ch >= cr componentwise, else 0."""
    if not (cl[0] >= ch[0] >= cr[0] and cl[1] >= ch[1] >= cr[1]):
        return 0.0
    a = U3[cl[0] - tau[0], tau[0] - cr[0], cl[0] - ch[0]]
    b = U3[cl[1] - tau[1], tau[1] - cr[1], cl[1] - ch[1]]
    return a * np.conj(b)


# --------------------------------------------------------------------------
# the site update (src/beam_splitter_parallel/Rank.py:32-206)
# --------------------------------------------------------------------------

@dataclass
class SectorResult:
    tau: tuple
    rows: np.ndarray      # indices into the left bond
    cols: np.ndarray      # indices into the right bond
    C: np.ndarray         # L x R contraction (no LR weight)
    V: np.ndarray = None
    S: np.ndarray = None
    W: np.ndarray = None


def contract_sectors(d, U3, CL, CC, CR, Glc, Gcr, LL=None, LC=None, LR=None):
    """theta-sector contractions C_tau for every tau, as a list of SectorResult.

    C_tau[l, r] = sum_c Glc[l, c] Gcr[c, r] u(cl, cc, cr; tau)  with rows
    cl >= tau, cols cr <= tau (src/mpo_sort.py:424-432 windows,
    src/update_MPO_kernel.cu:441-456,491-497,596-597).  Computed per
    (cl-block, cr-block) pair as sum_ch A[tau, ch] (Glc[cl,ch] @ Gcr[ch,cr]),
    which is the same sum regrouped.  If LL/LC/LR are given, rows/centres/cols
    whose Lambda is not > 0 are dropped (CPU-twin semantics, src/MPO.py:182-189).
    """
    bl = charge_blocks(CL, d)
    bc = charge_blocks(CC, d)
    br = charge_blocks(CR, d)

    def idx_of(blocks, lam):
        out = {}
        for q, (s, e) in blocks.items():
            ix = np.arange(s, e)
            if lam is not None:
                ix = ix[np.asarray(lam)[s:e] > 0]
            if ix.size:
                out[q] = ix
        return out

    il, ic, ir = idx_of(bl, LL), idx_of(bc, LC), idx_of(br, LR)
    t0r, t1r = sector_range(d, CL, CR)
    taus = [(t0, t1) for t0 in t0r for t1 in t1r]
    # rows / cols per tau in sorted (original) order
    lkeys = sorted(il)
    rkeys = sorted(ir)
    sectors = {}
    row_off = {}
    col_off = {}
    for tau in taus:
        lsel = [q for q in lkeys if q[0] >= tau[0] and q[1] >= tau[1]]
        rsel = [q for q in rkeys if q[0] <= tau[0] and q[1] <= tau[1]]
        if not lsel or not rsel:
            continue
        rows = np.concatenate([il[q] for q in lsel])
        cols = np.concatenate([ir[q] for q in rsel])
        offs, o = {}, 0
        for q in lsel:
            offs[q] = o
            o += il[q].size
        row_off[tau] = offs
        offs, o = {}, 0
        for q in rsel:
            offs[q] = o
            o += ir[q].size
        col_off[tau] = offs
        sectors[tau] = SectorResult(tau, rows, cols,
                                    np.zeros((rows.size, cols.size), np.complex128))
    if not sectors:
        return []
    for a in lkeys:
        A = il[a]
        for b in rkeys:
            if not (a[0] >= b[0] and a[1] >= b[1]):
                continue
            B = ir[b]
            chs = [c for c in ic if a[0] >= c[0] >= b[0] and a[1] >= c[1] >= b[1]]
            ts = [t for t in sectors if a[0] >= t[0] >= b[0] and a[1] >= t[1] >= b[1]]
            if not chs or not ts:
                continue
            X = np.stack([Glc[np.ix_(A, ic[c])] @ Gcr[np.ix_(ic[c], B)] for c in chs])
            G = np.array([[bs_weight(U3, a, c, b, t) for c in chs] for t in ts])
            Y = np.tensordot(G, X, axes=(1, 0))
            for k, t in enumerate(ts):
                ro, co = row_off[t][a], col_off[t][b]
                sectors[t].C[ro:ro + A.size, co:co + B.size] += Y[k]
    return [sectors[t] for t in taus if t in sectors]


def contract_sector_literal(d, U3, CL, CC, CR, Glc, Gcr, tau):
    """Element-by-element C_tau (pure loops): for tiny cases only, to pin
    ``contract_sectors`` to the kernel's per-element weight definition
    (src/update_MPO_kernel.cu:441-456,596-597)."""
    rows = [l for l in range(CL.shape[0]) if CL[l, 0] < d and CL[l, 1] < d
            and CL[l, 0] >= tau[0] and CL[l, 1] >= tau[1]]
    cols = [r for r in range(CR.shape[0]) if CR[r, 0] <= tau[0] and CR[r, 1] <= tau[1]]
    C = np.zeros((len(rows), len(cols)), np.complex128)
    for i, l in enumerate(rows):
        for j, r in enumerate(cols):
            acc = 0.0
            for c in range(CC.shape[0]):
                if CC[c, 0] >= d or CC[c, 1] >= d:
                    continue
                u = bs_weight(U3, CL[l], CC[c], CR[r], tau)
                acc += Glc[l, c] * Gcr[c, r] * u
            C[i, j] = acc
    return np.array(rows), np.array(cols), C


# Values within TIE_REL (relative) of the chi-th largest singular value are
# ties: exactly degenerate twins are the rule here, because the density
# matrix is Hermitian and sectors (a, b) and (b, a) have equal singular
# values; which twin survives a cut between them is decided by rounding
# noise in any implementation (argpartition on float32 in the reference,
# Rank.py:142).  Ties are broken by the lower concatenated index, on the CPU
# and on the GPU alike.
TIE_REL = 1e-6


def select_top(svals_per_sector, chi, tie_rel=None):
    """Global top-chi over all sectors' singular values (Rank.py:139-145).

    Deterministic stand-in for argpartition: keep num = min(total, chi)
    values: every value above the chi-th largest t by more than tie_rel * t,
    then the ties (|value - t| <= tie_rel * t) in concatenated order.  The
    kept ones are ordered by (sector order, position in sector) -- the stable
    lexsort by charge of Rank.py:182-196 with a defined within-sector order.
    Returns a list of (sector, position).
    """
    tie_rel = TIE_REL if tie_rel is None else tie_rel
    flat = []
    for s, S in enumerate(svals_per_sector):
        for p, v in enumerate(S):
            flat.append((float(v), s, p))
    num = min(len(flat), chi)
    if num == 0:
        return []
    t = sorted((f[0] for f in flat), reverse=True)[num - 1]
    hi, lo = t * (1 + tie_rel), t * (1 - tie_rel)
    above = [i for i, f in enumerate(flat) if f[0] > hi]
    band = [i for i, f in enumerate(flat) if lo <= f[0] <= hi]
    chosen = above + band[:num - len(above)]
    return sorted((flat[i][1], flat[i][2]) for i in chosen)


def _svd_left(T):
    """Left singular vectors and values of T (LAPACK gesdd, gesvd fallback)."""
    try:
        V, S, _ = np.linalg.svd(T, full_matrices=False)
    except np.linalg.LinAlgError:
        import scipy.linalg
        V, S, _ = scipy.linalg.svd(T, full_matrices=False, lapack_driver="gesvd")
    return V, S


# Singular values at or below DROP_REL * (largest of their sector) are
# treated as exact zeros and dropped with their vectors.  The GPU path gets
# its singular values from the Gram matrix, whose noise floor in sigma is
# ~sqrt(n eps) sigma_max; below it LAPACK's null-space completion and the
# Gram route's vectors are both arbitrary, and they carry no weight in any
# observable.  The reference keeps them (Rank.py:139-145).
DROP_REL = 1e-6


def site_update(d, chi, U3, LR, CL, CC, CR, Glc, Gcr, *, LL=None, LC=None,
                filter_lambda=False, drop_rel=DROP_REL):
    """One beam-splitter site update, RankProcess semantics
    (src/beam_splitter_parallel/Rank.py:32-206), complex128 throughout.

    Returns (new_charge[chi, 2] padded with d, new_lambda[chi] padded 0,
    Gamma0Out[chi, chi], Gamma1Out[chi, chi], num_lambda).
    """
    sec = contract_sectors(d, U3, CL, CC, CR, Glc, Gcr,
                           LL=LL if filter_lambda else None,
                           LC=LC if filter_lambda else None,
                           LR=LR if filter_lambda else None)
    for s in sec:
        T = s.C * np.asarray(LR, np.float64)[s.cols][None, :]     # Rank.py:115-116
        V, S = _svd_left(T)                                        # Rank.py:123
        keep = S > drop_rel * S[0] if (S.size and S[0] > 0) else np.zeros(S.size, bool)
        V, S = V[:, keep], S[keep]
        s.V, s.S = V, S
        s.W = V.conj().T @ s.C                                    # Rank.py:124
    kept = select_top([s.S for s in sec], chi)
    num = len(kept)
    new_charge = np.full((chi, 2), d, np.int32)
    new_lambda = np.zeros(chi, np.float64)
    G0 = np.zeros((Glc.shape[0], chi), np.complex128)
    G1 = np.zeros((chi, Gcr.shape[1]), np.complex128)
    for j, (si, p) in enumerate(kept):
        s = sec[si]
        new_charge[j] = s.tau
        new_lambda[j] = s.S[p]
        G0[s.rows, j] = s.V[:, p]
        G1[j, s.cols] = s.W[p, :]
    return new_charge, new_lambda, G0, G1, num


# --------------------------------------------------------------------------
# MPO state, initialisation and canonicalisation (src/MPO.py:21-156)
# --------------------------------------------------------------------------

@dataclass
class MPOState:
    d: int
    chi: int
    Gamma: np.ndarray
    Lambda: np.ndarray
    charge: np.ndarray
    edge: np.ndarray = field(default=None)

    @property
    def M(self):
        return self.Gamma.shape[0]

    def copy(self):
        return MPOState(self.d, self.chi, self.Gamma.copy(), self.Lambda.copy(),
                        self.charge.copy(), self.edge.copy())


def input_density(d: int, eta: float) -> np.ndarray:
    """Single photon through a pure-loss channel of transmission eta:
    rho = (1 - eta)|0><0| + eta|1><1| (src/MPO.py:41-44 with loss = 1 - eta)."""
    rho = np.zeros((d, d), np.float64)
    rho[0, 0] = 1.0 - eta
    rho[1, 1] = eta
    return rho


def raw_product_mpo(M, d, chi, input_modes, eta, post_selected=None):
    """Product-state MPO before canonicalisation, generalising
    src/MPO.py:50-109 to any set of occupied modes and any chi >= N + 1.

    Bond i carries the diagonal charges (c, c), c = 0..(photons in sites >= i);
    Gamma_i[(a,a), (b,b)] = rho_i[a-b, a-b]; Lambda = 1 on used slots.
    """
    inputs = sorted(set(int(m) for m in input_modes))
    rho_in = input_density(d, eta)
    rho_vac = np.zeros((d, d))
    rho_vac[0, 0] = 1.0
    n_right = [sum(1 for m in inputs if m >= i) for i in range(M + 1)]
    charge = np.full((M + 1, chi, 2), d, np.int32)
    slots = []
    for i in range(M + 1):
        if i == 0:
            cs = [post_selected] if post_selected is not None else list(range(min(n_right[0], d - 1) + 1))
        elif i == M:
            cs = [0]
        else:
            cs = list(range(min(n_right[i], d - 1) + 1))
        if len(cs) > chi:
            raise ValueError("chi too small for the product state")
        for k, c in enumerate(cs):
            charge[i, k] = (c, c)
        slots.append(cs)
    Gamma = np.zeros((M, chi, chi), np.complex128)
    for i in range(M):
        rho = rho_in if i in inputs else rho_vac
        for a_i, a in enumerate(slots[i]):
            for b_i, b in enumerate(slots[i + 1]):
                n = a - b
                if 0 <= n < d:
                    Gamma[i, a_i, b_i] = rho[n, n]
    Lambda = np.zeros((M - 1, chi), np.float64)
    for i in range(1, M):
        Lambda[i - 1, :len(slots[i])] = 1.0
    edge = np.ones(chi, np.float64)
    return MPOState(d, chi, Gamma, Lambda, charge, edge)


def apply_update(state: MPOState, mode: int, U3: np.ndarray, filter_lambda=False):
    """Update the pair (mode, mode+1) in place, with the master-side merge of
    src/beam_splitter_parallel/Full.py:113-137 (last-pair edge storage)."""
    M, d, chi = state.M, state.d, state.chi
    LL = state.edge if mode == 0 else state.Lambda[mode - 1]
    LC = state.Lambda[mode]
    LR = state.edge if mode == M - 2 else state.Lambda[mode + 1]
    ch, lam, G0, G1, num = site_update(
        d, chi, U3, LR, state.charge[mode], state.charge[mode + 1], state.charge[mode + 2],
        state.Gamma[mode], state.Gamma[mode + 1], LL=LL, LC=LC, filter_lambda=filter_lambda)
    state.charge[mode + 1] = ch
    state.Lambda[mode] = lam
    if mode == M - 2:
        w = min(chi, d * d)
        state.Gamma[M - 2][:, :w] = G0[:, :w]
        state.Gamma[M - 1][:w, 0] = G1[:w, 0]
    else:
        state.Gamma[mode] = G0
        state.Gamma[mode + 1] = G1
    return num


def canonicalise(state: MPOState) -> MPOState:
    """Right-to-left sweep of identity-phase updates (src/MPO.py:137-156),
    U3 = Rand_BS_MPS(0, d, 0)."""
    U3 = fock.canonicalisation_bs(state.d)
    for mode in range(state.M - 2, -1, -1):
        apply_update(state, mode, U3, filter_lambda=True)
    return state


def initialize(M, d, chi, input_modes, eta, post_selected=None) -> MPOState:
    """Initial MPO (src/MPO.py:21-134 generalised, src/beam_splitter_parallel/Full.py:44-63)."""
    st = raw_product_mpo(M, d, chi, input_modes, eta, post_selected)
    return canonicalise(st)


# --------------------------------------------------------------------------
# simulation driver (src/beam_splitter_parallel/Full.py:149-202)
# --------------------------------------------------------------------------

def layer_modes(M: int, layer: int, max_input: int):
    """(position, mode) pairs of a brick-wall layer that are not skipped by the
    light cone (Full.py:152-153: skip mode >= n_input_states + layer, i.e.
    mode > last occupied input mode + layer)."""
    out = []
    for i, mode in enumerate(range(layer % 2, M - 1, 2)):
        if mode > max_input + layer:
            continue
        out.append((i, mode))
    return out


def simulate(state: MPOState, mesh: np.ndarray, input_modes, layers=None, callback=None):
    """Apply the mesh layer by layer; returns per-layer trace and EE
    (Full.py:169-202: prob[M+1], EE[M-1, M+1])."""
    M, d = state.M, state.d
    max_input = max(input_modes)
    prob = np.zeros(M + 1)
    EE = np.zeros((M - 1, M + 1))
    prob[0] = trace(state)
    EE[:, 0] = entanglement_entropy(state.Lambda)
    for layer in (range(M) if layers is None else layers):
        for i, mode in layer_modes(M, layer, max_input):
            U3 = fock.bs_table(d, mesh[layer, i])
            apply_update(state, mode, U3)
            if callback is not None:
                callback(layer, i, mode, state)
        prob[layer + 1] = trace(state)
        EE[:, layer + 1] = entanglement_entropy(state.Lambda)
    return prob, EE


# --------------------------------------------------------------------------
# observables (src/summary.py:3-36) and marginals
# --------------------------------------------------------------------------

def _diag_mask(state: MPOState, bond: int) -> np.ndarray:
    ch = state.charge[bond]
    m = (ch[:, 0] == ch[:, 1]) & (ch[:, 0] < state.d)
    if 0 < bond < state.M:
        m &= state.Lambda[bond - 1] > 0
    return m


def right_envs(state: MPOState):
    """R[b] = vector over bond b slots: contraction of sites b..M-1 over the
    ket = bra diagonal (summary.py:11-21 chain, masked by Lambda > 0)."""
    M, chi = state.M, state.chi
    R = [None] * (M + 1)
    r = np.zeros(chi, np.complex128)
    r[0] = 1.0
    R[M] = r
    for b in range(M - 1, -1, -1):
        r = state.Gamma[b] @ (R[b + 1] * _diag_mask(state, b + 1))
        R[b] = r
    return R


def left_envs(state: MPOState):
    M, chi = state.M, state.chi
    L = [None] * (M + 1)
    L[0] = _diag_mask(state, 0).astype(np.complex128)
    for b in range(1, M + 1):
        L[b] = (L[b - 1] @ state.Gamma[b - 1]) * _diag_mask(state, b)
    return L


def trace(state: MPOState) -> float:
    """Total probability, src/summary.py:11-26 (complex value's real part)."""
    R = right_envs(state)
    return float(np.real(np.sum(R[0] * _diag_mask(state, 0))))


def marginals(state: MPOState) -> np.ndarray:
    """P[j, n] = probability of n photons in mode j (unnormalised: sums to the
    trace), from left/right environments over the ket = bra diagonal."""
    M, d = state.M, state.d
    L, R = left_envs(state), right_envs(state)
    P = np.zeros((M, d))
    for j in range(M):
        cl = state.charge[j][:, 0]
        cr = state.charge[j + 1][:, 0]
        lv = L[j]
        rv = R[j + 1] * _diag_mask(state, j + 1)
        occ = cl[:, None] - cr[None, :]
        E = lv[:, None] * state.Gamma[j] * rv[None, :]
        for n in range(d):
            P[j, n] = float(np.real(E[occ == n].sum()))
    return P


def pattern_probability(state: MPOState, pattern) -> float:
    """P(output photon-number pattern): the diagonal chain with the local
    occupation of site j restricted to pattern[j]."""
    M, chi = state.M, state.chi
    r = np.zeros(chi, np.complex128)
    r[0] = 1.0
    for b in range(M - 1, -1, -1):
        cl = state.charge[b][:, 0]
        cr = state.charge[b + 1][:, 0]
        keep = (cl[:, None] - cr[None, :]) == int(pattern[b])
        r = (state.Gamma[b] * keep) @ (r * _diag_mask(state, b + 1))
    return float(np.real(np.sum(r * _diag_mask(state, 0))))


def entanglement_entropy(Lambda: np.ndarray) -> np.ndarray:
    """src/summary.py:3-8,28-30: S = -sum p log2 p, p = Lambda^2 / sum Lambda^2.
    Zero entries add nothing (nansum over 0 log 0) and an all-zero bond gives
    0 (the reference's nansum over the 0 / 0 column)."""
    out = np.zeros(Lambda.shape[0])
    sq = np.asarray(Lambda, np.float64) ** 2
    for i in range(Lambda.shape[0]):
        tot = sq[i].sum()
        if tot <= 0:
            continue
        p = sq[i] / tot
        with np.errstate(divide="ignore", invalid="ignore"):
            out[i] = -np.nansum(p * np.log2(p))
    return out
