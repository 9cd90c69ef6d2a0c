"""GPU tests of the drop-in launchers (bosonsampling_amd.cuda_kernels), which
keep the reference's names and arguments (src/cuda_kernels.py), against the
FP64 oracle."""
import math

import numpy as np
import pytest

from oracle import fock, mpo

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.mark.parametrize("r", [0.0, 0.37, 0.9])
def test_rand_u_matches_reference_formula(gpu, r):
    """Same numpy RNG draws (alpha for t, beta for r; r == 0 draws beta only)
    and the reference's binomial formula (src/MPO.py:11-17)."""
    from bosonsampling_amd.cuda_kernels import Rand_U
    d, seed = 9, 1234
    np.random.seed(seed)
    U_r, U_i = Rand_U(d, r)
    got = U_r.cpu().numpy() + 1j * U_i.cpu().numpy()
    np.random.seed(seed)
    if r == 0:
        t, rr = 1.0, 0.0 * np.exp(1j * np.random.rand() * 2 * math.pi)
    else:
        t = math.sqrt(1 - r * r) * np.exp(1j * np.random.rand() * 2 * math.pi)
        rr = r * np.exp(1j * np.random.rand() * 2 * math.pi)
    want = fock.bs_table_reference(d, t, rr)
    assert np.abs(got - want).max() <= 1e-12
    # the RNG advanced by the same number of draws
    np.random.seed(seed)
    np.random.rand(1 if r == 0 else 2)
    nxt = np.random.rand()
    np.random.seed(seed)
    Rand_U(d, r)
    assert np.random.rand() == nxt


def _state():
    from bosonsampling_amd import mesh
    M, d, chi, inputs = 8, 5, 40, [0, 2, 4, 6]
    mm, _ = mesh.haar_clements_mesh(M, 7)
    st = mpo.initialize(M, d, chi, inputs, 0.7)
    mpo.simulate(st, mm, inputs, layers=range(4))
    return st, mm


def _change_table(CC, d):
    """The reference's centre change table (src/mpo_sort.py:57-61): the
    first index of every charge present in CC (including the end padding)."""
    starts = [0] + [i for i in range(1, CC.shape[0]) if (CC[i] != CC[i - 1]).any()]
    chc = np.array([[CC[i, 0] for i in starts], [CC[i, 1] for i in starts]], np.int32)
    return chc, np.array(starts, np.int32)


def _aligned(ch, G, axis, gaps):
    """Insert padding rows (charge -1, zero data) before the given positions,
    as the reference's Aligner does between charge blocks."""
    ch2, idx = [], []
    for i in range(ch.shape[0]):
        if i in gaps:
            ch2.append([-1, -1])
            idx.append(-1)
        ch2.append(list(ch[i]))
        idx.append(i)
    ch2 = np.array(ch2, np.int32)
    idx = np.array(idx)
    shape = list(G.shape)
    shape[axis] = idx.size
    G2 = np.zeros(shape, G.dtype)
    take = idx >= 0
    if axis == 0:
        G2[take] = G[idx[take]]
    else:
        G2[:, take] = G[:, idx[take]]
    return ch2, G2, idx


def test_update_mpo_every_sector(gpu):
    from bosonsampling_amd.cuda_kernels import Data, update_MPO
    st, mm = _state()
    d, mode = st.d, 2
    CL, CC, CR = st.charge[mode], st.charge[mode + 1], st.charge[mode + 2]
    Glc, Gcr = st.Gamma[mode], st.Gamma[mode + 1]
    U3 = fock.bs_table(d, mm[4, 1])
    ref = mpo.contract_sectors(d, U3, CL, CC, CR, Glc, Gcr)
    assert ref
    # aligned inputs with padding rows inside the left and right bonds
    CLa, Glca, il = _aligned(CL, Glc, 0, {3, 11, 20})
    CRa, Gcra, ir = _aligned(CR, Gcr, 1, {0, 7, 15})
    chc, chi_idx = _change_table(CC, d)
    print(f"update_MPO d={d} sectors={len(ref)} rows={CLa.shape[0]} cols={CRa.shape[0]}", flush=True)
    for s in ref:
        out = update_MPO(d, s.tau[0], s.tau[1], U3.real, U3.imag,
                         Data("Glc", True, Glca, [np.arange(CLa.shape[0]), np.arange(CC.shape[0])]),
                         Data("Gcr", True, Gcra, [np.arange(CC.shape[0]), np.arange(CRa.shape[0])]),
                         Data("CL", True, CLa, [0]), Data("CR", True, CRa, [0]), chc, chi_idx)
        got = out.data.cpu().numpy()
        want = np.zeros_like(got)
        pos_l = {v: k for k, v in enumerate(il) if v >= 0}
        pos_r = {v: k for k, v in enumerate(ir) if v >= 0}
        rl = np.array([pos_l[i] for i in s.rows])
        rr = np.array([pos_r[i] for i in s.cols])
        want[np.ix_(rl, rr)] = s.C
        scale = max(np.abs(s.C).max(), 1e-300)
        assert np.abs(got - want).max() <= TOL * scale, s.tau


def test_update_mps_pure_state(gpu):
    """One U(1) charge per bond: the MPO weights with a zero second charge.
    Unsorted charges on every bond; lambdas strictly positive."""
    from bosonsampling_amd.cuda_kernels import update_MPS
    rng = np.random.default_rng(5)
    d, nl, nc, nr = 6, 23, 31, 19
    CL = rng.integers(0, d, nl).astype(np.int32)
    CC = rng.integers(0, d, nc).astype(np.int32)
    CR = rng.integers(0, d, nr).astype(np.int32)
    Glc = rng.standard_normal((nl, nc)) + 1j * rng.standard_normal((nl, nc))
    Gcr = rng.standard_normal((nc, nr)) + 1j * rng.standard_normal((nc, nr))
    LL, LC, LR = (rng.uniform(0.1, 1.0, n) for n in (nl, nc, nr))
    B = fock.reference_bs_matrix(0.6 * np.exp(0.3j), 0.8 * np.exp(1.1j))
    U3 = fock.bs_table(d, B)
    # oracle on the sorted problem, one charge as (c, 0)
    ol, oc, orr = (np.argsort(x, kind="stable") for x in (CL, CC, CR))
    pair = lambda x: np.stack([x, np.zeros_like(x)], 1)
    ref = mpo.contract_sectors(d, U3, pair(CL[ol]), pair(CC[oc]), pair(CR[orr]),
                               Glc[ol][:, oc] * LC[oc][None, :], Gcr[oc][:, orr])
    print(f"update_MPS d={d} sectors={len(ref)}", flush=True)
    for s in ref:
        got = update_MPS(d, s.tau[0], U3.real, U3.imag, Glc, Gcr, LL, LC, LR, CL, CC, CR,
                         None).cpu().numpy()
        want = np.zeros((nl, nr), np.complex128)
        rows, cols = ol[s.rows], orr[s.cols]
        want[np.ix_(rows, cols)] = LL[rows][:, None] * s.C * LR[cols][None, :]
        assert np.abs(got - want).max() <= TOL * np.abs(want).max(), s.tau
