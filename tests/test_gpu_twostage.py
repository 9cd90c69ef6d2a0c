"""GPU parity of the two-stage reduction of the large Gram matrices
(bsmpo_set_two_stage: dense -> band of width 16 on the FP64 MFMA, band ->
tridiagonal by bulge chasing, back-transformation through both sets of
reflectors).  The site update it serves is src/beam_splitter_parallel/
Rank.py:122-127 (per-sector SVD, W = V^H C) with the top-chi truncation of
Rank.py:139-145; the oracle is oracle.mpo.site_update (numpy FP64, LAPACK).

- a synthetic d=13, chi=512 update whose sectors span n = 64 .. ~300 (every
  residue of n mod 16, panels with fewer than 16 reflectors at the end),
  against the oracle and against the one-stage reduction on the same input;
- the same input scaled by 1e-50 (Gram entries ~1e-200: the reduction must
  work on the exactly rescaled matrix);
- one update of layer 16 of the M=64, d=13, chi=1024 workload against the
  oracle on the same input;
- the synthetic chi=1024 update against its committed fixture;
- grouped complex128 against sequential plain runs of two M=64, chi=1024
  instances, both with the two-stage reduction.
Charges bit-exact; singular values, gauge-invariant projections, trace and
marginals to 1e-10 relative.
"""
import os

import numpy as np
import pytest

from oracle import fock, mpo

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-10
TS_MIN = 64


@pytest.fixture
def two_stage():
    from bosonsampling_amd._capi import check, lib
    prev = lib().bsmpo_two_stage_threshold()
    check(lib().bsmpo_set_two_stage(TS_MIN), "set_two_stage")
    assert lib().bsmpo_two_stage_threshold() == TS_MIN
    yield TS_MIN
    check(lib().bsmpo_set_two_stage(prev), "set_two_stage")


def _set_two_stage(n):
    from bosonsampling_amd._capi import check, lib
    check(lib().bsmpo_set_two_stage(n), "set_two_stage")


def _three_site(d, chi, CL, CC, CR, Glc, Gcr, LR, gpu, scale=1.0):
    from bosonsampling_amd import engine
    G = np.zeros((3, chi, chi), np.complex128)
    G[0], G[1] = Glc * scale, Gcr * scale
    Lam = np.zeros((2, chi))
    Lam[1] = LR
    charge = np.full((4, chi, 2), d, np.int32)
    charge[0], charge[1], charge[2] = CL, CC, CR
    return engine.DeviceMPO.from_host(G, Lam, charge, d, gpu)


def _sector_sizes(st, mode=0):
    sec = st.plan(mode).sectors()[0]
    return np.minimum(sec[:, 2], sec[:, 3])


def _synthetic(chi, seed=11):
    from synth import synthetic_update
    return synthetic_update(13, chi, (4, 12), (2, 10), (0, 8), seed)


def _run(d, chi, inp, U3, gpu, scale=1.0):
    st = _three_site(d, chi, *inp, gpu, scale=scale)
    st.update(0, U3)
    return st.to_host()


@pytest.mark.timeout(600)
def test_two_stage_matches_oracle_and_one_stage(gpu, two_stage):
    import torch
    from golden.make_golden import synth_bs
    d, chi = 13, 512
    inp = _synthetic(chi)
    U3 = torch.from_numpy(synth_bs(d)).to(gpu)
    n = _sector_sizes(_three_site(d, chi, *inp, gpu))
    ts = n[(n >= TS_MIN) & (n <= 528)]
    print(f"chi={chi}: {len(n)} sectors, {len(ts)} two-stage, n {ts.min()}..{ts.max()}, "
          f"{len(set(ts % 16))} residues mod 16", flush=True)
    assert len(ts) >= 20 and len(set(ts % 16)) == 16
    G2, L2, ch2 = _run(d, chi, inp, U3, gpu)
    _set_two_stage(0)
    G1, L1, ch1 = _run(d, chi, inp, U3, gpu)
    _set_two_stage(TS_MIN)
    CL, CC, CR, Glc, Gcr, LR = inp
    ch, lam, G0o, G1o, num = mpo.site_update(d, chi, synth_bs(d), LR, CL, CC, CR, Glc, Gcr)
    assert np.array_equal(ch2[1], ch) and np.array_equal(ch1[1], ch)
    assert np.abs(L2[0] - lam).max() <= TOL * lam.max()
    assert np.abs(L2[0] - L1[0]).max() <= TOL * lam.max()
    P_ref = G0o @ G1o
    for G in (G2, G1):
        P = G[0] @ G[1]
        assert np.abs(P - P_ref).max() <= TOL * np.abs(P_ref).max()


@pytest.mark.timeout(600)
def test_two_stage_is_scale_invariant(gpu, two_stage):
    """Every sector on the tridiagonal route (threshold 0; the Jacobi route
    works on the unscaled Gram matrix in LDS), the large ones two-stage."""
    import torch
    from bosonsampling_amd._capi import check, lib
    from golden.make_golden import synth_bs
    d, chi, c = 13, 512, 1e-50
    inp = _synthetic(chi, seed=12)
    U3 = torch.from_numpy(synth_bs(d)).to(gpu)
    check(lib().bsmpo_set_eig_threshold(0), "set_eig_threshold")
    try:
        G1, L1, ch1 = _run(d, chi, inp, U3, gpu)
        Gc, Lc, chc = _run(d, chi, inp, U3, gpu, scale=c)
    finally:
        check(lib().bsmpo_set_eig_threshold(96), "set_eig_threshold")
    assert np.array_equal(ch1[1], chc[1])
    assert np.all(np.isfinite(Lc[0])) and np.all(np.isfinite(Gc[0]))
    assert np.abs(Lc[0] / c ** 2 - L1[0]).max() <= TOL * L1[0].max()
    P1, Pc = G1[0] @ G1[1], Gc[0] @ Gc[1]
    assert np.abs(Pc / c ** 2 - P1).max() <= TOL * np.abs(P1).max()


@pytest.mark.timeout(600)
def test_two_stage_bench_update_matches_oracle(gpu, two_stage):
    from bosonsampling_amd import engine
    from test_gpu_scale import _bench_state
    M, d, chi, layer = 64, 13, 1024, 16
    print("bench state to layer 16", flush=True)
    st, mm, inputs = _bench_state(gpu, chi, layer)
    best = None
    for i, mode in engine.layer_modes(M, layer, max(inputs)):
        n = _sector_sizes(st, mode)
        k = int(((n >= TS_MIN) & (n <= 528)).sum())
        if best is None or k > best[0]:
            best = (k, i, mode)
    k, i, mode = best
    print(f"update mode {mode}: {k} two-stage sectors", flush=True)
    assert k >= 10
    G, L, ch = st.to_host()
    host = mpo.MPOState(d, chi, G, L, ch, np.ones(chi))
    st.update(mode, engine.bs_tables(d, mm[layer, i], gpu)[0])
    G2, L2, ch2 = st.to_host()
    mpo.apply_update(host, mode, fock.bs_table(d, mm[layer, i]))
    assert np.array_equal(ch2[mode + 1], host.charge[mode + 1])
    lam = host.Lambda[mode]
    assert np.abs(L2[mode] - lam).max() <= TOL * lam.max()
    P = G2[mode] @ G2[mode + 1]
    P_ref = host.Gamma[mode] @ host.Gamma[mode + 1]
    assert np.abs(P - P_ref).max() <= TOL * np.abs(P_ref).max()


@pytest.mark.timeout(600)
def test_two_stage_synthetic_chi1024_matches_fixture(gpu, two_stage):
    import torch
    from golden.make_golden import SYNTH_CASES, SYNTH_SEED, synth_bs, synth_projectors
    from synth import synthetic_update
    z = np.load(os.path.join(GOLDEN, "synth_d13.npz"))
    d, chi = 13, 1024
    inp = synthetic_update(d, chi, *SYNTH_CASES[chi], SYNTH_SEED)
    U3 = torch.from_numpy(synth_bs(d)).to(gpu)
    G2, L2, ch2 = _run(d, chi, inp, U3, gpu)
    assert np.array_equal(ch2[1], z[f"charge_{chi}"])
    lam = z[f"lambda_{chi}"]
    assert np.abs(L2[0] - lam).max() <= TOL * lam.max()
    X, Y = synth_projectors(chi)
    proj = (X.conj().T @ G2[0]) @ (G2[1] @ Y)
    ref = z[f"proj_{chi}"]
    assert np.abs(proj - ref).max() <= TOL * np.abs(ref).max()


@pytest.mark.timeout(900)
def test_two_stage_grouped_equals_sequential_m64(gpu, two_stage):
    from bosonsampling_amd import engine, mesh
    M, d, chi, inputs = 64, 13, 1024, list(range(0, 60, 5))
    meshes = [mesh.haar_clements_mesh(M, s)[0] for s in (100, 101)]
    init = engine.initialize(M, d, chi, inputs, 0.6, device=gpu)
    seq = [init.clone() for _ in meshes]
    for st, mm in zip(seq, meshes):
        print("sequential instance", flush=True)
        engine.simulate(st, mm, inputs, layers=range(18), observe=False)
    grp = [init.clone() for _ in meshes]
    print("grouped instances", flush=True)
    engine.simulate_batch(grp, meshes, inputs, layers=range(18), precision=engine.C128)
    for a, b in zip(seq, grp):
        ta, ma = a.observables()
        tb, mb = b.observables()
        assert np.array_equal(a.charge.cpu().numpy(), b.charge.cpu().numpy())
        assert abs(ta - tb) <= TOL * abs(ta)
        assert np.abs(ma - mb).max() <= TOL * abs(ta)
