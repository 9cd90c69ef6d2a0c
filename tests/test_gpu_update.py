"""GPU parity of the HIP site update against the FP64 oracle.

Every comparison starts both sides from the same host state.  Bit-exact
checks are used for integer outputs (charges, kept counts); the FP64 path is
held to 1e-10 relative on gauge-invariant quantities (singular values, the
two-site product Gamma0 @ Gamma1 over the new bond, trace and marginals).
"""
import numpy as np
import pytest

from oracle import fock, mpo

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(params=[96, 0], ids=["jacobi", "tridiag"])
def eig_route(request):
    """Run the test with the default eigensolver routing (n <= 96 Jacobi in
    LDS, larger n tridiagonal) and with every sector on the tridiagonal route."""
    from bosonsampling_amd._capi import check, lib
    check(lib().bsmpo_set_eig_threshold(request.param), "set_eig_threshold")
    yield request.param
    check(lib().bsmpo_set_eig_threshold(96), "set_eig_threshold")


def _oracle_state(M, d, chi, inputs, eta, seed, layers):
    from bosonsampling_amd import mesh
    mm, _ = mesh.haar_clements_mesh(M, seed)
    st = mpo.initialize(M, d, chi, inputs, eta)
    if layers:
        mpo.simulate(st, mm, inputs, layers=range(layers))
    return st, mm


def _upload(st):
    from bosonsampling_amd.engine import DeviceMPO
    return DeviceMPO.from_host(st.Gamma, st.Lambda, st.charge, st.d)


def test_bs_table(gpu):
    from bosonsampling_amd import engine, mesh
    rng = np.random.default_rng(7)
    for d in (2, 5, 9, 13):
        Bs = [mesh.haar_unitary(2, int(s)) for s in rng.integers(0, 1000, 4)]
        Bs.append(mesh.reference_bs_matrix(*mesh.reference_phases(11, 0.3)))
        U3 = engine.bs_tables(d, np.stack(Bs), gpu).cpu().numpy()
        for k, B in enumerate(Bs):
            ref = fock.bs_table(d, B)
            assert np.abs(U3[k] - ref).max() < 1e-13 * max(1.0, np.abs(ref).max()), (d, k)


@pytest.mark.parametrize("case", [(8, 5, 16, [0, 1, 2, 3], 0.5, 0, 3, 2),
                                  (10, 6, 48, [0, 2, 4, 6, 8], 0.7, 1, 5, 4)])
def test_contract_matches_oracle(gpu, case):
    import torch
    import ctypes as C
    from bosonsampling_amd import engine
    from bosonsampling_amd._capi import check, lib
    M, d, chi, inputs, eta, seed, layers, mode = case
    st, mm = _oracle_state(M, d, chi, inputs, eta, seed, layers)
    print(f"contract case M={M} d={d} chi={chi} mode={mode}", flush=True)
    U3h = fock.bs_table(d, mm[layers, (mode - layers % 2) // 2])
    ref = mpo.contract_sectors(d, U3h, st.charge[mode], st.charge[mode + 1], st.charge[mode + 2],
                               st.Gamma[mode], st.Gamma[mode + 1])
    plan = engine.Plan(d, chi, st.charge[mode], st.charge[mode + 1], st.charge[mode + 2])
    sec, coff, rows, cols = plan.sectors()
    assert sec.shape[0] == len(ref)
    dev = gpu
    Glc = torch.from_numpy(st.Gamma[mode]).to(dev)
    Gcr = torch.from_numpy(st.Gamma[mode + 1]).to(dev)
    LR = torch.from_numpy(st.Lambda[mode + 1] if mode < M - 2 else np.ones(chi)).to(dev)
    U3 = torch.from_numpy(U3h).to(dev)
    n = int(plan.info().c_elems)
    Cd = torch.empty(n, dtype=torch.complex128, device=dev)
    Td = torch.empty(n, dtype=torch.complex128, device=dev)
    work = torch.empty(plan.work_bytes, dtype=torch.uint8, device=dev)
    check(lib().bsmpo_contract(d, chi, C.c_void_p(plan.address), C.c_void_p(Glc.data_ptr()),
                               C.c_void_p(Gcr.data_ptr()), C.c_void_p(LR.data_ptr()),
                               C.c_void_p(U3.data_ptr()), C.c_void_p(Cd.data_ptr()),
                               C.c_void_p(Td.data_ptr()), None, None, 0,
                               C.c_void_p(work.data_ptr()), work.numel(), engine._stream()),
          "contract")
    Ch = Cd.cpu().numpy()
    Th = Td.cpu().numpy()
    lr = st.Lambda[mode + 1] if mode < M - 2 else np.ones(chi)
    for k, s in enumerate(ref):
        L, R = s.rows.size, s.cols.size
        got = Ch[coff[k]:coff[k] + L * R].reshape(L, R)
        scale = max(np.abs(s.C).max(), 1e-300)
        assert np.abs(got - s.C).max() <= 1e-12 * scale, (k, s.tau)
        gotT = Th[coff[k]:coff[k] + L * R].reshape(L, R)
        assert np.abs(gotT - s.C * lr[s.cols][None, :]).max() <= 1e-12 * scale


@pytest.mark.parametrize("case", [(8, 5, 16, [0, 1, 2, 3], 0.5, 0, 3, 2),
                                  (8, 5, 16, [0, 1, 2, 3], 0.5, 0, 5, 6),
                                  (10, 6, 48, [0, 2, 4, 6, 8], 0.7, 1, 5, 4)])
def test_site_update_matches_oracle(gpu, eig_route, case):
    import torch
    from bosonsampling_amd import engine
    M, d, chi, inputs, eta, seed, layers, mode = case
    st, mm = _oracle_state(M, d, chi, inputs, eta, seed, layers)
    print(f"site_update case M={M} d={d} chi={chi} mode={mode}", flush=True)
    B = mm[layers, (mode - layers % 2) // 2]
    dev_st = _upload(st)
    U3 = engine.bs_tables(d, B, gpu)[0]
    dev_st.update(mode, U3)
    G, Lam, ch = dev_st.to_host()
    mpo.apply_update(st, mode, fock.bs_table(d, B))
    # integer outputs: bit-exact
    assert np.array_equal(ch[mode + 1], st.charge[mode + 1])
    # singular values
    assert np.allclose(Lam[mode], st.Lambda[mode], rtol=TOL, atol=TOL * st.Lambda[mode].max())
    # gauge-invariant two-site product over the new bond
    P_gpu = G[mode] @ G[mode + 1]
    P_ref = st.Gamma[mode] @ st.Gamma[mode + 1]
    assert np.abs(P_gpu - P_ref).max() <= TOL * max(np.abs(P_ref).max(), 1e-300)


def test_observables_match_oracle(gpu):
    st, _ = _oracle_state(10, 6, 48, [0, 2, 4, 6, 8], 0.7, 1, 6)
    dev_st = _upload(st)
    tr, marg = dev_st.observables()
    assert abs(tr - mpo.trace(st)) <= 1e-12
    assert np.abs(marg - mpo.marginals(st)).max() <= 1e-12
    assert np.allclose(dev_st.entanglement_entropy(), mpo.entanglement_entropy(st.Lambda),
                       atol=1e-10, equal_nan=True)


@pytest.mark.parametrize("case", [(8, 5, 16, [0, 1, 2, 3], 0.5, 0),      # reference CPU size
                                  (10, 6, 48, [0, 2, 4, 6, 8], 0.7, 1)])
def test_simulate_matches_oracle(gpu, eig_route, case):
    from bosonsampling_amd import engine, mesh
    M, d, chi, inputs, eta, seed = case
    print(f"simulate case M={M} d={d} chi={chi}", flush=True)
    mm, _ = mesh.haar_clements_mesh(M, seed)
    st = mpo.initialize(M, d, chi, inputs, eta)
    dev_st = _upload(st)
    prob_ref, EE_ref = mpo.simulate(st, mm, inputs)
    prob, EE = engine.simulate(dev_st, mm, inputs)
    assert np.allclose(prob, prob_ref, rtol=TOL, atol=0)
    tr, marg = dev_st.observables()
    ref = mpo.marginals(st)
    assert np.abs(marg - ref).max() <= TOL * np.abs(ref).max()


def test_gpu_initialize_matches_oracle(gpu, eig_route):
    from bosonsampling_amd import engine
    M, d, chi, inputs, eta = 10, 6, 48, [0, 2, 4, 6, 8], 0.7
    st = mpo.initialize(M, d, chi, inputs, eta)
    dev_st = engine.initialize(M, d, chi, inputs, eta)
    tr, marg = dev_st.observables()
    assert abs(tr - 1.0) < 1e-12
    assert np.abs(marg - mpo.marginals(st)).max() < 1e-12
    _, Lam, ch = dev_st.to_host()
    for b in range(M - 1):
        assert np.allclose(np.sort(Lam[b])[::-1], np.sort(st.Lambda[b])[::-1], atol=1e-12)


def test_tridiagonal_route_on_a_larger_mesh(gpu):
    """Every sector on the tridiagonal route (threshold 0) against the default
    routing on a mesh large enough to hold zero-weight and tiny-weight
    sectors (M=24, six photons, d=7, chi=64, all 24 layers): the eigensolver
    routes differ only by rounding, so trace, marginals and the output
    probabilities agree to 1e-9."""
    from bosonsampling_amd import engine, mesh
    from bosonsampling_amd._capi import check, lib
    M, d, chi, inputs, eta = 24, 7, 64, [0, 4, 8, 12, 16, 20], 0.6
    mm, _ = mesh.haar_clements_mesh(M, 3)
    out = {}
    for thr in (96, 0):
        check(lib().bsmpo_set_eig_threshold(thr), "set_eig_threshold")
        st = engine.initialize(M, d, chi, inputs, eta)
        prob, _ = engine.simulate(st, mm, inputs)
        tr, marg = st.observables()
        out[thr] = (prob, tr, marg)
        print(f"threshold {thr}: trace {tr:.12f}", flush=True)
    check(lib().bsmpo_set_eig_threshold(96), "set_eig_threshold")
    (p96, t96, m96), (p0, t0, m0) = out[96], out[0]
    assert np.isfinite(t0) and abs(t0 - t96) <= 1e-9
    assert np.abs(m0 - m96).max() <= 1e-9
    assert np.allclose(p0, p96, rtol=1e-9, atol=1e-14)


def test_tridiagonal_route_is_scale_invariant(gpu):
    """A site update is homogeneous: Glc, Gcr -> c Glc, c Gcr gives T -> c^2 T,
    new singular values c^2 Lambda and the same bond charges and vectors.
    With c = 1e-50 the Gram matrices hold entries ~1e-200, whose squares
    underflow, so the tridiagonal route must work on an exactly rescaled
    matrix.  Every sector takes that route (threshold 0)."""
    import torch
    from bosonsampling_amd import engine
    from bosonsampling_amd._capi import check, lib
    M, d, chi, inputs, eta, seed, layers, mode = 10, 6, 48, [0, 2, 4, 6, 8], 0.7, 1, 5, 4
    st, mm = _oracle_state(M, d, chi, inputs, eta, seed, layers)
    B = mm[layers, (mode - layers % 2) // 2]
    c = 1e-50
    check(lib().bsmpo_set_eig_threshold(0), "set_eig_threshold")
    try:
        outs = []
        for scale in (1.0, c):
            dev_st = _upload(st)
            dev_st.Gamma[mode] *= scale
            dev_st.Gamma[mode + 1] *= scale
            U3 = engine.bs_tables(d, B, gpu)[0]
            dev_st.update(mode, U3)
            outs.append(dev_st.to_host())
    finally:
        check(lib().bsmpo_set_eig_threshold(96), "set_eig_threshold")
    (G1, L1, ch1), (Gc, Lc, chc) = outs
    assert np.array_equal(ch1[mode + 1], chc[mode + 1])
    assert np.all(np.isfinite(Lc[mode])) and np.all(np.isfinite(Gc[mode]))
    assert np.allclose(Lc[mode] / c ** 2, L1[mode], rtol=1e-10, atol=1e-10 * L1[mode].max())
    # the new left site carries no Lambda factor: equal up to the gauge of each
    # singular vector, so compare the projector onto each kept column block
    P1 = G1[mode] @ G1[mode + 1]
    Pc = Gc[mode] @ Gc[mode + 1]
    assert np.abs(Pc / c ** 2 - P1).max() <= 1e-10 * np.abs(P1).max()


def test_entanglement_entropy_on_device(gpu):
    """bsmpo_entanglement_entropy against the oracle (pinned by closed forms
    in test_oracle.py) on random bonds, a one-value bond and an all-zero bond."""
    import torch
    from bosonsampling_amd import engine
    rng = np.random.default_rng(12)
    M, chi, d = 6, 300, 5
    lam = rng.random((M - 1, chi)) * (rng.random((M - 1, chi)) < 0.6)
    lam[1] = 0.0
    lam[1, 17] = 2.5
    lam[3] = 0.0
    st = engine.DeviceMPO(M, d, chi, gpu)
    st.Lambda.copy_(torch.from_numpy(lam))
    S = st.entanglement_entropy()
    ref = mpo.entanglement_entropy(lam)
    assert np.abs(S - ref).max() <= 1e-12
    assert S[1] == 0.0 and S[3] == 0.0
