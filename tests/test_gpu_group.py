"""GPU tests of the grouped (one launch per stage) path and of the reduced-
precision sector contractions, against the per-update path and the FP64 oracle.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import fock, mpo

pytestmark = pytest.mark.gpu

CASE = (10, 6, 48, [0, 2, 4, 6, 8], 0.7)


def _meshes(M, seeds):
    from bosonsampling_amd import mesh
    return [mesh.haar_clements_mesh(M, s)[0] for s in seeds]


def test_group_matches_sequential_and_oracle(gpu):
    from bosonsampling_amd import engine
    M, d, chi, inputs, eta = CASE
    seeds = [1, 2, 3]
    print(f"group case M={M} d={d} chi={chi} instances={len(seeds)}", flush=True)
    meshes = _meshes(M, seeds)
    ref = mpo.initialize(M, d, chi, inputs, eta)
    batch = [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in seeds]
    seq = [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in seeds]
    n, _ = engine.simulate_batch(batch, meshes, inputs)
    assert n > 0
    for k, mm in enumerate(meshes):
        engine.simulate(seq[k], mm, inputs, observe=False)
        orc = ref.copy()
        mpo.simulate(orc, mm, inputs)
        tb, mb = batch[k].observables()
        ts, ms = seq[k].observables()
        mo = mpo.marginals(orc)
        assert abs(tb - ts) <= 1e-12 and np.abs(mb - ms).max() <= 1e-12
        assert abs(tb - mpo.trace(orc)) <= 1e-10 * abs(mpo.trace(orc))
        assert np.abs(mb - mo).max() <= 1e-10 * np.abs(mo).max()
        _, _, ch_b = batch[k].to_host()
        assert np.array_equal(ch_b, orc.charge)


def test_group_single_update_chunks(gpu):
    """A group split into several workspaces gives the same result."""
    from bosonsampling_amd import engine
    M, d, chi, inputs, eta = CASE
    meshes = _meshes(M, [4, 5])
    ref = mpo.initialize(M, d, chi, inputs, eta)
    a = [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in meshes]
    b = [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in meshes]
    engine.simulate_batch(a, meshes, inputs)
    engine.simulate_batch(b, meshes, inputs, runner=engine.GroupRunner(max_bytes=1))
    for x, y in zip(a, b):
        assert np.abs(x.observables()[1] - y.observables()[1]).max() <= 1e-12


@pytest.mark.parametrize("prec,tol", [(1, 1e-6), (2, 1e-4)], ids=["c64", "bf16x3"])
def test_lowp_contraction_error(gpu, prec, tol):
    import torch
    from bosonsampling_amd import engine, mesh
    from bosonsampling_amd._capi import check, lib
    M, d, chi, inputs, eta = CASE
    mm, _ = mesh.haar_clements_mesh(M, 1)
    st = mpo.initialize(M, d, chi, inputs, eta)
    mpo.simulate(st, mm, inputs, layers=range(5))
    mode = 4
    print(f"lowp contraction prec={prec} mode={mode}", flush=True)
    U3h = fock.bs_table(d, mm[5, 1])
    ref = mpo.contract_sectors(d, U3h, st.charge[mode], st.charge[mode + 1], st.charge[mode + 2],
                               st.Gamma[mode], st.Gamma[mode + 1])
    plan = engine.Plan(d, chi, st.charge[mode], st.charge[mode + 1], st.charge[mode + 2])
    sec, coff, rows, cols = plan.sectors()
    dev = gpu
    Glc = torch.from_numpy(st.Gamma[mode]).to(dev)
    Gcr = torch.from_numpy(st.Gamma[mode + 1]).to(dev)
    LR = torch.from_numpy(st.Lambda[mode + 1]).to(dev)
    U3 = torch.from_numpy(U3h).to(dev)
    n = int(plan.info().c_elems)
    Cd = torch.empty(n, dtype=torch.complex128, device=dev)
    Td = torch.empty(n, dtype=torch.complex128, device=dev)
    work = torch.empty(plan.work_bytes, dtype=torch.uint8, device=dev)
    check(lib().bsmpo_contract(d, chi, C.c_void_p(plan.address), C.c_void_p(Glc.data_ptr()),
                               C.c_void_p(Gcr.data_ptr()), C.c_void_p(LR.data_ptr()),
                               C.c_void_p(U3.data_ptr()), C.c_void_p(Cd.data_ptr()),
                               C.c_void_p(Td.data_ptr()), None, None, prec,
                               C.c_void_p(work.data_ptr()), work.numel(), engine._stream()),
          "contract")
    Ch = Cd.cpu().numpy()
    worst = 0.0
    for k, s in enumerate(ref):
        L, R = s.rows.size, s.cols.size
        got = Ch[coff[k]:coff[k] + L * R].reshape(L, R)
        scale = max(np.abs(s.C).max(), 1e-300)
        worst = max(worst, np.abs(got - s.C).max() / scale)
    print(f"max relative error {worst:.3e}")
    assert worst <= tol


@pytest.mark.parametrize("prec,tol,band", [(1, 1e-5, 1e-6), (2, 1e-4, 1e-4)],
                         ids=["c64", "bf16x3"])
def test_lowp_simulation_marginals(gpu, prec, tol, band):
    """Whole run with the FP32 / BF16x3 contraction (SVD, truncation and norm
    in FP64) against the FP64 oracle: marginals and trace.  The truncation
    tie band is set wider than the contraction's rounding noise, on both
    sides alike (BF16x3 errors are ~1e-5 relative)."""
    from bosonsampling_amd import engine, mesh
    from bosonsampling_amd._capi import check, lib
    M, d, chi, inputs, eta = CASE
    check(lib().bsmpo_set_tie_rel(band), "set_tie_rel")
    old = mpo.TIE_REL
    mpo.TIE_REL = band
    try:
        _lowp_run(prec, tol, M, d, chi, inputs, eta)
    finally:
        mpo.TIE_REL = old
        check(lib().bsmpo_set_tie_rel(1e-6), "set_tie_rel")


def _lowp_run(prec, tol, M, d, chi, inputs, eta):
    from bosonsampling_amd import engine, mesh
    mm, _ = mesh.haar_clements_mesh(M, 1)
    ref = mpo.initialize(M, d, chi, inputs, eta)
    dev_st = engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d)
    engine.simulate(dev_st, mm, inputs, precision=prec, observe=False)
    mpo.simulate(ref, mm, inputs)
    tr, marg = dev_st.observables()
    mo = mpo.marginals(ref)
    err_m = np.abs(marg - mo).max()
    err_t = abs(tr - mpo.trace(ref))
    print(f"prec={prec}: marginal error {err_m:.3e}, trace error {err_t:.3e}")
    assert err_m <= tol and err_t <= tol
