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


def test_launching_streams_agree(gpu):
    """Groups spread over two launching streams (several groups per layer,
    the first one capped) give the state of the one-stream run."""
    from bosonsampling_amd import engine
    M, d, chi, inputs, eta = CASE
    seeds = list(range(30, 36))
    meshes = _meshes(M, seeds)
    ref = mpo.initialize(M, d, chi, inputs, eta)
    runs = {}
    for ns in (1, 2):
        print(f"launching streams {ns}, {len(seeds)} instances", flush=True)
        sts = [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in seeds]
        runner = engine.GroupRunner(max_bytes=2 << 20, streams=ns)
        engine.simulate_batch(sts, meshes, inputs, runner=runner)
        assert runner.last_info["groups"] >= 3
        runs[ns] = sts
    for x, y in zip(runs[1], runs[2]):
        tx, mx = x.observables()
        ty, my = y.observables()
        assert abs(tx - ty) <= 1e-12 * abs(tx)
        assert np.abs(mx - my).max() <= 1e-12
        assert all(np.array_equal(a, b) for a, b in zip(x.charge_host, y.charge_host))


def test_tridiagonal_block_shapes_agree(gpu):
    """A many-sector group on the tridiagonal route: the sector list (sorted
    largest first) split between the 512-thread Householder blocks, launched
    first, and 256-thread blocks on a side stream gives the state of the run
    with one block shape.  Split 512 puts every sector of this case on the
    256-thread shape; split 8 sends the sectors with n > 8 to the 512-thread
    shape.  The shapes differ only in their reduction order."""
    from bosonsampling_amd import engine
    from bosonsampling_amd._capi import check, lib
    M, d, chi, inputs, eta = CASE
    seeds = list(range(20, 28))
    meshes = _meshes(M, seeds)
    ref = mpo.initialize(M, d, chi, inputs, eta)
    runs = {}
    check(lib().bsmpo_set_eig_threshold(0), "set_eig_threshold")
    try:
        for split in (512, 8):
            check(lib().bsmpo_set_tri_split(split), "set_tri_split")
            print(f"tridiagonal block-shape split {split}, {len(seeds)} instances", flush=True)
            sts = [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in seeds]
            engine.simulate_batch(sts, meshes, inputs, runner=engine.GroupRunner(streams=1))
            runs[split] = sts
    finally:
        check(lib().bsmpo_set_tri_split(256), "set_tri_split")
        check(lib().bsmpo_set_eig_threshold(96), "set_eig_threshold")
    # the last layer's group has many sectors on both sides of the split
    layer = M - 1
    sizes = []
    for st in runs[8]:
        ch = st.charge_host
        for _, mode in engine.layer_modes(M, layer, max(inputs)):
            sec = engine.Plan(d, chi, ch[mode], ch[mode + 1], ch[mode + 2]).sectors()[0]
            sizes += list(np.minimum(sec[:, 2], sec[:, 3]))
    sizes = np.array(sizes)
    assert len(sizes) >= 256 and (sizes > 8).sum() > 0 and (sizes <= 8).sum() > 0
    for x, y in zip(runs[512], runs[8]):
        tx, mx = x.observables()
        ty, my = y.observables()
        assert abs(tx - ty) <= 1e-10 * abs(tx)
        assert np.abs(mx - my).max() <= 1e-10
        assert all(np.array_equal(a, b) for a, b in zip(x.charge_host, y.charge_host))


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


def test_group_graph_replays_match_eager(gpu):
    """One layer's group captured as a HIP graph and replayed three times on
    restored inputs, with eager work and synchronisation in between; every
    replay equals the eager group run."""
    import torch
    from bosonsampling_amd import _capi, engine
    from bosonsampling_amd._capi import check, lib
    M, d, chi, inputs, eta = CASE
    meshes = _meshes(M, [6, 7])
    ref = mpo.initialize(M, d, chi, inputs, eta)
    mpo.simulate(ref, meshes[0], inputs, layers=range(2))
    layer = 2
    todo = engine.layer_modes(M, layer, max(inputs))

    def fresh():
        return [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in meshes]

    def updates(states):
        out = []
        for st, mm in zip(states, meshes):
            U3 = engine.bs_tables(d, np.stack([mm[layer, i] for i, _ in todo]), st.device)
            out += [(st, mode, U3[k]) for k, (i, mode) in enumerate(todo)]
        return out

    eager = fresh()
    engine.GroupRunner().run(updates(eager))
    want = [st.to_host() for st in eager]

    states = fresh()
    saved = [(st.Gamma.clone(), st.Lambda.clone(), st.charge.clone()) for st in states]
    ups = updates(states)
    plans = []
    for st, mode, U3 in ups:
        ch = st.charge_host
        plans.append(engine.Plan(d, chi, ch[mode], ch[mode + 1], ch[mode + 2]))
    n = len(plans)
    arr = (C.c_void_p * n)(*[p.address for p in plans])
    hb, wb = C.c_size_t(), C.c_size_t()
    check(lib().bsmpo_group_query(n, arr, C.byref(hb), C.byref(wb)), "group_query")
    work = torch.empty(wb.value, dtype=torch.uint8, device=gpu)
    host = torch.empty(hb.value, dtype=torch.uint8, pin_memory=True)
    io = (_capi.UpdateIO * n)()
    for j, (st, mode, U3) in enumerate(ups):
        LR = st.edge if mode == M - 2 else st.Lambda[mode + 1]
        io[j] = _capi.UpdateIO(st.Gamma[mode].data_ptr(), st.Gamma[mode + 1].data_ptr(),
                               LR.data_ptr(), U3.data_ptr(), st.charge[mode + 1].data_ptr(),
                               st.Lambda[mode].data_ptr(), st.num_lambda[mode].data_ptr(),
                               1 if mode == M - 2 else 0, 0)
    check(lib().bsmpo_group_build(n, arr, io, C.c_void_p(work.data_ptr()),
                                  C.c_void_p(host.data_ptr()), hb.value), "group_build")
    side = torch.cuda.Stream(device=gpu)
    side.wait_stream(torch.cuda.current_stream())
    ex = C.c_void_p()
    check(lib().bsmpo_group_graph_create(C.c_void_p(host.data_ptr()), 0, C.c_void_p(work.data_ptr()),
                                         wb.value, C.c_void_p(side.cuda_stream), C.byref(ex)),
          "group_graph_create")
    try:
        for rep in range(3):
            print(f"graph replay {rep}", flush=True)
            for st, (G, L, ch) in zip(states, saved):
                st.Gamma.copy_(G)
                st.Lambda.copy_(L)
                st.charge.copy_(ch)
            side.wait_stream(torch.cuda.current_stream())
            check(lib().bsmpo_graph_launch(ex, C.c_void_p(side.cuda_stream)), "graph_launch")
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            for st, (Gw, Lw, chw) in zip(states, want):
                G, L, ch = st.to_host()
                assert np.array_equal(ch, chw)
                assert np.abs(L - Lw).max() <= 1e-12 * np.abs(Lw).max()
                assert np.abs(G - Gw).max() <= 1e-12 * np.abs(Gw).max()
            _ = engine.GroupRunner().run(updates(fresh()))  # eager work between replays
    finally:
        check(lib().bsmpo_graph_destroy(ex), "graph_destroy")


def test_group_runner_with_graphs_matches_oracle(gpu):
    from bosonsampling_amd import engine
    M, d, chi, inputs, eta = CASE
    meshes = _meshes(M, [8, 9])
    ref = mpo.initialize(M, d, chi, inputs, eta)
    batch = [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in meshes]
    engine.simulate_batch(batch, meshes, inputs, runner=engine.GroupRunner(use_graph=True))
    for st, mm in zip(batch, meshes):
        orc = ref.copy()
        mpo.simulate(orc, mm, inputs)
        mo = mpo.marginals(orc)
        assert np.abs(st.observables()[1] - mo).max() <= 1e-10 * np.abs(mo).max()
