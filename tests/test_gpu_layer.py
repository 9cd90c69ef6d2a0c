"""GPU tests of layer programs: brick-wall layers planned on the device and
replayed as one captured HIP graph per layer (bsmpo_layer_*), against the
host-planned grouped path on the same states, and the FP64 oracle.
"""
import numpy as np
import pytest

from oracle import mpo

pytestmark = pytest.mark.gpu

CASE = (10, 6, 48, [0, 2, 4, 6, 8], 0.7)


def _meshes(M, seeds):
    from bosonsampling_amd import mesh
    return [mesh.haar_clements_mesh(M, s)[0] for s in seeds]


def _max_diff(a, b):
    ga, la, ca = a.to_host()
    gb, lb, cb = b.to_host()
    return (float(np.abs(ga - gb).max()), float(np.abs(la - lb).max()),
            bool(np.array_equal(ca, cb)))


def _same_truncation(a, b):
    """Two device states compared by truncation.same_truncation."""
    from truncation import same_truncation
    _, la, ca = a.to_host()
    _, lb, cb = b.to_host()
    return same_truncation(la, ca, lb, cb)


def test_layer_program_matches_grouped_and_oracle(gpu):
    from bosonsampling_amd import engine
    M, d, chi, inputs, eta = CASE
    seeds = [11, 12, 13]
    meshes = _meshes(M, seeds)
    ref = mpo.initialize(M, d, chi, inputs, eta)
    grouped = [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in seeds]
    graph = [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in seeds]
    print(f"layer program case M={M} d={d} chi={chi} instances={len(seeds)}", flush=True)
    engine.simulate_batch(grouped, meshes, inputs)
    runner = engine.LayerGraphRunner(graph, meshes, engine.C128, n_groups=2)
    n = runner.run(range(M))
    assert n == sum(len(runner.updates[x % 2]) for x in range(M))
    # capacities sized on layer 0 (a product state) are outgrown as the bond
    # dimension builds up: run() falls back and re-plans as needed
    print(f"fallbacks {runner.fallbacks}", flush=True)
    st, needs = runner.state()
    assert st.status == 0 and st.replays == M
    assert all(nd["status"] == 0 for nd in needs)
    for k, mm in enumerate(meshes):
        dg, dl, same_ch = _max_diff(grouped[k], graph[k])
        print(f"instance {k}: |dGamma| {dg:.3e} |dLambda| {dl:.3e} charges equal {same_ch}", flush=True)
        assert same_ch and dg <= 1e-13 and dl <= 1e-13
        orc = ref.copy()
        mpo.simulate(orc, mm, inputs)
        tb, mb = graph[k].observables()
        mo = mpo.marginals(orc)
        assert abs(tb - mpo.trace(orc)) <= 1e-10 * abs(mpo.trace(orc))
        assert np.abs(mb - mo).max() <= 1e-10 * np.abs(mo).max()
        assert np.array_equal(graph[k].charge_host, orc.charge)


def test_layer_program_over_capacity_changes_nothing_then_falls_back(gpu):
    """Capacities below what the first layer needs: that replay and the later
    ones leave the state as it was; run() grows the capacities and finishes."""
    from bosonsampling_amd import engine
    M, d, chi, inputs, eta = CASE
    seeds = [21, 22]
    meshes = _meshes(M, seeds)
    ref = mpo.initialize(M, d, chi, inputs, eta)
    grouped = [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in seeds]
    graph = [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in seeds]
    engine.simulate_batch(grouped, meshes, inputs, layers=range(2))
    engine.simulate_batch(graph, meshes, inputs, layers=range(2))
    before = [s.clone() for s in graph]
    runner = engine.LayerGraphRunner(graph, meshes, engine.C128, n_groups=2, headroom=1.0)
    w = runner.needs(0)
    runner._set_caps(w)
    runner.caps.work_bytes = max(1, w["work_bytes"] // 2)  # the first replay cannot fit
    runner._create()
    # first, the failing replays alone: nothing may change
    import ctypes as C
    import torch
    from bosonsampling_amd import _capi
    lib = _capi.lib()
    s = C.c_void_p(runner.stream.cuda_stream)
    _capi.check(lib.bsmpo_layer_seek(runner.prog, 2, s), "seek")
    for _ in range(3):
        _capi.check(lib.bsmpo_layer_launch(runner.prog, s), "launch")
    st, needs = runner.state()
    print(f"status {st.status} failed_at {st.failed_at} replays {st.replays}", flush=True)
    assert st.status == 1 and st.failed_at == 2 and st.replays == 5
    assert max(nd["work_bytes"] for nd in needs) > runner.caps.work_bytes
    for a, b in zip(before, graph):
        dg, dl, same = _max_diff(a, b)
        assert dg == 0.0 and dl == 0.0 and same
    torch.cuda.synchronize()
    # then run(): the first replay fails again, run() grows and finishes
    runner.run(range(2, M))
    assert runner.fallbacks >= 1
    engine.simulate_batch(grouped, meshes, inputs, layers=range(2, M))
    for a, b in zip(grouped, graph):
        dg, dl, same = _max_diff(a, b)
        assert same and dg <= 1e-13 and dl <= 1e-13


def test_layer_program_bench_scale(gpu):
    """M=64, N=12, d=13, chi=1024 (the benchmark's workload, two instances):
    layers 16-23 as device-planned graph replays against host-planned groups,
    complex128 throughout.  Charges and singular values as _same_truncation
    holds them (measured: charges bit-exact); trace and marginals to 1e-10
    relative: the Householder reduction's row sums are LDS atomics, so two
    runs of the tridiagonal route round differently (and singular vectors of
    near-degenerate values differ by a gauge rotation), whichever planner.
    (With a complex64 contraction such 1e-16 differences flip single FP32
    roundings of later layers, and two runs agree to ~1e-7 only.)"""
    from bosonsampling_amd import engine
    M, d, chi, eta = 64, 13, 1024, 0.6
    inputs = list(range(0, 60, 5))
    seeds = [100, 101]
    meshes = _meshes(M, seeds)
    try:
        states = [engine.initialize(M, d, chi, inputs, eta) for _ in seeds]
        print("bench-scale layer program: preparing layers 0-15", flush=True)
        engine.simulate_batch(states, meshes, inputs, layers=range(16))
        graph = [s.clone() for s in states]
        print("host-planned groups, layers 16-23", flush=True)
        engine.simulate_batch(states, meshes, inputs, layers=range(16, 24))
        print("device-planned graph replays, layers 16-23", flush=True)
        runner = engine.LayerGraphRunner(graph, meshes, engine.C128, n_groups=2)
        runner.run(range(16, 24))
        print(f"fallbacks {runner.fallbacks}", flush=True)
        for a, b in zip(states, graph):
            dg, dl, same = _max_diff(a, b)
            ta, ma = a.observables()
            tb, mb = b.observables()
            print(f"|dLambda| {dl:.3e} |dtrace| {abs(ta - tb):.3e} "
                  f"|dmarg| {np.abs(ma - mb).max():.3e} charges equal {same}", flush=True)
            _same_truncation(a, b)
            assert abs(ta - tb) <= 1e-10 * abs(ta)
            assert np.abs(ma - mb).max() <= 1e-10 * abs(ta)
    finally:
        engine.set_precision_rules(engine.C128)


def test_layer_autotune_and_geometry_knobs(gpu):
    """The autotuner's knobs (route split, Householder block-shape split,
    contraction tile order) change launch geometry only: every setting gives
    the state of the default one (tile order: bit-exact on the Jacobi route;
    the tridiagonal route to rounding), and autotune_layer sets the values it
    returns."""
    from bosonsampling_amd import engine
    M, d, chi, inputs, eta = CASE
    seeds = [31, 32]
    meshes = _meshes(M, seeds)
    ref = mpo.initialize(M, d, chi, inputs, eta)
    base = [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in seeds]
    engine.simulate_batch(base, meshes, inputs, layers=range(3))
    saved = {n: engine._knob(n) for n, _ in engine.LAYER_KNOBS}
    try:
        results = {}
        for thr, split, order in ((96, 256, 8), (96, 256, 2), (96, 256, 16), (0, 256, 8),
                                  (0, 384, 4), (0, 512, 16)):
            engine._knob("eig_threshold", thr)
            engine._knob("tri_split", split)
            engine._knob("tile_order", order)
            sts = [s.clone() for s in base]
            runner = engine.LayerGraphRunner(sts, meshes, engine.C128, n_groups=2)
            runner.run(range(3, M))
            results[(thr, split, order)] = sts
            print(f"knobs {thr} {split} {order}: fallbacks {runner.fallbacks}", flush=True)
        for key, sts in results.items():
            ref_key = (96, 256, 8) if key[0] == 96 else (0, 256, 8)
            for a, b in zip(results[ref_key], sts):
                dg, dl, same = _max_diff(a, b)
                ta, ma = a.observables()
                tb, mb = b.observables()
                if key[0] == 96:  # Jacobi route: deterministic, identical work per sector
                    assert same and dg == 0.0 and dl == 0.0
                else:
                    _same_truncation(a, b)
                assert abs(ta - tb) <= 1e-10 * abs(ta) and np.abs(ma - mb).max() <= 1e-10 * abs(ta)
        best, table = engine.autotune_layer(base, meshes, 3, precision=engine.C128, n_groups=2,
                                            repeats=1)
        print(f"autotuned {best}", flush=True)
        for name, cands in engine.LAYER_KNOBS:
            assert best[name] in cands and engine._knob(name) == best[name]
            assert set(table[name]) == set(cands)
    finally:
        for n, v in saved.items():
            engine._knob(n, v)


@pytest.mark.timeout(600)
def test_layer_program_chi2048(gpu):
    """The largest bond dimension a layer program plans: M=64, d=13,
    chi=2048 (seed 100), layers 14-16 as graph replays against host-planned
    groups from the same state, complex128.  Singular values to 1e-10 of each
    bond's largest, charges equal wherever both keep as many values
    (_same_truncation; measured: charges bit-exact in some runs, in others one
    bond where one run keeps a value of 1e-13 x the bond's largest at its
    sector's drop threshold, the LDS-atomic order of the row sums differing
    from run to run); trace and marginals to 1e-10 relative."""
    from bosonsampling_amd import engine
    M, d, chi = 64, 13, 2048
    inputs = list(range(0, 60, 5))
    meshes = _meshes(M, [100])
    st = engine.initialize(M, d, chi, inputs, 0.6)
    print("chi=2048 layer program: preparing layers 0-13", flush=True)
    engine.simulate_batch([st], meshes, inputs, layers=range(14))
    graph = st.clone()
    engine.simulate_batch([st], meshes, inputs, layers=range(14, 17))
    print("graph replays, layers 14-16", flush=True)
    runner = engine.LayerGraphRunner([graph], meshes, engine.C128, n_groups=2)
    runner.run(range(14, 17))
    dg, dl, same = _max_diff(st, graph)
    ta, ma = st.observables()
    tb, mb = graph.observables()
    flips = _same_truncation(st, graph)
    print(f"fallbacks {runner.fallbacks} |dLambda| {dl:.3e} |dtrace| {abs(ta - tb):.3e} "
          f"|dmarg| {np.abs(ma - mb).max():.3e} charges equal {same}, drop-threshold flips "
          f"at bonds {flips}", flush=True)
    assert abs(ta - tb) <= 1e-10 * abs(ta) and np.abs(ma - mb).max() <= 1e-10 * abs(ta)


def test_layer_program_light_cone_matches_golden(gpu):
    """The M=8 golden run (photons in modes 0-3, d=5, chi=16, 28 beam
    splitters; the first layers skip the pairs beyond the light cone) as
    layer-program replays of two copies, every layer observed: trace per
    layer and final marginals to 1e-10 of the oracle's fixture."""
    import os
    from bosonsampling_amd import engine
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "m8_sp4.npz"))
    inputs = [0, 1, 2, 3]
    states = [engine.DeviceMPO.from_host(g["init_Gamma"], g["init_Lambda"], g["init_charge"], 5)
              for _ in range(2)]
    meshes = [g["mesh"], g["mesh"]]
    runner = engine.LayerGraphRunner(states, meshes, engine.C128, n_groups=1, n_streams=1,
                                     input_modes=inputs)
    n, probs = engine.simulate_batch(states, meshes, inputs, runner=runner, observe=True)
    assert n == 2 * sum(len(engine.layer_modes(8, l, 3)) for l in range(8))
    for k in range(2):
        assert np.abs(np.array(probs[k]) - g["prob"][1:]).max() <= 1e-10
        tr, marg = states[k].observables()
        assert np.abs(marg - g["marginals"]).max() <= 1e-10


@pytest.mark.parametrize("prec_name", ["c64", "bf16x3"])
def test_layer_program_low_precision_matches_grouped(gpu, prec_name):
    """The reduced-precision contractions as layer programs: the state of the
    host-planned grouped run with the same precision and truncation rules,
    bit-exact (Jacobi route: every stage deterministic)."""
    from bosonsampling_amd import engine
    prec = {"c64": engine.C64, "bf16x3": engine.BF16X3}[prec_name]
    M, d, chi, inputs, eta = CASE
    seeds = [41, 42]
    meshes = _meshes(M, seeds)
    ref = mpo.initialize(M, d, chi, inputs, eta)
    engine.set_precision_rules(prec)
    try:
        grouped = [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in seeds]
        graph = [engine.DeviceMPO.from_host(ref.Gamma, ref.Lambda, ref.charge, d) for _ in seeds]
        engine.simulate_batch(grouped, meshes, inputs, precision=prec)
        runner = engine.LayerGraphRunner(graph, meshes, prec, n_groups=2, input_modes=inputs)
        engine.simulate_batch(graph, meshes, inputs, runner=runner)
        for a, b in zip(grouped, graph):
            dg, dl, same = _max_diff(a, b)
            print(f"{prec_name}: |dGamma| {dg:.3e} |dLambda| {dl:.3e} charges equal {same}", flush=True)
            assert same and dg == 0.0 and dl == 0.0
    finally:
        engine.set_precision_rules(engine.C128)
