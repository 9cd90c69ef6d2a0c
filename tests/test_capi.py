"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and the host planner reproduces the oracle's sector selection exactly."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import fock, mpo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "bsmpo.h")).read()
    return sorted(set(re.findall(r"\b(bsmpo_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from bosonsampling_amd import _capi
    L = _capi.lib()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing
    assert set(_declared()) == set(_capi.SIGNATURES)
    assert L.bsmpo_abi_version() >= 1


def _state(M=10, d=6, chi=48, inputs=(0, 2, 4, 6, 8), eta=0.7, seed=1, layers=5):
    from bosonsampling_amd import mesh
    mm, _ = mesh.haar_clements_mesh(M, seed)
    st = mpo.initialize(M, d, chi, list(inputs), eta)
    mpo.simulate(st, mm, list(inputs), layers=range(layers))
    return st, mm


@pytest.mark.parametrize("mode", [0, 3, 4, 8])
def test_planner_matches_oracle_sectors(mode):
    from bosonsampling_amd.engine import Plan
    st, mm = _state()
    d, chi = st.d, st.chi
    U3 = fock.bs_table(d, mm[0, 0])
    ref = mpo.contract_sectors(d, U3, st.charge[mode], st.charge[mode + 1], st.charge[mode + 2],
                               st.Gamma[mode], st.Gamma[mode + 1])
    p = Plan(d, chi, st.charge[mode], st.charge[mode + 1], st.charge[mode + 2])
    sec, coff, rows, cols = p.sectors()
    assert sec.shape[0] == len(ref)
    ro = co = 0
    for k, s in enumerate(ref):
        assert tuple(sec[k, :2]) == s.tau
        L, R = sec[k, 2], sec[k, 3]
        assert np.array_equal(rows[ro:ro + L], s.rows) and np.array_equal(cols[co:co + R], s.cols)
        ro, co = ro + L, co + R
    inf = p.info()
    assert inf.c_elems == sum(s.rows.size * s.cols.size for s in ref)
    assert inf.sv_total == sum(min(s.rows.size, s.cols.size) for s in ref)


def _alg_counts(d, CL, CC, CR):
    bl, bc, br = (mpo.charge_blocks(c, d) for c in (CL, CC, CR))
    n = lambda b, q: b[q][1] - b[q][0]
    f = sum(8 * n(bl, a) * n(bc, c) * n(br, b) for a in bl for c in bc for b in br
            if a[0] >= c[0] >= b[0] and a[1] >= c[1] >= b[1])
    return f


def test_planner_algorithmic_flops():
    from bosonsampling_amd.engine import Plan
    st, _ = _state()
    mode = 4
    p = Plan(st.d, st.chi, st.charge[mode], st.charge[mode + 1], st.charge[mode + 2])
    assert p.info().flops_alg == _alg_counts(st.d, st.charge[mode], st.charge[mode + 1],
                                             st.charge[mode + 2])
    assert p.info().flops_tau >= p.info().flops_alg


def test_single_sector_plan():
    from bosonsampling_amd.engine import Plan
    st, _ = _state()
    mode = 4
    full = Plan(st.d, st.chi, st.charge[mode], st.charge[mode + 1], st.charge[mode + 2]).sectors()[0]
    t = tuple(full[len(full) // 2, :2])
    one = Plan(st.d, st.chi, st.charge[mode], st.charge[mode + 1], st.charge[mode + 2], tau=t)
    sec = one.sectors()[0]
    assert sec.shape[0] == 1 and tuple(sec[0]) == tuple(full[len(full) // 2])


def test_planner_rejects_bad_charges():
    from bosonsampling_amd._capi import BsmpoError
    from bosonsampling_amd.engine import Plan
    d, chi = 4, 8
    ok = np.full((chi, 2), d, np.int32)
    ok[:3] = [[0, 0], [1, 1], [2, 2]]
    bad = ok.copy()
    bad[[0, 2]] = bad[[2, 0]]          # unsorted
    with pytest.raises(BsmpoError, match="sorted"):
        Plan(d, chi, bad, ok, ok)
    bad = ok.copy()
    bad[0] = [7, 0]                    # out of range
    with pytest.raises(BsmpoError, match="range"):
        Plan(d, chi, ok, bad, ok)


def test_site_update_rejects_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from bosonsampling_amd import engine
    from bosonsampling_amd._capi import BsmpoError
    with pytest.raises(BsmpoError):
        engine.DeviceMPO(4, 3, 8)


def test_launch_knobs_validate_their_range():
    """The process-wide launch knobs reject values outside their range with
    an error code and a message, and leave the setting as it was."""
    from bosonsampling_amd._capi import lib
    L = lib()
    assert L.bsmpo_set_tri_split(0) < 0 and b"split" in L.bsmpo_last_error()
    assert L.bsmpo_set_tri_split(513) < 0
    assert L.bsmpo_set_eig_threshold(97) < 0 and b"threshold" in L.bsmpo_last_error()
    assert L.bsmpo_set_tri_split(256) == 0
    assert L.bsmpo_set_eig_threshold(96) == 0


def test_drop_rule_knob_and_precision_rules():
    """The drop rule is a process-wide knob in [1e-8, 1e-3] (default 1e-6,
    the oracle's DROP_REL); a value outside it is refused and leaves the
    setting as it was.  The per-precision rules set both truncation knobs and
    are widest for split BF16x3."""
    from bosonsampling_amd import engine
    from bosonsampling_amd._capi import lib
    from oracle import mpo
    L = lib()
    assert L.bsmpo_drop_rel() == mpo.DROP_REL == 1e-6
    assert L.bsmpo_set_drop_rel(1e-2) < 0 and b"drop" in L.bsmpo_last_error()
    assert L.bsmpo_set_drop_rel(0.0) < 0
    assert L.bsmpo_drop_rel() == 1e-6
    try:
        assert engine.set_precision_rules(engine.BF16X3) == (1e-4, 1e-5)
        assert L.bsmpo_drop_rel() == 1e-5
        for p in (engine.C128, engine.C64):
            assert engine.PRECISION_RULES[p] == (mpo.TIE_REL, mpo.DROP_REL)
    finally:
        engine.set_precision_rules(engine.C128)
    assert L.bsmpo_drop_rel() == 1e-6


def test_group_runner_stream_count():
    """GroupRunner takes 1..4 launching streams; a graph-replay runner uses one."""
    from bosonsampling_amd import engine
    with pytest.raises(ValueError):
        engine.GroupRunner("cpu", streams=5)
    assert engine.GroupRunner("cpu", streams=3).streams == 3
    assert engine.GroupRunner("cpu", streams=3, use_graph=True).streams == 1


def test_plan_query_with_every_sector_on_the_tridiagonal_route():
    """Planning reads only the plan header when it sizes the single-update
    workspace; with every sector on the tridiagonal route (threshold 0) the
    back-transformation's per-sector items enter that size too."""
    from bosonsampling_amd import engine
    from bosonsampling_amd._capi import check, lib
    from synth import square_charges
    d, chi = 6, 48
    CL, CC, CR = square_charges(chi, 2, 5), square_charges(chi, 1, 4), square_charges(chi, 0, 3)
    sizes = {}
    try:
        for thr in (96, 0):
            check(lib().bsmpo_set_eig_threshold(thr), "set_eig_threshold")
            p = engine.Plan(d, chi, CL, CC, CR)
            sizes[thr] = p.work_bytes
    finally:
        check(lib().bsmpo_set_eig_threshold(96), "set_eig_threshold")
    c_bytes = 16 * int(p.info().c_elems)
    assert 2 * c_bytes < sizes[0] < 1000 * c_bytes and sizes[96] < 1000 * c_bytes


def test_two_stage_knob():
    """bsmpo_set_two_stage: 0 turns the two-stage reduction off, other values
    are raised to at least 64, values outside [0, 528] are rejected and leave
    the setting alone."""
    from bosonsampling_amd import _capi
    L = _capi.lib()
    prev = L.bsmpo_two_stage_threshold()
    try:
        assert L.bsmpo_set_two_stage(0) == 0 and L.bsmpo_two_stage_threshold() == 0
        assert L.bsmpo_set_two_stage(10) == 0 and L.bsmpo_two_stage_threshold() == 64
        assert L.bsmpo_set_two_stage(200) == 0 and L.bsmpo_two_stage_threshold() == 200
        assert L.bsmpo_set_two_stage(-1) < 0 and L.bsmpo_set_two_stage(529) < 0
        assert L.bsmpo_two_stage_threshold() == 200
    finally:
        L.bsmpo_set_two_stage(prev)


def test_two_stage_plan_query_on_large_sectors():
    """Planning a d=13, chi=512 update with the two-stage reduction on: the
    same sectors as with it off (the route changes only the eigensolver), and
    a workspace at least as large (the band and panel arrays share the
    tridiagonal work area)."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from synth import synthetic_update
    from bosonsampling_amd import _capi
    from bosonsampling_amd.engine import Plan
    L = _capi.lib()
    prev = L.bsmpo_two_stage_threshold()
    CL, CC, CR, *_ = synthetic_update(13, 512, (4, 12), (2, 10), (0, 8), 11)
    try:
        L.bsmpo_set_two_stage(0)
        p0 = Plan(13, 512, CL, CC, CR)
        L.bsmpo_set_two_stage(64)
        p1 = Plan(13, 512, CL, CC, CR)
    finally:
        L.bsmpo_set_two_stage(prev)
    s0, s1 = p0.sectors()[0], p1.sectors()[0]
    assert np.array_equal(s0, s1)
    n = np.minimum(s1[:, 2], s1[:, 3])
    assert ((n >= 64) & (n <= 528)).sum() > 100
    assert p1.work_bytes >= p0.work_bytes


def test_layer_program_query_and_needs_on_host():
    """Layer programs (device-planned graph replays): the spec is validated
    and the workspace sized without a device; bsmpo_layer_needs sums host
    plans into the capacities a device-planned group needs."""
    import ctypes as C
    from bosonsampling_amd import _capi
    from bosonsampling_amd.engine import Plan
    L = _capi.lib()
    st, _ = _state()
    d, chi = st.d, st.chi
    plans = [Plan(d, chi, st.charge[m], st.charge[m + 1], st.charge[m + 2]) for m in (0, 2, 4)]
    arr = (C.c_void_p * 3)(*[p.address for p in plans])
    need = _capi.LayerCaps()
    _capi.check(L.bsmpo_layer_needs(3, arr, C.byref(need)), "layer_needs")
    infos = [p.info() for p in plans]
    assert need.sec == sum(i.n_sectors for i in infos)
    assert need.tiles == sum(i.n_items for i in infos)
    assert need.var_bytes > 0 and need.work_bytes > 0 and need.gram > 0 and need.w > 0
    recs = (_capi.LayerUpdate * 2)()
    for j in range(2):
        recs[j] = _capi.LayerUpdate(*([8] * 9), 0, 0)  # non-NULL stand-ins: query reads no memory
    spec = _capi.LayerSpec()
    spec.d, spec.chi, spec.precision, spec.n_sets = d, chi, _capi.C64, 2
    spec.n_upd[0] = spec.n_upd[1] = 2
    spec.upd[0] = spec.upd[1] = C.cast(recs, C.c_void_p)
    spec.n_groups, spec.n_streams, spec.n_layers, spec.d_B = 2, 2, 4, 8
    spec.caps = need
    size = C.c_size_t()
    _capi.check(L.bsmpo_layer_query(C.byref(spec), C.byref(size)), "layer_query")
    assert size.value > 2 * need.work_bytes
    for field, bad in (("n_groups", 3), ("n_streams", 5), ("n_sets", 0), ("chi", 4096), ("d", 17)):
        old = getattr(spec, field)
        setattr(spec, field, bad)
        assert L.bsmpo_layer_query(C.byref(spec), C.byref(size)) < 0
        setattr(spec, field, old)
    recs[1].Glc = None
    assert L.bsmpo_layer_query(C.byref(spec), C.byref(size)) < 0
    assert L.bsmpo_layer_launch(None, None) < 0
