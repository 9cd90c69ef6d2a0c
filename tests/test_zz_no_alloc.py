"""The allocation-free contraction (bsmpo_contract_noalloc) against
bsmpo_contract with a caller workspace: the same kernels, so the packed C and
T must agree bit for bit.  Two calls in a row, the second on a smaller plan,
check that the library's scratch array carries nothing over; a plan larger
than the scratch is rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _contract(plan, args, precision, workspace):
    import torch
    from bosonsampling_amd import engine
    from bosonsampling_amd._capi import check, lib
    d, chi, Glc, Gcr, LR, U3 = args
    n = int(plan.info().c_elems)
    Cd = torch.full((n,), float("nan"), dtype=torch.complex128, device=Glc.device)
    Td = torch.full((n,), float("nan"), dtype=torch.complex128, device=Glc.device)
    ptrs = [C.c_void_p(t.data_ptr()) for t in (Glc, Gcr, LR, U3, Cd, Td)]
    if workspace:
        work = torch.empty(plan.work_bytes, dtype=torch.uint8, device=Glc.device)
        check(lib().bsmpo_contract(d, chi, C.c_void_p(plan.address), *ptrs, None, None, precision,
                                   C.c_void_p(work.data_ptr()), work.numel(), engine._stream()),
              "contract")
    else:
        check(lib().bsmpo_contract_noalloc(d, chi, C.c_void_p(plan.address), *ptrs, None, None,
                                           precision, engine._stream()), "contract_noalloc")
    torch.cuda.synchronize()
    return Cd.cpu().numpy(), Td.cpu().numpy()


def _case(gpu, d, chi, lq, cq, rq, seed):
    import torch
    from bosonsampling_amd import engine, mesh
    from oracle import fock
    from synth import synthetic_update
    CL, CC, CR, Glc, Gcr, LR = synthetic_update(d, chi, lq, cq, rq, seed)
    plan = engine.Plan(d, chi, CL, CC, CR)
    U3 = fock.bs_table(d, mesh.haar_unitary(2, seed))
    dev = [torch.from_numpy(np.ascontiguousarray(x)).to(gpu) for x in (Glc, Gcr, LR, U3)]
    return plan, (d, chi, *dev)


@pytest.mark.parametrize("precision", [0, 1, 2], ids=["c128", "c64", "bf16x3"])
def test_noalloc_matches_workspace_contract(gpu, precision):
    big = _case(gpu, 13, 1024, (4, 12), (2, 10), (0, 8), 5)
    small = _case(gpu, 9, 256, (2, 8), (1, 6), (0, 5), 6)
    for name, (plan, args) in (("chi=1024 d=13", big), ("chi=256 d=9", small)):
        print(f"noalloc {name} precision {precision}", flush=True)
        C1, T1 = _contract(plan, args, precision, True)
        C2, T2 = _contract(plan, args, precision, False)
        assert not np.isnan(C2).any() and not np.isnan(T2).any()
        assert np.array_equal(C1, C2) and np.array_equal(T1, T2), name


def test_noalloc_rejects_plan_above_scratch(gpu):
    from bosonsampling_amd import engine
    from bosonsampling_amd._capi import lib
    from synth import square_charges
    d, chi = 13, 4096
    CL = square_charges(chi, 4, 12)
    CC = square_charges(chi, 2, 10)
    CR = square_charges(chi, 0, 8)
    plan = engine.Plan(d, chi, CL, CC, CR)
    rc = lib().bsmpo_contract_noalloc(d, chi, C.c_void_p(plan.address), C.c_void_p(16),
                                      C.c_void_p(16), C.c_void_p(16), C.c_void_p(16),
                                      C.c_void_p(16), C.c_void_p(16), None, None, 0, None)
    assert rc < 0
    assert b"scratch" in lib().bsmpo_last_error()
