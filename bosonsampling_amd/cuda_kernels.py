"""Drop-in for the reference's kernel launchers (src/cuda_kernels.py).

Same function names, argument order and meaning as the reference, backed by
libbsmpo's gfx950 kernels (include/bsmpo.h) instead of cupy RawKernels:

  Rand_U(d, r)           src/cuda_kernels.py:18-50  -> bsmpo_bs_table
  Rand_MPO_U(d, r)       src/cuda_kernels.py:54-61  (ket (x) bra products)
  update_MPO(...)        src/cuda_kernels.py:101-127 -> bsmpo_contract (one sector)
  update_MPS(...)        src/cuda_kernels.py:72-90   -> bsmpo_contract (one charge)

Arrays may be numpy arrays or torch tensors (cupy is not part of this
stack); results are torch tensors on the current HIP device.  Differences
that are deliberate: every table and product is complex128 / float64 (the
reference computes in float32 and returns planar float32), and the
reference's padded, 8-aligned row layout is accepted but not required.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import engine
from ._capi import check, lib
from .mesh import reference_bs_matrix


class Data:
    """The reference's data holder (src/mpo_sort.py:6-36): array_type,
    aligned flag, data and the selected indices of each dimension."""

    def __init__(self, array_type: str, aligned: bool, data, idx_select: list):
        self.array_type = array_type
        self.aligned = aligned
        self.data = data
        self.idx_select = idx_select
        self.ndim = getattr(data, "ndim", None)

    def clone(self):
        return Data(self.array_type, self.aligned, self.data, self.idx_select)


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError("bosonsampling_amd.cuda_kernels needs a HIP device")
    return torch.device("cuda", torch.cuda.current_device())


def _t(x, dtype):
    if isinstance(x, torch.Tensor):
        return x.to(device=_dev(), dtype=dtype).contiguous()
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=_dev()).contiguous()


def _np(x, dtype=None):
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    x = np.asarray(x)
    return x.astype(dtype) if dtype is not None else x


def Rand_U(d: int, r: float):
    """Fock BS table for reflectivity |r| with two random phases drawn from
    numpy's global RNG (alpha for t, then beta for r; r == 0: t = 1 and only
    beta is drawn), as src/cuda_kernels.py:18-50.  Returns (U_r, U_i),
    float64 [d, d, d] on the device, U3[n1, n2, p] = <n1, n2|U|p, n1+n2-p>."""
    if r == 0:
        t = 1.0 + 0.0j
        rr = r * np.exp(1j * np.random.rand() * 2 * math.pi)
    else:
        t = math.sqrt(1 - r ** 2) * np.exp(1j * np.random.rand() * 2 * math.pi)
        rr = r * np.exp(1j * np.random.rand() * 2 * math.pi)
    U3 = engine.bs_tables(d, reference_bs_matrix(t, rr), _dev())[0]
    return U3.real.contiguous(), U3.imag.contiguous()


def Rand_MPO_U(d: int, r: float):
    """Density-matrix form of Rand_U: U (x) conj(U) as planar real / imaginary
    parts (src/cuda_kernels.py:54-61)."""
    U_r, U_i = Rand_U(d, r)
    return torch.kron(U_r, U_r) + torch.kron(U_i, U_i), torch.kron(U_i, U_r) - torch.kron(U_r, U_i)


def _centre_charges(change_charges_C, change_idx_C, chi):
    """Centre-bond charges from the reference's change table
    (src/mpo_sort.py:62-65): rows from change_idx_C[i] on carry
    change_charges_C[:, i] until the next change."""
    chc = _np(change_charges_C, np.int32).reshape(2, -1)
    idx = _np(change_idx_C, np.int64).reshape(-1)
    CC = np.zeros((chi, 2), np.int32)
    for i in range(idx.size):
        stop = idx[i + 1] if i + 1 < idx.size else chi
        CC[idx[i]:stop] = chc[:, i]
    return CC


def _compact(ch, d):
    """Indices of the rows of `ch` ([n, 2]) that carry a real charge (both
    components in [0, d); the reference pads with -1 when aligning and with
    d past the end), in stable lexicographic order."""
    ch = np.asarray(ch).reshape(-1, 2)
    ok = np.where((ch >= 0).all(1) & (ch < d).all(1))[0]
    key = ch[ok, 0].astype(np.int64) * (d + 1) + ch[ok, 1]
    return ok[np.argsort(key, kind="stable")]


def _contract(d, CL, CC, CR, tau, Glc, Gcr, U3, LL=None, LC=None, LR=None):
    """Sector tau of the contraction on the device for arbitrary row / centre
    / column charge lists (unsorted, padded or aligned).  The real rows are
    gathered into sorted, compact order, contracted by bsmpo_contract, and
    scattered back; every other entry of the [len(CL), len(CR)] results C
    (no LR) and T (with LL, LC, LR as given) is 0."""
    nL, nR = Glc.shape[0], Gcr.shape[1]
    ol, oc, orr = _compact(CL, d), _compact(CC, d), _compact(CR, d)
    Cout = torch.zeros((nL, nR), dtype=torch.complex128, device=_dev())
    Tout = torch.zeros((nL, nR), dtype=torch.complex128, device=_dev())
    L, K, R = ol.size, oc.size, orr.size
    if L == 0 or K == 0 or R == 0:
        return Cout, Tout
    chi = max(L, K, R)
    il, ic, ir = (torch.as_tensor(o, device=_dev()) for o in (ol, oc, orr))

    def pad_ch(ch, o):
        out = np.full((chi, 2), d, np.int32)
        out[:o.size] = np.asarray(ch).reshape(-1, 2)[o]
        return out

    def pad_vec(v, idx, n, fill):
        out = torch.full((chi,), fill, dtype=torch.float64, device=_dev())
        out[:n] = _t(v, torch.float64)[idx]
        return out

    Gl = torch.zeros((chi, chi), dtype=torch.complex128, device=_dev())
    Gr = torch.zeros((chi, chi), dtype=torch.complex128, device=_dev())
    Gl[:L, :K] = Glc[il][:, ic]
    Gr[:K, :R] = Gcr[ic][:, ir]
    lr = pad_vec(LR, ir, R, 0.0) if LR is not None else torch.ones(chi, dtype=torch.float64,
                                                                    device=_dev())
    ll = pad_vec(LL, il, L, 0.0) if LL is not None else None
    lc = pad_vec(LC, ic, K, 0.0) if LC is not None else None
    plan = engine.Plan(d, chi, pad_ch(CL, ol), pad_ch(CC, oc), pad_ch(CR, orr), tau=tau)
    sec, coff, rows, cols = plan.sectors()
    if sec.shape[0] == 0:
        return Cout, Tout
    n = int(plan.info().c_elems)
    Cd = torch.empty(n, dtype=torch.complex128, device=_dev())
    Td = torch.empty(n, dtype=torch.complex128, device=_dev())
    work = torch.empty(max(plan.work_bytes, 1), dtype=torch.uint8, device=_dev())
    vp = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    check(lib().bsmpo_contract(d, chi, C.c_void_p(plan.address), vp(Gl), vp(Gr), vp(lr), vp(U3),
                               vp(Cd), vp(Td), vp(ll), vp(lc), 0, vp(work), work.numel(),
                               engine._stream()), "contract")
    Ls, Rs = int(sec[0, 2]), int(sec[0, 3])
    rsel = il[torch.as_tensor(rows[:Ls].astype(np.int64), device=_dev())]
    csel = ir[torch.as_tensor(cols[:Rs].astype(np.int64), device=_dev())]
    Cout[rsel[:, None], csel[None, :]] = Cd[:Ls * Rs].view(Ls, Rs)
    Tout[rsel[:, None], csel[None, :]] = Td[:Ls * Rs].view(Ls, Rs)
    return Cout, Tout


def update_MPO(d, charge_c_0, charge_c_1, U_r, U_i, glc_obj, gcr_obj, cl_obj, cr_obj,
               change_charges_C, change_idx_C):
    """C_tau for the centre charge tau = (charge_c_0, charge_c_1)
    (src/cuda_kernels.py:101-127, weights of src/update_MPO_kernel.cu:441-497):
        C[l, r] = sum_c Glc[l, c] Gcr[c, r] U3[cl0-t0, t0-cr0, cl0-cc0] conj(U3[cl1-t1, t1-cr1, cl1-cc1])
    over cl >= cc >= cr componentwise.  glc_obj / gcr_obj / cl_obj / cr_obj
    are Data objects (anything with .data and .idx_select) holding the
    selected rows / columns; rows whose charge is -1 (alignment padding) or
    d (end padding) give zero rows, as in the reference.  The centre charges
    come from the change table (change_charges_C [2, n], change_idx_C [n]).
    Returns Data('T', True, C, [glc_obj.idx_select[0], gcr_obj.idx_select[1]])."""
    U3 = _table(U_r, U_i)
    Glc = _t(glc_obj.data, torch.complex128)
    Gcr = _t(gcr_obj.data, torch.complex128)
    CL = _np(cl_obj.data, np.int32).reshape(-1, 2)
    CR = _np(cr_obj.data, np.int32).reshape(-1, 2)
    if Glc.shape[0] != CL.shape[0] or Gcr.shape[1] != CR.shape[0] or Glc.shape[1] != Gcr.shape[0]:
        raise ValueError(f"update_MPO: shapes Glc {tuple(Glc.shape)} Gcr {tuple(Gcr.shape)} "
                         f"CL {CL.shape} CR {CR.shape} do not match")
    CC = _centre_charges(change_charges_C, change_idx_C, Glc.shape[1])
    Cm, _ = _contract(d, CL, CC, CR, (int(charge_c_0), int(charge_c_1)), Glc, Gcr, U3)
    return Data("T", True, Cm, [glc_obj.idx_select[0], gcr_obj.idx_select[1]])


def update_MPS(d, tau, U_r, U_i, Glc, Gcr, LL, LC, LR, CL, CC, CR, incC):
    """Pure-state theta for one charge tau (src/cuda_kernels.py:72-90,
    src/update_kernel.cu:594-595,706-715):
        T[l, r] = LL[l] LR[r] sum_c LC[c] Glc[l, c] Gcr[c, r] U3[cl-tau, tau-cr, cl-cc]
    for cl >= cc >= cr and cl >= tau >= cr (other entries are 0).  One U(1)
    charge per bond, run as the charge pair (c, 0), whose second weight factor
    U3[0, 0, 0] is 1.  incC (centre block starts) is implied by CC."""
    del incC
    U3 = _table(U_r, U_i)
    pair = lambda x: np.stack([x, np.zeros_like(x)], 1).astype(np.int32)
    cl, cc, cr = (pair(_np(x, np.int32).reshape(-1)) for x in (CL, CC, CR))
    _, T = _contract(d, cl, cc, cr, (int(tau), 0), _t(Glc, torch.complex128),
                     _t(Gcr, torch.complex128), U3, LL=LL, LC=LC, LR=LR)
    return T


def _table(U_r, U_i):
    return (_t(U_r, torch.float64) + 1j * _t(U_i, torch.float64)).to(torch.complex128).contiguous()
