"""GPU reverse mode against the FP64 oracle's adjoint (oracle/adjoint.py) and
against central finite differences of the GPU forward."""
import numpy as np
import pytest
import torch

from oracle import adjoint as oadj, fock, mpo
from bosonsampling_amd import mesh as meshmod

pytestmark = pytest.mark.gpu


def _params(M, seed):
    """(theta, phi) per brick-wall slot of the Haar/Clements mesh of `seed`."""
    U = meshmod.haar_unitary(M, seed)
    seq, _ = meshmod.clements_decompose(U)
    params = np.zeros((M, M // 2, 2))
    last = [-1] * M
    for (m, th, ph) in seq:
        lay = max(last[m], last[m + 1]) + 1
        if lay % 2 != m % 2:
            lay += 1
        params[lay, (m - lay % 2) // 2] = (th, ph)
        last[m] = last[m + 1] = lay
    return params


def _upload(st):
    from bosonsampling_amd.engine import DeviceMPO
    return DeviceMPO.from_host(st.Gamma, st.Lambda, st.charge, st.d)


def test_bs_table_vjp_matches_oracle(gpu):
    import ctypes as C
    from bosonsampling_amd import engine
    from bosonsampling_amd._capi import check, lib
    d = 7
    rng = np.random.default_rng(3)
    Bs = np.stack([meshmod.haar_unitary(2, s) for s in range(3)])
    G = rng.standard_normal((3, d, d, d)) + 1j * rng.standard_normal((3, d, d, d))
    Bd, Gd = torch.from_numpy(Bs).to(gpu), torch.from_numpy(G).to(gpu)
    out = torch.empty((3, 2, 2), dtype=torch.complex128, device=gpu)
    check(lib().bsmpo_bs_table_vjp(d, 3, C.c_void_p(Bd.data_ptr()), C.c_void_p(Gd.data_ptr()),
                                   C.c_void_p(out.data_ptr()), engine._stream()), "bs_table_vjp")
    got = out.cpu().numpy()
    for b in range(3):
        want = oadj.bs_table_vjp(d, Bs[b], G[b])
        assert np.abs(got[b] - want).max() <= 1e-12 * np.abs(want).max()


def test_observable_vjp_matches_oracle(gpu):
    from bosonsampling_amd.adjoint import Run
    M, d, chi, inputs = 6, 4, 16, [0, 1, 2]
    mm, _ = meshmod.haar_clements_mesh(M, 3)
    st = mpo.initialize(M, d, chi, inputs, 0.8)
    mpo.simulate(st, mm, inputs)
    pattern, w = [1, 0, 1, 0, 1, 0], np.linspace(0.1, 0.6, M)
    loss, P, nw, g = oadj.observable_loss(st, pattern, w)
    run = Run(_upload(st), inputs, pattern, w)
    gl, gP, gnw, gG = run.observable_vjp(_upload(st))
    print(f"observable vjp: loss {gl:.12f} vs {loss:.12f}", flush=True)
    assert abs(gl - loss) <= 1e-10 * abs(loss) and abs(gP - P) <= 1e-12 and abs(gnw - nw) <= 1e-12
    assert np.abs(gG.cpu().numpy() - g).max() <= 1e-10 * np.abs(g).max()


def test_site_update_vjp_gauge_invariant_functional(gpu):
    """f = Re<Z, G0' G1'> + p . Lambda' of the update's outputs: both sides
    take the gradient of the same gauge-invariant functional of their own
    outputs, so the input gradients must agree."""
    import ctypes as C
    from bosonsampling_amd import engine
    from bosonsampling_amd._capi import check, lib
    M, d, chi, inputs = 6, 4, 16, [0, 1, 2]
    mm, _ = meshmod.haar_clements_mesh(M, 5)
    st = mpo.initialize(M, d, chi, inputs, 0.8)
    mpo.simulate(st, mm, inputs, layers=range(3))
    mode = 2
    U3 = fock.bs_table(d, mm[3, 1])
    LR = st.Lambda[mode + 1]
    rng = np.random.default_rng(9)
    Z = rng.standard_normal((chi, chi)) + 1j * rng.standard_normal((chi, chi))
    p = rng.standard_normal(chi)
    # oracle
    after = st.copy()
    mpo.apply_update(after, mode, U3)
    G0, G1, lam = after.Gamma[mode], after.Gamma[mode + 1], after.Lambda[mode]
    want = oadj.site_update_vjp(d, chi, U3, LR, st.charge[mode], st.charge[mode + 1],
                                st.charge[mode + 2], st.Gamma[mode], st.Gamma[mode + 1],
                                Z @ G1.conj().T, G0.conj().T @ Z, p)
    # device
    dev = _upload(st)
    U3d = torch.from_numpy(U3).to(gpu)
    plan = engine.Plan(d, chi, st.charge[mode], st.charge[mode + 1], st.charge[mode + 2])
    pre = (dev.Gamma[mode].clone(), dev.Gamma[mode + 1].clone(), dev.Lambda[mode + 1].clone())
    dev.update(mode, U3d, plan=plan)
    G0d, G1d, lamd = dev.Gamma[mode].clone(), dev.Gamma[mode + 1].clone(), dev.Lambda[mode].clone()
    chn = dev.charge_host[mode + 1].copy()
    assert np.array_equal(chn, after.charge[mode + 1])
    Zd = torch.from_numpy(Z).to(gpu)
    gG0 = (Zd @ G1d.conj().T).contiguous()
    gG1 = (G0d.conj().T @ Zd).contiguous()
    gLam = torch.from_numpy(p).to(gpu)
    gLR = torch.zeros(chi, dtype=torch.float64, device=gpu)
    gU3 = torch.empty((d, d, d), dtype=torch.complex128, device=gpu)
    gGlc = torch.empty_like(gG0)
    gGcr = torch.empty_like(gG1)
    wb = C.c_size_t()
    check(lib().bsmpo_site_update_vjp_query(d, chi, C.c_void_p(plan.address), chn.ctypes.data,
                                            C.byref(wb)), "vjp_query")
    work = torch.empty(wb.value, dtype=torch.uint8, device=gpu)
    vp = lambda t: C.c_void_p(t.data_ptr())
    print(f"site_update_vjp mode={mode} work={wb.value}", flush=True)
    check(lib().bsmpo_site_update_vjp(d, chi, C.c_void_p(plan.address), chn.ctypes.data,
                                      vp(pre[0]), vp(pre[1]), vp(pre[2]), vp(U3d), vp(G0d),
                                      vp(G1d), vp(lamd), vp(gG0), vp(gG1), vp(gLam), vp(gGlc),
                                      vp(gGcr), vp(gLR), vp(gU3), 0, vp(work), work.numel(),
                                      engine._stream()), "site_update_vjp")
    got = [gGlc.cpu().numpy(), gGcr.cpu().numpy(), gLR.cpu().numpy(), gU3.cpu().numpy()]
    for name, a, b in zip(["gGlc", "gGcr", "gLR", "gU3"], got, want):
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
        print(f"  {name}: rel err {err:.2e}")
        assert err <= 1e-8, name


@pytest.mark.parametrize("ckpt", [1, 3])
def test_mesh_gradient_matches_oracle(gpu, ckpt):
    from bosonsampling_amd.adjoint import Run, mesh_loss
    M, d, chi, inputs, eta = 6, 4, 16, [0, 1, 2], 0.8
    params = _params(M, 5)
    pattern, w = [1, 0, 1, 0, 1, 0], np.linspace(0.2, -0.3, M)
    st0 = mpo.initialize(M, d, chi, inputs, eta)
    loss_o, grad_o = oadj.loss_and_grad(st0, params, inputs, pattern, w)
    run = Run(_upload(st0), inputs, pattern, w, checkpoint_every=ckpt)
    pt = torch.tensor(params, requires_grad=True)
    loss = mesh_loss(pt, run)
    loss.backward()
    g = pt.grad.numpy()
    err = np.abs(g - grad_o).max() / np.abs(grad_o).max()
    print(f"mesh gradient ckpt={ckpt}: loss {loss.item():.12f} vs {loss_o:.12f}, rel err {err:.2e}",
          flush=True)
    assert abs(loss.item() - loss_o) <= 1e-10 * abs(loss_o)
    assert err <= 1e-8


DIFF = dict(M=16, d=9, chi=128, inputs=list(range(8)), eta=0.8, seed=2,
            pattern=[1, 0] * 8, weights=np.array([0.0] * 8 + [0.05] * 8))
_FD = {}


def _diff_state(gpu):
    from bosonsampling_amd import engine
    c = DIFF
    return engine.initialize(c["M"], c["d"], c["chi"], c["inputs"], c["eta"], device=gpu)


def _central_differences(gpu):
    """Central differences (step 1e-6) of the complex128 forward of the
    differentiation workload at 64 angles: 32 random (layer, beam splitter)
    slots, theta and phi each.  Computed once per session and shared by the
    precisions under test."""
    if "fd" in _FD:
        return _FD["fd"]
    from bosonsampling_amd import engine
    from bosonsampling_amd.adjoint import Run
    c = DIFF
    M, params = c["M"], _params(c["M"], c["seed"])
    ref = Run(_diff_state(gpu), c["inputs"], c["pattern"], c["weights"], checkpoint_every=8,
              precision=0)

    def L(p):
        return ref.forward(oadj.mesh_from_params(M, p))[0]

    used = [(lay, i) for lay in range(M) for i, _ in engine.layer_modes(M, lay, 7)]
    rng = np.random.default_rng(0)
    pick = [used[j] for j in rng.choice(len(used), 32, replace=False)]
    h, fd = 1e-6, {}
    for lay, i in pick:
        for k in range(2):
            pp, pm = params.copy(), params.copy()
            pp[lay, i, k] += h
            pm[lay, i, k] -= h
            fd[(lay, i, k)] = (L(pp) - L(pm)) / (2 * h)
    print(f"central differences at {len(fd)} angles", flush=True)
    _FD["fd"] = fd
    return fd


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", [0, 1], ids=["c128", "c64"])
def test_diff_config_gradient_vs_finite_differences(gpu, prec):
    """M=16, photons in modes 0-7, d=9, chi=128, eta=0.8, Haar seed 2 mesh
    (120 beam splitters, 240 angles); L = -log P(pattern) + 0.05 sum_{j >= 8}
    <n_j>, the pattern one photon in each even mode; checkpoints every 8
    layers.  The adjoint gradient of the complex128 and complex64 forwards
    against central differences of the complex128 forward (step 1e-6) at 64
    angles, within 1e-5 of the largest gradient entry (the target; measured
    2.0e-8 and 2.7e-6)."""
    from bosonsampling_amd.adjoint import Run, mesh_loss
    c = DIFF
    params = _params(c["M"], c["seed"])
    run = Run(_diff_state(gpu), c["inputs"], c["pattern"], c["weights"], checkpoint_every=8,
              precision=prec)
    pt = torch.tensor(params, requires_grad=True)
    loss = mesh_loss(pt, run)
    loss.backward()
    g = pt.grad.numpy()
    print(f"diff config prec={prec}: loss {loss.item():.10f}", flush=True)
    assert np.isfinite(loss.item()) and np.all(np.isfinite(g))
    fd = _central_differences(gpu)
    scale = np.abs(g).max()
    errs = [abs(v - g[key]) / scale for key, v in fd.items()]
    print(f"prec={prec}: max |adjoint - fd| / max|grad| over {len(errs)} angles = {max(errs):.2e}")
    assert len(errs) >= 64 and max(errs) <= 1e-5


def test_diff_config_pattern_of_modes_8_to_15_is_truncated_away(gpu):
    """One photon in each of modes 8-15 (the pattern the differentiation
    workload names) has P = 0 in the chi = 128 MPO: truncation leaves no slot
    of charge (8, 8) on bond 8, though the exact value
    eta^8 |Perm(W[8:16, 0:8])|^2 is positive.  This is why the loss uses one
    photon in each even mode instead."""
    from bosonsampling_amd import engine
    c = DIFF
    M, d, chi = c["M"], c["d"], c["chi"]
    mm = oadj.mesh_from_params(M, _params(M, c["seed"]))
    st = _diff_state(gpu)
    engine.simulate(st, mm, c["inputs"], observe=False)
    G, L, ch = st.to_host()
    host = mpo.MPOState(d, chi, G, L, ch, np.ones(chi))
    pattern = [0] * 8 + [1] * 8
    assert mpo.pattern_probability(host, pattern) == 0.0
    assert not np.any((ch[8][:, 0] == 8) & (ch[8][:, 1] == 8))
    W = meshmod.mesh_unitary(mm)
    exact = c["eta"] ** 8 * abs(fock.permanent(W[8:16, 0:8])) ** 2
    print(f"exact P(modes 8-15) = {exact:.3e}", flush=True)
    assert exact > 1e-9
