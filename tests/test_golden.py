"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle must keep reproducing them (CPU), and the device path must match
them (GPU) without re-running the oracle."""
import os

import numpy as np
import pytest

from oracle import fock, mpo

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_oracle_bs_tables_golden():
    g = _load("bs_tables.npz")
    for k in range(3):
        t, r = g[f"tr{k}"]
        for d in (5, 9):
            assert np.abs(fock.bs_table(d, fock.reference_bs_matrix(t, r)) - g[f"U3_{k}_d{d}"]).max() < 1e-13


def test_oracle_m8_golden():
    g = _load("m8_sp4.npz")
    M, d, chi, inputs = 8, 5, 16, [0, 1, 2, 3]
    st = mpo.initialize(M, d, chi, inputs, 0.5)
    assert np.array_equal(st.charge, g["init_charge"])
    mpo.simulate(st, g["mesh"], inputs)
    assert abs(mpo.trace(st) - g["prob"][-1]) < 1e-12
    assert np.abs(mpo.marginals(st) - g["marginals"]).max() < 1e-12


@pytest.mark.gpu
def test_gpu_bs_tables_golden(gpu):
    from bosonsampling_amd import engine
    g = _load("bs_tables.npz")
    for k in range(3):
        t, r = g[f"tr{k}"]
        for d in (5, 9):
            got = engine.bs_tables(d, fock.reference_bs_matrix(t, r), gpu)[0].cpu().numpy()
            assert np.abs(got - g[f"U3_{k}_d{d}"]).max() < 1e-13


@pytest.mark.gpu
def test_gpu_site_update_golden(gpu):
    import torch
    from bosonsampling_amd import engine
    g = _load("m8_sp4.npz")
    d, mode = 5, int(g["upd_mode"])
    st = engine.DeviceMPO.from_host(g["upd_Gamma"], g["upd_Lambda"], g["upd_charge"], d)
    st.update(mode, torch.from_numpy(g["upd_U3"]).to(gpu))
    G, L, ch = st.to_host()
    n = int((g["upd_out_charge"][:, 0] < d).sum())
    assert np.array_equal(ch[mode + 1], g["upd_out_charge"])
    assert np.abs(L[mode][:n] - g["upd_out_Lambda"][:n]).max() <= 1e-10 * L[mode].max()
    pair = G[mode] @ G[mode + 1]
    assert np.abs(pair - g["upd_out_pair"]).max() <= 1e-10 * np.abs(pair).max()


@pytest.mark.gpu
def test_gpu_m8_run_golden(gpu):
    from bosonsampling_amd import engine
    g = _load("m8_sp4.npz")
    st = engine.DeviceMPO.from_host(g["init_Gamma"], g["init_Lambda"], g["init_charge"], 5)
    prob, EE = engine.simulate(st, g["mesh"], [0, 1, 2, 3])
    assert np.abs(prob - g["prob"]).max() <= 1e-10
    np.testing.assert_allclose(EE, g["EE"], rtol=0, atol=1e-8)
    tr, marg = st.observables()
    assert np.abs(marg - g["marginals"]).max() <= 1e-10
