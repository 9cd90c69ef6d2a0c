"""GPU tests of the drop-in worker (Full.py interface) and of the
rank-process site-update operator (Rank.py message layout) against the FP64
oracle."""
import numpy as np
import pytest

from oracle import fock, mpo

pytestmark = pytest.mark.gpu


def test_rank_process_matches_oracle(gpu):
    from bosonsampling_amd import engine, mesh
    from bosonsampling_amd.worker import rank_process
    M, d, chi, inputs = 8, 5, 40, [0, 1, 2, 3]
    mm, _ = mesh.haar_clements_mesh(M, 3)
    st = mpo.initialize(M, d, chi, inputs, 0.5)
    mpo.simulate(st, mm, inputs, layers=range(3))
    for mode, edge in [(2, False), (M - 2, True)]:
        print(f"rank_process mode={mode} edge={edge}", flush=True)
        r, seed = 0.41, 777
        LR = st.Lambda[mode + 1] if not edge else np.ones(chi)
        ch, lam, g0, g1 = rank_process(d, chi, st.Lambda[mode], LR, st.charge[mode],
                                       st.charge[mode + 1], st.charge[mode + 2],
                                       st.Gamma[mode], st.Gamma[mode + 1], r, seed, edge=edge)
        orc = st.copy()
        t, rr = mesh.reference_phases(seed, r)
        mpo.apply_update(orc, mode, fock.bs_table(d, mesh.reference_bs_matrix(t, rr)))
        n = lam.size
        assert np.array_equal(ch, orc.charge[mode + 1][:n])
        assert np.all(orc.charge[mode + 1][n:] == d)
        assert np.abs(lam - orc.Lambda[mode][:n]).max() <= 1e-10 * lam.max()
        want = orc.Gamma[mode] @ orc.Gamma[mode + 1]
        got = g0 @ g1
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()


def test_worker_single_photon_run(gpu):
    """Worker.ExperimentInit/Simulate on the M=8, 4-photon, d=5, chi=16 case
    with the reference's reflectivity/seed mesh; prob and EE against the
    oracle run on the same mesh."""
    from bosonsampling_amd import mesh
    from bosonsampling_amd.worker import Worker
    M, n, d, chi, loss, seed = 8, 4, 5, 16, 0.5, 0
    w = Worker()
    assert w.ExperimentInit("Single photon", M, n, None, d, chi, {"loss": loss}, seed)
    ok, prob, EE, REPar = w.Simulate()
    assert ok and prob.shape == (M + 1,) and EE.shape == (M - 1, M + 1) and REPar.shape == (M - 1, M + 1, 5)
    refl, seeds = mesh.reflectivity_and_seeds(M, seed)
    mm = mesh.reference_mesh(M, refl, seeds)
    st = mpo.initialize(M, d, chi, list(range(n)), 1 - loss)
    want = [mpo.trace(st)]
    ee = [mpo.entanglement_entropy(st.Lambda)]
    for layer in range(M):
        mpo.simulate(st, mm, list(range(n)), layers=[layer])
        want.append(mpo.trace(st))
        ee.append(mpo.entanglement_entropy(st.Lambda))
    print(f"worker prob {prob}", flush=True)
    assert np.abs(prob - np.array(want)).max() <= 1e-10
    np.testing.assert_allclose(EE, np.stack(ee, 1), rtol=0, atol=1e-8)
    with pytest.raises(NotImplementedError):
        Worker().ExperimentInit("Gaussian", M, n, None, d, chi, {"loss": 0.1, "r": 0.8}, 0)
