"""Writes the golden fixtures of the site-update path (run from the repo root:
`python tests/golden/make_golden.py`).

The reference holds no test vectors and its path cannot run here (cupy,
mpi4py, qutip and nvcc are absent), so every vector below comes from the
FP64 oracle after the oracle itself is pinned by tests/test_oracle.py
(closed forms, brute force, permanents).  The fixtures freeze those
outputs so that later changes to the oracle or the device path are checked
against fixed data on the GPU box, where the oracle is not re-derived.

  bs_tables.npz  U3 tables of the reference formula for a few (t, r), d = 5, 9
  m8_sp4.npz     M=8, single photons in modes 0-3, eta=0.5, d=5, chi=16,
                 Haar seed 0 Clements mesh (28 beam splitters): mesh, trace
                 and entropies per layer, final marginals, two pattern
                 probabilities, and one site update's input and output
                 (layer 4, mode 2) as gauge-invariant data
  m32_sp8.npz    M=32, single photons in modes 0,4,...,28, eta=0.7, d=9,
                 chi=256, Haar seed 1 Clements mesh (496 beam splitters):
                 trace and entropies per layer, final marginals and three
                 pattern probabilities (takes ~15 min of CPU)

  synth_d13.npz  one site update at d=13 on synthetic inputs (tests/synth.py,
                 seed 5) at chi=1024 and chi=2048, whose largest sectors
                 have n = 614 and n = 1228: the kept charges and singular
                 values and X^H (Gamma0 Gamma1) Y for fixed random X, Y
                 (gauge invariant; ~2 min of CPU)

`python tests/golden/make_golden.py m32` (or `synth`) writes only that file.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from bosonsampling_amd import mesh  # noqa: E402
from oracle import fock, mpo  # noqa: E402


def bs_tables():
    out = {}
    for k, (seed, r) in enumerate([(3, 0.42), (11, 0.0), (29, 0.93)]):
        t, rr = mesh.reference_phases(seed, r)
        out[f"tr{k}"] = np.array([t, rr])
        for d in (5, 9):
            out[f"U3_{k}_d{d}"] = fock.bs_table_reference(d, t, rr)
    np.savez_compressed(os.path.join(HERE, "bs_tables.npz"), **out)


def m8_sp4():
    M, d, chi, inputs, eta = 8, 5, 16, [0, 1, 2, 3], 0.5
    mm, _ = mesh.haar_clements_mesh(M, 0)
    st = mpo.initialize(M, d, chi, inputs, eta)
    init = st.copy()
    prob, EE = [mpo.trace(st)], [mpo.entanglement_entropy(st.Lambda)]
    snap = None
    for layer in range(M):
        if layer == 4:
            snap = st.copy()
        mpo.simulate(st, mm, inputs, layers=[layer])
        prob.append(mpo.trace(st))
        EE.append(mpo.entanglement_entropy(st.Lambda))
    mode = 2
    U3 = fock.bs_table(d, mm[4, 1])
    after = snap.copy()
    mpo.apply_update(after, mode, U3)
    np.savez_compressed(
        os.path.join(HERE, "m8_sp4.npz"),
        mesh=mm, prob=np.array(prob), EE=np.stack(EE, 1), marginals=mpo.marginals(st),
        patterns=np.array([[1, 1, 0, 0, 1, 0, 1, 0], [0, 0, 0, 0, 1, 1, 1, 1]]),
        pattern_prob=np.array([mpo.pattern_probability(st, [1, 1, 0, 0, 1, 0, 1, 0]),
                               mpo.pattern_probability(st, [0, 0, 0, 0, 1, 1, 1, 1])]),
        init_Gamma=init.Gamma, init_Lambda=init.Lambda, init_charge=init.charge,
        upd_Gamma=snap.Gamma, upd_Lambda=snap.Lambda, upd_charge=snap.charge, upd_mode=mode,
        upd_U3=U3, upd_out_charge=after.charge[mode + 1], upd_out_Lambda=after.Lambda[mode],
        upd_out_pair=after.Gamma[mode] @ after.Gamma[mode + 1])


def m32_sp8():
    M, d, chi, inputs, eta = 32, 9, 256, list(range(0, 32, 4)), 0.7
    mm, _ = mesh.haar_clements_mesh(M, 1)
    st = mpo.initialize(M, d, chi, inputs, eta)
    prob, EE = [mpo.trace(st)], [mpo.entanglement_entropy(st.Lambda)]
    for layer in range(M):
        mpo.simulate(st, mm, inputs, layers=[layer])
        prob.append(mpo.trace(st))
        EE.append(mpo.entanglement_entropy(st.Lambda))
        print(f"m32_sp8 layer {layer}: trace {prob[-1]:.12f}", flush=True)
    patterns = np.zeros((3, M), np.int64)
    patterns[0, inputs] = 1                    # every photon back in its input mode
    patterns[1, [m + 1 for m in inputs]] = 1   # shifted by one mode
    # patterns[2]: vacuum, P = (1 - eta)^8 before truncation
    np.savez_compressed(
        os.path.join(HERE, "m32_sp8.npz"),
        mesh=mm, prob=np.array(prob), EE=np.stack(EE, 1), marginals=mpo.marginals(st),
        patterns=patterns,
        pattern_prob=np.array([mpo.pattern_probability(st, p) for p in patterns]))


SYNTH_CASES = {1024: ((4, 12), (2, 10), (0, 8)), 2048: ((4, 12), (2, 10), (0, 8))}
SYNTH_SEED = 5


def synth_projectors(chi, k=8):
    rng = np.random.default_rng(99)
    X = rng.standard_normal((chi, k)) + 1j * rng.standard_normal((chi, k))
    Y = rng.standard_normal((chi, k)) + 1j * rng.standard_normal((chi, k))
    return X, Y


def synth_bs(d):
    return fock.bs_table(d, mesh.haar_unitary(2, 3))


def synth_d13():
    sys.path.insert(0, os.path.dirname(HERE))
    from synth import synthetic_update
    d, out = 13, {}
    for chi, (lq, cq, rq) in SYNTH_CASES.items():
        CL, CC, CR, Glc, Gcr, LR = synthetic_update(d, chi, lq, cq, rq, SYNTH_SEED)
        ch, lam, G0, G1, num = mpo.site_update(d, chi, synth_bs(d), LR, CL, CC, CR, Glc, Gcr)
        X, Y = synth_projectors(chi)
        out[f"charge_{chi}"] = ch
        out[f"lambda_{chi}"] = lam
        out[f"num_{chi}"] = num
        out[f"proj_{chi}"] = (X.conj().T @ G0) @ (G1 @ Y)
        print(f"synth_d13 chi={chi}: kept {num}", flush=True)
    np.savez_compressed(os.path.join(HERE, "synth_d13.npz"), **out)


if __name__ == "__main__":
    if sys.argv[1:] == ["m32"]:
        m32_sp8()
        sys.exit(0)
    if sys.argv[1:] == ["synth"]:
        synth_d13()
        sys.exit(0)
    bs_tables()
    m8_sp4()
    m32_sp8()
    synth_d13()
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
