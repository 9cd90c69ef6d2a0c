"""Drop-in site-update operator and simulation worker.

`rank_process` is the reference's per-update rank work with its message
layout (src/beam_splitter_parallel/Rank.py:32-206: LC, LR, CL, CC, CR, Glc,
Gcr, r, seed in; new_charge, new_Lambda, Gamma0Out, Gamma1Out out), run on
one GPU by bsmpo_site_update.  `Worker` fills the empty worker slot of the
reference's driver (run.py:37-43: `gpu_algorithm.Worker()`) with the
interface of FullWorker (src/beam_splitter_parallel/Full.py:19-202):
ExperimentInit(...) then Simulate() -> (success, prob, EE, REPar).  One
process drives one GPU; the master/rank message passing of the reference is
replaced by the MPO living in HBM for the whole run.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import engine, mesh
from ._capi import check, lib


def _rand_bs(r: float, seed: int) -> np.ndarray:
    """B = [[t, r], [-r*, t*]] from |r| and the integer seed, with the
    phases drawn as src/beam_splitter_parallel/Rank.py:56-57 +
    src/cuda_kernels.py:20-43 draw them."""
    t, rr = mesh.reference_phases(int(seed), float(r))
    return mesh.reference_bs_matrix(t, rr)


def rank_process(d: int, chi: int, LC, LR, CL, CC, CR, Glc, Gcr, r: float, seed: int,
                 edge: bool = False, precision: int = engine.C128, device="cuda"):
    """One beam-splitter site update with the reference rank's inputs and
    outputs (Rank.py:37-53 receives, :201-206 sends).

    LC, LR: [chi] (LC is received but, as in the reference, not used by the
    MPO update); CL, CC, CR: int [chi, 2], lexicographically sorted with
    padding d; Glc, Gcr: complex [chi, chi]; r, seed: beam-splitter
    reflectivity and phase seed.  edge: the pair is the last one (mode M-2),
    where only the first min(chi, d*d) columns / rows are kept
    (Full.py:130-132).

    Returns (new_charge int32 [num_lambda, 2], new_Lambda float64
    [num_lambda], Gamma0Out, Gamma1Out complex128 [chi, chi]) as numpy
    arrays; FP64 throughout (the reference sends float32 / complex64)."""
    del LC
    dev = torch.device(device)
    CL, CC, CR = (np.ascontiguousarray(x, np.int32).reshape(chi, 2) for x in (CL, CC, CR))
    if np.shape(Glc) != (chi, chi) or np.shape(Gcr) != (chi, chi) or np.shape(LR) != (chi,):
        raise ValueError(f"rank_process: expected Glc/Gcr [{chi},{chi}] and LR [{chi}], got "
                         f"{np.shape(Glc)} {np.shape(Gcr)} {np.shape(LR)}")
    plan = engine.Plan(d, chi, CL, CC, CR)
    g0 = torch.from_numpy(np.ascontiguousarray(Glc, np.complex128)).to(dev)
    g1 = torch.from_numpy(np.ascontiguousarray(Gcr, np.complex128)).to(dev)
    lr = torch.from_numpy(np.ascontiguousarray(LR, np.float64)).to(dev)
    U3 = engine.bs_tables(d, _rand_bs(r, seed), dev)[0]
    charge_c = torch.full((chi, 2), d, dtype=torch.int32, device=dev)
    lambda_c = torch.zeros(chi, dtype=torch.float64, device=dev)
    nl = torch.zeros(1, dtype=torch.int32, device=dev)
    work = torch.empty(max(plan.work_bytes, 1), dtype=torch.uint8, device=dev)
    vp = lambda t: C.c_void_p(t.data_ptr())
    check(lib().bsmpo_site_update(d, chi, C.c_void_p(plan.address), vp(g0), vp(g1), vp(lr),
                                  vp(U3), vp(charge_c), vp(lambda_c), vp(nl), 1 if edge else 0,
                                  int(precision), vp(work), work.numel(), engine._stream()),
          "site_update")
    n = int(nl.item())
    return (charge_c[:n].cpu().numpy(), lambda_c[:n].cpu().numpy(), g0.cpu().numpy(),
            g1.cpu().numpy())


class Worker:
    """Single-GPU worker with FullWorker's interface (Full.py:19-202).

    ExperimentInit(input_state_type, n_modes, n_input_states,
                   post_selected_photon_number, local_hilbert_space_dimension,
                   bond_dimension, parameters, seed=None)
      input_state_type 'Single photon' (parameters['loss'] is the
      probability of the vacuum, MPO.py:41-44: eta = 1 - loss); inputs in
      modes 0 .. n_input_states-1.  The beam splitters are the reference's
      Haar-induced reflectivities and per-BS phase seeds
      (src/interferometer.py:9-54) unless `mesh_override` ([M, M//2, 2, 2]
      complex matrices, e.g. a Clements mesh) is given.
    Simulate() -> (success, prob[M+1], EE[M-1, M+1], REPar[M-1, M+1, 5])
      prob[k] is the trace after k layers, EE the entanglement entropy of
      every bond (src/summary.py:3-26).  REPar stays zero as in the reference.
    """

    def __init__(self, precision: int = engine.C128, device="cuda"):
        self.precision = precision
        self.device = device
        self.state = None

    def ExperimentInit(self, input_state_type, n_modes, n_input_states, post_selected_photon_number,
                       local_hilbert_space_dimension, bond_dimension, parameters, seed=None,
                       mesh_override=None):
        if input_state_type != "Single photon":
            raise NotImplementedError(f"input_state_type {input_state_type!r}: this build "
                                      "simulates single-photon inputs only")
        M, d, chi = int(n_modes), int(local_hilbert_space_dimension), int(bond_dimension)
        self.n_modes, self.n_input_states = M, int(n_input_states)
        self.local_hilbert_space_dimension, self.bond_dimension = d, chi
        self.prob = np.zeros(M + 1)
        self.EE = np.zeros((M - 1, M + 1))
        self.REPar = np.zeros((M - 1, M + 1, 5))
        if mesh_override is not None:
            self.mesh = np.asarray(mesh_override, np.complex128)
        else:
            self.reflectivity, self.seeds = mesh.reflectivity_and_seeds(M, seed)
            self.mesh = mesh.reference_mesh(M, self.reflectivity, self.seeds)
        self.inputs = list(range(self.n_input_states))
        eta = 1.0 - float(parameters["loss"])
        self.state = engine.initialize(M, d, chi, self.inputs, eta,
                                       post_selected=post_selected_photon_number,
                                       device=self.device)
        self.prob[0] = self.state.observables()[0]
        self.initialization_successful = bool(self.prob[0] > 0)
        if not self.initialization_successful:
            print("Configuration yields invalid probabilities. Exiting")
        return self.initialization_successful

    def Simulate(self):
        if self.state is None:
            raise RuntimeError("Simulate() before ExperimentInit()")
        success = True
        if self.initialization_successful:
            prob, EE = engine.simulate(self.state, self.mesh, self.inputs,
                                       precision=self.precision)
            # layers with no update (beyond the light cone) keep the last values
            for k in range(1, self.n_modes + 1):
                if np.isnan(prob[k]):
                    prob[k], EE[:, k] = prob[k - 1], EE[:, k - 1]
            self.prob[:] = prob
            self.EE[:] = EE
        return success, self.prob, self.EE, self.REPar
