"""Experiment-queue driver with the reference's on-disk formats (run.py:81-216).

  experiments file : list of dicts with keys status ('incomplete' / 'in
                     progress'), input_state_type, n_modes, n_input_states,
                     post_selected_photon_number, local_hilbert_space_dimension,
                     bond_dimension, parameters (experiments_gen.py:57-79).
                     JSON here (`.json`); a `.pickle` list written by the
                     user's own generator is read as well.
  results          : Results/<type>/n_<M>_m_<n>_PS_<ps>_d_<d>_p_<parameters>_
                     + chi_<chi>_prob.npy (trace per layer), chi_<chi>_EE.npy
                     (entropy per bond and layer), chi.npy (bond dimensions run)
  bond dimension   : a configuration already run is re-run with chi doubled
                     when max(prob)/min(prob) - 1 over the positive entries
                     exceeds 0.1; it is skipped above chi 20000 or 200 modes.

One process drives one GPU (`python -m bosonsampling_amd.run --experiments
experiments.json`); several processes share the queue through a file lock.
"""
from __future__ import annotations

import argparse
import json
import os
import pickle
import time

import numpy as np

ERROR_LIMIT = 0.1
MAX_CHI = 20000
MAX_MODES = 200
# The library's tridiagonal eigensolver route takes sectors up to n = 2048 and a
# sector's n is at most chi, so chi doubling stops there: a larger chi could
# produce a sector the library rejects (bsmpo_site_update returns < 0).
LIB_MAX_CHI = 2048


def results_prefix(exp: dict, root: str = "Results") -> str:
    """run.py:133 file-name prefix."""
    return os.path.join(root, str(exp["input_state_type"]),
                        "n_{}_m_{}_PS_{}_d_{}_p_{}_".format(
                            exp["n_modes"], exp["n_input_states"],
                            exp["post_selected_photon_number"],
                            exp["local_hilbert_space_dimension"], exp["parameters"]))


def trace_error(prob: np.ndarray) -> float:
    """max / min - 1 over the positive traces (run.py:142-146)."""
    p = np.asarray(prob)
    p = p[p > 0]
    if p.size == 0:
        return float("inf")
    return float(p.max() / p.min() - 1)


def choose_bond_dimension(exp: dict, root: str = "Results"):
    """('run', chi) or ('skip', chi) following run.py:135-168."""
    prefix = results_prefix(exp, root)
    chi = int(exp["bond_dimension"])
    status = "run"
    if os.path.isfile(prefix + "chi.npy"):
        chi = int(np.max(np.load(prefix + "chi.npy")))
        prob = np.load(prefix + "chi_{}_prob.npy".format(chi))
        if trace_error(prob) > ERROR_LIMIT:
            chi *= 2
        else:
            status = "skip"
    if status == "run" and (chi > min(MAX_CHI, LIB_MAX_CHI) or int(exp["n_modes"]) > MAX_MODES):
        status = "skip"
    return status, chi


def save_results(exp: dict, chi: int, prob: np.ndarray, EE: np.ndarray, root: str = "Results"):
    """run.py:193-209: results are written only when some trace is positive."""
    prefix = results_prefix(exp, root)
    if not np.sum(np.asarray(prob) > 0) > 0:
        return False
    os.makedirs(os.path.dirname(prefix), exist_ok=True)
    chis = np.load(prefix + "chi.npy") if os.path.isfile(prefix + "chi.npy") else np.array([])
    np.save(prefix + "chi_{}_prob.npy".format(chi), prob)
    np.save(prefix + "chi_{}_EE.npy".format(chi), EE)
    np.save(prefix + "chi.npy", np.append(chis, chi))
    return True


def load_experiments(path: str, allow_pickle: bool = False) -> list:
    """JSON queue; a pickled queue (the reference's experiments.pickle) only
    with allow_pickle=True, since unpickling a file runs code from it."""
    if path.endswith(".json"):
        with open(path) as f:
            return json.load(f)
    if not allow_pickle:
        raise ValueError(f"{path}: not a .json queue; convert it, or pass allow_pickle=True "
                         "(--allow-pickle) for a pickled queue you trust")
    with open(path, "rb") as f:
        return pickle.load(f)


def store_experiments(path: str, experiments: list):
    if path.endswith(".json"):
        with open(path, "w") as f:
            json.dump(experiments, f, indent=1)
    else:
        with open(path, "wb") as f:
            pickle.dump(experiments, f)


def claim_next(path: str, start: int = 0, allow_pickle: bool = False):
    """Index of the first 'incomplete' experiment at or after `start`, marked
    'in progress' in the file under its lock (run.py:93-127), or None."""
    from filelock import FileLock
    with FileLock(path + ".lock"):
        exps = load_experiments(path, allow_pickle)
        for k in range(start, len(exps)):
            if exps[k].get("status") == "incomplete":
                exps[k]["status"] = "in progress"
                store_experiments(path, exps)
                return k, dict(exps[k])
    return None


def release(path: str, k: int, status: str, allow_pickle: bool = False):
    from filelock import FileLock
    with FileLock(path + ".lock"):
        exps = load_experiments(path, allow_pickle)
        exps[k]["status"] = status
        store_experiments(path, exps)


def run_queue(path: str, make_worker, root: str = "Results", seed=None, max_passes: int = 8,
              log=print, allow_pickle: bool = False):
    """Work through the queue: each claimed experiment is run at the bond
    dimension chosen from its earlier results, re-queued ('incomplete') while
    its trace error is too large, and marked 'complete' otherwise.  An
    experiment whose run raises is marked 'failed' (not left 'in progress')
    and the queue goes on."""
    ran = 0
    for _ in range(max_passes):
        k, progress = 0, False
        while True:
            got = claim_next(path, k, allow_pickle)
            if got is None:
                break
            k, exp = got
            status, chi = choose_bond_dimension(exp, root)
            if status == "skip":
                release(path, k, "complete", allow_pickle)
                k += 1
                continue
            t0 = time.time()
            try:
                w = make_worker()
                w.ExperimentInit(exp["input_state_type"], exp["n_modes"], exp["n_input_states"],
                                 exp["post_selected_photon_number"],
                                 exp["local_hilbert_space_dimension"], chi, exp["parameters"],
                                 seed)
                success, prob, EE, _ = w.Simulate()
            except Exception as e:  # noqa: BLE001 - any failure of one experiment
                log(f"experiment {k}: chi={chi} failed: {type(e).__name__}: {e}")
                release(path, k, "failed", allow_pickle)
                k += 1
                continue
            saved = save_results(exp, chi, prob, EE, root)
            err = trace_error(prob)
            log(f"experiment {k}: chi={chi} trace error {err:.3e} saved={saved} "
                f"{time.time() - t0:.1f}s")
            again = saved and err > ERROR_LIMIT and 2 * chi <= min(MAX_CHI, LIB_MAX_CHI)
            release(path, k, "incomplete" if again else "complete", allow_pickle)
            ran += 1
            progress = True
            k += 1
        if not progress:
            break
    return ran


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--experiments", default="experiments.json")
    ap.add_argument("--results", default="Results")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--precision", default="c128", choices=["c128", "c64", "bf16x3"])
    ap.add_argument("--allow-pickle", action="store_true",
                    help="read a pickled queue (runs code from the file: trusted files only)")
    args = ap.parse_args()
    from . import engine
    from .worker import Worker
    prec = {"c128": engine.C128, "c64": engine.C64, "bf16x3": engine.BF16X3}[args.precision]
    engine.set_precision_rules(prec)
    run_queue(args.experiments, lambda: Worker(prec), args.results, args.seed,
              allow_pickle=args.allow_pickle)


if __name__ == "__main__":
    main()
