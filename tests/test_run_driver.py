"""CPU tests of the experiment-queue driver: file names, the bond-dimension
doubling rule and the queue states (run.py:81-216 formats)."""
import json
import os

import numpy as np
import pytest

from bosonsampling_amd import run

EXP = {"status": "incomplete", "input_state_type": "Single photon", "n_modes": 8,
       "n_input_states": 4, "post_selected_photon_number": None,
       "local_hilbert_space_dimension": 5, "bond_dimension": 16,
       "parameters": {"loss": 0.5}}


def test_prefix_matches_reference_naming(tmp_path):
    p = run.results_prefix(EXP, str(tmp_path))
    assert p.endswith(os.path.join("Single photon",
                                   "n_8_m_4_PS_None_d_5_p_{'loss': 0.5}_"))


def test_trace_error():
    assert run.trace_error(np.array([0.0, 1.0, 1.05, 0.0])) == np.float64(1.05) - 1
    assert run.trace_error(np.zeros(3)) == float("inf")


def test_bond_dimension_policy(tmp_path):
    root = str(tmp_path)
    assert run.choose_bond_dimension(EXP, root) == ("run", 16)
    run.save_results(EXP, 16, np.array([1.0, 0.8, 0.7]), np.zeros((7, 3)), root)
    assert run.choose_bond_dimension(EXP, root) == ("run", 32)      # error 0.43 > 0.1
    run.save_results(EXP, 32, np.array([1.0, 0.99, 0.98]), np.zeros((7, 3)), root)
    assert run.choose_bond_dimension(EXP, root) == ("skip", 32)
    assert list(np.load(run.results_prefix(EXP, root) + "chi.npy")) == [16.0, 32.0]
    big = dict(EXP, bond_dimension=40000)
    assert run.choose_bond_dimension(big, root + "/other") == ("skip", 40000)
    # nothing positive: not saved
    assert not run.save_results(EXP, 64, np.zeros(3), np.zeros((7, 3)), root)


class FakeWorker:
    """Trace error shrinks with chi: 2.4 / chi."""

    def ExperimentInit(self, input_state_type, n_modes, n_input_states, ps, d, chi, parameters,
                       seed=None):
        self.chi, self.M = chi, n_modes

    def Simulate(self):
        x = 2.4 / self.chi
        prob = np.linspace(1.0, 1.0 / (1 + x), self.M + 1)
        return True, prob, np.zeros((self.M - 1, self.M + 1)), np.zeros((self.M - 1, self.M + 1, 5))


def test_queue_doubles_until_accurate(tmp_path):
    path = str(tmp_path / "experiments.json")
    with open(path, "w") as f:
        json.dump([EXP, dict(EXP, status="complete")], f)
    ran = run.run_queue(path, FakeWorker, str(tmp_path / "Results"), log=lambda *a: None)
    chis = np.load(run.results_prefix(EXP, str(tmp_path / "Results")) + "chi.npy")
    assert list(chis) == [16.0, 32.0]       # error 0.15 at chi 16, 0.075 at chi 32
    assert ran == 2
    with open(path) as f:
        assert [e["status"] for e in json.load(f)] == ["complete", "complete"]


class FailingWorker(FakeWorker):
    def Simulate(self):
        raise RuntimeError("sector too large for the tridiagonal route")


def test_failed_experiment_is_released(tmp_path):
    """A run that raises leaves its experiment 'failed', not 'in progress',
    and the queue goes on to the next one."""
    path = str(tmp_path / "experiments.json")
    with open(path, "w") as f:
        json.dump([EXP, dict(EXP, n_modes=8)], f)
    ran = run.run_queue(path, FailingWorker, str(tmp_path / "Results"), log=lambda *a: None)
    assert ran == 0
    with open(path) as f:
        assert [e["status"] for e in json.load(f)] == ["failed", "failed"]


def test_chi_doubling_stops_at_the_library_limit(tmp_path):
    root = str(tmp_path)
    exp = dict(EXP, bond_dimension=run.LIB_MAX_CHI)
    run.save_results(exp, run.LIB_MAX_CHI, np.array([1.0, 0.5, 0.4]), np.zeros((7, 3)), root)
    assert run.choose_bond_dimension(exp, root)[0] == "skip"


def test_pickled_queue_needs_opt_in(tmp_path):
    import pickle
    path = str(tmp_path / "experiments.pickle")
    with open(path, "wb") as f:
        pickle.dump([EXP], f)
    with pytest.raises(ValueError):
        run.load_experiments(path)
    assert run.load_experiments(path, allow_pickle=True)[0]["n_modes"] == EXP["n_modes"]
