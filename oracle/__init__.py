"""CPU FP64 oracle for the lossy boson-sampling MPO beam-splitter site update.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
package: it is imported by ``tests/``, by ``__graft_entry__.smoke()`` and by
the ``cpu_baseline`` leg of ``bench.py``, and there only as the checker (or as
the timed CPU baseline), never as the thing measured or shipped.

It is a numpy/scipy complex128 restatement of the reference algorithm
(sss441803/BosonSampling).  Each function cites the reference file:line it
follows.  The reference's own CPU path (``src/cpu_algorithm``) is an empty
file and the reference ships no tests, fixtures or golden vectors, so every
function here is pinned instead by closed forms, invariants, library routines
and brute force (see ``tests/test_oracle_*.py``):

* ``fock``   -- beam-splitter Fock matrix: reference binomial formula
  (src/MPO.py:6-18, src/u_kernel.cu:37-70) == general-U(2) formula ==
  permanent formula; unitarity per photon-number sector; HOM dip.
* ``mpo``    -- product-state init + canonicalisation (src/MPO.py:21-303),
  the site update (src/beam_splitter_parallel/Rank.py:32-206), the layer
  schedule (src/beam_splitter_parallel/Full.py:113-166) and observables
  (src/summary.py:3-36): pinned against ``brute`` (dense Fock density
  matrix, no truncation), against permanents (single-photon marginals
  <n_j> = eta * sum_i |W_ji|^2, pattern probabilities) and trace = 1.
* ``brute``  -- dense density-matrix simulator for M <= 5 modes.

Parity status: every function above has at least one such check; none is
"parity unpinned".  Bitwise reference rounding (complex64 cuSOLVER output)
and argpartition tie order are gauge/rounding and are not reproduced.
"""
