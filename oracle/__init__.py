"""FP64 CPU oracle for TreeTime's Felsenstein-pruning path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``treetime_amd`` imports this package;
only ``tests/``, ``__graft_entry__.smoke()`` (as the checker) and the
``cpu_baseline`` leg of ``bench.py`` may use it.  It is a plain numpy
restatement of the reference algorithm (TreeTime 0.12.1), written to mirror
the reference's per-node call structure so it can also be timed as the CPU
baseline.

Modules
-------
``fp64``      alphabets, GTR, profiles, compression, marginal/joint passes,
              ancestral likelihood, branch-length helpers.

Every function cites the reference file:line it restates.  Pinning of each
function (golden vectors, brute force, closed forms) lives in
``tests/test_oracle_*.py``; see the pin table in the module docstring of
``oracle/fp64.py``.
"""
