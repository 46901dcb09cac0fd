"""Compact digests of the FP64 oracle's results on the full-size workloads.

The oracle needs minutes per large job, so its results on the full-size
workloads are computed once on the CPU (tests/golden/make_config_golden.py)
and committed as a small fixture; the GPU tests rebuild the same seeded jobs,
check that the inputs hash to the same digest, and compare the engine's
results with the fixture:

  * total and per-pattern log-likelihood (a seeded sample of patterns);
  * marginal profile, subtree and outgroup rows at a seeded sample of
    (internal node, pattern) pairs;
  * every reconstructed state, through one CRC32 per group of 32 consecutive
    node rows of the state matrix, with the near-tie positions (FP64 margin
    <= 1e-5) blanked to 255.
"""
import hashlib
import zlib

import numpy as np

from treetime_amd import synth

MARGIN = 1e-5
N_PATTERN_SAMPLE = 256
N_ROW_SAMPLE = 256

# the jobs of the 256-job mix held to the oracle: 16 ranks spread over the
# size range (tips x sites), smallest to largest
MIX_SAMPLE = 16


def mix_sample_jobs():
    specs = [synth.mix_job_spec(j) for j in range(256)]
    order = sorted(range(256), key=lambda j: (specs[j]["n_tips"] * specs[j]["L"], j))
    ranks = np.round(np.linspace(0, 255, MIX_SAMPLE)).astype(int)
    return [order[r] for r in ranks]


def config_specs():
    """name -> job spec of every workload with an oracle digest."""
    out = {}
    for name in ("large-5000", "large-5000-deep", "aa-1000"):
        n_tips, L, depth, n, seed, pa, pg = synth.WORKLOADS[name]
        out[name] = dict(n_tips=n_tips, L=L, depth=depth, n_states=n, seed=seed, p_ambig=pa, p_gap=pg)
    for j in mix_sample_jobs():
        out["mix%03d" % j] = synth.mix_job_spec(j)
    return out


def make_prepared(spec):
    """The job and its prepared (compressed) form, as the engine gets it."""
    from treetime_amd.prepare import prepare

    job = synth.make_job(spec["n_tips"], spec["L"], spec["depth"], spec["n_states"], spec["seed"],
                         spec.get("p_ambig", 0.0), spec.get("p_gap", 0.0))
    prep = prepare(job["parent"], job["branch_length"], job["chars"], job["gtr"], job["ambiguous"])
    return job, prep


def prepared_mix_job(j):
    """Job j of the 256-job mix, prepared (a pool worker's task)."""
    _, prep = make_prepared(synth.mix_job_spec(j))
    prep.pop("full_to_compressed", None)
    return prep


def input_hash(prep):
    h = hashlib.sha256()
    for a, dt in ((prep["parent"], np.int32), (prep["t"], np.float64), (prep["tip_codes"], np.uint8),
                  (prep["multiplicity"], np.float64)):
        h.update(np.ascontiguousarray(a, dtype=dt).tobytes())
    return h.hexdigest()


def internal_nodes(parent):
    is_tip = np.ones(len(parent), dtype=bool)
    is_tip[parent[parent >= 0]] = False
    return np.where(~is_tip)[0]


def samples(parent, P, seed):
    """Seeded pattern sample and (internal node, pattern) sample."""
    rng = np.random.default_rng(seed)
    pats = np.sort(rng.choice(P, size=min(P, N_PATTERN_SAMPLE), replace=False))
    ints = internal_nodes(parent)
    nodes = ints[rng.integers(0, len(ints), N_ROW_SAMPLE)]
    rows = rng.integers(0, P, N_ROW_SAMPLE)
    return pats, nodes, rows


CRC_GROUP = 32


def state_crcs(states, ties):
    """CRC32 of every group of CRC_GROUP node rows of the state matrix (uint8
    [n_nodes][P]) with the flat positions ``ties`` (node * P + pattern) set
    to 255."""
    s = np.array(states, dtype=np.uint8, copy=True)
    s.reshape(-1)[np.asarray(ties, dtype=np.int64)] = 255
    return np.array([zlib.crc32(s[g:g + CRC_GROUP].tobytes()) for g in range(0, len(s), CRC_GROUP)],
                    dtype=np.uint32)
