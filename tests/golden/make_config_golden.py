"""Writes tests/golden/configs_golden.npz: the FP64 oracle's results on the
full-size workloads (5000 tips x 30000 sites at both depths, the 20-state
1000 x 5000 job, and 16 jobs of the 256-job mix spread over its size range),
as compact digests (tests/config_digest.py).  Run on the CPU:

    python tests/golden/make_config_golden.py [--workers 2]

Each job takes the oracle up to a few minutes and ~20 GB at the largest."""
import argparse
import gc
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
sys.path.insert(0, os.path.dirname(TESTS))
sys.path.insert(0, TESTS)

import config_digest as CD  # noqa: E402
from helpers import build_oracle  # noqa: E402
from treetime_amd import synth  # noqa: E402

OUT = os.path.join(HERE, "configs_golden.npz")


def digest(item):
    name, spec = item
    t0 = time.time()
    job, prep = CD.make_prepared(spec)
    parent = prep["parent"]
    P = prep["tip_codes"].shape[1]
    n_nodes = len(parent)
    alphabet = synth.ALPHABET_OF[spec["n_states"]]
    pats, snodes, srows = CD.samples(parent, P, spec["seed"])
    out = {"hash": np.array(CD.input_hash(prep)), "P": np.array(P), "n_nodes": np.array(n_nodes),
           "pats": pats, "snodes": snodes, "srows": srows}

    ta, nodes, tips = build_oracle(prep, alphabet, job["gtr"])
    out["m_logL"] = np.array(ta.marginal())
    out["m_slh"] = ta.sequence_LH[pats]
    out["m_prof"] = np.stack([nodes[v].marginal_profile[p] for v, p in zip(snodes, srows)])
    out["m_sub"] = np.stack([nodes[v].marginal_subtree_LH[p] for v, p in zip(snodes, srows)])
    out["m_outg"] = np.stack([nodes[v].marginal_outgroup_LH[p] if nodes[v].up is not None else ta._root_prior()[p]
                              for v, p in zip(snodes, srows)])
    states = np.full((n_nodes, P), 255, dtype=np.uint8)
    ties = []
    for v, nd in enumerate(nodes):
        if nd.is_terminal():
            continue
        states[v] = nd.seq_idx
        s = np.sort(nd.marginal_profile, axis=1)
        ties.append(v * P + np.nonzero(s[:, -1] - s[:, -2] <= CD.MARGIN)[0])
    out["m_ties"] = np.concatenate(ties).astype(np.int64)
    out["m_crc"] = CD.state_crcs(states, out["m_ties"])
    del ta, nodes, tips, states
    gc.collect()

    ta, nodes, tips = build_oracle(prep, alphabet, job["gtr"])
    out["j_logL"] = np.array(ta.joint())
    out["j_slh"] = ta.joint_sequence_LH[pats]
    states = np.full((n_nodes, P), 255, dtype=np.uint8)
    ties = []
    for v, nd in enumerate(nodes):
        if nd.is_terminal():
            continue
        states[v] = nd.seq_idx
        ties.append(v * P + np.nonzero(nd.path_margin <= CD.MARGIN)[0])
    out["j_ties"] = np.concatenate(ties).astype(np.int64)
    out["j_crc"] = CD.state_crcs(states, out["j_ties"])
    del ta, nodes, tips, states
    gc.collect()
    print(f"{name}: tips {spec['n_tips']} L {spec['L']} P {P}: marginal ties {len(out['m_ties'])}, "
          f"joint ties {len(out['j_ties'])}, {time.time() - t0:.0f} s", flush=True)
    return name, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=2)
    ap.add_argument("--only", default="", help="comma-separated names (default: all)")
    args = ap.parse_args()
    specs = CD.config_specs()
    names = args.only.split(",") if args.only else list(specs)
    items = sorted(((k, specs[k]) for k in names), key=lambda kv: -kv[1]["n_tips"] * kv[1]["L"])
    res = {}
    if os.path.exists(OUT):
        with np.load(OUT) as z:
            res = {k: z[k] for k in z.files}
    with mp.get_context("fork").Pool(args.workers, maxtasksperchild=1) as pool:
        for name, out in pool.imap_unordered(digest, items):
            for k, v in out.items():
                res[f"{name}/{k}"] = v
            np.savez_compressed(OUT, **res)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
