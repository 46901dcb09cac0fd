"""Writes tests/golden/oracle_small.npz: inputs and FP64 oracle outputs of two
seeded small configurations (a 4-state and a 20-state job), the golden
vectors the tests compare the oracle and the HIP engine against.

Run from the repository root: python tests/golden/make_oracle_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import fp64 as O  # noqa: E402
from treetime_amd import synth  # noqa: E402
from treetime_amd.prepare import prepare  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_small.npz")


def case(tag, n_tips, L, n_states, seed, depth):
    job = synth.make_job(n_tips, L, depth, n_states, seed, 0.01, 0.005)
    prep = prepare(job["parent"], job["branch_length"], job["chars"], job["gtr"], job["ambiguous"])
    alphabet = "nuc_nogap" if n_states == 4 else "aa_nogap"
    g = job["gtr"]
    parent = prep["parent"]
    names = ["n%d" % i for i in range(len(parent))]
    root, nodes = O.build_tree(parent, prep["t"], names)
    is_tip = np.ones(len(parent), dtype=bool)
    is_tip[parent[parent >= 0]] = False
    chars = prep["tip_codes"].view("S1").astype("U1")
    comp = {names[v]: chars[k] for k, v in enumerate(np.where(is_tip)[0])}
    ta = O.OracleTreeAnc(root, O.GTRModel(alphabet, mu=g.mu, pi=g.Pi, W=g.W), comp, prep["multiplicity"],
                         full_length=1e300)
    lh = ta.marginal()
    internal = [i for i, nd in enumerate(nodes) if not nd.is_terminal()]
    prof = np.stack([nodes[i].marginal_profile for i in internal])
    mstates = np.stack([nodes[i].seq_idx for i in internal])
    seq_lh = ta.sequence_LH.copy()
    jl = ta.joint()
    jstates = np.stack([nodes[i].seq_idx for i in internal])
    jmargin = np.stack([nodes[i].path_margin for i in internal])
    return {
        f"{tag}_parent": parent, f"{tag}_t": prep["t"], f"{tag}_tip_codes": prep["tip_codes"],
        f"{tag}_multiplicity": prep["multiplicity"], f"{tag}_pi": g.Pi, f"{tag}_W": g.W, f"{tag}_mu": np.array(g.mu),
        f"{tag}_eigvals": g.eigenvals, f"{tag}_v": g.v, f"{tag}_v_inv": g.v_inv,
        f"{tag}_internal": np.array(internal), f"{tag}_logL": np.array(lh), f"{tag}_sequence_LH": seq_lh,
        f"{tag}_profiles": prof, f"{tag}_marginal_states": mstates.astype(np.uint8),
        f"{tag}_joint_logL": np.array(jl), f"{tag}_joint_states": jstates.astype(np.uint8),
        f"{tag}_joint_margin": jmargin,
    }


if __name__ == "__main__":
    data = {}
    data.update(case("nuc", 40, 400, 4, 501, 0.05))
    data.update(case("aa", 20, 120, 20, 502, 0.3))
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
