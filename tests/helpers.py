"""Shared test helpers: build the oracle's view of a synthetic job."""
import numpy as np

from oracle import fp64 as O
from treetime_amd import synth
from treetime_amd.prepare import prepare


def oracle_gtr(gtr, alphabet):
    """Oracle model with the same (already normalised) parameters."""
    return O.GTRModel(alphabet, mu=gtr.mu, pi=gtr.Pi, W=gtr.W)


def build_oracle(prep, alphabet, gtr, names=None):
    """OracleTreeAnc over the compressed patterns of ``prep``."""
    parent = prep["parent"]
    n = len(parent)
    is_tip = np.ones(n, dtype=bool)
    is_tip[parent[parent >= 0]] = False
    tips = np.where(is_tip)[0]
    if names is None:
        names = ["n%d" % i for i in range(n)]
    # branch lengths are the already-floored effective lengths; full length
    # is set huge so the oracle's own floor is a no-op
    root, nodes = O.build_tree(parent, prep["t"], names)
    chars = prep["tip_codes"].view("S1").astype("U1")
    comp = {names[v]: chars[k] for k, v in enumerate(tips)}
    og = oracle_gtr(gtr, alphabet)
    ta = O.OracleTreeAnc(root, og, comp, prep["multiplicity"], full_length=1e300)
    return ta, nodes, tips


def small_job(n_tips, L, n_states=4, seed=1, depth=0.02, p_ambig=0.01, p_gap=0.005):
    job = synth.make_job(n_tips, L, depth, n_states, seed, p_ambig, p_gap)
    prep = prepare(job["parent"], job["branch_length"], job["chars"], job["gtr"], job["ambiguous"])
    alphabet = "nuc_nogap" if n_states == 4 else "aa_nogap"
    return job, prep, alphabet
