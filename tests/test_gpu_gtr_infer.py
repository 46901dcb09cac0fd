"""GTR inference statistics on the GPU (tt_mutation_counts,
tt_branch_state_counts) and TreeAnc.infer_gtr against the oracle.

Tolerances: marginal statistics come from FP32 profiles (<= 2e-7 absolute
from the oracle's), so n_ij and T_i agree to 1e-6 relative and the inferred
model to 1e-5; joint statistics are exact integer-weighted counts, so the
inferred model agrees to 1e-10."""
import sys

import numpy as np
import pytest

import treetime_amd as tt
from helpers import build_oracle, oracle_gtr, small_job
from oracle import fp64 as O
from treetime_amd import synth
from treetime_amd.prepare import device_tree

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("n_tips,L,seed", [(60, 700, 91), (200, 1500, 92)])
def test_mutation_counts_match_oracle(gpu_engine, n_tips, L, seed):
    say(f"[gtr-infer] marginal counts tips={n_tips} L={L} seed={seed}")
    job, prep, alphabet = small_job(n_tips, L, 4, seed, depth=0.05)
    dev = device_tree(gpu_engine, prep)
    dev.marginal()
    nij, Ti, pb = dev.mutation_counts(per_branch=True)
    ta, nodes, tips = build_oracle(prep, alphabet, job["gtr"])
    ta.marginal()
    rn, rt = ta.mutation_counts_marginal()
    en = np.max(np.abs(nij - rn)) / np.max(rn)
    et = np.max(np.abs(Ti - rt) / rt)
    say(f"  n_ij rel err {en:.2e}  T_i rel err {et:.2e}")
    assert en < 1e-6 and et < 1e-6
    assert np.allclose(pb.sum(axis=0)[:16].reshape(4, 4), nij, rtol=1e-12)
    dev.close()


def _facade_and_oracle(seed):
    job = synth.make_job(80, 800, 0.05, 4, seed, 0.02, 0.01)
    parent, bl = job["parent"], job["branch_length"]
    clades = [tt.Clade(branch_length=float(bl[i])) for i in range(len(parent))]
    for i, p in enumerate(parent):
        if p >= 0:
            clades[p].clades.append(clades[i])
    for k, v in enumerate(job["tips"]):
        clades[v].name = job["names"][k]
    aln = {job["names"][k]: job["chars"][k].tobytes().decode() for k in range(len(job["tips"]))}
    ta = tt.TreeAnc(tree=tt.Tree(clades[0]), aln=aln, gtr=job["gtr"], verbose=0)
    nodes = list(ta.tree.find_clades(order="preorder"))
    idx = {id(x): i for i, x in enumerate(nodes)}
    par = [-1] * len(nodes)
    for i, x in enumerate(nodes):
        for c in x.clades:
            par[idx[id(c)]] = i
    root, onodes = O.build_tree(par, [ta._branch_length_to_gtr(x) for x in nodes], [x.name for x in nodes])
    g = ta.gtr
    orc = O.OracleTreeAnc(root, O.GTRModel("nuc_nogap", mu=g.mu, pi=g.Pi, W=g.W), ta.data.compressed_alignment,
                          ta.data.multiplicity(), full_length=ta.data.full_length)
    return ta, orc, onodes


def test_facade_infer_gtr_marginal():
    """TreeAnc.infer_gtr(marginal=True) (treeanc.py:1500-1632) equals
    GTR.infer on the oracle's statistics."""
    say("[gtr-infer] facade, marginal")
    ta, orc, onodes = _facade_and_oracle(93)
    ta.infer_ancestral_sequences(marginal=True)
    root_seq = ta.tree.root.cseq.copy()
    g = ta.infer_gtr(marginal=True, pc=5.0)
    orc.marginal()
    nij, Ti = orc.mutation_counts_marginal()
    alpha = list(ta.gtr.alphabet)
    root_state = np.array([np.sum((root_seq == a) * ta.data.multiplicity()) for a in alpha])
    ref = O.gtr_infer("nuc_nogap", nij, Ti, root_state, pc=5.0)
    say(f"  W max rel diff {np.max(np.abs(g.W - ref.W)) / ref.W.max():.2e}, "
        f"pi max abs diff {np.max(np.abs(g.Pi - ref.Pi)):.2e}")
    assert np.allclose(g.W, ref.W, rtol=1e-5, atol=1e-8)
    assert np.allclose(g.Pi, ref.Pi, rtol=1e-5, atol=1e-8)
    assert g.mu == 1.0  # normalized_rate
    # the new model is used by the next reconstruction
    nd = ta.infer_ancestral_sequences(marginal=True, infer_gtr=True)
    assert nd >= 0 and np.isfinite(ta.sequence_LH())


def test_facade_infer_gtr_joint():
    """infer_gtr after a joint reconstruction: counts of state changes per
    branch and time in each state (treeanc.py:1573-1589) from the device
    pair counts equal those from the sequences."""
    say("[gtr-infer] facade, joint")
    ta, orc, onodes = _facade_and_oracle(94)
    ta.infer_ancestral_sequences(marginal=False)
    seqs = {id(c): c.cseq.copy() for c in ta.tree.find_clades()}
    g = ta.infer_gtr(marginal=False, pc=2.0, normalized_rate=False)
    alpha = list(ta.gtr.alphabet)
    idx = {a: i for i, a in enumerate(alpha)}
    mult = ta.data.multiplicity()
    n = len(alpha)
    nij, Ti = np.zeros((n, n)), np.zeros(n)
    for c in ta.tree.find_clades():
        if c.up is None:
            continue
        t = ta._branch_length_to_gtr(c)
        for a, d, w in zip(seqs[id(c.up)], seqs[id(c)], mult):
            if d not in idx:
                continue
            i, j = idx[d], idx[a]
            Ti[i] += t * w
            if i != j:
                nij[i, j] += w
                Ti[j] += 0.5 * t * w
                Ti[i] -= 0.5 * t * w
    root_state = np.array([np.sum((seqs[id(ta.tree.root)] == a) * mult) for a in alpha])
    ref = O.gtr_infer("nuc_nogap", nij, Ti, root_state, pc=2.0)
    assert np.allclose(g.W, ref.W, rtol=1e-10, atol=0)
    assert np.allclose(g.Pi, ref.Pi, rtol=1e-10, atol=0)
    assert abs(g.mu - ref.mu) <= 1e-10 * ref.mu
