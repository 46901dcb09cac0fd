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


def small_job(n_tips, L, n_states=4, seed=1, depth=0.02, p_ambig=0.01, p_gap=0.005, p_iupac=0.0):
    job = synth.make_job(n_tips, L, depth, n_states, seed, p_ambig, p_gap, p_iupac)
    prep = prepare(job["parent"], job["branch_length"], job["chars"], job["gtr"], job["ambiguous"])
    return job, prep, synth.ALPHABET_OF[n_states]


def collapse_job(job, frac, seed, root_extra=True):
    """Multifurcating variant of a synthetic job: a fraction ``frac`` of the
    internal non-root nodes is removed, their children attached to the
    nearest kept ancestor with the removed branch length added (a polytomy
    wherever a node is removed); ``root_extra`` also removes one internal
    child of the root, giving the root three or more children.  Tip rows
    keep their order."""
    parent, bl = job["parent"], job["branch_length"].astype(float).copy()
    n = len(parent)
    is_tip = np.ones(n, dtype=bool)
    is_tip[parent[parent >= 0]] = False
    internal = [v for v in range(n) if not is_tip[v] and parent[v] >= 0]
    rng = np.random.default_rng(seed)
    drop = set(int(x) for x in rng.choice(internal, size=int(frac * len(internal)), replace=False))
    if root_extra:
        root = int(np.where(parent < 0)[0][0])
        kids = [v for v in internal if parent[v] == root]
        if kids:
            drop.add(kids[0])
    keep = [v for v in range(n) if v not in drop]
    idx = {v: i for i, v in enumerate(keep)}
    new_parent = np.full(len(keep), -1, dtype=np.int32)
    new_bl = np.zeros(len(keep))
    for v in keep:
        p, add = int(parent[v]), 0.0
        while p >= 0 and p in drop:
            add += bl[p]
            p = int(parent[p])
        new_parent[idx[v]] = idx[p] if p >= 0 else -1
        new_bl[idx[v]] = bl[v] + add
    out = dict(job)
    out["parent"], out["branch_length"] = new_parent, new_bl
    out["tips"] = np.array([idx[int(v)] for v in job["tips"]])
    return out


def polytomy_job(n_tips, L, n_states=4, seed=1, frac=0.3, depth=0.05, p_ambig=0.01, p_gap=0.005, p_iupac=0.0):
    job = collapse_job(synth.make_job(n_tips, L, depth, n_states, seed, p_ambig, p_gap, p_iupac), frac, seed + 1)
    prep = prepare(job["parent"], job["branch_length"], job["chars"], job["gtr"], job["ambiguous"])
    return job, prep, synth.ALPHABET_OF[n_states]


def sparse_case(seed, nseq, L, amb="N", restate=0.05):
    """A reference + differences alignment with every class of position of
    vcf_utils.py:694-718: private and shared variants, positions where every
    sequence differs (with and without the ambiguous character), positions
    ambiguous in a few sequences, ambiguous reference positions, and a few
    entries that restate the reference character.  Each sequence's dict is
    in shuffled position order.

    Returns (names, diffs, ref (L,) chars, dense (nseq, L) chars)."""
    rng = np.random.default_rng(seed)
    ref = rng.choice(np.array(list("ACGT" + amb + "-")), size=L, p=[0.24, 0.24, 0.24, 0.24, 0.02, 0.02])
    A = np.repeat(ref[None, :], nseq, axis=0)
    mut = rng.random((nseq, L)) < 0.03
    A[mut] = rng.choice(np.array(list("ACGT-")), size=int(mut.sum()))
    cols = rng.choice(L, size=max(1, L // 10), replace=False)
    third = len(cols) // 3
    A[:, cols[:third]] = rng.choice(np.array(list("ACGT")), size=third)[None, :]       # all differ, one allele
    sub = cols[third:2 * third]
    A[:, sub] = rng.choice(np.array(list("ACGT")), size=len(sub))[None, :]
    hit = rng.random((nseq, len(sub))) < 0.3
    blk = A[:, sub]
    blk[hit] = amb                                                                   # allele + ambiguous
    A[:, sub] = blk
    A[:, cols[2 * third:2 * third + 2]] = amb                                        # all ambiguous
    amb_hit = rng.random((nseq, L)) < 0.01
    A[amb_hit] = amb
    names = ["s%d" % i for i in range(nseq)]
    diffs = {}
    for i, nm in enumerate(names):
        pos = np.flatnonzero(A[i] != ref)
        extra = np.flatnonzero((A[i] == ref) & (rng.random(L) < restate / 10))
        pos = np.concatenate((pos, extra))
        rng.shuffle(pos)
        diffs[nm] = {int(p): str(A[i, p]) for p in pos}
    return names, diffs, ref, A


def optimize_tree_case(seed, n_tips=40, L=300, depth=0.02, scale=1.7):
    """Oracle tree + compressed alignment of a synthetic job, branch lengths
    scaled away from their optimum."""
    from treetime_amd import synth

    job = synth.make_job(n_tips, L, depth, 4, seed, 0.01, 0.0)
    names = ["n%d" % i for i in range(len(job["parent"]))]
    root, nodes = O.build_tree(job["parent"], job["branch_length"] * scale, names)
    tip_names = [names[v] for v in job["tips"]]
    chars = job["chars"].astype("U1")
    comp, mult, _, _ = O.compress_alignment(tip_names, {nm: chars[k] for k, nm in enumerate(tip_names)}, "N")
    g = job["gtr"]
    og = O.GTRModel("nuc_nogap", mu=g.mu, pi=g.Pi, W=g.W)
    return O.OracleTreeAnc(root, og, comp, mult, full_length=L), job, names
