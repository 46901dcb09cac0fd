"""Synthetic workloads: Kingman-coalescent trees, random GTR models and
alignments simulated down the tree.

Tree: with k lineages the waiting time is Exp(k(k-1)/2) and a uniformly
random pair merges; the ultrametric tree is scaled to the requested
root-to-tip depth.  Sequences: the root is drawn from pi and every child
site from column P(t)[:, parent state] (the sampling of SeqGen.evolve,
treetime/seqgen.py:38-67).  Tip noise replaces characters by the ambiguous
character, a gap or a partial ambiguity code.
"""
import numpy as np

from .gtr import GTR
from .seq_utils import alphabets


def kingman_tree(n_tips, depth, rng):
    """Returns (parent, branch_length) with tips 0..n-1 and internal nodes
    n..2n-2 (root last); ultrametric with root height ``depth``."""
    n_nodes = 2 * n_tips - 1
    parent = np.full(n_nodes, -1, dtype=np.int32)
    height = np.zeros(n_nodes)
    lineages = list(range(n_tips))
    t = 0.0
    nxt = n_tips
    while len(lineages) > 1:
        k = len(lineages)
        t += rng.exponential(2.0 / (k * (k - 1)))
        i, j = rng.choice(k, size=2, replace=False)
        a, b = lineages[i], lineages[j]
        for x in sorted((i, j), reverse=True):
            lineages[x] = lineages[-1]
            lineages.pop()
        parent[a] = parent[b] = nxt
        height[nxt] = t
        lineages.append(nxt)
        nxt += 1
    height *= depth / height[-1]
    bl = np.zeros(n_nodes)
    nonroot = parent >= 0
    bl[nonroot] = height[parent[nonroot]] - height[nonroot]
    return parent, bl


def preorder_renumber(parent, bl):
    """Renumber nodes in preorder (root 0, children in input order) so that
    the children of each node have ascending indices."""
    n = len(parent)
    kids = [[] for _ in range(n)]
    root = int(np.where(parent < 0)[0][0])
    for i, p in enumerate(parent):
        if p >= 0:
            kids[p].append(i)
    order = []
    stack = [root]
    while stack:
        v = stack.pop()
        order.append(v)
        stack.extend(reversed(kids[v]))
    new = np.empty(n, dtype=np.int64)
    new[np.array(order)] = np.arange(n)
    parent2 = np.full(n, -1, dtype=np.int32)
    nonroot = parent >= 0
    parent2[new[nonroot]] = new[parent[nonroot]]
    bl2 = np.empty(n)
    bl2[new] = bl
    return parent2, bl2, new


def random_gtr(n_states, rng, alpha, sigma):
    """pi ~ Dirichlet(alpha), exchangeabilities ~ LogNormal(0, sigma),
    W symmetric and scaled so that pi^T W pi = 1 (rate 1 per unit length)."""
    pi = rng.dirichlet(alpha * np.ones(n_states))
    W = np.zeros((n_states, n_states))
    iu = np.triu_indices(n_states, 1)
    W[iu] = rng.lognormal(0.0, sigma, size=len(iu[0]))
    W = W + W.T
    W /= float(pi @ W @ pi)
    return pi, W


def simulate(parent, bl, gtr, L, rng):
    """(n_nodes, L) uint8 state indices of a sequence evolved down the tree."""
    n = len(parent)
    kids = [[] for _ in range(n)]
    root = int(np.where(parent < 0)[0][0])
    for i, p in enumerate(parent):
        if p >= 0:
            kids[p].append(i)
    states = np.empty((n, L), dtype=np.uint8)
    cum_pi = np.cumsum(gtr.Pi)
    states[root] = np.minimum(np.searchsorted(cum_pi, rng.random(L) * cum_pi[-1], side="right"),
                              gtr.n_states - 1)
    stack = list(kids[root])
    while stack:
        v = stack.pop()
        stack.extend(kids[v])
        P = gtr.expQt(bl[v])
        ps = states[parent[v]]
        u = rng.random(L, dtype=np.float32)
        child = ps.copy()
        move = np.nonzero(u >= P[ps, ps])[0]
        if move.size:
            Q = P.copy()
            np.fill_diagonal(Q, 0.0)
            cq = np.cumsum(Q, axis=0)
            pj = ps[move]
            u2 = rng.random(move.size) * cq[-1, pj]
            new = (u2[:, None] >= cq[:, pj].T).sum(axis=1)
            child[move] = np.minimum(new, gtr.n_states - 1)
        states[v] = child
    return states


# partial ambiguity codes of each alphabet (seq_utils.py:28-122): nucleotide
# IUPAC sets, amino-acid B (D/N) and Z (E/Q)
IUPAC_CODES = {4: b"RYSWKMDHBV", 5: b"RYSWKMDHBV", 20: b"BZ"}
ALPHABET_OF = {4: "nuc_nogap", 5: "nuc", 20: "aa_nogap"}


def make_job(n_tips, L, depth, n_states, seed, p_ambig=0.0, p_gap=0.0, p_iupac=0.0):
    """One synthetic job.  Returns a dict with the tree (preorder numbering),
    the GTR model, and the tip alignment as a (n_tips, L) char matrix whose
    rows follow ascending tip node index.  n_states 4 ('nuc_nogap'), 5
    ('nuc', the gap is a state) or 20 ('aa_nogap').  Tip noise: a fraction
    p_ambig of the characters becomes the ambiguous character, p_gap a gap
    and p_iupac a random partial ambiguity code of the alphabet."""
    rng = np.random.default_rng(seed)
    parent, bl = kingman_tree(n_tips, depth, rng)
    parent, bl, _ = preorder_renumber(parent, bl)
    if n_states in (4, 5):
        pi, W = random_gtr(n_states, rng, 5.0, 0.5)
        amb = "N"
    elif n_states == 20:
        pi, W = random_gtr(20, rng, 2.0, 1.0)
        amb = "X"
    else:
        raise ValueError("n_states must be 4, 5 or 20")
    alphabet = ALPHABET_OF[n_states]
    gtr = GTR.custom(mu=1.0, pi=pi, W=W, alphabet=alphabet)
    alpha = alphabets[alphabet]
    states = simulate(parent, bl, gtr, L, rng)
    is_tip = np.ones(len(parent), dtype=bool)
    is_tip[parent[parent >= 0]] = False
    tips = np.where(is_tip)[0]
    chars = alpha.astype("S1")[states[tips]]
    if p_ambig or p_gap or p_iupac:
        u = rng.random(chars.shape)
        chars[u < p_ambig] = amb.encode()
        chars[(u >= p_ambig) & (u < p_ambig + p_gap)] = b"-"
        if p_iupac:
            sel = (u >= p_ambig + p_gap) & (u < p_ambig + p_gap + p_iupac)
            codes = np.frombuffer(IUPAC_CODES[n_states], dtype="S1")
            chars[sel] = codes[rng.integers(0, len(codes), size=int(sel.sum()))]
    names = ["t%05d" % i for i in range(len(tips))]
    return {"parent": parent, "branch_length": bl, "tips": tips, "names": names, "gtr": gtr,
            "chars": chars, "ambiguous": amb, "L": L, "seed": seed, "true_states": states}


# ---------------------------------------------------------------------------
# named workloads
# ---------------------------------------------------------------------------
def log_uniform_int(rng, lo, hi):
    return int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))


def mix_job_spec(j, base_seed=3000):
    """Job j of the uneven nucleotide mix: tips log-uniform in [50, 4000],
    sites log-uniform in [1000, 30000], depth 0.02, 1% N and 0.5% gaps at
    tips, own seed."""
    seed = base_seed + j
    rng = np.random.default_rng(seed)
    n_tips = log_uniform_int(rng, 50, 4000)
    L = log_uniform_int(rng, 1000, 30000)
    return dict(n_tips=n_tips, L=L, depth=0.02, n_states=4, seed=seed, p_ambig=0.01, p_gap=0.005)


WORKLOADS = {
    # name: (n_tips, L, depth, n_states, seed, p_ambig, p_gap)
    "plain-200": (200, 1700, 0.02, 4, 0, 0.01, 0.005),
    "large-5000": (5000, 30000, 1e-3, 4, 1000, 0.02, 0.0),
    "large-5000-deep": (5000, 30000, 1e-2, 4, 1001, 0.02, 0.0),
    "aa-1000": (1000, 5000, 0.3, 20, 2000, 0.0, 0.0),
}


def prepared_job(spec):
    """Worker-friendly: generate + prepare one job; only numpy arrays out."""
    from .prepare import prepare

    job = make_job(spec["n_tips"], spec["L"], spec["depth"], spec["n_states"], spec["seed"],
                   spec.get("p_ambig", 0.0), spec.get("p_gap", 0.0))
    prep = prepare(job["parent"], job["branch_length"], job["chars"], job["gtr"], job["ambiguous"])
    prep.pop("full_to_compressed", None)
    prep["spec"] = spec
    return prep
