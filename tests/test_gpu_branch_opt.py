"""Joint branch-length optimisation on the GPU (tt_optimize_branch_lengths)
against the oracle's state_pair + optimal_t_compressed (scipy Brent) per
branch, and through the TreeAnc facade.

Tolerances: state-pair counts are exact (integer-valued multiplicities).
The optimum's position is ill-conditioned: near it f(s) - f* ~ f''/2 ds^2, so
FP64 rounding of f (~1e-16 |f|, here summation order) hides ds up to
~sqrt(2e-16 |f| / f''), a few 1e-9 in s for one mutation in 800 sites --
both Brent searches stop anywhere in that flat window.  So the objective
values at the two optima must agree to 1e-12 relative (both are optimal to
FP64 precision), and the lengths to 1e-6 relative + 1e-10 absolute."""
import sys

import numpy as np
import pytest

from helpers import oracle_gtr, small_job
from treetime_amd.prepare import device_tree

pytestmark = pytest.mark.gpu

T_RTOL, T_ATOL = 1e-6, 1e-10
F_RTOL = 1e-12


def same_optimum(og, pairs, mult, t_dev, t_ref):
    """Objective (gtr.py:714-747) at both optima agrees to F_RTOL."""
    if not len(pairs):
        return True
    f_dev = -og.prob_t_compressed(pairs, mult, t_dev, return_log=True)
    f_ref = -og.prob_t_compressed(pairs, mult, t_ref, return_log=True)
    return abs(f_dev - f_ref) <= F_RTOL * max(1.0, abs(f_ref))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def code_states(alphabet_chars):
    other = -1 if len(alphabet_chars) < 10 else 1
    out = np.full(256, other, dtype=np.int16)
    for i, a in enumerate(alphabet_chars):
        out[ord(a)] = i
    return out


def oracle_sweep(prep, job, alphabet, states, tips_reconstructed):
    og = oracle_gtr(job["gtr"], alphabet)
    alpha = np.array([str(a) for a in og.alphabet])
    parent = prep["parent"]
    n = len(parent)
    is_tip = np.ones(n, dtype=bool)
    is_tip[parent[parent >= 0]] = False
    tip_rank = np.cumsum(is_tip) - 1
    chars = prep["tip_codes"].view("S1").astype("U1")

    def seq(v):
        if is_tip[v] and not tips_reconstructed:
            return chars[tip_rank[v]]
        return alpha[states[v]]

    out = {}
    for v in range(n):
        if parent[v] < 0:
            continue
        pairs, mult = og.state_pair(seq(parent[v]), seq(v), pattern_multiplicity=prep["multiplicity"])
        cnt = np.zeros((og.n_states, og.n_states))
        if len(pairs):
            cnt[pairs[:, 0], pairs[:, 1]] = mult
        out[v] = (og.optimal_t_compressed(pairs, mult), cnt, pairs, mult)
    return out, alpha, og


@pytest.mark.parametrize("n_tips,L,n_states,seed,joint,tips", [
    (60, 800, 4, 21, True, False),
    (60, 800, 4, 22, False, False),
    (40, 500, 4, 23, True, True),
    (30, 300, 20, 24, True, False),
])
def test_sweep_matches_oracle(gpu_engine, n_tips, L, n_states, seed, joint, tips):
    say(f"[branch-opt] tips={n_tips} L={L} n={n_states} seed={seed} joint={joint} reconstruct_tips={tips}")
    job, prep, alphabet = small_job(n_tips, L, n_states, seed, depth=0.05)
    dev = device_tree(gpu_engine, prep)
    if joint:
        dev.joint(reconstruct_tips=tips)
    else:
        dev.marginal(reconstruct_tips=tips)
    states = dev.states()
    ref, alpha, og = oracle_sweep(prep, job, alphabet, states, tips)
    t, st, cnt = dev.optimize_branch_lengths(code_states(list(alpha)), -1, 4.0, 1e-10)
    worst = 0.0
    for v, (t_ref, c_ref, pairs, mult) in ref.items():
        assert st[v] == 0, (v, st[v])
        assert np.array_equal(np.trunc(cnt[v]), c_ref), v
        assert same_optimum(og, pairs, mult, t[v], t_ref), (v, t[v], t_ref)
        err = abs(t[v] - t_ref)
        if t_ref > 1e-6:
            worst = max(worst, err / t_ref)
        assert err <= T_RTOL * t_ref + T_ATOL, (v, t[v], t_ref)
    say(f"  {len(ref)} branches, max rel diff of optimal t {worst:.2e} (t > 1e-6)")
    dev.close()


def test_sweep_requires_reconstruction(gpu_engine):
    from treetime_amd.errors import MissingDataError

    job, prep, alphabet = small_job(20, 200, 4, 31)
    dev = device_tree(gpu_engine, prep)
    dev.loglik()
    with pytest.raises(MissingDataError):
        dev.optimize_branch_lengths(code_states(list("ACGT")))
    dev.close()


def test_facade_optimize_matches_oracle():
    """TreeAnc.optimize_branch_len after a joint reconstruction equals the
    per-branch oracle optimum (treeanc.py:1176-1270), and branch_state holds
    the oracle's state-pair counts."""
    import treetime_amd as tt
    from oracle import fp64 as O
    from treetime_amd import synth

    job = synth.make_job(60, 700, 0.05, 4, 43, 0.02, 0.01)
    parent, bl = job["parent"], job["branch_length"]
    clades = [tt.Clade(branch_length=float(bl[i])) for i in range(len(parent))]
    for i, p in enumerate(parent):
        if p >= 0:
            clades[p].clades.append(clades[i])
    for k, v in enumerate(job["tips"]):
        clades[v].name = job["names"][k]
    aln = {job["names"][k]: job["chars"][k].tobytes().decode() for k in range(len(job["tips"]))}
    ta = tt.TreeAnc(tree=tt.Tree(clades[0]), aln=aln, gtr=job["gtr"], verbose=0)
    ta.infer_ancestral_sequences(marginal=False)
    seqs = {id(c): c.cseq.copy() for c in ta.tree.find_clades()}
    assert ta.optimize_branch_len(store_old=True) == "success"
    g = ta.gtr
    og = O.GTRModel("nuc_nogap", mu=g.mu, pi=g.Pi, W=g.W)
    mult = ta.data.multiplicity()
    worst = 0.0
    for c in ta.tree.find_clades(order="postorder"):
        if c.up is None:
            continue
        pairs, m = og.state_pair(seqs[id(c.up)], seqs[id(c)], pattern_multiplicity=mult)
        t_ref = max(0, og.optimal_t_compressed(pairs, m))
        assert same_optimum(og, pairs, m, c.branch_length, t_ref), (c.name, c.branch_length, t_ref)
        assert abs(c.branch_length - t_ref) <= T_RTOL * t_ref + T_ATOL, (c.name, c.branch_length, t_ref)
        assert np.array_equal(c.branch_state["pair"], pairs)
        assert np.array_equal(c.branch_state["multiplicity"], m)
        assert c.mutation_length == c.branch_length
        if t_ref > 1e-6:
            worst = max(worst, abs(c.branch_length - t_ref) / t_ref)
    say(f"[branch-opt] facade: max rel diff {worst:.2e} (t > 1e-6)")
    # the new lengths are used by the next reconstruction
    ta.infer_ancestral_sequences(marginal=True)
    assert np.isfinite(ta.sequence_LH())


# ---------------------------------------------------------------- marginal
# The device profiles are FP32 (<= 2e-7 absolute from the FP64 oracle's), so
# the objective differs from the oracle's by ~1e-7 relative and its optimum
# moves by up to ~1e-5 relative on flat optima: lengths agree to 1e-4
# relative + 1e-8 absolute, and the device optimum is optimal for the
# oracle's own objective to 1e-9 relative.
TM_RTOL, TM_ATOL, FM_RTOL = 1e-4, 1e-8, 1e-9


def _profile_objective(og, pp, pc, mult, t):
    s = np.sqrt(t)
    return -og.prob_t_profiles((pp, pc), mult, t, return_log=True) + np.exp(s ** 4 / 10000)


@pytest.mark.parametrize("n_tips,L,seed,poly", [(50, 600, 71, False), (25, 300, 72, False), (60, 500, 74, True)])
def test_marginal_sweep_matches_oracle(gpu_engine, n_tips, L, seed, poly):
    from helpers import build_oracle, polytomy_job
    from oracle.fp64 import normalize_profile

    say(f"[branch-opt marginal] tips={n_tips} L={L} seed={seed} multifurcating={poly}")
    if poly:
        job, prep, alphabet = polytomy_job(n_tips, L, 4, seed, 0.3)
    else:
        job, prep, alphabet = small_job(n_tips, L, 4, seed, depth=0.05)
    dev = device_tree(gpu_engine, prep)
    dev.marginal()
    t, st, pair = dev.optimize_branch_lengths_marginal(-1, 4.0, 1e-10)
    ta, nodes, tips = build_oracle(prep, alphabet, job["gtr"])
    ta.marginal()
    og = oracle_gtr(job["gtr"], alphabet)
    mult = prep["multiplicity"]
    worst = 0.0
    for v, nd in enumerate(nodes):
        if nd.up is None:
            continue
        pp, pc = nd.marginal_outgroup_LH, nd.marginal_subtree_LH
        t_ref = og.optimal_t_compressed((pp, pc), mult, profiles=True, tol=1e-10)
        assert st[v] == 0, (v, st[v])
        f_dev = _profile_objective(og, pp, pc, mult, t[v])
        f_ref = _profile_objective(og, pp, pc, mult, t_ref)
        assert f_dev - f_ref <= FM_RTOL * abs(f_ref), (v, f_dev, f_ref)
        assert abs(t[v] - t_ref) <= TM_RTOL * t_ref + TM_ATOL, (v, t[v], t_ref)
        if t_ref > 1e-6:
            worst = max(worst, abs(t[v] - t_ref) / t_ref)
    # the bifurcating root's pair (treeanc.py:1318-1331)
    root = [nd for nd in nodes if nd.up is None][0]
    if len(root.clades) == 2:
        assert pair[1] == 0
        n1, n2 = root.clades
        pp = normalize_profile(n2.marginal_subtree_LH * og.Pi)[0]
        pc = n1.marginal_subtree_LH
        t_ref = og.optimal_t_compressed((pp, pc), mult, profiles=True, tol=1e-10)
        assert abs(pair[0] - t_ref) <= TM_RTOL * t_ref + TM_ATOL, (pair, t_ref)
    else:
        assert pair[1] == -1
    say(f"  {len(nodes) - 1} branches + root pair, max rel diff of optimal t {worst:.2e} (t > 1e-6)")
    # grid of prob_t_profiles values (BranchLenInterpolator, marginal mode)
    grid = np.array([-0.1, 0.0, 1e-4, 0.003, 0.05, 0.4])
    lh = dev.branch_profile_lh(grid)
    for v, nd in enumerate(nodes):
        if nd.up is None:
            assert np.all(lh[v] == 0)
            continue
        ref = np.array([og.prob_t_profiles((nd.marginal_outgroup_LH, nd.marginal_subtree_LH), mult, x,
                                           return_log=True) for x in grid])
        # t > 0: 1e-6.  t = 0 reduces to pc . pp, which for a tip whose
        # character has an outgroup probability near the reference's 1e-12
        # clamp (treeanc.py:895-896) depends on which side of the clamp the
        # FP32 and FP64 profiles fall: 1e-3 there.
        assert np.allclose(lh[v][grid != 0], ref[grid != 0], rtol=1e-6, atol=1e-6), (v, lh[v], ref)
        assert np.allclose(lh[v][grid == 0], ref[grid == 0], rtol=1e-3), (v, lh[v], ref)
    dev.close()


def test_marginal_sweep_requires_fresh_profiles(gpu_engine):
    from treetime_amd.errors import MissingDataError

    job, prep, alphabet = small_job(20, 200, 4, 73)
    dev = device_tree(gpu_engine, prep)
    dev.joint()
    with pytest.raises(MissingDataError):
        dev.optimize_branch_lengths_marginal()
    dev.marginal()
    dev.optimize_branch_lengths_marginal()
    dev.set_branch_lengths(np.array([1], np.int32), np.array([0.01]))
    with pytest.raises(MissingDataError):
        dev.branch_profile_lh([0.01])
    dev.close()


def test_facade_optimize_tree_marginal():
    """TreeAnc.optimize_tree_marginal (treeanc.py:1297-1358): the likelihood
    does not decrease, and optimal_marginal_branch_length equals the
    oracle's optimum on the facade's own profiles."""
    import treetime_amd as tt
    from oracle import fp64 as O
    from treetime_amd import synth

    job = synth.make_job(40, 500, 0.05, 4, 75, 0.02, 0.01)
    parent, bl = job["parent"], job["branch_length"]
    clades = [tt.Clade(branch_length=float(bl[i]) * 1.5) for i in range(len(parent))]
    for i, p in enumerate(parent):
        if p >= 0:
            clades[p].clades.append(clades[i])
    for k, v in enumerate(job["tips"]):
        clades[v].name = job["names"][k]
    aln = {job["names"][k]: job["chars"][k].tobytes().decode() for k in range(len(job["tips"]))}
    ta = tt.TreeAnc(tree=tt.Tree(clades[0]), aln=aln, gtr=job["gtr"], verbose=0)
    ta.infer_ancestral_sequences(marginal=True)
    lh0 = ta.sequence_LH()
    # one branch against the oracle run on the facade's tree and alignment
    nodes = list(ta.tree.find_clades(order="preorder"))
    idx = {id(x): i for i, x in enumerate(nodes)}
    par = [-1] * len(nodes)
    for i, x in enumerate(nodes):
        for c in x.clades:
            par[idx[id(c)]] = i
    root, onodes = O.build_tree(par, [ta._branch_length_to_gtr(x) for x in nodes], [x.name for x in nodes])
    g = ta.gtr
    og = O.GTRModel("nuc_nogap", mu=g.mu, pi=g.Pi, W=g.W)
    orc = O.OracleTreeAnc(root, og, ta.data.compressed_alignment, ta.data.multiplicity(),
                          full_length=ta.data.full_length)
    orc.marginal()
    k = 7
    got = ta.optimal_marginal_branch_length(nodes[k])
    ref = og.optimal_t_compressed((onodes[k].marginal_outgroup_LH, onodes[k].marginal_subtree_LH),
                                  ta.data.multiplicity(), profiles=True, tol=1e-10)
    assert abs(got - ref) <= TM_RTOL * ref + TM_ATOL, (got, ref)
    assert ta.optimize_tree_marginal(max_iter=4) == "success"
    lh1 = ta.sequence_LH()
    say(f"[branch-opt marginal] facade: logL {lh0:.3f} -> {lh1:.3f}")
    assert lh1 >= lh0 - 1e-6


@pytest.mark.parametrize("seed,n_tips,L,depth,scale", [(6, 60, 400, 0.2, 0.3), (3, 60, 400, 0.1, 3.0),
                                                       (3, 40, 300, 0.02, 1.7)])
def test_facade_optimize_tree_matches_oracle(gpu_engine, seed, n_tips, L, depth, scale):
    """TreeAnc.optimize_tree (treeanc.py:1384-1495: joint reconstruction,
    joint branch lengths, pruning of zero-length internal branches, until no
    character changes) against the oracle's iteration: the same pruned
    nodes, the same topology, every branch length to the optimum tolerance
    and the final joint log-likelihood to 1e-7 relative."""
    import treetime_amd as tt
    from helpers import optimize_tree_case
    from oracle import fp64 as O

    say(f"[optimize_tree] seed={seed} tips={n_tips} L={L} depth={depth} scale={scale}")
    orc, job, names = optimize_tree_case(seed, n_tips, L, depth, scale)
    parent, bl = job["parent"], job["branch_length"] * scale
    clades = [tt.Clade(name=names[i], branch_length=float(bl[i])) for i in range(len(parent))]
    for i, p in enumerate(parent):
        if p >= 0:
            clades[p].clades.append(clades[i])
    chars = job["chars"].astype("U1")
    aln = {names[v]: "".join(chars[k]) for k, v in enumerate(job["tips"])}
    g = job["gtr"]
    ta = tt.TreeAnc(tree=tt.Tree(clades[int(np.flatnonzero(parent < 0)[0])]), aln=aln, verbose=0,
                    gtr=tt.GTR.custom(alphabet="nuc_nogap", pi=g.Pi, W=g.W, mu=g.mu), engine=gpu_engine,
                    keep_node_order=True)
    assert ta.optimize_tree(max_iter=5) == "success"
    total = orc.optimize_tree(max_iter=5)
    assert orc.min_path_margin > 1e-5        # no joint decision of the oracle's run is a near tie
    say(f"  oracle: {orc.n_iter} iterations, {len(orc.pruned)} pruned, logL {total:.6f}; "
        f"device logL {ta.tree.unconstrained_sequence_LH:.6f}")
    o_nodes = {nd.name: nd for nd in O.preorder(orc.root)}
    f_nodes = {nd.name: nd for nd in ta.tree.find_clades()}
    assert set(o_nodes) == set(f_nodes)
    worst = 0.0
    for name, nd in f_nodes.items():
        on = o_nodes[name]
        assert (nd.up.name if nd.up is not None else None) == (on.up.name if on.up is not None else None), name
        assert [c.name for c in nd.clades] == [c.name for c in on.clades], name
        if nd.up is None:
            continue
        t_ref = on.branch_length
        assert abs(nd.branch_length - t_ref) <= T_RTOL * t_ref + T_ATOL, (name, nd.branch_length, t_ref)
        if t_ref > 1e-6:
            worst = max(worst, abs(nd.branch_length - t_ref) / t_ref)
    say(f"  max rel diff of branch lengths {worst:.2e} (t > 1e-6)")
    assert abs(ta.tree.unconstrained_sequence_LH - total) <= 1e-7 * abs(total)
