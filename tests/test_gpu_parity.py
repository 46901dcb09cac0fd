"""Parity of the HIP passes (through the C ABI) with the FP64 oracle.

Tolerances (FP32 storage, FP64 arithmetic and normaliser sums):
  total logL            relative error <= 1e-7   (measured ~1e-10)
  per-pattern logL      error <= 5e-7 * max(1, |logL_p|)  (FP32 arithmetic)
  marginal profiles     max absolute error <= 2e-6
  argmax states         bit-exact wherever the FP64 margin between the two
                        best states exceeds 1e-5 (joint: the smallest decision
                        margin on the path from the root)
"""
import sys

import numpy as np
import pytest

from helpers import build_oracle, polytomy_job, small_job
from treetime_amd.engine import Engine
from treetime_amd.errors import MissingDataError
from treetime_amd.prepare import device_tree

pytestmark = pytest.mark.gpu

LOGL_RTOL = 1e-7
SITE_RTOL = 5e-7
PROF_ATOL = 2e-6
MARGIN = 1e-5


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def site_err(got, ref):
    return float(np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref))))


def marginal_margin(prof):
    s = np.sort(prof, axis=-1)
    return s[..., -1] - s[..., -2]


CASES = [
    # name, n_tips, L, n_states, seed, depth, p_ambig, p_gap
    ("c0-like", 200, 1700, 4, 0, 0.02, 0.01, 0.005),
    ("deep", 64, 600, 4, 5, 0.5, 0.02, 0.01),
    ("tiny", 3, 40, 4, 2, 0.1, 0.0, 0.0),
    ("aa", 60, 300, 20, 3, 0.3, 0.0, 0.0),
    ("p-not-multiple-of-64", 37, 97, 4, 9, 0.05, 0.05, 0.05),
    ("big", 2000, 3000, 4, 11, 0.02, 0.01, 0.005),
    # multifurcating trees (30% of internal nodes collapsed, root with >= 3 children)
    ("poly-nuc", 150, 900, 4, 31, 0.05, 0.01, 0.005),
    ("poly-aa", 50, 300, 20, 32, 0.3, 0.0, 0.0),
    ("poly-big", 1500, 1200, 4, 33, 0.02, 0.01, 0.005),
    # the reference's default alphabet 'nuc': five states, the gap is a state
    ("nuc5", 150, 900, 5, 41, 0.05, 0.01, 0.02),
    ("poly-nuc5", 120, 700, 5, 42, 0.05, 0.01, 0.02),
    ("nuc5-deep", 64, 500, 5, 43, 0.5, 0.02, 0.03),
    # tips carrying every partial ambiguity code (multi-hot tip rows):
    # IUPAC R Y S W K M D H B V for nucleotides, B Z (and X) for amino acids
    ("iupac-nuc", 200, 1200, 4, 44, 0.05, 0.01, 0.005, 0.05),
    ("iupac-nuc5", 100, 800, 5, 45, 0.05, 0.01, 0.01, 0.05),
    ("iupac-aa", 60, 400, 20, 46, 0.3, 0.02, 0.0, 0.05),
    ("poly-iupac-aa", 50, 300, 20, 47, 0.3, 0.02, 0.0, 0.05),
]


def make_case(name, n_tips, L, n_states, seed, depth, pa, pg, pi=0.0):
    if name.startswith("poly"):
        return polytomy_job(n_tips, L, n_states, seed, 0.3, depth, pa, pg, pi)
    return small_job(n_tips, L, n_states, seed, depth, pa, pg, pi)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_marginal_parity(gpu_engine, case):
    name, n_tips, L, n_states, seed, depth, pa, pg = case[:8]
    say(f"[marginal] case={name} n_tips={n_tips} L={L} n={n_states} seed={seed}")
    job, prep, alphabet = make_case(*case)
    ta, nodes, tips = build_oracle(prep, alphabet, job["gtr"])
    ref = ta.marginal()
    dt = device_tree(gpu_engine, prep)
    got = dt.marginal()
    slh = dt.sequence_lh()
    prof = dt.profiles()
    states = dt.states()
    perr = max(float(np.max(np.abs(prof[i] - nd.marginal_profile))) for i, nd in enumerate(nodes)
               if not nd.is_terminal())
    say(f"  logL rel err {abs(got - ref) / abs(ref):.2e}  site err {site_err(slh, ta.sequence_LH):.2e}  "
        f"profile max abs err {perr:.2e}")
    assert abs(got - ref) <= LOGL_RTOL * abs(ref), (got, ref)
    assert site_err(slh, ta.sequence_LH) <= SITE_RTOL
    for i, nd in enumerate(nodes):
        if nd.is_terminal():
            assert (states[i] == 255).all()
            continue
        err = np.max(np.abs(prof[i] - nd.marginal_profile))
        assert err <= PROF_ATOL, (i, err)
        ok = marginal_margin(nd.marginal_profile) > MARGIN
        assert np.array_equal(states[i][ok], nd.seq_idx[ok]), i
    dt.close()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_joint_parity(gpu_engine, case):
    name, n_tips, L, n_states, seed, depth, pa, pg = case[:8]
    say(f"[joint] case={name} n_tips={n_tips} L={L} n={n_states} seed={seed}")
    job, prep, alphabet = make_case(*case)
    ta, nodes, tips = build_oracle(prep, alphabet, job["gtr"])
    ref = ta.joint()
    dt = device_tree(gpu_engine, prep)
    got = dt.joint()
    slh = dt.sequence_lh()
    say(f"  logL rel err {abs(got - ref) / abs(ref):.2e}  site err {site_err(slh, ta.joint_sequence_LH):.2e}")
    assert abs(got - ref) <= LOGL_RTOL * abs(ref), (got, ref)
    assert site_err(slh, ta.joint_sequence_LH) <= SITE_RTOL
    states = dt.states()
    for i, nd in enumerate(nodes):
        if nd.is_terminal():
            assert (states[i] == 255).all()
            continue
        ok = nd.path_margin > MARGIN
        assert np.array_equal(states[i][ok], nd.seq_idx[ok]), i
    dt.close()


def test_reconstruct_tips(gpu_engine):
    say("[tips] reconstruct_tip_states marginal + joint")
    job, prep, alphabet = small_job(40, 300, 4, 13, 0.05, 0.05, 0.02)
    ta, nodes, tips = build_oracle(prep, alphabet, job["gtr"])
    ta.marginal(reconstruct_tip_states=True)
    dt = device_tree(gpu_engine, prep)
    dt.marginal(reconstruct_tips=True)
    prof = dt.profiles()
    states = dt.states()
    for i in tips:
        nd = nodes[i]
        assert np.max(np.abs(prof[i] - nd.marginal_profile)) <= PROF_ATOL
        ok = marginal_margin(nd.marginal_profile) > MARGIN
        assert np.array_equal(states[i][ok], nd.seq_idx[ok])
    ta.joint(reconstruct_tip_states=True)
    dt.joint(reconstruct_tips=True)
    states = dt.states()
    for i in tips:
        ok = nodes[i].path_margin > MARGIN
        assert np.array_equal(states[i][ok], nodes[i].seq_idx[ok])
    dt.close()


def test_loglik_and_branch_update(gpu_engine):
    say("[update] loglik + set_branch_lengths + update_branch")
    job, prep, alphabet = small_job(80, 500, 4, 21, 0.05, 0.01, 0.0)
    dt = device_tree(gpu_engine, prep)
    base = dt.loglik()
    ta, nodes, tips = build_oracle(prep, alphabet, job["gtr"])
    assert abs(base - ta.marginal()) <= LOGL_RTOL * abs(base)
    node = int(tips[3])
    newt = prep["t"][node] * 1.7
    got = dt.update_branch(node, newt)
    prep2 = dict(prep)
    prep2["t"] = prep["t"].copy()
    prep2["t"][node] = newt
    ta2, _, _ = build_oracle(prep2, alphabet, job["gtr"])
    ref = ta2.marginal()
    assert abs(got - ref) <= LOGL_RTOL * abs(ref)
    dt.close()


def test_repeat_runs_identical(gpu_engine):
    """Graph replays give bit-identical results (deterministic reductions)."""
    say("[repeat] graph replays")
    job, prep, alphabet = small_job(150, 900, 4, 4)
    dt = device_tree(gpu_engine, prep)
    a = [dt.marginal() for _ in range(3)]
    b = [dt.joint() for _ in range(2)]
    c = dt.marginal()
    assert a[0] == a[1] == a[2] == c
    assert b[0] == b[1]
    dt.close()


def test_budget_rejection():
    say("[budget] admission control")
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from treetime_amd._native import BudgetError

    eng = Engine(0, 1 << 20)  # 1 MiB
    job, prep, alphabet = small_job(200, 1700, 4, 0)
    with pytest.raises(BudgetError):
        device_tree(eng, prep)
    used, budget = eng.usage()
    assert used == 0 and budget == 1 << 20
    eng.close()


@pytest.mark.parametrize("tips_flag", [False, True])
def test_batch_matches_plain(gpu_engine, tips_flag):
    """One persistent launch over uneven trees equals tt_marginal per tree:
    identical profiles and states, logL within 1e-12 (summation order only)."""
    from treetime_amd.engine import batch_marginal

    say(f"[batch] 7 uneven trees (one multifurcating), reconstruct_tips={tips_flag}")
    specs = [(50, 1000, 1), (300, 200, 2), (7, 65, 3), (120, 3000, 4), (3, 10, 5), (900, 640, 6)]
    jobs = [small_job(n, L, 4, seed) for n, L, seed in specs] + [polytomy_job(200, 700, 4, 7, 0.3)]
    trees = [device_tree(gpu_engine, prep) for _, prep, _ in jobs]
    plain = [t.marginal(reconstruct_tips=tips_flag) for t in trees]
    plain_prof = [t.profiles() for t in trees]
    plain_states = [t.states() for t in trees]
    plain_slh = [t.sequence_lh() for t in trees]
    for rep in range(2):
        got = batch_marginal(gpu_engine, trees, reconstruct_tips=tips_flag)
        for i, t in enumerate(trees):
            assert abs(got[i] - plain[i]) <= 1e-12 * abs(plain[i]), (i, got[i], plain[i])
            assert np.array_equal(t.states(), plain_states[i])
            assert np.array_equal(t.profiles(), plain_prof[i])
            assert np.allclose(t.sequence_lh(), plain_slh[i], rtol=0, atol=1e-12)
    # oracle check of one member
    job, prep, alphabet = jobs[3]
    ta, nodes, tips = build_oracle(prep, alphabet, job["gtr"])
    ref = ta.marginal()
    assert abs(got[3] - ref) <= LOGL_RTOL * abs(ref)
    for t in trees:
        t.close()


@pytest.fixture
def block_batches(monkeypatch):
    """Plans built while this fixture is active use the persistent kernel
    (TT_BATCH_MODE=block), whatever the tuner would pick."""
    monkeypatch.setenv("TT_BATCH_MODE", "block")


def test_batch_run_modes(gpu_engine, block_batches):
    """tt_batch_run: 'loglik' (persistent launch without the preorder) equals
    tt_loglik per tree and leaves profiles stale; 'joint' equals tt_joint per
    tree; 'marginal' equals tt_batch_marginal."""
    from treetime_amd.engine import batch_info, batch_marginal, batch_run

    say("[batch] tt_batch_run loglik / joint / marginal on 6 trees")
    specs = [(50, 1000, 11), (300, 200, 12), (7, 65, 13), (120, 3000, 14), (3, 10, 15)]
    jobs = [small_job(n, L, 4, seed) for n, L, seed in specs] + [polytomy_job(150, 500, 4, 16, 0.3)]
    trees = [device_tree(gpu_engine, prep) for _, prep, _ in jobs]
    plain_ll = [t.loglik() for t in trees]
    plain_slh = [t.sequence_lh() for t in trees]
    for rep in range(2):
        got = batch_run(gpu_engine, trees, "loglik")
        for i, t in enumerate(trees):
            assert abs(got[i] - plain_ll[i]) <= 1e-12 * abs(plain_ll[i]), (i, got[i], plain_ll[i])
            assert np.allclose(t.sequence_lh(), plain_slh[i], rtol=0, atol=1e-12)
    assert batch_info(gpu_engine, trees, kind="loglik")["blocks_per_cu"] >= 0
    with pytest.raises(MissingDataError):  # profiles are stale after a loglik pass
        trees[0].optimize_branch_lengths_marginal(-1, 4.0, 1e-10)
    for tips in (False, True):
        joint_plain = [t.joint(reconstruct_tips=tips) for t in trees]
        joint_states = [t.states() for t in trees]
        for rep in range(2):
            got = batch_run(gpu_engine, trees, "joint", reconstruct_tips=tips)
            for i, t in enumerate(trees):
                assert abs(got[i] - joint_plain[i]) <= 1e-12 * abs(joint_plain[i]), (tips, i, got[i], joint_plain[i])
                assert np.array_equal(t.states(), joint_states[i]), (tips, i)
    # freeing the joint buffers drops the joint plan; the next batched joint rebuilds it
    for t in trees:
        t.trim()
    got = batch_run(gpu_engine, trees, "joint")
    assert np.allclose(got, [t.joint() for t in trees], rtol=1e-12, atol=0)
    got_m = batch_run(gpu_engine, trees, "marginal")
    ref_m = batch_marginal(gpu_engine, trees)
    assert np.array_equal(got_m, ref_m)
    for i in range(len(trees)):
        assert abs(got_m[i] - plain_ll[i]) <= 1e-12 * abs(plain_ll[i])
    for t in trees:
        t.close()



@pytest.mark.parametrize("tiles", ["1", "2"])
@pytest.mark.parametrize("n_states", [4, 5])
@pytest.mark.parametrize("wps", ["0", "6", "8"])
def test_batch_joint_register_budgets(gpu_engine, block_batches, monkeypatch, wps, n_states, tiles):
    """Every register budget and item size of the batched joint kernel (the
    tuner's candidates, pinned one at a time): states identical to tt_joint,
    logL to 1e-12, with and without tip reconstruction, on binary and
    multifurcating trees.  The 8-wave budget exists for alphabets of up to 4
    states only; pinning it for 5 states is refused with an error naming the
    budget."""
    from treetime_amd.engine import batch_info, batch_run

    say(f"[batch] joint kernel, waves per SIMD {wps}, n={n_states}, tiles per item {tiles}")
    monkeypatch.setenv("TT_JOINT_WAVES_PER_SIMD", wps)
    monkeypatch.setenv("TT_BATCH_TILES", tiles)
    jobs = [small_job(n, L, n_states, seed) for n, L, seed in [(90, 700, 81), (12, 130, 82)]]
    jobs.append(polytomy_job(120, 400, n_states, 83, 0.3))
    if n_states == 5 and wps == "8":
        trees = [device_tree(gpu_engine, prep) for _, prep, _ in jobs]
        with pytest.raises(ValueError, match="waves per SIMD"):
            batch_run(gpu_engine, trees, "joint")
        for t in trees:
            t.close()
        return
    trees = [device_tree(gpu_engine, prep) for _, prep, _ in jobs]
    for tips in (False, True):
        plain = [t.joint(reconstruct_tips=tips) for t in trees]
        plain_states = [t.states() for t in trees]
        got = batch_run(gpu_engine, trees, "joint", reconstruct_tips=tips)
        info = batch_info(gpu_engine, trees, kind="joint", reconstruct_tips=tips)
        assert info["waves_per_simd"] == int(wps) and info["tiles"] == int(tiles)
        for i, t in enumerate(trees):
            assert abs(got[i] - plain[i]) <= 1e-12 * abs(plain[i]), (wps, tips, i, got[i], plain[i])
            assert np.array_equal(t.states(), plain_states[i]), (wps, tips, i)
    for t in trees:
        t.close()

def test_tip_states_reset_after_reconstruction(gpu_engine):
    """Tip rows of the states hold 255 after a pass without tip
    reconstruction, also when an earlier pass (plain or batched) filled them."""
    from treetime_amd.engine import batch_marginal

    say("[tips] tip states reset to 255")
    jobs = [small_job(n, L, 4, seed) for n, L, seed in [(30, 200, 61), (70, 500, 62)]]
    trees = [device_tree(gpu_engine, prep) for _, prep, _ in jobs]

    def tip_rows(t):
        st = t.states()
        is_tip = np.ones(t.n_nodes, dtype=bool)
        is_tip[t.parent[t.parent >= 0]] = False
        return st[is_tip]

    t0 = trees[0]
    t0.marginal(reconstruct_tips=True)
    assert (tip_rows(t0) < 4).all()
    t0.marginal()
    assert (tip_rows(t0) == 255).all()
    t0.joint(reconstruct_tips=True)
    assert (tip_rows(t0) < 4).all()
    t0.joint()
    assert (tip_rows(t0) == 255).all()
    batch_marginal(gpu_engine, trees, reconstruct_tips=True)
    assert all((tip_rows(t) < 4).all() for t in trees)
    batch_marginal(gpu_engine, trees)
    assert all((tip_rows(t) == 255).all() for t in trees)
    trees[1].marginal(reconstruct_tips=True)
    batch_marginal(gpu_engine, trees)
    assert all((tip_rows(t) == 255).all() for t in trees)
    for t in trees:
        t.close()


def test_engine_lifetime_and_growing_updates():
    """An engine with open trees refuses to close; branch-length updates of
    growing size (temporary buffers re-allocated) give the oracle's logL."""
    from treetime_amd.errors import MissingDataError

    say("[engine] lifetime, growing updates")
    eng = Engine(0, 0)
    job, prep, alphabet = small_job(60, 400, 4, 81)
    dt = device_tree(eng, prep)
    with pytest.raises(MissingDataError):
        eng.close()
    t = prep["t"].copy()
    rng = np.random.default_rng(5)
    for k in (1, 7, 40, len(t) - 1):
        nodes = rng.choice(np.nonzero(prep["parent"] >= 0)[0], size=k, replace=False).astype(np.int32)
        t[nodes] *= rng.uniform(0.5, 2.0, size=k)
        dt.set_branch_lengths(nodes, t[nodes])
        got = dt.loglik()
        p2 = dict(prep)
        p2["t"] = t
        ta, _, _ = build_oracle(p2, alphabet, job["gtr"])
        ref = ta.marginal()
        assert abs(got - ref) <= LOGL_RTOL * abs(ref), (k, got, ref)
    used = eng.usage()[0]
    dt.close()
    assert eng.usage()[0] < used
    eng.close()



def test_batch_run_twenty_states(gpu_engine, block_batches):
    """tt_batch_run on 20-state trees against the FP64 oracle. The batched
    kernel multiplies by P row by row while the plain 20-state passes use
    matrix-core tiles, so the two GPU paths agree only to FP32 rounding; each
    is held to the oracle's tolerances."""
    from treetime_amd.engine import batch_run

    say("[batch] 20-state trees: marginal / loglik / joint against the oracle")
    specs = [(40, 300, 31), (9, 70, 32), (120, 500, 33)]
    jobs = [small_job(n, L, 20, seed, 0.3, 0.0, 0.0) for n, L, seed in specs]
    trees = [device_tree(gpu_engine, prep) for _, prep, _ in jobs]
    oracles = [build_oracle(prep, alphabet, job["gtr"]) for job, prep, alphabet in jobs]
    ref_m = [ta.marginal() for ta, _, _ in oracles]
    got = batch_run(gpu_engine, trees, "marginal")
    for i, t in enumerate(trees):
        assert abs(got[i] - ref_m[i]) <= LOGL_RTOL * abs(ref_m[i]), (i, got[i], ref_m[i])
        prof, states = t.profiles(), t.states()
        for j, nd in enumerate(oracles[i][1]):
            if nd.is_terminal():
                continue
            assert np.max(np.abs(prof[j] - nd.marginal_profile)) <= PROF_ATOL, (i, j)
            ok = marginal_margin(nd.marginal_profile) > MARGIN
            assert np.array_equal(states[j][ok], nd.seq_idx[ok]), (i, j)
    got = batch_run(gpu_engine, trees, "loglik")
    for i, t in enumerate(trees):
        assert abs(got[i] - ref_m[i]) <= LOGL_RTOL * abs(ref_m[i]), (i, got[i], ref_m[i])
        assert site_err(t.sequence_lh(), oracles[i][0].sequence_LH) <= SITE_RTOL
    got = batch_run(gpu_engine, trees, "joint")
    for i, t in enumerate(trees):
        ta, nodes, _ = oracles[i]
        ref = ta.joint()
        assert abs(got[i] - ref) <= LOGL_RTOL * abs(ref), (i, got[i], ref)
        states = t.states()
        for j, nd in enumerate(nodes):
            if nd.is_terminal():
                continue
            ok = nd.path_margin > MARGIN
            assert np.array_equal(states[j][ok], nd.seq_idx[ok]), (i, j)
    for t in trees:
        t.close()


def _edge_jobs():
    """Degenerate shapes: two tips, one site, all-ambiguous and all-gap
    columns, identical sequences (constant columns only), zero and very long
    branches, a 20-state pair."""
    from treetime_amd.prepare import prepare

    def reprep(job):
        prep = prepare(job["parent"], job["branch_length"], job["chars"], job["gtr"], job["ambiguous"])
        return job, prep, "nuc_nogap" if job["gtr"].n_states == 4 else "aa_nogap"

    out = [("two-tips-one-site", small_job(2, 1, 4, 71, 0.05, 0.0, 0.0)),
           ("two-tips-aa", small_job(2, 50, 20, 72, 0.3, 0.0, 0.0))]
    job = small_job(20, 100, 4, 73, 0.05)[0]
    job["chars"][:, 0] = b"N"
    job["chars"][:, 1] = b"-"
    job["chars"][:5, 2] = b"N"
    job["chars"][3, :] = b"N"
    out.append(("ambiguous-columns-and-row", reprep(job)))
    job = small_job(30, 200, 4, 74, 0.05, 0.0, 0.0)[0]
    job["chars"][:] = job["chars"][0]
    out.append(("identical-sequences", reprep(job)))
    job = small_job(40, 300, 4, 75, 0.05, 0.0, 0.0)[0]
    bl = job["branch_length"].copy()
    bl[1::7] = 0.0
    bl[2::11] = 5.0
    job["branch_length"] = bl
    out.append(("zero-and-long-branches", reprep(job)))
    return out


def test_edge_shapes(gpu_engine):
    """Degenerate inputs through the plain and the batched passes, against
    the oracle at the stated tolerances."""
    from treetime_amd.engine import batch_run

    jobs = _edge_jobs()
    trees, refs = [], []
    for name, (job, prep, alphabet) in jobs:
        say(f"[edge] {name}: P={prep['tip_codes'].shape[1]} nodes={len(prep['parent'])}")
        ta, nodes, tips = build_oracle(prep, alphabet, job["gtr"])
        dt = device_tree(gpu_engine, prep)
        ref = ta.marginal()
        got = dt.marginal()
        assert abs(got - ref) <= LOGL_RTOL * abs(ref), (name, got, ref)
        assert site_err(dt.sequence_lh(), ta.sequence_LH) <= SITE_RTOL, name
        prof, states = dt.profiles(), dt.states()
        for i, nd in enumerate(nodes):
            if nd.is_terminal():
                continue
            assert np.max(np.abs(prof[i] - nd.marginal_profile)) <= PROF_ATOL, (name, i)
            ok = marginal_margin(nd.marginal_profile) > MARGIN
            assert np.array_equal(states[i][ok], nd.seq_idx[ok]), (name, i)
        ref_j = ta.joint()
        got_j = dt.joint()
        assert abs(got_j - ref_j) <= LOGL_RTOL * abs(ref_j), (name, got_j, ref_j)
        states = dt.states()
        for i, nd in enumerate(nodes):
            if not nd.is_terminal():
                ok = nd.path_margin > MARGIN
                assert np.array_equal(states[i][ok], nd.seq_idx[ok]), (name, i)
        trees.append(dt)
        refs.append((ref, ref_j))
    nuc = [i for i, (_, (job, _, _)) in enumerate(jobs) if job["gtr"].n_states == 4]
    for kind, col in (("marginal", 0), ("loglik", 0), ("joint", 1)):
        got = batch_run(gpu_engine, [trees[i] for i in nuc], kind)
        for g, i in zip(got, nuc):
            assert abs(g - refs[i][col]) <= LOGL_RTOL * abs(refs[i][col]), (kind, jobs[i][0], g, refs[i][col])
    for t in trees:
        t.close()


def test_joint_logL_is_score_of_decoded_states(gpu_engine):
    """Size-independent property at the 5000-tip x 30000-site size: the
    joint per-pattern logL is the FP64 log-likelihood of the decoded states
    themselves (the oracle's ancestral_likelihood, treeanc.py:718-750, on the
    device's states with tips reconstructed), to 1e-9 relative per pattern:
    the two sides build P(t) in different operation orders, and log P_jj of
    a short branch (P_jj ~ 1 - 1e-6) turns P's ~1e-16 absolute rounding into
    ~1e-10 relative.  A missing or misindexed term is off by >= 1e-6."""
    from treetime_amd import synth
    from treetime_amd.prepare import prepare

    say("[joint] logL = score of the decoded states, 5000 tips x 30000 sites")
    job = synth.make_job(5000, 30000, 1e-3, 4, 1, 0.02, 0.0)
    prep = prepare(job["parent"], job["branch_length"], job["chars"], job["gtr"], job["ambiguous"])
    ta, nodes, tips = build_oracle(prep, "nuc_nogap", job["gtr"])
    dt = device_tree(gpu_engine, prep)
    total = dt.joint(reconstruct_tips=True)
    slh = dt.sequence_lh()
    states = dt.states()
    assert (states < 4).all()
    seq = {nd.name: states[i].astype(int) for i, nd in enumerate(nodes)}
    ref = ta.ancestral_likelihood(seq)
    err = float(np.max(np.abs(slh - ref) / np.abs(ref)))
    say(f"  P={len(slh)} max per-pattern rel diff {err:.2e}")
    assert err <= 1e-9
    assert abs(total - float(np.sum(prep["multiplicity"] * ref))) <= 1e-9 * abs(total)
    dt.close()


# trees the one-launch pass takes (large trees stay on the level kernels)
ONE_LAUNCH_CASES = [c for c in CASES if c[3] in (4, 5) and c[1] <= 1000]


@pytest.mark.parametrize("case", ONE_LAUNCH_CASES, ids=[c[0] for c in ONE_LAUNCH_CASES])
@pytest.mark.parametrize("reuse", ["0", "1"], ids=["rows", "reused"])
@pytest.mark.parametrize("stage", ["1", "0"], ids=["staged", "streamed"])
@pytest.mark.parametrize("presep", ["", "0"], ids=["pre-region-auto", "pre-region-shared"])
def test_one_launch_pass_equals_level_kernels(gpu_engine, case, stage, reuse, presep, monkeypatch):
    """The one-launch pass with LDS-resident rows (TT_TILE=1), with its
    records staged in LDS or streamed from memory, and with one LDS slot per
    internal row or slots reused by lifetime (subtree rows re-read from
    memory in the preorder), and with the preorder's staged records in a
    region of their own where that fits or replacing the postorder's at the
    root step, gives the level kernels' profiles, subtree rows
    and states bit for bit, and their logL to rounding, with and without
    reconstructed tips and for the log-likelihood pass."""
    say(f"[one-launch] case={case[0]} stage={stage} reuse={reuse} pre-region={presep or 'auto'}")
    job, prep, alphabet = make_case(*case)
    monkeypatch.setenv("TT_TILE", "0")
    lv = device_tree(gpu_engine, prep)
    monkeypatch.setenv("TT_TILE", "1")
    monkeypatch.setenv("TT_TILE_STAGE", stage)
    monkeypatch.setenv("TT_TILE_REUSE", reuse)
    monkeypatch.setenv("TT_TILE_PRE_SEP", presep)
    one = device_tree(gpu_engine, prep)
    assert lv.pass_shape() == (0, False, False)
    tp, staged, reused = one.pass_shape()
    assert tp > 0 and staged == (stage == "1") and reused == (reuse == "1"), (tp, staged, reused)
    for tips in (False, True):
        a, b = lv.marginal(reconstruct_tips=tips), one.marginal(reconstruct_tips=tips)
        assert abs(a - b) <= 1e-12 * abs(a), (tips, a, b)
        assert site_err(one.sequence_lh(), lv.sequence_lh()) <= 1e-12
        assert np.array_equal(one.profiles(), lv.profiles())
        assert np.array_equal(one.profiles(1), lv.profiles(1))
        assert np.array_equal(one.states(), lv.states())
    a, b = lv.loglik(), one.loglik()
    assert abs(a - b) <= 1e-12 * abs(a)
    # new branch lengths (set, and changed on the root path) reach the one-launch pass
    nodes = lv.nonroot_nodes()[:3]
    for dt in (lv, one):
        dt.set_branch_lengths(nodes, dt.t_cur[nodes] * 1.7)
    a, b = lv.marginal(), one.marginal()
    assert abs(a - b) <= 1e-12 * abs(a)
    assert np.array_equal(one.profiles(), lv.profiles())
    node = int(lv.nonroot_nodes()[-1])
    ua, ub = lv.update_branch(node, 0.031), one.update_branch(node, 0.031)
    assert abs(ua - ub) <= 1e-12 * abs(ua)
    a, b = lv.marginal(), one.marginal()
    assert abs(a - b) <= 1e-12 * abs(a) and abs(a - ua) <= 1e-11 * abs(a)
    assert np.array_equal(one.profiles(), lv.profiles())
    assert np.array_equal(one.states(), lv.states())
    lv.close()
    one.close()


@pytest.mark.parametrize("case", ONE_LAUNCH_CASES, ids=[c[0] for c in ONE_LAUNCH_CASES])
@pytest.mark.parametrize("reuse", ["0", "1"], ids=["rows", "reused"])
@pytest.mark.parametrize("jstage", ["1", "0"], ids=["staged", "streamed"])
def test_one_launch_joint_equals_level_kernels(gpu_engine, case, reuse, jstage, monkeypatch):
    """The one-launch joint pass (postorder max-product rows in LDS, root,
    backtrack and score in the same launch; descriptors staged in LDS or
    streamed from memory) gives the level kernels' states,
    message rows and argmax rows bit for bit and their joint logL to
    rounding, with and without reconstructed tips, after new branch lengths,
    and a backtrack from given root states after it equals theirs."""
    say(f"[one-launch joint] case={case[0]} reuse={reuse} descriptors staged={jstage}")
    job, prep, alphabet = make_case(*case)
    monkeypatch.setenv("TT_TILE", "0")
    lv = device_tree(gpu_engine, prep)
    monkeypatch.setenv("TT_TILE", "1")
    monkeypatch.setenv("TT_TILE_REUSE", reuse)
    monkeypatch.setenv("TT_TILE_JOINT_STAGE", jstage)
    one = device_tree(gpu_engine, prep)
    assert not lv.joint_one_launch() and one.joint_one_launch()
    parent = np.asarray(lv.parent)
    internal = np.isin(np.arange(len(parent)), parent[parent >= 0])
    for tips in (False, True, False):
        a, b = lv.joint(reconstruct_tips=tips), one.joint(reconstruct_tips=tips)
        assert abs(a - b) <= 1e-12 * abs(a), (tips, a, b)
        assert site_err(one.sequence_lh(), lv.sequence_lh()) <= 1e-12
        assert np.array_equal(one.states(), lv.states()), tips
        la, ca = lv.joint_debug()
        lb, cb = one.joint_debug()
        # argmax rows exist for every node with a branch: internal ones always,
        # tips when their states are reconstructed
        rows = (parent >= 0) & (internal | tips)
        assert np.array_equal(la, lb) and np.array_equal(ca[rows], cb[rows]), tips
    nodes = lv.nonroot_nodes()[:3]
    for dt in (lv, one):
        dt.set_branch_lengths(nodes, dt.t_cur[nodes] * 1.7)
    a, b = lv.joint(), one.joint()
    assert abs(a - b) <= 1e-12 * abs(a)
    assert np.array_equal(one.states(), lv.states())
    # the marginal pass of the same tree still equals the level kernels' after a joint pass
    a, b = lv.marginal(), one.marginal()
    assert abs(a - b) <= 1e-12 * abs(a)
    assert np.array_equal(one.profiles(), lv.profiles())
    a, b = lv.joint(), one.joint()
    rs = (np.arange(lv.P) % lv.n).astype(np.uint8)
    a, b = lv.joint_from_root(rs), one.joint_from_root(rs)
    assert abs(a - b) <= 1e-12 * abs(a)
    assert np.array_equal(one.states(), lv.states())
    lv.close()
    one.close()
