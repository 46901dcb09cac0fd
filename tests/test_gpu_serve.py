"""Served request stream on the GPU: each result is the same alone and in the
mix, root-path branch updates equal full recomputation, and the final state
matches the FP64 oracle."""
import sys

import numpy as np
import pytest

from helpers import build_oracle, small_job
from treetime_amd.prepare import device_tree
from treetime_amd.serve import Request, Server, make_stream, summarize

pytestmark = pytest.mark.gpu

SPECS = [(40, 300, 1), (150, 900, 2), (7, 60, 3), (400, 1200, 4), (90, 2000, 5)]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def fresh(engine, jobs):
    return [device_tree(engine, prep) for _, prep, _ in jobs]


def copy_reqs(reqs):
    return [Request(r.kind, r.tree, r.arrival, r.node, r.scale, r.rid) for r in reqs]


def test_serve_same_alone_and_in_mix(gpu_engine):
    say("[serve] mixed stream vs alone vs scratch")
    jobs = [small_job(n, L, 4, seed) for n, L, seed in SPECS]
    rng = np.random.default_rng(11)
    trees = fresh(gpu_engine, jobs)
    reqs = make_stream(trees, 120, 2000.0, rng)
    mixed = Server(trees, engine=gpu_engine, slots=4).run(copy_reqs(reqs), realtime=True)
    final_t = [t.t_cur.copy() for t in trees]
    mixed_slh = [t.sequence_lh() for t in trees]
    for t in trees:
        t.close()

    trees = fresh(gpu_engine, jobs)
    alone = Server(trees, engine=gpu_engine, slots=1).run(copy_reqs(reqs), realtime=False)
    for t in trees:
        t.close()

    trees = fresh(gpu_engine, jobs)
    scratch = Server(trees, engine=gpu_engine, slots=4, scratch=True).run(copy_reqs(reqs), realtime=False)

    assert all(not r.error for r in mixed + alone + scratch)
    worst_mix = max(abs(a.result - b.result) / abs(b.result) for a, b in zip(mixed, alone))
    worst_scr = max(abs(a.result - b.result) / abs(b.result) for a, b in zip(mixed, scratch))
    say(f"  mix vs alone max rel diff {worst_mix:.1e}; root-path vs full recompute {worst_scr:.1e}")
    assert worst_mix == 0.0          # same per-tree order, deterministic kernels
    assert worst_scr < 1e-11
    # final state of each tree against the FP64 oracle
    for k, (job, prep, alphabet) in enumerate(jobs):
        p2 = dict(prep)
        p2["t"] = final_t[k]
        ta, _, _ = build_oracle(p2, alphabet, job["gtr"])
        ref = ta.marginal()
        got = scratch_tree_total = trees[k].loglik()
        assert abs(got - ref) <= 1e-7 * abs(ref), (k, got, ref)
        assert np.max(np.abs(mixed_slh[k] - ta.sequence_LH) / np.maximum(1, np.abs(ta.sequence_LH))) < 5e-7
        del scratch_tree_total
    s = summarize(mixed, max(r.t_done for r in mixed))
    say(f"  served {s['requests']} requests: p50 {s['p50_ms']:.2f} ms, p99 {s['p99_ms']:.2f} ms")
    for t in trees:
        t.close()


def test_serve_budget_admission(gpu_engine):
    """Joint requests are admitted only while the engine's budget has room;
    their buffers are returned on completion; one that can never fit fails
    alone."""
    say("[serve] HBM budget admission")
    from treetime_amd.engine import Engine

    jobs = [small_job(n, L, 4, seed) for n, L, seed in SPECS[:3]]
    probe = Engine(0, 0)
    trees = fresh(probe, jobs)
    resident = probe.usage()[0]
    jb = [t.joint_bytes() for t in trees]
    for t in trees:
        t.close()
    probe.close()
    budget = resident + max(jb[:2]) + 4096 + 65536
    eng = Engine(0, budget)
    trees = fresh(eng, jobs)
    reqs = [Request("joint", 0, rid=0), Request("joint", 1, rid=1), Request("marginal", 2, rid=2),
            Request("joint", 0, rid=3)]
    out = Server(trees, engine=eng, slots=4, hbm_budget=budget).run(reqs, realtime=False)
    assert all(not r.error for r in out), [r.error for r in out]
    assert eng.usage()[0] <= budget
    for t in trees:
        t.close()
    eng.close()


def test_batch_members_equal_own_passes(gpu_engine):
    """tt_batch_run_members over subsets of a pool gives each member what its
    own pass gives (states and profiles identical, logL to 1e-12), leaves
    non-members untouched and is ordered after work already queued on a
    member's stream; tt_get_total after tt_batch_run equals the logL that
    call returned."""
    from treetime_amd.engine import batch_run, batch_run_members

    say("[serve] batched member runs vs each tree's own pass")
    jobs = [small_job(n, L, 4, seed) for n, L, seed in SPECS]
    pool = fresh(gpu_engine, jobs)
    ref = fresh(gpu_engine, jobs)
    for kind in ("marginal", "loglik", "joint"):
        logL = batch_run(gpu_engine, pool, kind=kind)  # builds and tunes the pool's plan
        assert [t.total() for t in pool] == list(logL), kind
        for members in ([0, 2, 4], [3], [1, 3]):
            others = [i for i in range(len(pool)) if i not in members]
            before = {i: (pool[i].states(), pool[i].total()) for i in others}
            batch_run_members(gpu_engine, pool, members, kind=kind)
            for i in members:
                want = getattr(ref[i], kind)()
                got = pool[i].total()
                assert abs(got - want) <= 1e-12 * abs(want), (kind, members, i, got, want)
                if kind != "loglik":
                    assert np.array_equal(pool[i].states(), ref[i].states()), (kind, i)
                if kind == "marginal":
                    assert np.array_equal(pool[i].profiles(), ref[i].profiles()), (kind, i)
                if kind == "loglik":
                    np.testing.assert_allclose(pool[i].sequence_lh(), ref[i].sequence_lh(), rtol=1e-12)
            for i in others:
                assert np.array_equal(pool[i].states(), before[i][0]) and pool[i].total() == before[i][1], (kind, i)
    # ordering: a marginal pass queued on tree 0's own stream, then a joint member run
    pool[0].marginal(want_logL=False)
    batch_run_members(gpu_engine, pool, [0], kind="joint")
    ref[0].joint()
    assert np.array_equal(pool[0].states(), ref[0].states())
    with pytest.raises(ValueError):
        batch_run_members(gpu_engine, pool, [0, 0], kind="marginal")
    with pytest.raises(ValueError):
        batch_run_members(gpu_engine, pool, [len(pool)], kind="marginal")
    for t in pool + ref:
        t.close()


def test_serve_device_batching_same_results(gpu_engine):
    """With device batching the served results equal the one-at-a-time run
    (to 1e-12: batched and per-tree passes sum logL in different orders), and
    batched launches were made."""
    say("[serve] device-batched stream vs alone")
    jobs = [small_job(n, L, 4, seed) for n, L, seed in SPECS]
    rng = np.random.default_rng(12)
    trees = fresh(gpu_engine, jobs)
    reqs = make_stream(trees, 150, 2000.0, rng, mix=(0.3, 0.4, 0.3))
    srv = Server(trees, engine=gpu_engine, slots=8, device_batch=2)
    batched = srv.run(copy_reqs(reqs), realtime=False)
    for t in trees:
        t.close()
    trees = fresh(gpu_engine, jobs)
    alone = Server(trees, engine=gpu_engine, slots=1).run(copy_reqs(reqs), realtime=False)
    for t in trees:
        t.close()
    assert all(not r.error for r in batched + alone), [r.error for r in batched + alone if r.error]
    worst = max(abs(a.result - b.result) / abs(b.result) for a, b in zip(batched, alone))
    n_fused = sum(1 for r in batched if r.meta.get("batched"))
    say(f"  {srv.batches} batched launches carried {n_fused} requests; max rel diff {worst:.1e}")
    assert srv.batches >= 1 and n_fused >= 2
    assert worst <= 1e-12
