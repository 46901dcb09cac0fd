"""The full-size workloads on the GPU, held to the FP64 oracle.

* 5000 tips x 30000 sites at root-to-tip depth 1e-3 and 1e-2, and the
  20-state 1000 x 5000 job: marginal and joint passes through the C ABI.
* The 256-job mix, all jobs resident: the batched marginal, log-likelihood
  and joint passes equal the per-tree plain passes (profiles and states
  bit-identical, logL to 1e-12), and 16 jobs spread over the size range,
  the largest included, match the oracle.
* A served stream over a 64-tree pool of the mix under a 200 GB budget:
  each result is the same alone and in the mix, and every tree's final
  log-likelihood equals a fresh full pass.

The oracle needs minutes per large job, so its results come from
tests/golden/configs_golden.npz (tests/golden/make_config_golden.py, run on
the CPU); every test first checks that the job it rebuilt hashes to the
inputs the fixture was made from.  Tolerances as in test_gpu_parity: logL
relative 1e-7, per-pattern logL 5e-7 * max(1, |logL_p|), profile rows 2e-6
absolute, states bit-exact wherever the FP64 margin exceeds 1e-5.
"""
import multiprocessing as mp
import os
import sys

import numpy as np
import pytest

import config_digest as CD
from treetime_amd import synth
from treetime_amd.prepare import device_tree

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "configs_golden.npz")
LOGL_RTOL = 1e-7
SITE_RTOL = 5e-7
ROW_ATOL = 2e-6


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def golden(name):
    with np.load(GOLDEN) as z:
        return {k.split("/", 1)[1]: z[k] for k in z.files if k.startswith(name + "/")}


def site_err(got, ref):
    return float(np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref))))


def check_inputs(prep, g, name):
    assert CD.input_hash(prep) == str(g["hash"]), f"{name}: rebuilt job differs from the fixture's inputs"


def check_marginal(dt, total, g, name):
    ref = float(g["m_logL"])
    slh = dt.sequence_lh()
    prof, sub, outg = dt.profiles(0), dt.profiles(1), dt.profiles(2)
    n, r = g["snodes"], g["srows"]
    e_prof = float(np.max(np.abs(prof[n, r] - g["m_prof"])))
    e_sub = float(np.max(np.abs(sub[n, r] - g["m_sub"])))
    e_out = float(np.max(np.abs(outg[n, r] - g["m_outg"])))
    say(f"  {name} marginal: logL rel err {abs(total - ref) / abs(ref):.1e}, site err "
        f"{site_err(slh[g['pats']], g['m_slh']):.1e}, rows {e_prof:.1e} / {e_sub:.1e} / {e_out:.1e}")
    assert abs(total - ref) <= LOGL_RTOL * abs(ref)
    assert site_err(slh[g["pats"]], g["m_slh"]) <= SITE_RTOL
    assert max(e_prof, e_sub, e_out) <= ROW_ATOL
    bad = np.nonzero(CD.state_crcs(dt.states(), g["m_ties"]) != g["m_crc"])[0]
    assert len(bad) == 0, f"{name}: marginal states differ in node groups {bad[:8]}"


def check_joint(dt, total, g, name):
    ref = float(g["j_logL"])
    slh = dt.sequence_lh()
    say(f"  {name} joint: logL rel err {abs(total - ref) / abs(ref):.1e}, site err "
        f"{site_err(slh[g['pats']], g['j_slh']):.1e}")
    assert abs(total - ref) <= LOGL_RTOL * abs(ref)
    assert site_err(slh[g["pats"]], g["j_slh"]) <= SITE_RTOL
    bad = np.nonzero(CD.state_crcs(dt.states(), g["j_ties"]) != g["j_crc"])[0]
    assert len(bad) == 0, f"{name}: joint states differ in node groups {bad[:8]}"


@pytest.mark.parametrize("name", ["large-5000", "large-5000-deep", "aa-1000"])
def test_full_size_config(gpu_engine, name):
    say(f"[config] {name}")
    g = golden(name)
    _, prep = CD.make_prepared(CD.config_specs()[name])
    check_inputs(prep, g, name)
    dt = device_tree(gpu_engine, prep)
    check_marginal(dt, dt.marginal(), g, name)
    check_joint(dt, dt.joint(), g, name)
    dt.close()


# ---------------------------------------------------------------------------
# the 256-job mix
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def mix_preps():
    say("[mix] generating the 256 jobs")
    order = sorted(range(256), key=lambda j: -synth.mix_job_spec(j)["n_tips"] * synth.mix_job_spec(j)["L"])
    with mp.get_context("spawn").Pool(min(16, os.cpu_count() or 1)) as pool:
        res = pool.map(CD.prepared_mix_job, order, chunksize=1)
    out = [None] * 256
    for j, p in zip(order, res):
        out[j] = p
    return out


@pytest.mark.timeout(900)
def test_mix_batched_equals_plain_and_oracle(gpu_engine, mix_preps):
    from treetime_amd.engine import batch_run

    trees = [device_tree(gpu_engine, p) for p in mix_preps]
    say("[mix] 256 trees resident")
    sample = CD.mix_sample_jobs()
    for kind in ("marginal", "loglik", "joint"):
        got = batch_run(gpu_engine, trees, kind=kind)
        if kind != "loglik":  # the 16 sampled jobs against the oracle
            for j in sample:
                name = "mix%03d" % j
                g = golden(name)
                check_inputs(mix_preps[j], g, name)
                (check_marginal if kind == "marginal" else check_joint)(trees[j], got[j], g, name)
        worst = 0.0
        for k, t in enumerate(trees):  # every job against its own plain pass
            b_slh = t.sequence_lh()
            b_states = t.states() if kind != "loglik" else None
            b_prof = t.profiles(0) if kind == "marginal" else None
            plain = {"marginal": t.marginal, "loglik": t.loglik, "joint": t.joint}[kind]()
            worst = max(worst, abs(plain - got[k]) / abs(plain))
            assert abs(plain - got[k]) <= 1e-12 * abs(plain), (kind, k, plain, got[k])
            assert site_err(b_slh, t.sequence_lh()) <= 1e-12, (kind, k)
            if b_states is not None:
                assert np.array_equal(t.states(), b_states), (kind, k)
            if b_prof is not None:
                assert np.array_equal(t.profiles(0), b_prof), (kind, k)
            del b_prof
        say(f"  {kind}: batched == plain over 256 jobs (logL max rel diff {worst:.1e})")
    for t in trees:
        t.close()


@pytest.mark.timeout(600)
def test_mix_served_pool(gpu_engine, mix_preps):
    """A 70/20/10 branch/marginal/joint stream over 64 trees of the mix under
    a 200 GB admission budget: the same results alone and in the mix, and
    each tree's final root-path logL equals a full pass."""
    from treetime_amd.serve import Request, Server, make_stream, summarize

    pool = mix_preps[:64]
    budget = 200e9

    def fresh():
        return [device_tree(gpu_engine, p) for p in pool]

    def copy_reqs(reqs):
        return [Request(r.kind, r.tree, r.arrival, r.node, r.scale, r.rid) for r in reqs]

    trees = fresh()
    reqs = make_stream(trees, 400, 4000.0, np.random.default_rng(64))
    mixed = Server(trees, engine=gpu_engine, slots=16, hbm_budget=budget).run(copy_reqs(reqs), realtime=True)
    full = [t.loglik() for t in trees]  # a full postorder at the final lengths
    for t in trees:
        t.close()
    trees = fresh()
    alone = Server(trees, engine=gpu_engine, slots=1, hbm_budget=budget).run(copy_reqs(reqs), realtime=False)
    for t in trees:
        t.close()
    assert all(not r.error for r in mixed + alone), [r.error for r in mixed + alone if r.error][:3]
    worst = max(abs(a.result - b.result) / abs(b.result) for a, b in zip(mixed, alone))
    last = {}
    for r in mixed:
        last[r.tree] = r
    # trees whose last request was a root-path branch update: its logL
    # against a full postorder at the final lengths
    ends = [k for k, r in last.items() if r.kind == "branch"]
    assert ends
    fin = max(abs(last[k].result - full[k]) / abs(full[k]) for k in ends)
    s = summarize(mixed, max(r.t_done for r in mixed))
    say(f"[serve-pool] 64 trees, {s['requests']} requests: p50 {s['p50_ms']:.2f} ms p99 {s['p99_ms']:.2f} ms; "
        f"mix vs alone {worst:.1e}; final root-path vs full pass {fin:.1e}")
    assert worst == 0.0
    assert fin < 1e-11
