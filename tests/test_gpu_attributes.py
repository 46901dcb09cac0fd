"""The reference's node attributes beyond the profiles, from the device,
against the FP64 oracle (whose attributes are pinned by enumeration in
tests/test_oracle_pins.py):

* marginal_subtree_LH, marginal_outgroup_LH (treeanc.py:857-899), and with
  debug=True marginal_log_Lx and marginal_subtree_LH_prefactor (803-810);
* get_branch_mutation_matrix (1085-1120) and the GTR-inference counts for
  the 20-state alphabet (1556-1572);
* joint_Lx / joint_Cx with debug=True (1066-1077), and joint reconstruction
  with a sampled root state (sample_from_profile='root' / True, 1022-1037),
  shared and site-specific models.

Tolerances: rows 2e-6 absolute (FP32 storage); log-valued rows (log
messages, prefactors, joint log messages) 5e-7 * max(1, |value|) + 2e-6;
argmax tables bit-exact where the oracle's margin exceeds 1e-5; sampled root
states identical where the draw lies more than 1e-5 from a cumulative
boundary."""
import sys

import numpy as np
import pytest

from helpers import build_oracle, polytomy_job, small_job
from treetime_amd.prepare import device_tree

pytestmark = pytest.mark.gpu

ROW_ATOL = 2e-6
MARGIN = 1e-5

CASES = [
    # name, n_tips, L, n_states, seed, depth, p_ambig, p_gap, p_iupac
    ("nuc", 120, 900, 4, 61, 0.05, 0.01, 0.005, 0.02),
    ("poly-nuc5", 90, 600, 5, 62, 0.05, 0.01, 0.02, 0.02),
    ("aa", 50, 300, 20, 63, 0.3, 0.02, 0.0, 0.02),
    ("poly-aa", 40, 250, 20, 64, 0.3, 0.02, 0.0, 0.0),
]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def make(case):
    name, n_tips, L, n, seed, depth, pa, pg, pi = case
    if name.startswith("poly"):
        return polytomy_job(n_tips, L, n, seed, 0.3, depth, pa, pg, pi)
    return small_job(n_tips, L, n, seed, depth, pa, pg, pi)


def log_err(got, ref):
    """max of |got - ref| - 5e-7 * max(1, |ref|) (<= 2e-6 passes)."""
    ok = np.isfinite(ref)
    return float(np.max(np.abs(got[ok] - ref[ok]) - 5e-7 * np.maximum(1.0, np.abs(ref[ok]))))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_marginal_node_rows(gpu_engine, case):
    say(f"[attributes] marginal rows case={case[0]}")
    job, prep, alphabet = make(case)
    ta, nodes, tips = build_oracle(prep, alphabet, job["gtr"])
    ta.marginal()
    dt = device_tree(gpu_engine, prep)
    dt.marginal(keep_normalisers=True)
    sub, outg, logm = dt.profiles(1), dt.profiles(2), dt.profiles(3)
    pref = dt.subtree_prefactor()
    e_sub = max(float(np.max(np.abs(sub[i] - nd.marginal_subtree_LH))) for i, nd in enumerate(nodes))
    e_out = max(float(np.max(np.abs(outg[i] - nd.marginal_outgroup_LH))) for i, nd in enumerate(nodes))
    e_log = max(log_err(logm[i], nd.marginal_log_Lx) for i, nd in enumerate(nodes) if nd.up is not None)
    e_pref = max(log_err(pref[i], nd.marginal_subtree_LH_prefactor) for i, nd in enumerate(nodes))
    say(f"  subtree {e_sub:.1e}  outgroup {e_out:.1e}  log msg {e_log:.1e}  prefactor {e_pref:.1e}")
    assert e_sub <= ROW_ATOL and e_out <= ROW_ATOL
    assert e_log <= 2e-6 and e_pref <= 2e-6
    # get_branch_mutation_matrix of internal and tip branches
    for i in [1, 2, len(nodes) // 2, len(nodes) - 1]:
        if nodes[i].up is None:
            continue
        m = dt.branch_mutation_matrix(i)
        ref = ta.branch_mutation_matrix(nodes[i])
        assert float(np.max(np.abs(m - ref))) <= ROW_ATOL, i
    # statistics for every alphabet size (20 states included)
    nij, Ti = dt.mutation_counts()
    rn, rt = ta.mutation_counts_marginal()
    en = float(np.max(np.abs(nij - rn)) / np.max(rn))
    et = float(np.max(np.abs(Ti - rt) / np.maximum(rt, 1e-300)))
    say(f"  mutation counts: n_ij rel err {en:.1e}  T_i rel err {et:.1e}")
    assert en < 1e-6 and et < 1e-6
    dt.close()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_joint_debug_rows(gpu_engine, case):
    say(f"[attributes] joint rows case={case[0]}")
    job, prep, alphabet = make(case)
    ta, nodes, tips = build_oracle(prep, alphabet, job["gtr"])
    ta.joint()
    dt = device_tree(gpu_engine, prep)
    dt.joint()
    lx, cx = dt.joint_debug()
    e_lx = max(log_err(lx[i], nd.joint_Lx) for i, nd in enumerate(nodes))
    say(f"  joint_Lx {e_lx:.1e}")
    assert e_lx <= 2e-6
    for i, nd in enumerate(nodes):
        if nd.up is None:
            continue
        sure = nd.joint_margin > MARGIN
        assert np.array_equal(cx[i][sure], nd.joint_Cx[sure]), i
    dt.close()


@pytest.mark.parametrize("case", CASES[:3], ids=[c[0] for c in CASES[:3]])
def test_joint_root_sampling(gpu_engine, case):
    """The root's joint profile, the draw of the root states with a seeded
    generator, and the backtrack and sequence_LH from the drawn states."""
    say(f"[attributes] joint root sampling case={case[0]}")
    job, prep, alphabet = make(case)
    ta, nodes, tips = build_oracle(prep, alphabet, job["gtr"])
    ref_total = ta.joint(root_rng=np.random.default_rng(17))
    r = ta.root
    dt = device_tree(gpu_engine, prep)
    dt.joint()
    got = dt.joint_root()
    normed = r.joint_Lx - r.joint_Lx.max(axis=1)[:, None]
    ok = normed > -80.0  # below, exp() of the FP32 row underflows (the draw cannot land there)
    assert float(np.max(np.abs(got[ok] - normed[ok]))) <= 2e-6
    # the draw, as TreeAnc._ml_anc_joint makes it
    prof = np.exp(got)
    prof = np.einsum("ai,a->ai", prof, 1.0 / prof.sum(axis=1))
    u = np.random.default_rng(17).random(size=prof.shape[0])
    idx = np.argmax(prof.cumsum(axis=1).T >= u, axis=0)
    sure = r.sample_margin > MARGIN
    assert np.array_equal(idx[sure], r.seq_idx[sure])
    say(f"  {int((~sure).sum())} of {len(sure)} draws within 1e-5 of a boundary")
    # finish from the oracle's draws
    total = dt.joint_from_root(r.seq_idx.astype(np.uint8))
    assert abs(total - ref_total) <= 1e-7 * abs(ref_total)
    slh = dt.sequence_lh()
    assert np.all(np.abs(slh - ta.joint_sequence_LH) <= 5e-7 * np.maximum(1.0, np.abs(ta.joint_sequence_LH)))
    states = dt.states()
    for i, nd in enumerate(nodes):
        if nd.is_terminal():
            continue
        sure = nd.path_margin > MARGIN
        assert np.array_equal(states[i][sure], nd.seq_idx[sure]), i
    dt.close()


def test_site_joint_root_sampling_and_rows(gpu_engine):
    """Root sampling, joint_Lx and joint_Cx under a site-specific model."""
    from test_gpu_site import device, site_case

    say("[attributes] site-specific joint sampling and rows")
    prep, g, ta, nodes = site_case(60, 400, 7, True)
    ref_total = ta.joint(root_rng=np.random.default_rng(23))
    r = ta.root
    dt = device(gpu_engine, prep, g)
    dt.joint()
    normed = r.joint_Lx - r.joint_Lx.max(axis=1)[:, None]
    assert float(np.max(np.abs(dt.joint_root() - normed))) <= 1e-9
    lx, cx = dt.joint_debug()
    e_lx = max(float(np.max(np.abs(lx[i] - nd.joint_Lx) / np.maximum(1.0, np.abs(nd.joint_Lx))))
               for i, nd in enumerate(nodes))
    assert e_lx <= 1e-8, e_lx
    for i, nd in enumerate(nodes):
        if nd.up is not None:
            sure = nd.joint_margin > MARGIN
            assert np.array_equal(cx[i][sure], nd.joint_Cx[sure]), i
    total = dt.joint_from_root(r.seq_idx.astype(np.uint8))
    assert abs(total - ref_total) <= 1e-9 * abs(ref_total)
    states = dt.states()
    for i, nd in enumerate(nodes):
        if not nd.is_terminal():
            sure = nd.path_margin > MARGIN
            assert np.array_equal(states[i][sure], nd.seq_idx[sure]), i
    dt.close()


def test_joint_from_root_refusals(gpu_engine):
    job, prep, alphabet = small_job(30, 200, 4, 65)
    dt = device_tree(gpu_engine, prep)
    with pytest.raises(Exception):
        dt.joint_root()  # no joint pass yet
    dt.marginal()
    with pytest.raises(Exception):
        dt.joint_from_root(np.zeros(dt.P, dtype=np.uint8))
    dt.joint()
    with pytest.raises(ValueError):
        dt.joint_from_root(np.full(dt.P, 7, dtype=np.uint8))  # state >= n
    with pytest.raises(ValueError):
        dt.joint_from_root(np.zeros(dt.P, dtype=np.uint8), reconstruct_tips=True)  # flags differ
    dt.close()
