"""Site-specific GTR models on the device (tt_tree_set_site_model and the
tt_site.hip kernels) against the oracle's SiteGTRModel
(gtr_site_specific.py), which is pinned to scipy expm, to the standard
model and to brute force in tests/test_oracle_pins.py.

Tolerances: marginal logL 1e-7 relative, per-pattern logL
5e-7 * max(1, |logL_p|), profiles 2e-6 absolute (FP32 rows and eigenbasis
arithmetic, FP64 normaliser products); states bit-exact where the oracle's
margin exceeds 1e-5.  Joint (FP64 messages, per-lane FP64 log P): logL
1e-9 relative, states bit-exact where the path margin exceeds 1e-5."""
import sys

import numpy as np
import pytest

import treetime_amd as tt
from oracle import fp64 as O
from treetime_amd import synth
from treetime_amd.engine import DeviceTree, batch_marginal, code_lut
from treetime_amd.prepare import prepare

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def site_case(n_tips, L, seed, approximate, depth=0.05, alphabet="nuc_nogap"):
    job = synth.make_job(n_tips, L, depth, 4, seed, 0.01, 0.005)
    prep = prepare(job["parent"], job["branch_length"], job["chars"], job["gtr"], job["ambiguous"], compress=False)
    g = tt.GTR_site_specific.random(L=L, alphabet=alphabet, rng=np.random.default_rng(seed + 100))
    g.approximate = approximate
    og = O.SiteGTRModel(alphabet, mu=g.mu, pi=g.Pi, W=g.W, approximate=approximate)
    parent = prep["parent"]
    names = ["n%d" % i for i in range(len(parent))]
    root, nodes = O.build_tree(parent, prep["t"], names)
    is_tip = np.ones(len(parent), dtype=bool)
    is_tip[parent[parent >= 0]] = False
    tips = np.where(is_tip)[0]
    chars = prep["tip_codes"].view("S1").astype("U1")
    comp = {names[v]: chars[k] for k, v in enumerate(tips)}
    ta = O.OracleTreeAnc(root, og, comp, prep["multiplicity"], full_length=1e300)
    return prep, g, ta, nodes


def device(engine, prep, g):
    lut = code_lut(g.profile_map, g.n_states)
    return DeviceTree(engine, prep["parent"], prep["t"], prep["tip_codes"], lut, prep["multiplicity"],
                      g.Pi.mean(axis=1), None, None, None, None, site=g.device_model())


@pytest.mark.parametrize("n_tips,L,seed,approximate", [(60, 400, 1, True), (60, 400, 2, False),
                                                       (300, 1500, 3, True)])
def test_site_marginal_matches_oracle(gpu_engine, n_tips, L, seed, approximate):
    say(f"[site] marginal tips={n_tips} L={L} seed={seed} approximate={approximate}")
    prep, g, ta, nodes = site_case(n_tips, L, seed, approximate)
    dev = device(gpu_engine, prep, g)
    lh = dev.marginal()
    ref = ta.marginal()
    rel = abs(lh - ref) / abs(ref)
    say(f"  logL {lh:.10f} oracle {ref:.10f} rel {rel:.2e}")
    assert rel < 1e-7
    seq = dev.sequence_lh()
    assert np.all(np.abs(seq - ta.sequence_LH) <= 5e-7 * np.maximum(1.0, np.abs(ta.sequence_LH)))
    prof = dev.profiles()
    states = dev.states()
    worst = 0.0
    for i, nd in enumerate(nodes):
        if nd.is_terminal():
            continue
        worst = max(worst, float(np.max(np.abs(prof[i] - nd.marginal_profile))))
        srt = np.sort(nd.marginal_profile, axis=1)
        sure = srt[:, -1] - srt[:, -2] > 1e-5
        assert np.array_equal(states[i][sure], nd.seq_idx[sure])
    say(f"  profiles max abs diff {worst:.2e}")
    assert worst < 2e-6
    dev.close()


@pytest.mark.parametrize("approximate", [True, False])
def test_site_joint_matches_oracle(gpu_engine, approximate):
    say(f"[site] joint approximate={approximate}")
    prep, g, ta, nodes = site_case(80, 500, 4, approximate)
    dev = device(gpu_engine, prep, g)
    lh = dev.joint()
    ref = ta.joint()
    say(f"  joint logL {lh:.10f} oracle {ref:.10f}")
    assert abs(lh - ref) <= 1e-9 * abs(ref)
    states = dev.states()
    for i, nd in enumerate(nodes):
        if nd.is_terminal():
            continue
        sure = nd.path_margin > 1e-5
        assert np.array_equal(states[i][sure], nd.seq_idx[sure])
    dev.close()


def test_site_model_switch_and_refusals(gpu_engine):
    """tt_tree_set_model after a site model restores the standard kernels;
    batched launches and root-path updates refuse site-specific trees."""
    prep, g, ta, nodes = site_case(40, 300, 5, True)
    dev = device(gpu_engine, prep, g)
    dev.marginal()
    with pytest.raises(ValueError):
        batch_marginal(gpu_engine, [dev])
    with pytest.raises(ValueError):
        dev.update_branch(int(dev.nonroot_nodes()[0]), 0.1)
    std = synth.make_job(40, 300, 0.05, 4, 5, 0.01, 0.005)["gtr"]
    dev.set_model(std.eigenvals, std.v, std.v_inv, std.mu)
    lh_switched = dev.marginal()
    lut = code_lut(std.profile_map, std.n_states)
    fresh = DeviceTree(gpu_engine, prep["parent"], prep["t"], prep["tip_codes"], lut, prep["multiplicity"], std.Pi,
                       std.eigenvals, std.v, std.v_inv, std.mu)
    lh_fresh = fresh.marginal()
    say(f"  switched {lh_switched:.10f} fresh {lh_fresh:.10f}")
    # the tree descriptor's pi was the sites' mean, the fresh one the model's
    std_oracle = O.OracleTreeAnc(ta.root, O.GTRModel("nuc_nogap", mu=std.mu, pi=std.Pi, W=std.W), ta.aln,
                                 prep["multiplicity"], full_length=1e300)
    assert abs(lh_fresh - std_oracle.marginal()) < 1e-7 * abs(lh_fresh)
    fresh.close()
    dev.close()



def test_joint_buffers_follow_the_model_kind(gpu_engine):
    """Joint passes across model switches: the joint buffers are reallocated
    between the site-specific (FP64 log) and the shared-model (FP32
    product-space) layouts, and each joint equals a tree built directly in
    that mode."""
    say("[site] joint across site -> shared -> site model switches")
    prep, g, ta, nodes = site_case(40, 300, 8, True)
    dev = device(gpu_engine, prep, g)
    lj_site = dev.joint()
    st_site = dev.states()
    std = synth.make_job(40, 300, 0.05, 4, 9, 0.01, 0.005)["gtr"]
    dev.set_model(std.eigenvals, std.v, std.v_inv, std.mu)
    lj_std = dev.joint(reconstruct_tips=True)
    st_std = dev.states()
    lut = code_lut(g.profile_map, g.n_states)
    fresh = DeviceTree(gpu_engine, prep["parent"], prep["t"], prep["tip_codes"], lut, prep["multiplicity"],
                       g.Pi.mean(axis=1), std.eigenvals, std.v, std.v_inv, std.mu)
    assert abs(lj_std - fresh.joint(reconstruct_tips=True)) <= 1e-12 * abs(lj_std)
    assert np.array_equal(st_std, fresh.states())
    fresh.close()
    dev.set_site_model(*g.device_model())
    assert abs(dev.joint() - lj_site) <= 1e-12 * abs(lj_site)
    assert np.array_equal(dev.states(), st_site)
    dev.close()

def test_site_propagate_profile_primitive():
    rng = np.random.default_rng(6)
    g = tt.GTR_site_specific.random(L=50, alphabet="nuc", rng=rng)
    og = O.SiteGTRModel("nuc", mu=g.mu, pi=g.Pi, W=g.W)
    prof = rng.random((50, 5))
    for t in (0.01, 0.7, 30.0):
        assert np.allclose(g.propagate_profile(prof, t), og.propagate_profile(prof, t), rtol=1e-12, atol=1e-14)
        assert np.allclose(g.propagate_profile(prof, t, return_log=True),
                           og.propagate_profile(prof, t, return_log=True), rtol=1e-12, atol=1e-12)
        assert np.allclose(g.evolve(prof, t), og.evolve(prof, t), rtol=1e-12, atol=1e-14)


def test_site_facade_treeanc(gpu_engine):
    """TreeAnc with a GTR_site_specific model (compress=False, as the
    reference requires, treeanc.py:186-187) equals the oracle."""
    job = synth.make_job(50, 300, 0.05, 4, 7, 0.0, 0.0)
    parent, bl = job["parent"], job["branch_length"]
    clades = [tt.Clade(branch_length=float(bl[i])) for i in range(len(parent))]
    for i, p in enumerate(parent):
        if p >= 0:
            clades[p].clades.append(clades[i])
    for k, v in enumerate(job["tips"]):
        clades[v].name = job["names"][k]
    aln = {job["names"][k]: job["chars"][k].tobytes().decode() for k in range(len(job["tips"]))}
    g = tt.GTR_site_specific.random(L=300, alphabet="nuc_nogap", rng=np.random.default_rng(8))
    with pytest.raises(TypeError):
        tt.TreeAnc(tree=tt.Tree(clades[0]), aln=aln, gtr=g, verbose=0)
    ta = tt.TreeAnc(tree=tt.Tree(clades[0]), aln=aln, gtr=g, compress=False, verbose=0, engine=gpu_engine)
    ta.infer_ancestral_sequences(marginal=True)
    lh = ta.sequence_LH()
    nodes = list(ta.tree.find_clades(order="preorder"))
    idx = {id(x): i for i, x in enumerate(nodes)}
    par = [-1] * len(nodes)
    for i, x in enumerate(nodes):
        for c in x.clades:
            par[idx[id(c)]] = i
    root, onodes = O.build_tree(par, [ta._branch_length_to_gtr(x) for x in nodes], [x.name for x in nodes])
    og = O.SiteGTRModel("nuc_nogap", mu=g.mu, pi=g.Pi, W=g.W)
    orc = O.OracleTreeAnc(root, og, ta.data.compressed_alignment, ta.data.multiplicity(), full_length=1e300)
    ref = orc.marginal()
    say(f"[site] facade marginal {lh:.10f} oracle {ref:.10f}")
    assert abs(lh - ref) < 1e-7 * abs(ref)
    ta.infer_ancestral_sequences(marginal=False)
    assert abs(ta.tree.sequence_joint_LH - orc.joint()) < 1e-9 * abs(ref)
    ta._drop_device()


@pytest.mark.parametrize("site_model", [True, False])
def test_site_counts_match_oracle(gpu_engine, site_model):
    """tt_site_counts: per-pattern statistics summed over branches, marginal
    (pair posteriors; the pattern's own P(t) under a site model, the branch's
    P(t) otherwise) and joint (state changes), against the oracle."""
    say(f"[site] counts site_model={site_model}")
    prep, g, ta, nodes = site_case(70, 300, 9, True)
    if site_model:
        dev = device(gpu_engine, prep, g)
    else:
        std = synth.make_job(70, 300, 0.05, 4, 9, 0.01, 0.005)["gtr"]
        dev = DeviceTree(gpu_engine, prep["parent"], prep["t"], prep["tip_codes"], code_lut(std.profile_map, 4),
                         prep["multiplicity"], std.Pi, std.eigenvals, std.v, std.v_inv, std.mu)
        ta = O.OracleTreeAnc(ta.root, O.GTRModel("nuc_nogap", mu=std.mu, pi=std.Pi, W=std.W), ta.aln,
                             prep["multiplicity"], full_length=1e300)
    dev.marginal()
    ta.marginal()
    n_d, t_d = dev.site_counts(True)
    n_o, t_o = ta.site_counts_marginal()
    en = np.max(np.abs(np.transpose(n_d, (1, 2, 0)) - n_o))
    et = np.max(np.abs(t_d.T - t_o) / np.maximum(1e-300, np.abs(t_o).max()))
    say(f"  marginal: n_ija max abs diff {en:.2e}, T_ia max rel diff {et:.2e}")
    assert en < 2e-6 and et < 1e-6
    dev.joint()
    ta.joint()
    cs = np.full(256, -1, dtype=np.int16)
    for i, c in enumerate("ACGT"):
        cs[ord(c)] = i
    n_d, t_d = dev.site_counts(False, cs)
    states = dev.states()
    sure = np.ones(prep["tip_codes"].shape[1], dtype=bool)
    for i, nd in enumerate(nodes):
        if not nd.is_terminal():
            sure &= nd.path_margin > 1e-5
    n_o, t_o = ta.site_counts_joint()
    assert np.array_equal(np.transpose(n_d, (1, 2, 0))[:, :, sure], n_o[:, :, sure])
    assert np.allclose(t_d.T[:, sure], t_o[:, sure], rtol=1e-12, atol=1e-12)
    assert states is not None
    dev.close()


@pytest.mark.parametrize("marginal", [True, False])
def test_site_facade_infer_gtr(gpu_engine, marginal):
    """TreeAnc.infer_gtr(site_specific=True) (treeanc.py:1500-1632) equals
    GTR_site_specific.infer on the oracle's statistics."""
    say(f"[site] facade infer_gtr site_specific marginal={marginal}")
    job = synth.make_job(60, 250, 0.08, 4, 10, 0.0, 0.0)
    parent, bl = job["parent"], job["branch_length"]
    clades = [tt.Clade(branch_length=float(bl[i])) for i in range(len(parent))]
    for i, p in enumerate(parent):
        if p >= 0:
            clades[p].clades.append(clades[i])
    for k, v in enumerate(job["tips"]):
        clades[v].name = job["names"][k]
    aln = {job["names"][k]: job["chars"][k].tobytes().decode() for k in range(len(job["tips"]))}
    ta = tt.TreeAnc(tree=tt.Tree(clades[0]), aln=aln, gtr=job["gtr"], compress=False, verbose=0, engine=gpu_engine)
    with pytest.raises(TypeError):
        tt.TreeAnc(tree=tt.Tree(clades[0]), aln=aln, gtr=job["gtr"], verbose=0,
                   engine=gpu_engine).infer_gtr(site_specific=True)
    ta.infer_ancestral_sequences(marginal=marginal)
    nodes = list(ta.tree.find_clades(order="preorder"))
    idx = {id(x): i for i, x in enumerate(nodes)}
    par = [-1] * len(nodes)
    for i, x in enumerate(nodes):
        for c in x.clades:
            par[idx[id(c)]] = i
    root, onodes = O.build_tree(par, [ta._branch_length_to_gtr(x) for x in nodes], [x.name for x in nodes])
    g0 = ta.gtr
    orc = O.OracleTreeAnc(root, O.GTRModel("nuc_nogap", mu=g0.mu, pi=g0.Pi, W=g0.W), ta.data.compressed_alignment,
                          ta.data.multiplicity(), full_length=1e300)
    if marginal:
        orc.marginal()
        n_o, t_o = orc.site_counts_marginal()
        root_state = orc.root.marginal_profile.T
    else:
        orc.joint()
        n_o, t_o = orc.site_counts_joint()
        root_state = np.eye(4)[:, orc.root.seq_idx]
    g = ta.infer_gtr(marginal=marginal, site_specific=True, pc=1.0)
    assert isinstance(g, tt.GTR_site_specific)
    ref = O.site_gtr_infer("nuc_nogap", n_o, t_o, root_state, pc=1.0)
    ref_mu = ref.mu / ref.average_rate().mean()
    dpi = np.max(np.abs(g.Pi - ref.Pi))
    dw = np.max(np.abs(g.W - ref.W)) / ref.W.max()
    say(f"  pi max abs diff {dpi:.2e}, W max rel diff {dw:.2e}")
    tol = 1e-5 if marginal else 1e-10
    assert dpi < tol and dw < tol
    assert np.allclose(g.mu, ref_mu, rtol=tol)
    # the inferred site-specific model drives the next reconstruction
    ta.infer_ancestral_sequences(marginal=True)
    assert np.isfinite(ta.sequence_LH())
    ta._drop_device()


def test_site_marginal_branch_sweep_matches_oracle(gpu_engine):
    """tt_optimize_branch_lengths_marginal and tt_branch_profile_lh with a
    site-specific model: prob_t_profiles with a (n, n, L) expQt (gtr.py:950-952)
    and the interpolated P(t) at every Brent evaluation; tolerances as for
    the standard model (optimum 1e-4 relative + 1e-8, objective no worse than
    the oracle's optimum by 1e-9 relative, grid 1e-6)."""
    from oracle.fp64 import normalize_profile

    prep, g, ta, nodes = site_case(40, 300, 11, True)
    dev = device(gpu_engine, prep, g)
    dev.marginal()
    t, st, pair = dev.optimize_branch_lengths_marginal(-1, 4.0, 1e-10)
    ta.marginal()
    og = ta.gtr
    mult = prep["multiplicity"]

    def objective(pp, pc, x):
        return -og.prob_t_profiles((pp, pc), mult, x, return_log=True) + np.exp(x ** 2 / 10000)

    worst = 0.0
    for v, nd in enumerate(nodes):
        if nd.up is None:
            continue
        pp, pc = nd.marginal_outgroup_LH, nd.marginal_subtree_LH
        t_ref = og.optimal_t_compressed((pp, pc), mult, profiles=True, tol=1e-10)
        assert st[v] == 0, (v, st[v])
        assert objective(pp, pc, t[v]) - objective(pp, pc, t_ref) <= 1e-9 * abs(objective(pp, pc, t_ref))
        assert abs(t[v] - t_ref) <= 1e-4 * t_ref + 1e-8, (v, t[v], t_ref)
        if t_ref > 1e-6:
            worst = max(worst, abs(t[v] - t_ref) / t_ref)
    root = [nd for nd in nodes if nd.up is None][0]
    if len(root.clades) == 2:
        n1, n2 = root.clades
        pp = normalize_profile(n2.marginal_subtree_LH * og.Pi.T)[0]
        t_ref = og.optimal_t_compressed((pp, n1.marginal_subtree_LH), mult, profiles=True, tol=1e-10)
        assert pair[1] == 0 and abs(pair[0] - t_ref) <= 1e-4 * t_ref + 1e-8, (pair, t_ref)
    say(f"[site] marginal sweep: {len(nodes) - 1} branches, max rel diff of optimal t {worst:.2e}")
    grid = np.array([-0.1, 1e-4, 0.003, 0.05, 0.4, 30.0])
    lh = dev.branch_profile_lh(grid)
    for v, nd in enumerate(nodes):
        if nd.up is None:
            continue
        ref = np.array([og.prob_t_profiles((nd.marginal_outgroup_LH, nd.marginal_subtree_LH), mult, x,
                                           return_log=True) for x in grid])
        assert np.allclose(lh[v], ref, rtol=1e-6, atol=1e-6), (v, lh[v], ref)
    dev.close()


def test_site_five_states_with_gaps(gpu_engine):
    """Alphabet 'nuc' (gap as a fifth state): marginal, joint and the
    gap-weighted branch objective (gtr.py:953-959) against the oracle."""
    say("[site] 5 states")
    job = synth.make_job(50, 300, 0.05, 4, 12, 0.01, 0.05)
    prep = prepare(job["parent"], job["branch_length"], job["chars"], job["gtr"], job["ambiguous"], compress=False)
    g = tt.GTR_site_specific.random(L=300, alphabet="nuc", rng=np.random.default_rng(112))
    og = O.SiteGTRModel("nuc", mu=g.mu, pi=g.Pi, W=g.W)
    parent = prep["parent"]
    names = ["n%d" % i for i in range(len(parent))]
    root, nodes = O.build_tree(parent, prep["t"], names)
    is_tip = np.ones(len(parent), dtype=bool)
    is_tip[parent[parent >= 0]] = False
    chars = prep["tip_codes"].view("S1").astype("U1")
    comp = {names[v]: chars[k] for k, v in enumerate(np.where(is_tip)[0])}
    ta = O.OracleTreeAnc(root, og, comp, prep["multiplicity"], full_length=1e300)
    dev = device(gpu_engine, prep, g)
    lh, ref = dev.marginal(), ta.marginal()
    say(f"  marginal {lh:.10f} oracle {ref:.10f}")
    assert abs(lh - ref) < 1e-7 * abs(ref)
    lj, rj = dev.joint(), ta.joint()
    assert abs(lj - rj) < 1e-9 * abs(rj)
    dev.marginal()
    t, st, _ = dev.optimize_branch_lengths_marginal(4, 4.0, 1e-10)
    mult = prep["multiplicity"]
    for v, nd in enumerate(nodes[:40]):
        if nd.up is None:
            continue
        pp, pc = nd.marginal_outgroup_LH, nd.marginal_subtree_LH
        t_ref = og.optimal_t_compressed((pp, pc), mult, profiles=True, tol=1e-10)
        assert abs(t[v] - t_ref) <= 1e-4 * t_ref + 1e-8, (v, t[v], t_ref)
    dev.close()


def test_site_model_argument_errors(gpu_engine):
    """tt_tree_set_site_model refuses malformed models (ValueError) and
    leaves the tree usable."""
    prep, g, ta, nodes = site_case(20, 60, 13, True)
    dev = device(gpu_engine, prep, g)
    eig, V, Vi, mu, pi, grid, rs = g.device_model()
    with pytest.raises(ValueError):
        dev.set_site_model(eig, V, Vi, mu, pi, grid[:1], rs)          # one grid point
    with pytest.raises(ValueError):
        dev.set_site_model(eig, V, Vi, mu, pi, grid[::-1].copy(), rs)  # grid not increasing
    bad_pi = pi.copy()
    bad_pi[3, 1] = 0.0
    with pytest.raises(ValueError):
        dev.set_site_model(eig, V, Vi, mu, bad_pi, grid, rs)          # zero frequency (log pi at the root)
    with pytest.raises(ValueError):
        dev.set_site_model(eig, V, Vi, -mu, pi, grid, rs)             # negative rate
    dev.set_site_model(eig, V, Vi, mu, pi, grid, rs)
    assert abs(dev.marginal() - ta.marginal()) < 1e-7 * abs(ta.marginal())
    dev.close()


def test_batch_run_refuses_bad_mode(gpu_engine):
    import ctypes as C

    from treetime_amd import _native

    prep, g, ta, nodes = site_case(10, 40, 14, True)
    dev = device(gpu_engine, prep, g)
    arr = (C.c_void_p * 1)(dev.handle.value)
    assert gpu_engine.lib.tt_batch_run(gpu_engine.handle, arr, 1, 0x40, None, None) == _native.TT_EINVAL
    assert gpu_engine.lib.tt_batch_run(gpu_engine.handle, arr, 1, 0x30, None, None) == _native.TT_EINVAL
    with pytest.raises(ValueError):   # site-specific trees are not batched
        from treetime_amd.engine import batch_run
        batch_run(gpu_engine, [dev], "loglik")
    dev.close()
