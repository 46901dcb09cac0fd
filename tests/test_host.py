"""CPU tests of the host-side logic: compression, Newick, facade set-up,
error types, synthetic generators and the level schedule inputs."""
import numpy as np
import pytest

import treetime_amd as tt
from oracle import fp64 as O
from treetime_amd import synth
from treetime_amd.prepare import effective_lengths, fill_overhangs
from treetime_amd.sequence_data import compress_columns


def _codes(cols):
    return np.array([[ord(c[i]) for c in cols] for i in range(len(cols[0]))], dtype=np.uint8)


@pytest.mark.parametrize("seed", range(5))
def test_vectorised_compression_equals_oracle(seed):
    rng = np.random.default_rng(seed)
    N, L = 9, 300
    chars = np.array(list("ACGTN-R"))
    A = rng.choice(chars, size=(N, L), p=[0.3, 0.3, 0.2, 0.1, 0.05, 0.03, 0.02])
    # plenty of constant and near-constant columns
    A[:, ::3] = "A"
    A[rng.integers(0, N, size=L // 7), np.arange(0, L, 7)[: L // 7]] = "N"
    names = ["s%d" % i for i in range(N)]
    comp, mult, f2c, _ = O.compress_alignment(names, {nm: A[i] for i, nm in enumerate(names)}, "N")
    pat, c = compress_columns(A.view(np.uint32).astype(np.uint8), "N")
    assert np.array_equal(pat, f2c)
    assert np.array_equal(np.bincount(pat), mult)
    for i, nm in enumerate(names):
        assert "".join(c[i].view("S1").astype("U1")) == "".join(comp[nm])


def test_compression_rules_vectorised():
    cols = ["AAAA", "AANA", "ACAA", "ACAA", "A-AA", "NNNN", "CCCC", "NNNN", "ANCA", "N-NN"]
    pat, c = compress_columns(_codes(cols), "N")
    assert pat.tolist() == [0, 0, 1, 2, 3, 4, 5, 4, 6, 7]
    assert "".join(chr(x) for x in c[:, 7]) == "----"


def test_sequence_data_additional_constant_sites():
    d = tt.sequence_data.SequenceData({"a": "ACGTACGT", "b": "ACGAACGT", "c": "ACGTTCGT"}, sequence_length=20)
    assert d.full_length == 20 and d.multiplicity().sum() == 20
    assert d.additional_constant_sites == 12


def test_fill_overhangs_matches_reference_rule():
    rows = ["--AC-G--", "--------", "ACGT", "A--T"]
    for r in rows:
        got = fill_overhangs(np.array([[ord(x) for x in r]], dtype=np.uint8), "N")
        ref = O.fill_overhangs(np.array(list(r)), "N")
        assert "".join(chr(x) for x in got[0]) == "".join(ref)


def test_newick_roundtrip_and_ladderize():
    t = tt.Tree.from_newick("((A:0.1,(B:0.2,C:0.3)0.95:0.4)X:0.5,D:0.6);")
    assert [c.name for c in t.get_terminals()] == ["A", "B", "C", "D"]
    inner = t.root.clades[0].clades[1]
    assert inner.name is None and inner.confidence == 0.95
    t.ladderize()
    assert [c.name for c in t.get_terminals()] == ["D", "A", "B", "C"]
    t2 = tt.Tree.from_newick(t.to_newick())
    assert [c.name for c in t2.get_terminals()] == ["D", "A", "B", "C"]
    assert abs(t2.total_branch_length() - 2.1) < 1e-12


def test_postorder_preorder_orders():
    t = tt.Tree.from_newick("((A,B)X,(C,D)Y)R;")
    assert [n.name for n in t.find_clades("postorder")] == ["A", "B", "X", "C", "D", "Y", "R"]
    assert [n.name for n in t.find_clades("preorder")] == ["R", "X", "A", "B", "Y", "C", "D"]
    assert [n.name for n in t.find_clades("level")] == ["R", "X", "Y", "A", "B", "C", "D"]


def test_treeanc_setup_errors():
    with pytest.raises(tt.MissingDataError):
        tt.TreeAnc(tree=tt.Tree.from_newick("(A:0.1,B:0.2);"), aln={"A": "AC", "B": "AC"}, verbose=0)
    with pytest.raises(TypeError):
        tt.TreeAnc(tree=None)
    t = tt.TreeAnc(tree=tt.Tree.from_newick("((A:0.1,B:0.2):0.1,C:0.3);"), aln={"A": "ACG", "B": "ACT", "C": "ACG"},
                   gtr="JC69", verbose=0)
    with pytest.raises(tt.UnknownMethodError):
        t.infer_ancestral_sequences(method="bogus")
    with pytest.raises(NotImplementedError):
        t.infer_ancestral_sequences(method="fitch")
    assert [n.name for n in t.tree.get_nonterminals()] == ["NODE_0000000", "NODE_0000001"]
    with pytest.raises(tt.MissingDataError):
        tt.TreeAnc(tree=tt.Tree.from_newick("((A:0.1,B:0.2):0.1,(C:0.3,D:0.1):0.1);"), aln={"A": "ACG"}, verbose=0)


def test_gtr_host_matches_oracle():
    rng = np.random.default_rng(9)
    for alpha, n in [("nuc_nogap", 4), ("nuc", 5), ("aa_nogap", 20)]:
        pi = rng.dirichlet(np.ones(n) * 3)
        W = rng.lognormal(0, 0.5, (n, n))
        W = W + W.T
        g = tt.GTR.custom(mu=0.7, pi=pi, W=W, alphabet=alpha)
        o = O.GTRModel(alpha, mu=0.7, pi=pi, W=W)
        assert abs(g.mu - o.mu) < 1e-14 and np.allclose(g.W, o.W) and np.allclose(g.Pi, o.Pi)
        for t in [0.0, 0.01, 0.7]:
            assert np.allclose(g.expQt(t), o.expQt(t), atol=1e-14)
        with pytest.raises(ValueError):
            g.expQt(-1e6)


def test_gtr_standard_models():
    jc = tt.GTR.standard("JC69", alphabet="nuc_nogap")
    assert abs(jc.average_rate() - 1.0) < 1e-12
    P = jc.expQt(0.3)
    assert np.allclose(np.diag(P), 0.25 + 0.75 * np.exp(-0.4))
    for m in ["K80", "F81", "HKY85", "T92", "TN93"]:
        g = tt.GTR.standard(m)
        assert abs(g.Q.sum(0)).sum() < 1e-12
    with pytest.raises(KeyError):
        tt.GTR.standard("nope")


def test_gtr_from_file_fixture():
    import os

    g = tt.GTR.from_file(os.path.join(os.path.dirname(__file__), "golden", "gtr_fixture.txt"))
    assert np.allclose(g.Pi, [0.3088, 0.1897, 0.2335, 0.2581, 0.0099])
    assert abs(g.mu - 1.0) < 1e-4


def test_state_pair_matches_oracle():
    g = tt.GTR.standard("JC69", alphabet="nuc")
    o = O.GTRModel("nuc", mu=1.0, pi=np.ones(5), W=np.ones((5, 5)))
    rng = np.random.default_rng(1)
    a = rng.choice(list("ACGT-N"), 200)
    b = rng.choice(list("ACGT-N"), 200)
    m = rng.random(200) * 3
    for ig in [False, True]:
        p1, c1 = g.state_pair(a, b, m, ignore_gaps=ig)
        p2, c2 = o.state_pair(a, b, m, ignore_gaps=ig)
        d1 = {tuple(x): y for x, y in zip(p1.tolist(), c1.tolist())}
        d2 = {tuple(x): y for x, y in zip(p2.tolist(), c2.tolist())}
        assert d1 == d2


def test_kingman_and_simulation_shapes():
    rng = np.random.default_rng(0)
    parent, bl = synth.kingman_tree(50, 0.02, rng)
    assert (parent >= 0).sum() == 98 and (bl >= 0).all()
    p2, b2, _ = synth.preorder_renumber(parent, bl)
    assert p2[0] == -1 and all(p2[i] < i for i in range(1, len(p2)))
    job = synth.make_job(30, 200, 0.05, 4, 3, 0.01, 0.005)
    assert job["chars"].shape == (30, 200)
    # ultrametric: every tip at depth 0.05
    depth = np.zeros(len(job["parent"]))
    for i in range(1, len(job["parent"])):
        depth[i] = depth[job["parent"][i]] + job["branch_length"][i]
    assert np.allclose(depth[job["tips"]], 0.05)


def test_effective_lengths_floor():
    t = effective_lengths([0.0, 1e-9, 0.3], 1000)
    assert np.allclose(t, [1e-6, 1e-6, 0.3])


def test_gtr_infer_matches_oracle():
    """The facade's GTR.infer (host fixed point, gtr.py:491-599) equals the
    oracle's restatement on the same statistics."""
    from treetime_amd.gtr import GTR

    rng = np.random.default_rng(9)
    for alphabet, n in [("nuc_nogap", 4), ("nuc", 5)]:
        nij = rng.uniform(0, 50, (n, n))
        Ti = rng.uniform(10, 100, n)
        root = rng.uniform(0, 30, n)
        g = GTR.infer(nij, Ti, root, pc=1.0, alphabet=alphabet)
        o = O.gtr_infer(alphabet, nij, Ti, root, pc=1.0)
        assert np.allclose(g.W, o.W, rtol=1e-12, atol=0)
        assert np.allclose(g.Pi, o.Pi, rtol=1e-12, atol=0)
        assert abs(g.mu - o.mu) <= 1e-12 * o.mu


def test_gtr_site_specific_host_model_equals_oracle():
    """GTR_site_specific set-up (gtr_site_specific.py:43-116, 312-371) on
    the host: P(t) exact and interpolated equal the oracle's SiteGTRModel,
    the rate matrices have zero column sums, device_model's per-site layout
    rebuilds P(t)."""
    rng = np.random.default_rng(31)
    g = tt.GTR_site_specific.random(L=12, alphabet="nuc", rng=rng)
    o = O.SiteGTRModel("nuc", mu=g.mu, pi=g.Pi, W=g.W)
    for t in (0.0, 0.003, 0.05, 1.3, 40.0):
        assert np.allclose(g.expQt(t), o.expQt(t), atol=1e-14)
    assert np.abs(g.Q.sum(axis=0)).max() < 1e-14
    assert np.allclose(g.average_rate().mean(), 1.0)       # random() scales mu to avg_mu = 1
    eig, V, Vi, mu, pi, grid, rs = g.device_model()
    t = 0.21
    P = np.einsum("aik,ak,akj->aij", V, np.exp(t * mu[:, None] * eig), Vi)
    assert np.allclose(np.moveaxis(P, 0, -1), g._expQt(t), atol=1e-14)
    assert np.allclose(pi.T, g.Pi) and len(grid) == 61 and abs(grid[-1] * rs - 10) < 1e-12
    g.approximate = False
    assert len(g.device_model()[5]) == 0


def test_gtr_site_specific_argument_errors():
    g = tt.GTR_site_specific(seq_len=3, alphabet="nuc_nogap")
    with pytest.raises(ValueError):
        g.assign_rates(mu=np.ones(4), pi=np.ones((4, 3)))          # rate vector length != sites
    with pytest.raises(ValueError):
        g.assign_rates(pi=np.ones(5))                                # frequencies of the wrong length
    with pytest.raises(ValueError):
        g.assign_rates(W=np.ones((3, 3)))                            # W of the wrong size
    g.assign_rates(mu=2.0, pi=np.array([1.0, 2, 3, 4]))             # one pi for all sites
    assert g.Pi.shape == (4, 3) and np.allclose(g.Pi[:, 0], [0.1, 0.2, 0.3, 0.4])
    c = tt.GTR_site_specific.custom(mu=np.ones(5), pi=np.ones((4, 5)), alphabet="nuc_nogap")
    assert c.seq_len == 5 and c.is_site_specific


def test_site_treeanc_requires_uncompressed_alignment():
    g = tt.GTR_site_specific.random(L=3, alphabet="nuc_nogap", rng=np.random.default_rng(1))
    with pytest.raises(TypeError):
        tt.TreeAnc(tree=tt.Tree.from_newick("((A:0.1,B:0.2):0.1,C:0.3);"),
                   aln={"A": "ACG", "B": "ACT", "C": "ACG"}, gtr=g, verbose=0)


def test_batch_entry_points_reject_unknown_kind():
    """batch_info / batch_run name the pass kind; anything else is refused
    before any device call."""
    from treetime_amd.engine import batch_info, batch_run

    with pytest.raises(ValueError):
        batch_info(None, [], kind="posterior")
    with pytest.raises(ValueError):
        batch_run(None, [], "posterior")


# ------------------------------------------------------ sparse alignments --
@pytest.mark.parametrize("seed,nseq,L", [(0, 7, 400), (1, 12, 2000), (2, 3, 150), (3, 40, 5000)])
def test_sparse_compression_equals_oracle(seed, nseq, L):
    """The vectorised sparse path (vcf_utils.py:657-749 and the sparse
    branch of sequence_data.py:362-413) equals the oracle's position loop:
    patterns, their order, multiplicities, both maps (position order within
    each pattern included), non-reference and inferred constant sites."""
    from helpers import sparse_case
    from treetime_amd.sequence_data import SequenceData

    names, diffs, ref, A = sparse_case(seed, nseq, L)
    comp, mult, f2c, c2f, nonref, inferred = O.compress_sparse_alignment(names, diffs, ref, "N")
    sd = SequenceData(diffs, ref="".join(ref), ambiguous="N")
    assert sd.is_sparse and sd.full_length == L
    assert np.array_equal(sd.multiplicity(), mult)
    assert np.array_equal(sd.full_to_compressed_sequence_map, f2c)
    assert sorted(sd.compressed_to_full_sequence_map) == sorted(c2f)
    for p in c2f:
        assert np.array_equal(sd.compressed_to_full_sequence_map[p], c2f[p])
    for nm in names:
        assert np.array_equal(sd.compressed_alignment[nm], comp[nm])
    assert np.array_equal(sd.nonref_positions, nonref)
    assert sd.inferred_const_sites == inferred


def test_sparse_edge_cases():
    from treetime_amd.sequence_data import SequenceData

    ref = "ACGTACGT"
    # no differences at all: one pattern per reference character
    sd = SequenceData({"a": {}, "b": {}}, ref=ref)
    assert sd.compressed_length == 4 and sd.multiplicity().tolist() == [2, 2, 2, 2]
    assert sd.nonref_positions.tolist() == [] and sd.inferred_const_sites == []
    # one sequence: every mentioned position differs everywhere -> constant
    sd = SequenceData({"a": {1: "T", 2: "N"}}, ref=ref, ambiguous="N")
    assert sd.compressed_length == 4 and sd.inferred_const_sites == [2]
    assert sd.full_to_compressed_sequence_map.tolist() == [0, 3, 0, 3, 0, 1, 2, 3]
    # compress=False: the reference with the differences, one column each
    diffs = {"a": {1: "T"}, "b": {1: "G", 6: "N"}}
    sd = SequenceData(diffs, ref=ref, compress=False)
    assert sd.compressed_length == 8
    assert "".join(sd.compressed_alignment["b"]) == "AGGTACNT"
    with pytest.raises(tt.MissingDataError):
        SequenceData(diffs)                                  # no reference
    with pytest.raises(ValueError):
        SequenceData({"a": {8: "T"}}, ref=ref)               # outside the reference
    with pytest.raises(ValueError):
        SequenceData({"a": {1: "TT"}}, ref=ref)              # not a character
    with pytest.raises(ValueError):
        SequenceData(diffs, ref=ref, additional_constant_sites=5)
    # sequence_length is superseded by the reference's length
    assert SequenceData(diffs, ref=ref, sequence_length=20).full_length == 8


def test_sparse_sequence_helpers():
    """compressed_to_sparse_sequence / full_to_sparse_sequence / sparse
    differences (sequence_data.py:466-511, 541-568): a tip's compressed
    sequence maps back to its own differences, minus the ambiguous
    characters that compression rewrote and the restated reference
    characters."""
    from helpers import sparse_case
    from treetime_amd.sequence_data import SequenceData

    names, diffs, ref, A = sparse_case(5, 9, 600)
    sd = SequenceData(diffs, ref="".join(ref), ambiguous="N")
    exp_full = {}
    for i, nm in enumerate(names):
        full = sd.compressed_to_full_sequence(sd.compressed_alignment[nm])
        exp_full[nm] = full
        got = sd.compressed_to_sparse_sequence(sd.compressed_alignment[nm])
        want = {p: full[p] for p in sd.nonref_positions if full[p] != ref[p]}
        assert {int(k): str(v) for k, v in got.items()} == {int(k): str(v) for k, v in want.items()}
        # the expansion differs from the sequence only where compression
        # rewrote an ambiguous character or the position is in no pattern
        dropped = (A == "N").all(axis=0)
        changed = np.flatnonzero((full != A[i]) & ~dropped)
        assert all(A[i, p] == "N" for p in changed)
        # full_to_sparse_sequence: differences of the expanded sequence
        fs = sd.full_to_sparse_sequence(full)
        assert sorted(int(p) for p in fs) == np.flatnonzero(full != ref).tolist()
    # differences between two compressed sequences, and compressed vs sparse
    a, b = names[0], names[1]
    d = sd.differences(sd.compressed_alignment[a], sd.compressed_alignment[b])
    mism = [p for p in sd.nonref_positions if exp_full[a][p] != exp_full[b][p]]
    assert [int(x[1]) for x in d] == [int(p) for p in mism]
    assert all(x[0] == exp_full[a][x[1]] and x[2] == exp_full[b][x[1]] for x in d)
    d2 = sd.differences(sd.compressed_alignment[a], diffs[b], seq2_compressed=False)
    want = sorted({p for p in diffs[b]} | {int(p) for p in sd.nonref_positions if exp_full[a][p] != ref[p]})
    want = [p for p in want if diffs[b].get(p, ref[p]) != (exp_full[a][p] if p in sd.nonref_positions else ref[p])]
    assert [int(x[1]) for x in d2] == want
    with pytest.raises(NotImplementedError):
        import os
        import tempfile
        with tempfile.NamedTemporaryFile(suffix=".vcf", delete=False) as f:
            f.write(b"##fileformat=VCFv4.2\n")
        try:
            SequenceData(f.name, ref="".join(ref))
        finally:
            os.unlink(f.name)
