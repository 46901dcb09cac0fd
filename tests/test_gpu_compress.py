"""Pattern compression on the device (tt_compress_alignment) against the
oracle's column-by-column restatement of make_compressed_alignment
(sequence_data.py:325-464) and the overhang fill (seq_utils.py:196-202).

All outputs are integers and bytes: the bar is bit-exact equality.  Large
alignments (5000 x 30000) are checked against the vectorised numpy
compression, itself pinned to the oracle on the CPU, and through
size-independent properties: expanding the patterns through the column map
gives back the alignment (ambiguous characters of constant columns
rewritten), multiplicities are the column counts, and every variable column
has a pattern of its own."""
import sys

import numpy as np
import pytest

from oracle import fp64 as O
from treetime_amd.prepare import fill_overhangs
from treetime_amd.sequence_data import SequenceData, compress_columns

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _rule_alignment(rng, N, L, chars="ACGTN-R", p=(0.3, 0.3, 0.2, 0.1, 0.05, 0.03, 0.02)):
    A = rng.choice(np.array(list(chars)), size=(N, L), p=p)
    A[:, ::3] = "A"                                   # constant columns
    A[rng.integers(0, N, size=L // 7), np.arange(0, L, 7)[: L // 7]] = "N"   # one state + N
    A[:, 5::11] = "N"                                 # all-ambiguous columns
    A[:, 1::13] = "-"                                 # gap columns
    if N > 2:
        A[0, : L // 10] = "-"                         # overhangs
        A[1, -L // 9:] = "-"
        A[2, :] = "-"                                 # a row of gaps only
    return A


def _oracle(A, amb, fill):
    names = ["s%d" % i for i in range(A.shape[0])]
    rows = [O.fill_overhangs(A[i].copy(), amb) if fill else A[i] for i in range(A.shape[0])]
    comp, mult, f2c, _ = O.compress_alignment(names, dict(zip(names, rows)), amb)
    return np.array([comp[nm] for nm in names]), mult, f2c


def _codes(A):
    return np.ascontiguousarray(A).view(np.uint32).astype(np.uint8).reshape(A.shape)


@pytest.mark.parametrize("N,L,seed,fill", [(9, 300, 1, False), (9, 300, 2, True), (40, 2001, 3, True),
                                           (3, 17, 4, True), (1, 50, 5, False), (64, 1024, 6, False)])
def test_compression_matches_oracle(gpu_engine, N, L, seed, fill):
    say(f"[compress] oracle N={N} L={L} seed={seed} fill={fill}")
    rng = np.random.default_rng(seed)
    A = _rule_alignment(rng, N, L)
    comp, mult, f2c = _oracle(A, "N", fill)
    pat, got, m, _ = gpu_engine.compress_alignment(_codes(A), "N", fill_overhangs=fill)
    assert np.array_equal(pat, f2c)
    assert np.array_equal(m, mult)
    assert np.array_equal(got, comp.view(np.uint32).astype(np.uint8).reshape(comp.shape))


def test_compression_rules_by_hand(gpu_engine):
    """Expected values written out by hand for each rule of
    sequence_data.py:387-402."""
    cols = ["AAAA", "AANA", "ACAA", "ACAA", "A-AA", "NNNN", "CCCC", "NNNN", "ANCA", "N-NN"]
    A = np.array([[ord(c[i]) for c in cols] for i in range(4)], dtype=np.uint8)
    pat, comp, mult, _ = gpu_engine.compress_alignment(A, "N")
    assert pat.tolist() == [0, 0, 1, 2, 3, 4, 5, 4, 6, 7]
    assert mult.tolist() == [2, 1, 1, 1, 2, 1, 1, 1]
    assert "".join(chr(x) for x in comp[:, 0]) == "AAAA"      # N rewritten to A
    assert "".join(chr(x) for x in comp[:, 7]) == "----"      # '-' + N is constant '-'
    assert "".join(chr(x) for x in comp[:, 6]) == "ANCA"      # variable: kept as is


def test_no_ambiguous_character(gpu_engine):
    rng = np.random.default_rng(7)
    A = _rule_alignment(rng, 12, 500)
    pat, comp, mult, _ = gpu_engine.compress_alignment(_codes(A), None)
    rp, rc = compress_columns(_codes(A), None)
    assert np.array_equal(pat, rp) and np.array_equal(comp, rc)
    assert np.array_equal(mult, np.bincount(rp))


def test_char_counts_and_capacity(gpu_engine):
    rng = np.random.default_rng(8)
    A = _rule_alignment(rng, 20, 999)
    codes = _codes(A)
    _, _, _, counts = gpu_engine.compress_alignment(codes, "N", fill_overhangs=True, char_counts=True)
    assert np.array_equal(counts, np.bincount(fill_overhangs(codes, "N").ravel(), minlength=256))
    # capacity below P is an error (ValueError), not a truncated result
    import ctypes as C

    from treetime_amd import _native
    P = C.c_int64()
    out = np.empty(20 * 4, dtype=np.uint8)
    pat = np.empty(999, dtype=np.int32)
    mult = np.empty(4)
    rc = gpu_engine.lib.tt_compress_alignment(gpu_engine.handle, codes.ctypes.data, 20, 999, 999, ord("N"), 0, 4,
                                              C.byref(P), pat.ctypes.data, out.ctypes.data, mult.ctypes.data, None,
                                              None)
    assert rc == _native.TT_ERANGE and P.value > 4


def test_device_pointers_and_pitch(gpu_engine):
    """Device-resident input at a padded pitch (16-byte path) and at an odd
    pitch (byte path) equal the host-buffer call."""
    import torch

    rng = np.random.default_rng(9)
    N, L = 33, 4099
    A = _codes(_rule_alignment(rng, N, L))
    ref_pat, ref_comp, ref_mult, _ = gpu_engine.compress_alignment(A, "N", fill_overhangs=True)
    P = ref_comp.shape[1]
    for pitch in (4112, 4099, 4101):
        buf = torch.full((N, pitch), ord("Z"), dtype=torch.uint8)
        buf[:, :L] = torch.from_numpy(A)
        d = buf.cuda()
        pat = torch.empty(L, dtype=torch.int32, device="cuda")
        comp = torch.empty(N * L, dtype=torch.uint8, device="cuda")
        mult = torch.empty(L, dtype=torch.float64, device="cuda")
        got = gpu_engine.compress_alignment_device(d, L, "N", pat, comp, mult, fill_overhangs=True)
        torch.cuda.synchronize()
        assert got == P
        assert np.array_equal(pat.cpu().numpy(), ref_pat)
        assert np.array_equal(comp[:N * P].cpu().numpy().reshape(N, P), ref_comp)
        assert np.array_equal(mult[:P].cpu().numpy(), ref_mult)


@pytest.mark.parametrize("N,L,depth,seed", [(5000, 30000, 1e-3, 11), (1000, 5000, 0.3, 12)])
def test_large_alignment_properties(gpu_engine, N, L, depth, seed):
    """Alignment shapes of the large single-tree workloads: equal to the
    numpy compression and to the size-independent properties."""
    say(f"[compress] large N={N} L={L}")
    rng = np.random.default_rng(seed)
    # mostly constant columns with sparse mutations, N and overhangs
    base = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=L)
    A = np.repeat(base[None, :], N, axis=0)
    k = int(depth * N * L * 0.05) + 1
    A[rng.integers(0, N, k), rng.integers(0, L, k)] = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), size=k)
    k = N * L // 50
    A[rng.integers(0, N, k), rng.integers(0, L, k)] = ord("N")
    for r in rng.integers(0, N, 50):
        A[r, : rng.integers(1, 300)] = ord("-")
        A[r, L - rng.integers(1, 300):] = ord("-")
    pat, comp, mult, counts = gpu_engine.compress_alignment(A, "N", fill_overhangs=True, char_counts=True)
    F = fill_overhangs(A, "N")
    rp, rc = compress_columns(F, "N")
    assert np.array_equal(pat, rp)
    assert np.array_equal(comp, rc)
    assert np.array_equal(mult, np.bincount(rp).astype(float))
    assert np.array_equal(counts, np.bincount(F.ravel(), minlength=256))
    # properties that hold at any size
    expanded = comp[:, pat]
    amb = F == ord("N")
    assert np.array_equal(expanded[~amb], F[~amb])
    P = comp.shape[1]
    assert mult.sum() == L and np.all(np.diff(np.maximum.accumulate(pat)) >= 0)
    first = np.full(P, L)
    np.minimum.at(first, pat, np.arange(L))
    assert np.all(np.diff(first) > 0)     # patterns numbered in order of first column
    const = (comp == comp[:1]).all(axis=0)
    assert np.all(mult[~const] == 1)      # variable columns are private
    say(f"  P={P} ({int(const.sum())} constant patterns)")


def test_sequence_data_on_device_equals_host(gpu_engine):
    rng = np.random.default_rng(13)
    A = _rule_alignment(rng, 15, 700, chars="ACGTN-R")
    aln = {"s%d" % i: "".join(A[i]) for i in range(15)}
    for extra in (None, 1000):
        h = SequenceData(aln, sequence_length=extra)
        d = SequenceData(aln, sequence_length=extra, engine=gpu_engine)
        assert np.array_equal(h.full_to_compressed_sequence_map, d.full_to_compressed_sequence_map)
        assert np.array_equal(h.multiplicity(), d.multiplicity())
        assert np.array_equal(h.compressed_codes, d.compressed_codes)
        for k in aln:
            assert np.array_equal(h.compressed_alignment[k], d.compressed_alignment[k])


# ------------------------------------------------------ sparse alignments --
def test_sparse_on_device_equals_host(gpu_engine):
    """Sparse compression with the variable columns compressed on the device
    equals the vectorised host path (itself equal to the oracle's position
    loop on the CPU): bit-exact patterns, multiplicities and maps."""
    from helpers import sparse_case

    for seed, nseq, L in [(11, 7, 400), (12, 400, 40000)]:
        say(f"[compress] sparse device vs host nseq={nseq} L={L}")
        names, diffs, ref, A = sparse_case(seed, nseq, L)
        if L <= 400:
            comp, mult, f2c, c2f, nonref, inferred = O.compress_sparse_alignment(names, diffs, ref, "N")
        host = SequenceData(diffs, ref="".join(ref), ambiguous="N")
        dev = SequenceData(diffs, ref="".join(ref), ambiguous="N", engine=gpu_engine)
        assert np.array_equal(dev.compressed_codes, host.compressed_codes)
        assert np.array_equal(dev.multiplicity(), host.multiplicity())
        assert np.array_equal(dev.full_to_compressed_sequence_map, host.full_to_compressed_sequence_map)
        for p, v in host.compressed_to_full_sequence_map.items():
            assert np.array_equal(dev.compressed_to_full_sequence_map[p], v)
        assert np.array_equal(dev.nonref_positions, host.nonref_positions)
        assert dev.inferred_const_sites == host.inferred_const_sites
        if L <= 400:
            assert np.array_equal(dev.multiplicity(), mult)
            assert np.array_equal(dev.full_to_compressed_sequence_map, f2c)


def _named_tree(parent, bl, tip_names):
    from treetime_amd.tree import Clade, Tree

    n = len(parent)
    nodes = [Clade(branch_length=float(bl[i])) for i in range(n)]
    k = 0
    is_tip = np.ones(n, dtype=bool)
    is_tip[parent[parent >= 0]] = False
    for i in range(n):
        if is_tip[i]:
            nodes[i].name = tip_names[k]
            k += 1
        else:
            nodes[i].name = "node_%d" % i
        if parent[i] >= 0:
            nodes[parent[i]].clades.append(nodes[i])
    return Tree(root=nodes[int(np.flatnonzero(parent < 0)[0])])


def test_sparse_treeanc_equals_dense(gpu_engine):
    """TreeAnc on reference + differences against TreeAnc on the expanded
    dense alignment: same log-likelihood (1e-7 relative), the same
    reconstructed states at every position that belongs to a pattern, and
    the sparse output format of get_reconstructed_alignment
    (treeanc.py:1739-1746)."""
    import treetime_amd as tt
    from treetime_amd import synth

    say("[compress] sparse TreeAnc vs dense")
    job = synth.make_job(300, 3000, 0.05, 4, 21, p_ambig=0.01)
    chars = job["chars"].astype("U1")
    L = chars.shape[1]
    ref = np.array([max("ACGT", key=lambda c: int((chars[:, j] == c).sum())) for j in range(L)])
    ref[::97] = "N"                      # a few ambiguous reference positions
    chars[:, 5::211] = "N"               # and all-ambiguous columns
    names = ["t%d" % i for i in range(chars.shape[0])]
    diffs = {nm: {int(p): str(chars[i, p]) for p in np.flatnonzero(chars[i] != ref)} for i, nm in enumerate(names)}
    dense = {nm: "".join(chars[i]) for i, nm in enumerate(names)}
    g = job["gtr"]
    res = {}
    for kind, aln, ref_arg in [("sparse", diffs, "".join(ref)), ("dense", dense, None)]:
        tree = _named_tree(job["parent"], job["branch_length"], names)
        gtr = tt.GTR.custom(alphabet="nuc_nogap", pi=g.Pi, W=g.W, mu=g.mu)
        t = tt.TreeAnc(tree=tree, aln=aln, ref=ref_arg, gtr=gtr, engine=gpu_engine, verbose=0, fill_overhangs=False)
        t.infer_ancestral_sequences(marginal=True)
        res[kind] = t
    ts, td = res["sparse"], res["dense"]
    assert ts.data.is_sparse and not td.data.is_sparse
    lh_s, lh_d = ts.sequence_LH(), td.sequence_LH()
    say(f"  logL sparse {lh_s:.10f} dense {lh_d:.10f}")
    assert abs(lh_s - lh_d) <= 1e-7 * abs(lh_d)
    out = ts.get_reconstructed_alignment()
    assert set(out) == {"sequences", "reference", "positions", "inferred_const_sites"}
    assert np.array_equal(out["positions"], ts.data.nonref_positions)
    keep = ~(chars == "N").all(axis=0)     # positions in some pattern
    dn = {n.name: n for n in td.tree.find_clades()}
    for n in ts.tree.find_clades():
        if n.is_terminal():
            continue
        full = ref.copy()
        for p, c in out["sequences"][n.name].items():
            full[p] = c
        want = td.sequence(dn[n.name], as_string=False)
        assert np.array_equal(full[keep], want[keep]), n.name
    # tip mutations name the tip's own character
    for n in ts.tree.get_terminals()[:20]:
        for a, p, b in n.mutations:
            assert b == diffs[n.name].get(p, ref[p])
