"""Alignment container and pattern compression.

Implements the compression rules of SequenceData.make_compressed_alignment
(treetime/sequence_data.py:325-464) with whole-array numpy operations instead
of a per-column Python loop:

* a column with exactly one state plus the ambiguous character has the
  ambiguous characters rewritten to that state (sequence_data.py:387-394);
* a column with more than one distinct character gets a private pattern,
  identical variable columns are NOT merged (sequence_data.py:396-402);
* constant columns merge, patterns are numbered in order of first
  appearance, multiplicity = number of columns (sequence_data.py:445-447);
* additional constant sites are appended per character composition
  (sequence_data.py:415-442).

``SequenceData(..., engine=e)`` compresses on the device instead
(``Engine.compress_alignment``, the same rules in HIP kernels); the result is
identical.

Sparse alignments -- a reference sequence plus, per sequence, a dict
position -> character of its differences (the form read_vcf produces) --
follow process_alignment_dictionary (treetime/vcf_utils.py:657-749) and the
is_sparse branch of make_compressed_alignment (sequence_data.py:362-413):
constant positions are classified from the differences alone and give one
pattern per character in character order, then the variable positions, in
order of first mention, are compressed by the dense rules (on the device when
an engine is given) and their constant columns merged into those patterns.
Parsing VCF files is not part of this package.
"""
import os

import numpy as np

from .errors import MissingDataError
from .seq_utils import alphabets, guess_alphabet, seq2array


def read_fasta(path):
    names, seqs, cur = [], [], []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            if line.startswith(">"):
                if names:
                    seqs.append("".join(cur))
                names.append(line[1:].split()[0])
                cur = []
            else:
                cur.append(line)
    if names:
        seqs.append("".join(cur))
    return dict(zip(names, seqs))


def _is_sparse(in_aln):
    """A dict name -> dict(position -> character) is a sparse alignment."""
    return isinstance(in_aln, dict) and bool(in_aln) and all(isinstance(v, dict) for v in in_aln.values())


def _as_sequence_dict(in_aln):
    """Accept a dict name -> sequence, a FASTA path, or any iterable of
    records with .id/.name and .seq (e.g. a Bio MultipleSeqAlignment)."""
    if isinstance(in_aln, dict):
        if any(isinstance(v, dict) for v in in_aln.values()):
            raise MissingDataError("SequenceData: a sparse alignment (dicts of differences) needs a reference "
                                   "sequence (ref=...)")
        return dict(in_aln)
    if isinstance(in_aln, str):
        if not os.path.isfile(in_aln):
            raise MissingDataError("SequenceData: loading alignment failed... " + in_aln)
        if in_aln.lower().endswith((".vcf", ".vcf.gz")):
            raise NotImplementedError("VCF files are not parsed here: pass the dict of differences per sequence "
                                      "and the reference sequence (ref=...)")
        return read_fasta(in_aln)
    try:
        out = {}
        for rec in in_aln:
            name = rec.id if getattr(rec, "id", None) and "<unknown" not in rec.id else rec.name
            out[name] = str(rec.seq)
        return out
    except TypeError as e:
        raise MissingDataError("SequenceData: loading alignment failed... " + str(in_aln)) from e


def compress_columns(A, ambiguous):
    """Pattern compression of a (N, L) uint8 character matrix.

    Returns (pattern index per column (L,), compressed (N, P) uint8).
    """
    N, L = A.shape
    amb = 256 if ambiguous is None else ord(ambiguous)
    is_amb = A == amb
    lo = np.where(is_amb, 255, A).min(axis=0).astype(np.int64)
    hi = np.where(is_amb, 0, A).max(axis=0).astype(np.int64)
    n_amb = is_amb.sum(axis=0)
    all_amb = n_amb == N
    constant = all_amb | (lo == hi)
    letter = np.where(all_amb, amb, lo)
    pos = np.arange(L, dtype=np.int64)
    keys = np.where(constant, letter, 256 + pos)
    uniq, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty(len(uniq), dtype=np.int64)
    rank[order] = np.arange(len(uniq))
    pat = rank[inv.reshape(-1)]
    first_col = first[order]
    comp = A[:, first_col].copy()
    const_p = constant[first_col]
    comp[:, const_p] = letter[first_col][const_p].astype(np.uint8)
    return pat, comp


def _char_codes(values, what):
    """Single characters -> uint8 codes (ASCII only)."""
    arr = np.asarray(values if isinstance(values, np.ndarray) else list(values))
    if arr.size == 0:
        return np.zeros(0, dtype=np.uint8)
    if arr.dtype.kind == "S" and arr.dtype.itemsize == 1:
        return arr.view(np.uint8).reshape(arr.shape)
    if arr.dtype.kind != "U" or arr.dtype.itemsize != 4:
        raise ValueError("{}: expected single characters".format(what))
    codes = arr.view(np.uint32).reshape(arr.shape)
    if codes.max(initial=0) > 255:
        raise ValueError("{}: non-ASCII characters".format(what))
    return codes.astype(np.uint8)


_MASK = ord(".")  # the placeholder process_alignment_dictionary writes over variable positions


def sparse_alignment_codes(names, diffs, L):
    """Flatten dicts of differences into (row, position, code) arrays, in
    the order the dicts are iterated (sequence by sequence)."""
    rows, pos, base = [], [], []
    for i, k in enumerate(names):
        d = diffs[k]
        if not d:
            continue
        p = np.fromiter(d.keys(), dtype=np.int64, count=len(d))
        if p.min() < 0 or p.max() >= L:
            raise ValueError("sparse alignment: position outside the reference in sequence {!r}".format(k))
        rows.append(np.full(len(d), i, dtype=np.int64))
        pos.append(p)
        base.append(_char_codes(list(d.values()), "sparse alignment"))
    if not rows:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.uint8)
    return np.concatenate(rows), np.concatenate(pos), np.concatenate(base)


def classify_sparse(nseq, pos, base, ref_codes, amb):
    """process_alignment_dictionary (treetime/vcf_utils.py:657-749) over the
    flattened differences (amb: code of the ambiguous character, -1 if none).

    Returns (variable positions in order of first mention, the reference
    with each non-reference constant position set to its allele (int64),
    sorted non-reference positions, positions made constant up to the
    ambiguous character, in order of first mention)."""
    ref_mod = ref_codes.astype(np.int64)
    if len(pos) == 0:
        return np.zeros(0, np.int64), ref_mod, np.zeros(0, np.int64), []
    upos, first, inv, cnt = np.unique(pos, return_index=True, return_inverse=True, return_counts=True)
    inv = inv.reshape(-1)
    is_amb = base == amb
    n_nonamb = np.bincount(inv, weights=~is_amb, minlength=len(upos)).astype(np.int64)
    has_amb = np.bincount(inv, weights=is_amb, minlength=len(upos)) > 0
    by_pos = np.argsort(inv, kind="stable")
    starts = np.concatenate(([0], np.cumsum(cnt)[:-1]))
    b = base[by_pos].astype(np.int64)
    amb_sorted = is_amb[by_pos]
    lo = np.minimum.reduceat(np.where(amb_sorted, 256, b), starts)
    hi = np.maximum.reduceat(np.where(amb_sorted, -1, b), starts)
    n_letters = np.where(n_nonamb == 0, 0, np.where(lo == hi, 1, 2))   # distinct non-ambiguous characters, capped
    full = cnt == nseq                                  # every sequence differs from the reference
    nonref_const = full & (n_letters <= 1)
    amb_const = ~full & (n_letters == 0)
    variable = ~(nonref_const | amb_const)
    inferred = (nonref_const & has_amb) | amb_const
    ref_mod[upos[nonref_const]] = np.where(n_letters == 0, amb, lo)[nonref_const]
    mention = np.argsort(first, kind="stable")
    var_pos = upos[mention][variable[mention]]
    inferred_sites = [int(x) for x in upos[mention][inferred[mention]]]
    return var_pos, ref_mod, upos, inferred_sites


def _groups(values):
    """value -> positions holding it (ascending), for an int array."""
    if len(values) == 0:
        return {}
    order = np.argsort(values, kind="stable")
    v = values[order]
    cuts = np.flatnonzero(np.diff(v)) + 1
    return dict(zip(v[np.concatenate(([0], cuts))].tolist(), np.split(order, cuts)))


def compress_sparse(names, diffs, ref_codes, amb_char, engine=None):
    """Sparse compression (vcf_utils.py:657-749, sequence_data.py:362-413,
    444-464).

    Returns (compressed codes (N, P) uint8, multiplicity (P,) float,
    full_to_compressed (L,) int64, compressed_to_full dict p -> positions,
    nonref_positions (sorted int64), inferred constant sites list).
    Positions whose constant state is the ambiguous character belong to no
    pattern and map to pattern 0, as in the reference."""
    nseq = len(names)
    L = len(ref_codes)
    amb = -1 if amb_char is None else ord(amb_char)
    rows, pos, base = sparse_alignment_codes(names, diffs, L)
    var_pos, ref_mod, nonref, inferred = classify_sparse(nseq, pos, base, ref_codes, amb)
    # one constant pattern per character of the modified reference, in
    # character order, taken before the variable positions are masked; the
    # ambiguous character and characters left with no position get none
    states = np.unique(ref_mod).tolist()
    ref_mod[var_pos] = _MASK
    groups = _groups(ref_mod)
    const_chars = [c for c in states if c != amb and c in groups]
    const_lists = [groups[c] for c in const_chars]
    n_const = len(const_chars)
    const_index = {c: i for i, c in enumerate(const_chars)}
    pat_full = np.zeros(L, dtype=np.int64)
    for p, lst in enumerate(const_lists):
        pat_full[lst] = p
    cols = [np.repeat(np.array(const_chars, dtype=np.uint8)[None, :], nseq, axis=0)] if n_const else []
    var_lists = {}
    P = n_const
    V = len(var_pos)
    if V:
        # the variable columns: the reference with each sequence's differences
        col_of = np.full(L, -1, dtype=np.int64)
        col_of[var_pos] = np.arange(V)
        M = np.repeat(ref_codes[var_pos][None, :], nseq, axis=0)
        sel = col_of[pos] >= 0
        M[rows[sel], col_of[pos[sel]]] = base[sel]
        if engine is not None:
            pv, comp_v, _, _ = engine.compress_alignment(M, amb_char)
        else:
            pv, comp_v = compress_columns(M, amb_char)
        # constant columns join the constant pattern of their character;
        # the others (and new constant characters) follow in order of first
        # appearance
        constant_v = (comp_v == comp_v[:1, :]).all(axis=0)
        remap = np.empty(comp_v.shape[1], dtype=np.int64)
        fresh = []
        for q in range(comp_v.shape[1]):
            c = int(comp_v[0, q])
            if constant_v[q] and c in const_index:
                remap[q] = const_index[c]
            else:
                remap[q] = P
                P += 1
                fresh.append(q)
        if fresh:
            cols.append(comp_v[:, fresh])
        fp = remap[pv]
        pat_full[var_pos] = fp
        grp = np.argsort(fp, kind="stable")
        fs = fp[grp]
        cuts = np.flatnonzero(np.diff(fs)) + 1
        var_lists = dict(zip(fs[np.concatenate(([0], cuts))].tolist(), np.split(var_pos[grp], cuts)))
    comp = np.concatenate(cols, axis=1) if cols else np.zeros((nseq, 0), dtype=np.uint8)
    c2f = {}
    for p in range(P):
        parts = ([const_lists[p]] if p < n_const else []) + ([var_lists[p]] if p in var_lists else [])
        c2f[p] = np.concatenate(parts).astype(int)
    mult = np.array([len(c2f[p]) for p in range(P)], dtype=float)
    return comp, mult, pat_full, c2f, nonref, inferred


class SequenceData:
    """Dense or sparse alignment + compressed pattern representation
    (treetime/sequence_data.py:12-128)."""

    def __init__(self, aln, ref=None, logger=None, convert_upper=True, sequence_length=None, compress=True,
                 word_length=1, sequence_type=None, fill_overhangs=True, seq_multiplicity=None, ambiguous=None,
                 engine=None, **kwargs):
        self.logger = logger if logger else (lambda *a, **k: None)
        self.engine = engine
        self.convert_upper = convert_upper
        self.compress = compress
        self.word_length = word_length
        self.fill_overhangs = fill_overhangs
        self.ambiguous = ambiguous
        self.sequence_type = sequence_type
        self.seq_multiplicity = seq_multiplicity or {}
        self.additional_constant_sites = kwargs.get("additional_constant_sites", 0)
        self.is_sparse = False
        self.nonref_positions = None
        self.inferred_const_sites = None
        self._ref = None
        self._full_length = int(sequence_length) if sequence_length else None
        self._aln = None
        self.likely_alphabet = None
        self.ref = ref
        self.aln = aln

    @property
    def ref(self):
        """The reference sequence as a character array (sequence_data.py:264-296)."""
        return self._ref

    @ref.setter
    def ref(self, in_ref):
        if in_ref is None:
            return
        if isinstance(in_ref, str) and os.path.isfile(in_ref):
            recs = read_fasta(in_ref)
            if not recs:
                raise TypeError("SequenceData.ref: reference sequence file %s could not be parsed, fasta is "
                                "supported." % in_ref)
            in_ref = next(iter(recs.values()))
        self._ref = seq2array(in_ref, word_length=self.word_length, fill_overhangs=False)
        if not self._full_length:
            self._full_length = int(self._ref.shape[0])

    @property
    def aln(self):
        return self._aln

    @aln.setter
    def aln(self, in_aln):
        self._aln = None
        if in_aln is None:
            return
        if _is_sparse(in_aln):
            if self._ref is None:
                raise MissingDataError("SequenceData: a sparse alignment (dicts of differences) needs a reference "
                                       "sequence (ref=...)")
            self.is_sparse = True
            self._check_alphabet([self._ref])
            self._aln = in_aln
            if self._full_length != len(self._ref):
                self.logger("SequenceData.aln: specified sequence length doesn't match reference length, ignoring "
                            "sequence length.", 1, warn=True)
            self._full_length = len(self._ref)
            self.sequence_names = list(self._aln.keys())
            self.make_compressed_alignment()
            return
        self.is_sparse = False
        raw = _as_sequence_dict(in_aln)
        seqs = list(raw.values())
        self._check_alphabet([seq2array(s) for s in seqs[:50]])
        self._aln = {k: seq2array(v, convert_upper=self.convert_upper, fill_overhangs=self.fill_overhangs,
                                  ambiguous=self.ambiguous) for k, v in raw.items()}
        lengths = {len(v) for v in self._aln.values()}
        if len(lengths) != 1:
            raise MissingDataError("SequenceData: sequences have different lengths")
        aln_len = lengths.pop()
        if self._full_length:
            if self._full_length < aln_len:
                raise AttributeError("SequenceData.aln: specified sequence length is smaller than alignment length!")
            self.additional_constant_sites = max(0, self._full_length - aln_len)
        else:
            self._full_length = aln_len
        self.sequence_names = list(self._aln.keys())
        self.make_compressed_alignment()

    def _check_alphabet(self, seqs):
        """sequence_data.py:308-323."""
        self.likely_alphabet = guess_alphabet(seqs)
        if self.ambiguous is None:
            self.ambiguous = "N" if self.likely_alphabet == "nuc" else "X"

    @property
    def full_length(self):
        return self._full_length

    @property
    def compressed_length(self):
        return self._compressed_length

    def multiplicity(self, mask=None):
        return self._multiplicity if mask is None else self._multiplicity * mask

    def make_compressed_alignment(self):
        names = self.sequence_names
        if self.is_sparse:
            return self._make_sparse_compressed_alignment()
        A = np.array([self._aln[k] for k in names])
        L = A.shape[1]
        codes = A.view(np.uint32).reshape(A.shape)
        if codes.max(initial=0) > 255:
            raise ValueError("non-ASCII characters in the alignment")
        codes = codes.astype(np.uint8)
        extra = self.additional_constant_sites if self.compress else 0
        byte_counts = None
        if not self.compress:
            pat = np.arange(L)
            comp = codes
        elif self.engine is not None:
            pat, comp, _, byte_counts = self.engine.compress_alignment(codes, self.ambiguous, char_counts=bool(extra))
        else:
            pat, comp = compress_columns(codes, self.ambiguous)
        cols = [comp]
        pat_full = [pat]
        if extra:
            chars = [c for c in alphabets[self.likely_alphabet + "_nogap"] if c not in (self.ambiguous, "-")]
            if byte_counts is not None:
                counts = {c: int(byte_counts[ord(c)]) for c in chars}
            else:
                counts = {c: int(np.sum(A == c)) for c in chars}
            total = sum(counts.values())
            plan = [(c, int(np.round(extra * m / total))) for c, m in counts.items()]
            left = extra
            P = comp.shape[1]
            const_letter = {}
            for p in range(P):
                col = comp[:, p]
                if (col == col[0]).all():
                    const_letter.setdefault(int(col[0]), p)
            for c, m in plan:
                if c == plan[-1][0]:
                    m = left
                if m:
                    code = ord(c)
                    if code in const_letter:
                        p = const_letter[code]
                    else:
                        p = P
                        P += 1
                        const_letter[code] = p
                        cols.append(np.full((len(names), 1), code, dtype=np.uint8))
                    pat_full.append(np.full(m, p, dtype=np.int64))
                    left -= m
        comp = np.concatenate(cols, axis=1)
        pat = np.concatenate(pat_full)
        self.full_to_compressed_sequence_map = pat
        self._multiplicity = np.bincount(pat, minlength=comp.shape[1]).astype(float)
        order = np.argsort(pat, kind="stable")
        bounds = np.cumsum(np.bincount(pat, minlength=comp.shape[1]))[:-1]
        self._set_patterns(comp, dict(enumerate(np.split(order, bounds))))

    def _set_patterns(self, comp, c2f):
        self.compressed_to_full_sequence_map = c2f
        self.compressed_codes = comp  # (N, P) uint8, rows in sequence_names order
        chars = comp.view("S1").astype("U1")
        self.compressed_alignment = {k: chars[i] for i, k in enumerate(self.sequence_names)}
        self._compressed_length = comp.shape[1]

    def _make_sparse_compressed_alignment(self):
        """sequence_data.py:362-413 with vcf_utils.py:657-749; uncompressed,
        the reference with each sequence's differences, one column each."""
        names = self.sequence_names
        ref_codes = _char_codes(self._ref, "reference sequence")
        L = len(ref_codes)
        if self.additional_constant_sites:
            raise ValueError("SequenceData: additional constant sites are not supported for sparse alignments")
        if not self.compress:
            rows, pos, base = sparse_alignment_codes(names, self._aln, L)
            comp = np.repeat(ref_codes[None, :], len(names), axis=0)
            comp[rows, pos] = base
            self.nonref_positions = np.unique(pos)
            self.inferred_const_sites = []
            self.full_to_compressed_sequence_map = np.arange(L)
            self._multiplicity = np.ones(L)
            self._set_patterns(comp, {p: np.array([p]) for p in range(L)})
            return
        comp, mult, f2c, c2f, nonref, inferred = compress_sparse(names, self._aln, ref_codes, self.ambiguous,
                                                                 engine=self.engine)
        self.nonref_positions = nonref
        self.inferred_const_sites = inferred
        self.full_to_compressed_sequence_map = f2c
        self._multiplicity = mult
        self._set_patterns(comp, c2f)

    def full_to_sparse_sequence(self, sequence):
        """Differences of a full-length sequence from the reference
        (sequence_data.py:466-487)."""
        if self._ref is None:
            raise TypeError("SequenceData: sparse sequences can only be constructed when a reference sequence is "
                            "defined")
        aseq = sequence if isinstance(sequence, np.ndarray) else seq2array(sequence, fill_overhangs=False)
        return {p: aseq[p] for p in np.where(self._ref != aseq)[0]}

    def compressed_to_sparse_sequence(self, sequence):
        """Differences of a compressed sequence from the reference at the
        non-reference positions (sequence_data.py:489-511)."""
        if self._ref is None:
            raise TypeError("SequenceData: sparse sequences can only be constructed when a reference sequence is "
                            "defined")
        nonref = self.nonref_positions
        vals = np.asarray(sequence)[self.full_to_compressed_sequence_map[nonref]]
        mism = vals != self._ref[nonref]
        return dict(zip(nonref[mism], vals[mism]))

    def compressed_to_full_sequence(self, sequence, include_additional_constant_sites=False, as_string=False):
        L = self.full_length if include_additional_constant_sites else self.full_length - self.additional_constant_sites
        tmp = np.asarray(sequence)[self.full_to_compressed_sequence_map[:L]]
        return "".join(tmp.astype("U")) if as_string else tmp

    def differences(self, seq1, seq2, seq1_compressed=True, seq2_compressed=True, mask=None):
        """(state in seq1, position, state in seq2) of every difference, by
        position (sequence_data.py:541-568); sparse alignments compare dicts
        of differences from the reference."""
        if self.is_sparse:
            if seq1_compressed:
                seq1 = self.compressed_to_sparse_sequence(seq1)
            if seq2_compressed:
                seq2 = self.compressed_to_sparse_sequence(seq2)
            diffs = []
            for p in set(seq1.keys()).union(seq2.keys()):
                r = self._ref[p]
                a, b = seq1.get(p, r), seq2.get(p, r)
                if a != b:
                    diffs.append((a, p, b))
            return sorted(diffs, key=lambda x: x[1])
        if seq1_compressed:
            seq1 = self.compressed_to_full_sequence(seq1)
        if seq2_compressed:
            seq2 = self.compressed_to_full_sequence(seq2)
        sel = seq1 != seq2
        if mask is not None:
            sel &= mask > 0
        return [(seq1[p], int(p), seq2[p]) for p in np.where(sel)[0]]
