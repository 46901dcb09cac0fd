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
identical.  Sparse (VCF) alignments are out of scope.
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


def _as_sequence_dict(in_aln):
    """Accept a dict name -> sequence, a FASTA path, or any iterable of
    records with .id/.name and .seq (e.g. a Bio MultipleSeqAlignment)."""
    if isinstance(in_aln, dict):
        if in_aln and isinstance(next(iter(in_aln.values())), dict):
            raise NotImplementedError("sparse (VCF) alignments are not supported")
        return dict(in_aln)
    if isinstance(in_aln, str):
        if not os.path.isfile(in_aln):
            raise MissingDataError("SequenceData: loading alignment failed... " + in_aln)
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


class SequenceData:
    """Dense alignment + compressed pattern representation."""

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
        self.ref = None
        self._full_length = int(sequence_length) if sequence_length else None
        self._aln = None
        self.likely_alphabet = None
        if ref is not None:
            raise NotImplementedError("reference-based (sparse) alignments are not supported")
        self.aln = aln

    @property
    def aln(self):
        return self._aln

    @aln.setter
    def aln(self, in_aln):
        self._aln = None
        if in_aln is None:
            return
        raw = _as_sequence_dict(in_aln)
        seqs = list(raw.values())
        self.likely_alphabet = guess_alphabet([seq2array(s) for s in seqs[:50]])
        if self.ambiguous is None:
            self.ambiguous = "N" if self.likely_alphabet == "nuc" else "X"
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
        self.compressed_to_full_sequence_map = dict(enumerate(np.split(order, bounds)))
        self.compressed_codes = comp  # (N, P) uint8, rows in sequence_names order
        chars = comp.view("S1").astype("U1")
        self.compressed_alignment = {k: chars[i] for i, k in enumerate(names)}
        self._compressed_length = comp.shape[1]

    def compressed_to_full_sequence(self, sequence, include_additional_constant_sites=False, as_string=False):
        L = self.full_length if include_additional_constant_sites else self.full_length - self.additional_constant_sites
        tmp = np.asarray(sequence)[self.full_to_compressed_sequence_map[:L]]
        return "".join(tmp.astype("U")) if as_string else tmp

    def differences(self, seq1, seq2, seq1_compressed=True, seq2_compressed=True, mask=None):
        if seq1_compressed:
            seq1 = self.compressed_to_full_sequence(seq1)
        if seq2_compressed:
            seq2 = self.compressed_to_full_sequence(seq2)
        sel = seq1 != seq2
        if mask is not None:
            sel &= mask > 0
        return [(seq1[p], int(p), seq2[p]) for p in np.where(sel)[0]]
