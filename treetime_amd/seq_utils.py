"""Alphabets, profile maps and host-side profile utilities.

Same names, signatures and tables as treetime/seq_utils.py (alphabets 19-24,
profile_maps 28-122, seq2array 152-204, seq2prof 207-229, prof2seq 232-276,
normalize_profile 279-307).  The tables are generated from the IUPAC
definitions instead of being spelled out; their contents equal the
reference's.
"""
import numpy as np

alphabet_synonyms = {
    "nuc": "nuc", "nucleotide": "nuc", "aa": "aa", "aminoacid": "aa",
    "nuc_nogap": "nuc_nogap", "nucleotide_nogap": "nuc_nogap",
    "aa_nogap": "aa_nogap", "aminoacid_nogap": "aa_nogap",
    "DNA": "nuc", "DNA_nogap": "nuc_nogap",
}

_NUC = "ACGT"
_AA = "ACDEFGHIKLMNPQRSTVWY"
alphabets = {
    "nuc": np.array(list(_NUC + "-")),
    "nuc_nogap": np.array(list(_NUC)),
    "aa": np.array(list(_AA + "*-")),
    "aa_nogap": np.array(list(_AA)),
}

_NUC_AMBIG = {"R": "AG", "Y": "CT", "S": "CG", "W": "AT", "K": "GT", "M": "AC",
              "D": "AGT", "H": "ACT", "B": "CGT", "V": "ACG"}
_AA_AMBIG = {"B": "DN", "Z": "EQ"}


def _make_profile_map(name):
    chars = list(alphabets[name])
    n = len(chars)
    eye = np.eye(n)
    pm = {c: eye[i].copy() for i, c in enumerate(chars)}
    nuc = name.startswith("nuc")
    full = ["N", "X"] if nuc else ["X"]
    for c in full:
        pm[c] = np.ones(n)
    for c, members in (_NUC_AMBIG if nuc else _AA_AMBIG).items():
        pm[c] = eye[[chars.index(m) for m in members]].sum(axis=0)
    if name == "nuc_nogap":
        pm["-"] = np.ones(n)  # gaps carry no information without a gap state
    return pm


profile_maps = {k: _make_profile_map(k) for k in alphabets}


def extend_profile(gtr, aln, logger=None):
    """Unknown characters get an all-ones profile (seq_utils.py:126-136)."""
    seen = set()
    for seq in aln:
        seen.update(np.unique(seq).tolist())
    for c in sorted(seen):
        if c not in gtr.profile_map:
            gtr.profile_map[c] = np.ones(gtr.n_states)
            if logger:
                logger("WARNING: character %s is unknown. Treating it as missing information" % c, 1, warn=True)


def guess_alphabet(aln):
    total = 0
    nuc = 0
    for seq in aln:
        seq = np.asarray(seq)
        total += len(seq)
        nuc += int(np.isin(seq, list("acgtACGT-N")).sum())
    return "nuc" if nuc > 0.9 * total else "aa"


def seq2array(seq, word_length=1, convert_upper=False, fill_overhangs=False, ambiguous="N"):
    """Sequence -> numpy char array; with fill_overhangs leading and trailing
    gaps become the ambiguous character (seq_utils.py:152-204)."""
    if isinstance(seq, str):
        s = seq
    elif isinstance(seq, np.ndarray):
        s = "".join(seq.astype("U"))
    elif hasattr(seq, "seq"):
        s = str(seq.seq)
    else:
        s = str(seq)
    if convert_upper:
        s = s.upper()
    if word_length == 1:
        arr = np.array(list(s))
    else:
        if len(s) % word_length:
            raise ValueError("sequence length has to be multiple of word length")
        arr = np.array([s[i * word_length:(i + 1) * word_length] for i in range(len(s) // word_length)])
    if fill_overhangs:
        nongap = np.where(arr != "-")[0]
        if len(nongap):
            arr[: nongap[0]] = ambiguous
            arr[nongap[-1] + 1:] = ambiguous
        else:
            arr[:] = ambiguous
    return arr


def seq2prof(seq, profile_map):
    """Character array -> (L, n) profile (seq_utils.py:207-229)."""
    return np.array([profile_map[k] for k in seq])


def normalize_profile(in_profile, log=False, return_offset=True):
    """Row-normalise a profile; with log=True the input is log-probabilities
    and the offset is log(row sum) + row max (seq_utils.py:279-307)."""
    if log:
        pref = in_profile.max(axis=1)
        prof = np.exp(in_profile - pref[:, None])
    else:
        pref = 0.0
        prof = in_profile
    norm = prof.sum(axis=1)
    out = prof / norm[:, None]
    return out, ((np.log(norm) + pref) if return_offset else None)


def prof2seq(profile, gtr, sample_from_prof=False, normalize=True, rng=None):
    """Profile -> (sequence, chosen values, indices); argmax takes the first
    maximum, sampling draws from the row's cumulative distribution
    (seq_utils.py:232-276)."""
    if rng is None:
        rng = np.random.default_rng()
    tmp = normalize_profile(profile, return_offset=False)[0] if normalize else profile
    if sample_from_prof:
        cum = tmp.cumsum(axis=1)
        u = rng.random(size=cum.shape[0])
        idx = np.argmax(cum >= u[:, None], axis=1)
    else:
        idx = tmp.argmax(axis=1)
    seq = gtr.alphabet[idx]
    return seq, tmp[np.arange(tmp.shape[0]), idx], idx
