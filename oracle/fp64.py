"""FP64 numpy restatement of TreeTime's pruning path (TEST INFRASTRUCTURE ONLY).

This module is the checker for the HIP engine in ``treetime_amd``.  It is never
imported by the product.  It restates, in plain FP64 numpy with the reference's
per-node loop structure, the following reference code (TreeTime 0.12.1, paths
relative to the reference checkout):

==========================  =================================================
oracle function             reference
==========================  =================================================
``profile_map``             treetime/seq_utils.py:19-122 (alphabets, maps)
``normalize_profile``       treetime/seq_utils.py:279-307
``prof2seq``                treetime/seq_utils.py:232-276 (argmax branch)
``seq2prof``                treetime/seq_utils.py:207-229
``fill_overhangs``          treetime/seq_utils.py:196-202
``avg_transition``          treetime/gtr.py:7-11
``GTRModel``                treetime/gtr.py:235-283 (assign_rates),
                            612-629 (_eig_single_site), 1027-1067 (expQt),
                            965-1025 (propagate/evolve), 922-963
                            (prob_t_profiles), 717-749 (prob_t_compressed),
                            631-712 (state_pair), 816-920 (optimal_t_compressed)
``SiteGTRModel``            treetime/gtr_site_specific.py:43-116 (assign_rates),
                            312-329 (_eig), 331-371 (expQt and its
                            interpolator), 376-437 (propagate/evolve),
                            491-495 (average_rate)
``compress_alignment``      treetime/sequence_data.py:325-464
``compress_sparse_alignment`` treetime/vcf_utils.py:657-749,
                            treetime/sequence_data.py:362-413 (sparse)
``OracleTreeAnc.marginal``  treetime/treeanc.py:762-932
``OracleTreeAnc.joint``     treetime/treeanc.py:934-1080 (root_rng: the
                            root draw of 1022-1037, seq_utils.py:257-269)
``ancestral_likelihood``    treetime/treeanc.py:718-750
``OracleTreeAnc.optimize_tree`` treetime/treeanc.py:1384-1473 (joint mode),
                            1176-1270 (optimize_joint_lengths), 1475-1495
                            (prune_short_branches)
``branch_t``                treetime/treeanc.py:752-760, config.py:7
==========================  =================================================

Pinning (tests/test_oracle_*.py):
* expQt: scipy.linalg.expm, column sums, detailed balance, P(0)=I,
  Chapman-Kolmogorov, JC69 closed form, v.v_inv = I (reference test
  test/test_treetime.py:83), Q column sums (test_treetime.py:60).
* GTR text fixture test/test_sequence_evolution_model.txt values
  (test_treetime.py:53-75).
* marginal logL: brute-force enumeration on tiny trees; sum over all n^3
  tip patterns of exp(logL) == 1 (test_treetime.py:139-155); pulley
  principle (re-rooting invariance).
* marginal profiles: exact posteriors by enumeration on <=5-tip trees.
* joint: brute-force max over internal states (test_treetime.py:241-307);
  ancestral_likelihood == sequence_LH (test_treetime.py:160-238).
* node attributes by enumeration on a 5-tip tree: marginal_outgroup_LH,
  marginal_log_Lx + marginal_subtree_LH_prefactor, subtree likelihood
  (prefactor x marginal_subtree_LH); joint_Lx / joint_Cx of every node; the
  root draw of joint(root_rng=...) with fixed draws, and the argmax states
  and sequence_LH given the drawn root.
* compression: hand-built columns hitting every rule of
  sequence_data.py:387-402.
* sparse compression: hand-built reference + differences hitting every
  class of vcf_utils.py:694-718; on random inputs, every position's column
  equals the dense compression's column of the expanded alignment and the
  pattern multiplicities agree (dense compression pinned above), apart from
  the positions whose constant state is the ambiguous character, which the
  sparse path leaves out of every pattern.
* prob_t_profiles: one-hot profiles reduce to prob_t_compressed.
* optimize_tree: at its end the joint states are a fixed point (one more
  reconstruction changes nothing), every branch sits at the optimum of its
  end states (checked against a dense grid), no prunable branch is left,
  each pruned node carried its parent's sequence, the tips are untouched
  and the joint log-likelihood never decreased across the iterations.
* optimal_t_compressed: JC69 closed form t = -3/4 ln(1 - 4/3 p).
* SiteGTRModel: exact P(t) of every site = scipy.linalg.expm of the site's
  rate matrix; the interpolated P(t) equals the exact one on grid points and
  the mean of the neighbours half-way; a model whose sites share pi and mu
  gives the standard model's likelihoods; sum over all n^3 tip patterns of
  exp(logL) == 1 for every site's model; joint = brute-force max.  The grid
  points themselves (gtr_site_specific.py:336-343) are restated, not pinned.
"""
import numpy as np

# treetime/config.py:3-7, 34-36
BIG_NUMBER = 1e10
TINY_NUMBER = 1e-12
SUPERTINY_NUMBER = 1e-24
MIN_BRANCH_LENGTH = 1e-3
MAX_BRANCH_LENGTH = 4.0

# ---------------------------------------------------------------------------
# alphabets and profile maps (treetime/seq_utils.py:19-122)
# ---------------------------------------------------------------------------
_NUC = "ACGT"
_AA20 = "ACDEFGHIKLMNPQRSTVWY"
ALPHABETS = {
    "nuc": np.array(list(_NUC + "-")),
    "nuc_nogap": np.array(list(_NUC)),
    "aa": np.array(list(_AA20 + "*-")),
    "aa_nogap": np.array(list(_AA20)),
}
# IUPAC nucleotide ambiguity sets
_IUPAC = {"R": "AG", "Y": "CT", "S": "CG", "W": "AT", "K": "GT", "M": "AC",
          "D": "AGT", "H": "ACT", "B": "CGT", "V": "ACG"}


def profile_map(alphabet):
    """Character -> profile vector, as the reference's tables define them."""
    alpha = list(ALPHABETS[alphabet])
    n = len(alpha)
    m = {}
    for i, c in enumerate(alpha):
        v = np.zeros(n)
        v[i] = 1.0
        m[c] = v
    if alphabet.startswith("nuc"):
        for c in "NX":
            m[c] = np.ones(n)
        for c, states in _IUPAC.items():
            v = np.zeros(n)
            for s in states:
                v[alpha.index(s)] = 1.0
            m[c] = v
        if alphabet == "nuc_nogap":
            m["-"] = np.ones(n)
    else:
        m["X"] = np.ones(n)
        for c, states in {"B": "DN", "Z": "EQ"}.items():
            v = np.zeros(n)
            for s in states:
                v[alpha.index(s)] = 1.0
            m[c] = v
    return m


def ambiguous_char(alphabet):
    """treetime/gtr.py:88-97: 'N' if an all-ones 'N' exists, else the first all-ones char."""
    pm = profile_map(alphabet)
    n = len(ALPHABETS[alphabet])
    amb = [c for c, v in pm.items() if v.sum() == n]
    if not amb:
        return None
    return "N" if "N" in amb else amb[0]


def seq2prof(seq, pmap):
    """treetime/seq_utils.py:207-229 (unknown characters map to all-ones, as
    extend_profile at seq_utils.py:126-136 arranges)."""
    n = len(next(iter(pmap.values())))
    ones = np.ones(n)
    return np.array([pmap.get(c, ones) for c in seq], dtype=float)


def fill_overhangs(seq, ambiguous):
    """treetime/seq_utils.py:196-202: leading/trailing '-' become the ambiguous char."""
    seq = np.array(seq, copy=True)
    nongap = np.where(seq != "-")[0]
    if len(nongap):
        seq[: nongap[0]] = ambiguous
        seq[nongap[-1] + 1:] = ambiguous
    else:
        seq[:] = ambiguous
    return seq


def normalize_profile(in_profile, log=False, return_offset=True):
    """treetime/seq_utils.py:279-307."""
    if log:
        pref = in_profile.max(axis=1)
        prof = np.exp(in_profile - pref[:, None])
    else:
        pref = 0.0
        prof = in_profile
    norm = prof.sum(axis=1)
    out = prof * (1.0 / norm)[:, None]
    return out, ((np.log(norm) + pref) if return_offset else None)


def prof2seq(profile, alphabet_chars, normalize=True):
    """treetime/seq_utils.py:232-276, argmax branch (first maximum wins)."""
    tmp = normalize_profile(profile, return_offset=False)[0] if normalize else profile
    idx = tmp.argmax(axis=1)
    return alphabet_chars[idx], tmp[np.arange(tmp.shape[0]), idx], idx


# ---------------------------------------------------------------------------
# GTR model (treetime/gtr.py)
# ---------------------------------------------------------------------------
def avg_transition(W, pi, gap_index=None):
    """treetime/gtr.py:7-11."""
    r = float(pi @ W @ pi)
    if gap_index is None:
        return r
    return (r - np.sum(pi * W[:, gap_index]) * pi[gap_index]) / (1 - pi[gap_index])


def eig_single_site(W, p):
    """treetime/gtr.py:612-629: eigensystem of the symmetrised rate matrix;
    returns (eigenvalues, v, v_inv) with P(t) = v diag(exp(mu t lam)) v_inv."""
    sp = np.sqrt(p)
    sym = W * np.outer(sp, sp)
    np.fill_diagonal(sym, -np.sum(W * p, axis=1))
    lam, U = np.linalg.eigh(sym)
    tv = U.T * sp                       # rows: eigenvectors scaled by sqrt(pi)
    norm1 = np.abs(tv).sum(axis=1)
    return lam, tv.T / norm1, (U * norm1).T / sp


class GTRModel:
    """Reversible substitution model: rates, eigensystem and P(t).

    Mirrors GTR.assign_rates (gtr.py:235-283): pi is normalised, the diagonal
    of W is zeroed, W is divided by the average rate and mu multiplied by it.
    The eigensystem follows _eig_single_site (gtr.py:612-629).
    """

    def __init__(self, alphabet="nuc_nogap", mu=1.0, pi=None, W=None):
        self.alphabet_name = alphabet
        self.alphabet = ALPHABETS[alphabet]
        self.profile_map = profile_map(alphabet)
        self.ambiguous = ambiguous_char(alphabet)
        chars = list(self.alphabet)
        self.gap_index = chars.index("-") if "-" in chars else None
        self.n_states = len(chars)
        self.assign_rates(mu, pi, W)

    def assign_rates(self, mu=1.0, pi=None, W=None):
        n = self.n_states
        pi_arr = np.array(pi, dtype=float) if (pi is not None and len(pi) == n) else np.ones(n)
        self.Pi = pi_arr / pi_arr.sum()
        if W is None or np.shape(W) != (n, n):
            Wm = np.ones((n, n))
            np.fill_diagonal(Wm, 0.0)
        else:
            Wm = np.array(W, dtype=float)
        np.fill_diagonal(Wm, 0.0)
        avg = avg_transition(Wm, self.Pi, gap_index=self.gap_index)
        self.W = Wm / avg
        self.mu = mu * avg
        self._eig()

    def _eig(self):
        self.eigenvals, self.v, self.v_inv = eig_single_site(self.W, self.Pi)

    @property
    def Q(self):
        """treetime/gtr.py:131-143."""
        q = (self.W * self.Pi).T
        np.fill_diagonal(q, 0.0)
        np.fill_diagonal(q, -q.sum(axis=0))
        return q

    def _exp_lt(self, t):
        """treetime/gtr.py:1027-1049."""
        lv = self.mu * t * self.eigenvals
        if np.any(lv > 10):
            raise ValueError("exp(Q*t): positive eigenvalue or branch length too large")
        return np.exp(lv)

    def expQt(self, t):
        """treetime/gtr.py:1051-1067; P[i, j] = Pr(child=i | parent=j)."""
        return np.maximum(0.0, (self.v * self._exp_lt(t)) @ self.v_inv)

    def propagate_profile(self, profile, t, return_log=False):
        """treetime/gtr.py:965-995: message to the parent."""
        res = profile @ self.expQt(t)
        return np.log(res) if return_log else res

    def evolve(self, profile, t, return_log=False):
        """treetime/gtr.py:997-1025: child distribution given the parent profile."""
        res = profile @ self.expQt(t).T
        return np.log(res) if return_log else res

    def prob_t_profiles(self, profile_pair, multiplicity, t, return_log=False, ignore_gaps=True):
        """treetime/gtr.py:922-963 (non site-specific branch)."""
        if t < 0:
            logP = -BIG_NUMBER
        else:
            Qt = self.expQt(t)
            pp, pc = profile_pair
            res = np.sum((pc @ Qt) * pp, axis=1)
            if ignore_gaps and self.gap_index is not None:
                ng = (1 - pp[:, self.gap_index]) * (1 - pc[:, self.gap_index])
                logP = np.sum(multiplicity * np.log(res + SUPERTINY_NUMBER) * ng)
            else:
                logP = np.sum(multiplicity * np.log(res + SUPERTINY_NUMBER))
        return logP if return_log else np.exp(logP)

    def prob_t_compressed(self, seq_pair, multiplicity, t, return_log=False):
        """treetime/gtr.py:717-749; seq_pair rows are (parent_idx, child_idx)."""
        if t < 0:
            logP = -BIG_NUMBER
        else:
            lq = np.log(np.maximum(self.expQt(t), SUPERTINY_NUMBER))
            logP = np.sum(lq[seq_pair[:, 1], seq_pair[:, 0]] * multiplicity)
        return logP if return_log else np.exp(logP)

    def state_pair(self, seq_p, seq_ch, pattern_multiplicity=None, ignore_gaps=False):
        """treetime/gtr.py:631-712: multiplicity-weighted (parent, child) state
        pair counts, truncated to int as the reference does at gtr.py:709-712."""
        if pattern_multiplicity is None:
            pattern_multiplicity = np.ones(len(seq_p))
        idx = {c: i for i, c in enumerate(self.alphabet)}
        n = self.n_states
        counts = np.zeros((n, n))
        for a, b, w in zip(seq_p, seq_ch, pattern_multiplicity):
            if n >= 10:
                # large alphabets: characters outside the alphabet count as state 1
                # (gtr.py:685-691 initialises the index array with ones)
                i, j = idx.get(a, 1), idx.get(b, 1)
            elif a in idx and b in idx:
                i, j = idx[a], idx[b]
            else:
                continue
            if ignore_gaps and self.gap_index is not None and (i == self.gap_index or j == self.gap_index):
                continue
            counts[i, j] += w
        pairs, mult = [], []
        for i in range(n):
            for j in range(n):
                if counts[i, j]:
                    pairs.append((i, j))
                    mult.append(counts[i, j])
        return np.array(pairs, dtype=int).reshape(-1, 2), np.array(mult, dtype=int)

    def optimal_t_compressed(self, seq_pair, multiplicity, profiles=False, tol=1e-10):
        """treetime/gtr.py:816-920: Brent in s = sqrt(t)."""
        from scipy.optimize import minimize_scalar

        def neg(s):
            if profiles:
                return -self.prob_t_profiles(seq_pair, multiplicity, s**2, return_log=True) + np.exp(s**4 / 10000)
            return -self.prob_t_compressed(seq_pair, multiplicity, s**2, return_log=True)

        if profiles:
            hd = 1 - np.sum(multiplicity * np.sum(seq_pair[0] * seq_pair[1], axis=1)) / np.sum(multiplicity)
        else:
            hd = np.sum(multiplicity[seq_pair[:, 1] != seq_pair[:, 0]]) / np.sum(multiplicity)
        opt = minimize_scalar(neg, bracket=[-np.sqrt(MAX_BRANCH_LENGTH), np.sqrt(hd), np.sqrt(MAX_BRANCH_LENGTH)],
                              tol=tol, method="brent")
        new_len = opt["x"] ** 2
        if "success" in opt and opt["success"] is not True and opt["success"] != True:  # noqa: E712
            new_len = hd
        return new_len


def gtr_infer(alphabet, nij, Ti, root_state, fixed_pi=None, pc=1.0, gap_limit=0.01):
    """treetime/gtr.py:491-599 (GTR.infer): fixed point of
    n_ij + pc = mu pi_i W_ij (T_j + pc + root_state); returns a GTRModel."""
    nij = np.array(nij, dtype=float)
    Ti = np.asarray(Ti, dtype=float)
    root_state = np.asarray(root_state, dtype=float)
    probe = GTRModel(alphabet)
    gap = probe.gap_index
    pcm = pc * np.ones_like(nij)
    np.fill_diagonal(pcm, 0.0)
    np.fill_diagonal(nij, 0.0)
    pi = np.ones_like(Ti) if fixed_pi is None else np.array(fixed_pi, dtype=float)
    pi /= pi.sum()
    old = np.zeros_like(Ti)
    W = np.ones_like(nij)
    mu = (nij.sum() + pc) / (Ti.sum() + pc)
    it = 0
    while np.linalg.norm(old - pi) > 1e-5 and it < 40:
        it += 1
        old = pi.copy()
        W = (nij + nij.T + 2 * pcm) / mu / (np.outer(pi, Ti) + np.outer(Ti, pi) + TINY_NUMBER + 2 * pcm)
        np.fill_diagonal(W, 0)
        W = W / avg_transition(W, pi, gap_index=gap)
        if fixed_pi is None:
            pi = (np.sum(nij + pcm, axis=1) + root_state) / (
                TINY_NUMBER + mu * W.dot(Ti) + root_state.sum() + np.sum(pcm, axis=1))
            pi /= pi.sum()
            mu = (nij.sum() + pc) / (np.sum(pi * W.dot(Ti)) + pc)
        else:
            mu = (nij.sum() + pc) / (np.sum(pi * W.dot(pi)) * Ti.sum() + pc)
    if gap is not None:  # set unconditionally (gtr.py:585-596); only the warning is conditional
        pi[gap] = gap_limit
        pi /= pi.sum()
    return GTRModel(alphabet, mu=mu, pi=pi, W=W)


def site_gtr_infer(alphabet, sub_ija, T_ia, root_state, pc=1.0, gap_limit=0.01, Nit=30, dp=1e-5):
    """treetime/gtr_site_specific.py:208-310 (GTR_site_specific.infer):
    one W, per-site pi and mu; returns a SiteGTRModel (approximate=True)."""
    q = len(ALPHABETS[alphabet])
    L = sub_ija.shape[-1]
    n_ija = np.array(sub_ija, dtype=float)
    for i in range(q):
        n_ija[i, i, :] = 0.0
    n_ij = n_ija.sum(axis=-1)
    m_ia = n_ija.sum(axis=1) + root_state + pc
    n_a = n_ija.sum(axis=(0, 1)) + pc
    Lam = root_state.sum(axis=0) + q * pc
    p_old, p = np.zeros((q, L)), np.ones((q, L)) / q
    mu = np.ones(L)
    W = np.ones((q, q)) - np.eye(q)
    it = 0
    while np.sqrt(((p_old - p) ** 2).sum()) > dp and it < Nit:
        it += 1
        p_old = p.copy()
        S = (mu[None, None, :] * p[:, None, :] * T_ia[None, :, :]).sum(axis=-1)
        W = (n_ij + n_ij.T + pc) / (S + S.T + pc)
        avg_pi = p.mean(axis=-1)
        r = avg_pi @ W @ avg_pi
        W = W / r
        mu = mu * r
        p = m_ia / (mu * (W @ T_ia) + Lam)
        p = p / p.sum(axis=0)
        mu = n_a / (pc + (p * (W @ T_ia)).sum(axis=0))
    gap = list(ALPHABETS[alphabet]).index("-") if "-" in ALPHABETS[alphabet] else None
    if gap is not None:
        p[gap, :] = gap_limit          # every site, as the reference does
        p = p / p.sum(axis=0)
    return SiteGTRModel(alphabet, mu=mu, pi=p, W=W)


def gtr_from_text(text):
    """Parse the reference's GTR text format (treetime/gtr.py:174-233)."""
    lines = text.splitlines()
    mu, alphabet, pi, W = None, [], [], None
    i = 0
    while i < len(lines):
        s = lines[i].strip()
        if s.startswith("Substitution rate (mu):"):
            mu = float(s.split(":")[1])
        elif s.startswith("Equilibrium frequencies (pi_i):"):
            i += 1
            while i < len(lines) and lines[i].strip():
                a, b = lines[i].split(":")
                alphabet.append(a.strip())
                pi.append(float(b))
                i += 1
        elif s.startswith("Symmetrized rates from j->i (W_ij):"):
            i += 2
            n = len(pi)
            W = np.ones((n, n))
            r = 0
            while i < len(lines) and lines[i].strip():
                vals = lines[i].split()
                W[r, :] = [float(x) for x in vals[1:n + 1]]
                r += 1
                i += 1
        i += 1
    name = {5: "nuc", 4: "nuc_nogap", 22: "aa", 20: "aa_nogap"}[len(alphabet)]
    assert list(ALPHABETS[name]) == alphabet
    return GTRModel(name, mu=mu, pi=np.array(pi), W=W)


class SiteGTRModel:
    """Site-specific GTR (treetime/gtr_site_specific.py): one W for all
    sites, frequencies Pi (n, L) and rates mu (L,) per site.

    assign_rates (43-116): Pi columns normalised, W symmetrised with a zero
    diagonal and divided by the average of pi_a^T W pi_a over sites, mu
    multiplied by it.  expQt(t) has shape (n, n, L), [i, j, a] =
    Pr(child = i | parent = j) at site a; with approximate=True and
    t * rate_scale < 10 it is interpolated linearly between the grid points
    around t (331-371, scipy interp1d 'linear' rules), else exact.
    """

    def __init__(self, alphabet="nuc_nogap", mu=1.0, pi=None, W=None, approximate=True):
        self.alphabet_name = alphabet
        self.alphabet = ALPHABETS[alphabet]
        self.profile_map = profile_map(alphabet)
        self.ambiguous = ambiguous_char(alphabet)
        self.n_states = n = len(self.alphabet)
        chars = list(self.alphabet)
        self.gap_index = chars.index("-") if "-" in chars else None
        self.approximate = approximate
        pi = np.array(pi, dtype=float)
        L = pi.shape[1]
        self.L = L
        self.Pi = pi / pi.sum(axis=0)
        Wm = 0.5 * (np.array(W, dtype=float) + np.array(W, dtype=float).T)
        np.fill_diagonal(Wm, 0.0)
        avg = np.einsum("ia,ij,ja", self.Pi, Wm, self.Pi) / L
        self.W = Wm / avg
        self.mu = (mu * np.ones(L) if np.isscalar(mu) else np.array(mu, dtype=float)) * avg
        self.lam = np.zeros((L, n))
        self.V = np.zeros((L, n, n))
        self.Vi = np.zeros((L, n, n))
        for a in range(L):
            self.lam[a], self.V[a], self.Vi[a] = eig_single_site(self.W, self.Pi[:, a])
        self.rate_scale = self.average_rate().mean()
        self.t_grid = (1.0 / self.rate_scale) * np.concatenate((
            np.linspace(0, 0.1, 11)[:-1], np.linspace(0.1, 1, 21)[:-1],
            np.linspace(1, 5, 21)[:-1], np.linspace(5, 10, 11)))
        self._grid_P = np.stack([self._expQt(t) for t in self.t_grid], axis=0)

    def average_rate(self):
        """gtr_site_specific.py:491-495."""
        return self.mu * np.einsum("ia,ij,ja->a", self.Pi, self.W, self.Pi)

    def _expQt(self, t):
        """gtr_site_specific.py:350-365, per site V diag(exp(mu t lam)) V^-1."""
        e = np.exp(t * self.mu[:, None] * self.lam)                    # (L, n)
        P = np.einsum("aik,ak,akj->aij", self.V, e, self.Vi)
        return np.moveaxis(P, 0, -1)

    def expQt(self, t):
        if self.approximate and t * self.rate_scale < 10:
            g = self.t_grid
            hi = int(np.clip(np.searchsorted(g, t), 1, len(g) - 1))
            lo = hi - 1
            slope = (self._grid_P[hi] - self._grid_P[lo]) / (g[hi] - g[lo])
            return slope * (t - g[lo]) + self._grid_P[lo]
        return self._expQt(t)

    def propagate_profile(self, profile, t, return_log=False):
        """gtr_site_specific.py:376-406."""
        res = np.einsum("ai,ija->aj", profile, self.expQt(t))
        return np.log(np.maximum(TINY_NUMBER, res)) if return_log else np.maximum(0, res)

    def evolve(self, profile, t, return_log=False):
        """gtr_site_specific.py:408-437."""
        res = np.einsum("ai,jia->aj", profile, self.expQt(t))
        return np.log(res) if return_log else res

    def prob_t_profiles(self, profile_pair, multiplicity, t, return_log=False, ignore_gaps=True):
        """treetime/gtr.py:922-963, site-specific branch (Qt of shape (n, n, L))."""
        if t < 0:
            logP = -BIG_NUMBER
        else:
            pp, pc = profile_pair
            res = np.einsum("ai,ija,aj->a", pc, self.expQt(t), pp)
            if ignore_gaps and self.gap_index is not None:
                ng = (1 - pp[:, self.gap_index]) * (1 - pc[:, self.gap_index])
                logP = np.sum(multiplicity * np.log(res + SUPERTINY_NUMBER) * ng)
            else:
                logP = np.sum(multiplicity * np.log(res + SUPERTINY_NUMBER))
        return logP if return_log else np.exp(logP)

    optimal_t_compressed = GTRModel.optimal_t_compressed


# ---------------------------------------------------------------------------
# alignment compression (treetime/sequence_data.py:325-464)
# ---------------------------------------------------------------------------
def compress_alignment(names, seqs, ambiguous, additional_constant_sites=0, nogap_alphabet=None):
    """Column-by-column restatement of make_compressed_alignment.

    Returns (compressed: dict name -> (P,) char array, multiplicity (P,) f64,
    full_to_compressed (L_full,) int, compressed_to_full: dict p -> positions).
    """
    aln_t = np.array([seqs[k] for k in names]).T
    patterns = {}
    columns = []
    for pos in range(aln_t.shape[0]):
        col = np.copy(aln_t[pos])
        uniq = list(np.unique(col))
        if len(uniq) == 2 and ambiguous in uniq:
            other = [c for c in uniq if c != ambiguous][0]
            col[col == ambiguous] = other
            uniq = [other]
        key = "".join(col.astype("U"))
        if len(uniq) > 1:
            key += "_%d" % pos
        if key not in patterns:
            patterns[key] = (len(columns), [pos])
            columns.append(col)
        else:
            patterns[key][1].append(pos)
    full_len = aln_t.shape[0] + additional_constant_sites
    if additional_constant_sites:
        chars = [c for c in nogap_alphabet if c not in (ambiguous, "-")]
        counts = {c: np.sum(aln_t == c) for c in chars}
        total = np.sum(list(counts.values()))
        per_char = [(c, int(np.round(additional_constant_sites * m / total))) for c, m in counts.items()]
        left = additional_constant_sites
        pos = aln_t.shape[0]
        for c, m in per_char:
            if c == per_char[-1][0]:
                m = left
            key = c * len(names)
            if m:
                plist = list(range(pos, pos + m))
                if key in patterns:
                    patterns[key][1].extend(plist)
                else:
                    patterns[key] = (len(columns), plist)
                    columns.append(np.array(list(key)))
                pos += m
                left -= m
    mult = np.zeros(len(patterns))
    f2c = np.zeros(full_len, dtype=int)
    c2f = {}
    for p, plist in patterns.values():
        mult[p] = len(plist)
        f2c[np.array(plist)] = p
        c2f[p] = np.array(plist, dtype=int)
    comp = np.array(columns).T
    return {k: comp[i] for i, k in enumerate(names)}, mult, f2c, c2f


def compress_sparse_alignment(names, diffs, ref, ambiguous):
    """Restatement of the sparse (reference + differences) compression:
    process_alignment_dictionary (treetime/vcf_utils.py:657-749) followed by
    the is_sparse branch of make_compressed_alignment
    (treetime/sequence_data.py:362-413, 444-464).

    diffs: dict name -> {position: character}; ref: (L,) char array.
    Returns (compressed dict name -> (P,) char array, multiplicity (P,),
    full_to_compressed (L,), compressed_to_full dict p -> positions,
    nonref_positions (sorted), inferred constant sites (list)).
    """
    nseq = len(names)
    inv_map = {}
    for k in names:
        for pos, b in diffs[k].items():
            inv_map.setdefault(pos, []).append(b)
    nonref_positions = np.sort(list(inv_map.keys()))
    const_up_to_amb, nonref_const, nonref_alleles, variable = [], [], [], []
    for pos, bs in inv_map.items():
        bases = list(np.unique(bs))
        if len(bs) == nseq:
            if (len(bases) <= 2 and ambiguous in bases) or len(bases) == 1:
                nonref_const.append(pos)
                nonref_alleles.append(bases[0] if len(bases) == 1 else [x for x in bases if x != ambiguous][0])
                if ambiguous in bases:
                    const_up_to_amb.append(pos)
            else:
                variable.append(pos)
        elif len(bases) == 1 and bases[0] == ambiguous:
            const_up_to_amb.append(pos)
        else:
            variable.append(pos)
    ref_mod = np.copy(ref)
    ref_mod[nonref_const] = nonref_alleles
    states = np.unique(ref_mod)          # taken before the variable sites are masked
    ref_mod[variable] = "."
    columns, patterns = [], {}
    for base in states:
        if base == ambiguous:
            continue
        where = list(np.where(ref_mod == base)[0])
        if where:
            patterns[base * nseq] = [len(columns), where]
            columns.append(np.repeat(base, nseq))
    for pos in variable:
        col = np.array([diffs[k][pos] if pos in diffs[k] else ref[pos] for k in names])
        uniq = list(np.unique(col))
        if len(uniq) == 2 and ambiguous in uniq:
            other = [c for c in uniq if c != ambiguous][0]
            col[col == ambiguous] = other
            uniq = [other]
        key = "".join(col.astype("U"))
        if len(uniq) > 1:
            key += "_%d" % pos
        if key not in patterns:
            patterns[key] = [len(columns), [pos]]
            columns.append(col)
        else:
            patterns[key][1].append(pos)
    mult = np.zeros(len(patterns))
    f2c = np.zeros(len(ref), dtype=int)
    c2f = {}
    for p, plist in patterns.values():
        mult[p] = len(plist)
        f2c[np.array(plist)] = p
        c2f[p] = np.array(plist, dtype=int)
    comp = np.array(columns).T
    return ({k: comp[i] for i, k in enumerate(names)}, mult, f2c, c2f, nonref_positions, const_up_to_amb)


# ---------------------------------------------------------------------------
# tree + passes (treetime/treeanc.py)
# ---------------------------------------------------------------------------
class ONode:
    """Minimal clade: name, branch_length, ordered children, parent link."""

    def __init__(self, name=None, branch_length=0.0):
        self.name = name
        self.branch_length = branch_length
        self.clades = []
        self.up = None

    def is_terminal(self):
        return not self.clades


def build_tree(parent, branch_length, names):
    """Nodes from a parent array (root has parent -1); children keep index order."""
    nodes = [ONode(names[i], float(branch_length[i])) for i in range(len(parent))]
    root = None
    for i, p in enumerate(parent):
        if p < 0:
            root = nodes[i]
        else:
            nodes[p].clades.append(nodes[i])
            nodes[i].up = nodes[p]
    return root, nodes


def postorder(root):
    out, stack = [], [(root, False)]
    while stack:
        n, done = stack.pop()
        if done or n.is_terminal():
            out.append(n)
        else:
            stack.append((n, True))
            for c in reversed(n.clades):
                stack.append((c, False))
    return out


def preorder(root):
    out, stack = [], [root]
    while stack:
        n = stack.pop()
        out.append(n)
        stack.extend(reversed(n.clades))
    return out


class OracleTreeAnc:
    """Marginal and joint ML reconstruction over a compressed alignment."""

    def __init__(self, root, gtr, compressed, multiplicity, full_length):
        self.root = root
        self.gtr = gtr
        self.aln = compressed
        self.mult = np.asarray(multiplicity, dtype=float)
        self.L = len(self.mult)
        self.one_mutation = 1.0 / full_length

    def branch_t(self, node):
        """treetime/treeanc.py:752-760."""
        return max(MIN_BRANCH_LENGTH * self.one_mutation, node.branch_length)

    def _root_prior(self):
        """treetime/treeanc.py:817-820: (L, n) root prior."""
        pi = self.gtr.Pi
        return np.repeat([pi], self.L, axis=0) if pi.ndim == 1 else np.copy(pi.T)

    def _tip_profile(self, leaf):
        if leaf.name in self.aln:
            return seq2prof(self.aln[leaf.name], self.gtr.profile_map)
        return np.ones((self.L, self.gtr.n_states))

    # -- marginal ------------------------------------------------------------
    def postorder_marginal(self):
        """treetime/treeanc.py:840-878."""
        for leaf in postorder(self.root):
            if leaf.is_terminal():
                leaf.marginal_subtree_LH = self._tip_profile(leaf)
                leaf.marginal_subtree_LH_prefactor = np.zeros(self.L)
        for node in postorder(self.root):
            if node.is_terminal():
                continue
            acc = np.zeros((self.L, self.gtr.n_states))
            node.marginal_subtree_LH_prefactor = np.zeros(self.L)
            for ch in node.clades:
                ch.marginal_log_Lx = self.gtr.propagate_profile(ch.marginal_subtree_LH, self.branch_t(ch), return_log=True)
                acc += ch.marginal_log_Lx
                node.marginal_subtree_LH_prefactor += ch.marginal_subtree_LH_prefactor
            node.marginal_subtree_LH, off = normalize_profile(acc, log=True)
            node.marginal_subtree_LH_prefactor += off

    def total_LH(self):
        """treetime/treeanc.py:814-838."""
        r = self.root
        r.marginal_outgroup_LH = self._root_prior()
        r.marginal_profile, pre = normalize_profile(r.marginal_outgroup_LH * r.marginal_subtree_LH)
        self.sequence_LH = r.marginal_subtree_LH_prefactor + pre
        self.total_sequence_LH = float((self.sequence_LH * self.mult).sum())
        _, _, r.seq_idx = prof2seq(r.marginal_profile, self.gtr.alphabet, normalize=False)

    def preorder_marginal(self, reconstruct_tip_states=False):
        """treetime/treeanc.py:880-932 (division trick for the outgroup)."""
        for node in preorder(self.root):
            if node.up is None:
                continue
            node.marginal_outgroup_LH, _ = normalize_profile(
                np.log(np.maximum(TINY_NUMBER, node.up.marginal_profile)) - node.marginal_log_Lx,
                log=True, return_offset=False)
            if node.is_terminal() and not reconstruct_tip_states:
                continue
            msg = self.gtr.evolve(node.marginal_outgroup_LH, self.branch_t(node))
            node.marginal_profile, _ = normalize_profile(node.marginal_subtree_LH * msg, return_offset=False)
            _, _, node.seq_idx = prof2seq(node.marginal_profile, self.gtr.alphabet, normalize=False)

    def marginal(self, reconstruct_tip_states=False):
        """treetime/treeanc.py:762-812."""
        self.postorder_marginal()
        self.total_LH()
        self.preorder_marginal(reconstruct_tip_states)
        return self.total_sequence_LH

    # -- joint ---------------------------------------------------------------
    def joint(self, reconstruct_tip_states=False, root_rng=None):
        """treetime/treeanc.py:934-1080.  Also records, per node, the decision
        margin (best minus second-best log value) behind each chosen state.
        root_rng: draw the root states from the root's joint profile with
        root_rng.random(L) (sample_from_profile='root' or True,
        treeanc.py:1022-1037, prof2seq sampling at seq_utils.py:257-269) and
        record, per pattern, the distance of the draw from the nearest
        cumulative boundary (root.sample_margin)."""
        n = self.gtr.n_states
        for node in postorder(self.root):
            if node.up is None:
                continue
            logP = np.log(np.maximum(TINY_NUMBER, self.gtr.expQt(self.branch_t(node))))
            if node.is_terminal():
                with np.errstate(divide="ignore"):
                    msg = np.log(np.maximum(self._tip_profile(node), TINY_NUMBER))
                msg[np.isnan(msg) | np.isinf(msg)] = -BIG_NUMBER
            else:
                msg = np.sum(np.stack([c.joint_Lx for c in node.clades], axis=0), axis=0)
            node.joint_Lx = np.zeros((self.L, n))
            node.joint_Cx = np.zeros((self.L, n), dtype=np.uint16)
            node.joint_margin = np.zeros((self.L, n))
            for j in range(n):
                col = logP[:, j]
                cand = (col.T if col.ndim == 2 else col[None, :]) + msg
                node.joint_Cx[:, j] = cand.argmax(axis=1)
                node.joint_Lx[:, j] = cand.max(axis=1)
                srt = np.sort(cand, axis=1)
                node.joint_margin[:, j] = srt[:, -1] - srt[:, -2]
        r = self.root
        msg = np.sum(np.stack([c.joint_Lx for c in r.clades], axis=0), axis=0)
        r.joint_Lx = msg + np.log(self._root_prior())
        normed = r.joint_Lx - r.joint_Lx.max(axis=1)[:, None]
        if root_rng is None:
            _, _, idx = prof2seq(np.exp(normed), self.gtr.alphabet)
            srt = np.sort(r.joint_Lx, axis=1)
            r.path_margin = srt[:, -1] - srt[:, -2]
        else:
            prof, _ = normalize_profile(np.exp(normed), return_offset=False)
            cum = prof.cumsum(axis=1).T
            u = root_rng.random(size=cum.shape[1])
            idx = np.argmax(cum >= u, axis=0)
            r.sample_margin = np.min(np.abs(cum - u[None, :]), axis=0)
            r.path_margin = np.full(self.L, np.inf)  # the root state is drawn, not decided
        self.joint_sequence_LH = r.joint_Lx[np.arange(self.L), idx]
        self.total_joint_LH = float((self.joint_sequence_LH * self.mult).sum())
        r.seq_idx = idx
        for node in preorder(self.root):
            if node.up is None or (node.is_terminal() and not reconstruct_tip_states):
                continue
            pidx = node.up.seq_idx
            node.seq_idx = node.joint_Cx[np.arange(self.L), pidx].astype(int)
            node.path_margin = np.minimum(node.up.path_margin, node.joint_margin[np.arange(self.L), pidx])
        return self.total_joint_LH

    # -- iterative joint reconstruction + branch lengths ------------------------
    def _seq_chars(self, node):
        if node.is_terminal():
            return np.asarray(self.aln[node.name])
        return self.gtr.alphabet[node.seq_idx]

    def optimize_joint_lengths(self):
        """treetime/treeanc.py:1176-1270: every branch set to the optimum of
        the state pairs at its ends (gaps ignored), never below zero."""
        for node in postorder(self.root):
            if node.up is None:
                continue
            pairs, m = self.gtr.state_pair(self._seq_chars(node.up), self._seq_chars(node),
                                           pattern_multiplicity=self.mult, ignore_gaps=True)
            node.branch_length = max(0.0, self.gtr.optimal_t_compressed(pairs, m))

    def prune_short_branches(self):
        """treetime/treeanc.py:1475-1495: an internal branch shorter than a
        tenth of one mutation with probability above 0.1 at length zero
        (identical ends, gaps ignored) is removed, its children moved up."""
        pruned = []
        for node in preorder(self.root):
            if node.up is None or node.is_terminal():
                continue
            if node.branch_length < 0.1 * self.one_mutation:
                pairs, m = self.gtr.state_pair(self._seq_chars(node.up), self._seq_chars(node),
                                               pattern_multiplicity=self.mult, ignore_gaps=True)
                if self.gtr.prob_t_compressed(pairs, m, 0.0) > 0.1:
                    node.up.clades = [k for k in node.up.clades if k is not node] + node.clades
                    for c in node.clades:
                        c.up = node.up
                    pruned.append(node.name)
        return pruned

    def optimize_tree(self, max_iter=5, prune_short=True):
        """treetime/treeanc.py:1384-1473, joint branch-length mode with
        joint ML sequences: reconstruct, optimise the lengths, then repeat
        (prune, reconstruct, stop when no internal character changed,
        optimise) at most max_iter times.  Records the smallest joint
        decision margin over all reconstructions (min_path_margin), the
        number of iterations and the pruned nodes; returns the final joint
        log-likelihood."""
        def internal():
            return [nd for nd in preorder(self.root) if nd.up is not None and not nd.is_terminal()]

        def margin():
            return min(float(np.min(nd.path_margin)) for nd in preorder(self.root) if not nd.is_terminal())

        self.joint()
        self.min_path_margin = margin()
        self.optimize_joint_lengths()
        self.pruned = []
        n = 0
        while n < max_iter:
            n += 1
            if prune_short:
                self.pruned += self.prune_short_branches()
            prev = {id(nd): nd.seq_idx.copy() for nd in internal()}
            self.joint()
            self.min_path_margin = min(self.min_path_margin, margin())
            n_diff = sum(int((nd.seq_idx != prev[id(nd)]).sum()) for nd in internal())
            if n_diff < 1:
                break
            self.optimize_joint_lengths()
        self.n_iter = n
        return self.total_joint_LH

    # -- likelihood of one realisation ----------------------------------------
    def ancestral_likelihood(self, seq_idx):
        """treetime/treeanc.py:718-750 with sequences given as state indices
        (dict node-name -> (P,) int)."""
        out = np.zeros(self.L)
        for node in postorder(self.root):
            if node.up is None:
                out += np.log(self.gtr.Pi[seq_idx[node.name]])
                continue
            lq = np.log(self.gtr.expQt(node.branch_length))
            out += lq[seq_idx[node.name], seq_idx[node.up.name]]
        return out

    # -- GTR inference statistics ------------------------------------------------
    def branch_mutation_matrix(self, node):
        """treetime/treeanc.py:1085-1120 (get_branch_mutation_matrix, compressed):
        (P, n, n) joint distribution of (child state i, parent state j) per
        pattern, after a marginal pass."""
        pp, pc = node.marginal_outgroup_LH, node.marginal_subtree_LH
        E = self.gtr.expQt(self.branch_t(node)) + SUPERTINY_NUMBER
        M = np.einsum("ai,aj,ija->aij" if E.ndim == 3 else "ai,aj,ij->aij", pc, pp, E)
        return M / M.sum(axis=2).sum(axis=1)[:, None, None]

    def mutation_counts_marginal(self):
        """treetime/treeanc.py:1556-1572 summed over patterns: n_ij (i derived,
        j ancestral) and the time spent in each state, T_i."""
        n = self.gtr.n_states
        nij, Ti = np.zeros((n, n)), np.zeros(n)
        for node in preorder(self.root):
            if node.up is None:
                continue
            M = self.branch_mutation_matrix(node)
            t = self.branch_t(node)
            w = self.mult[:, None]
            Ti += 0.5 * t * (M.sum(axis=1) * w).sum(axis=0)
            Ti += 0.5 * t * (M.sum(axis=2) * w).sum(axis=0)
            nij += (M * self.mult[:, None, None]).sum(axis=0)
        return nij, Ti

    def site_counts_marginal(self):
        """treetime/treeanc.py:1556-1572 per pattern (site_specific=True):
        n_ija (n, n, P) and T_ia (n, P) summed over branches."""
        n = self.gtr.n_states
        n_ija, T_ia = np.zeros((n, n, self.L)), np.zeros((n, self.L))
        for node in preorder(self.root):
            if node.up is None:
                continue
            st = np.transpose(self.branch_mutation_matrix(node), (1, 2, 0))   # (i, j, a)
            t = self.branch_t(node)
            T_ia += 0.5 * t * st.sum(axis=0) * self.mult
            T_ia += 0.5 * t * st.sum(axis=1) * self.mult
            n_ija += st * self.mult
        return n_ija, T_ia

    def site_counts_joint(self):
        """treetime/treeanc.py:1573-1589 per pattern after a reconstruction
        (node.seq_idx; tips without one use their characters, characters
        outside the alphabet are skipped)."""
        n = self.gtr.n_states
        index = {c: i for i, c in enumerate(self.gtr.alphabet)}
        n_ija, T_ia = np.zeros((n, n, self.L)), np.zeros((n, self.L))
        cols = np.arange(self.L)
        for node in preorder(self.root):
            if node.up is None:
                continue
            if hasattr(node, "seq_idx"):
                cs = np.asarray(node.seq_idx, dtype=int)
            else:
                chars = self.aln.get(node.name, np.array([""] * self.L))
                cs = np.array([index.get(c, -1) for c in chars])
            ps = np.asarray(node.up.seq_idx, dtype=int)
            t = self.branch_t(node)
            mut = (cs >= 0) & (cs != ps)
            np.add.at(n_ija, (cs[mut], ps[mut], cols[mut]), self.mult[mut])
            np.add.at(T_ia, (ps[mut], cols[mut]), 0.5 * t * self.mult[mut])
            np.add.at(T_ia, (cs[mut], cols[mut]), -0.5 * t * self.mult[mut])
            ok = cs >= 0
            np.add.at(T_ia, (cs[ok], cols[ok]), t * self.mult[ok])
        return n_ija, T_ia
