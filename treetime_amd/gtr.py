"""GTR substitution model.

Same surface as treetime/gtr.py: the model set-up (assign_rates 235-283,
_eig_single_site 612-629, expQt 1027-1067) stays FP64 on the host; the
profile operations propagate_profile (965-995), evolve (997-1025) and
prob_t_profiles (922-963) run as HIP kernels through the C ABI.
"""
import ctypes as C

import numpy as np

from . import _native
from . import config as ttconf
from .errors import MissingDataError
from .seq_utils import alphabet_synonyms, alphabets, profile_maps


def avg_transition(W, pi, gap_index=None):
    """pi^T W pi, optionally excluding the gap state (gtr.py:7-11)."""
    r = float(pi @ W @ pi)
    if gap_index is None:
        return r
    return (r - np.sum(pi * W[:, gap_index]) * pi[gap_index]) / (1 - pi[gap_index])


def _device():
    import torch  # device buffers only

    if not torch.cuda.is_available():
        raise _native.NativeLibraryError("no HIP device visible: the GTR profile kernels need a GPU")
    return torch


class GTR:
    """General time-reversible model of character evolution."""

    def __init__(self, alphabet="nuc", prof_map=None, logger=None):
        self.debug = False
        self.is_site_specific = False
        if isinstance(alphabet, str):
            if alphabet not in alphabet_synonyms:
                raise AttributeError("Unknown alphabet type specified")
            name = alphabet_synonyms[alphabet]
            self.alphabet = alphabets[name]
            self.profile_map = {k: v.copy() for k, v in profile_maps[name].items()}
        else:
            self.alphabet = np.array(alphabet)
            if prof_map is None:
                self.profile_map = {s: x for s, x in zip(self.alphabet, np.eye(len(self.alphabet)))}
            else:
                self.profile_map = dict(prof_map)
        self.state_index = {s: i for i, s in enumerate(self.alphabet)}
        self.logger = logger if logger is not None else (lambda *a, **k: None)
        self.n_states = len(self.alphabet)
        self.ambiguous = None
        self.gap_index = None
        self.assign_gap_and_ambiguous()
        self.v = self.v_inv = self.eigenvals = None
        self.assign_rates()

    def assign_gap_and_ambiguous(self):
        """gtr.py:88-103: the ambiguous character has an all-ones profile
        ('N' preferred); the gap state is '-' if the alphabet has one."""
        n = len(self.alphabet)
        amb = [c for c, x in self.profile_map.items() if x.sum() == n]
        self.ambiguous = ("N" if "N" in amb else amb[0]) if amb else None
        self.gap_index = self.state_index.get("-")

    # --- parameters ---------------------------------------------------------
    @property
    def mu(self):
        return self._mu

    @mu.setter
    def mu(self, value):
        self.assign_rates(mu=value, pi=self.Pi, W=self.W)

    @property
    def Pi(self):
        return self._Pi

    @Pi.setter
    def Pi(self, value):
        self.assign_rates(mu=self.mu, pi=value, W=self.W)

    @property
    def W(self):
        return self._W

    @W.setter
    def W(self, value):
        self.assign_rates(mu=self.mu, pi=self.Pi, W=value)

    @property
    def Q(self):
        """Rate matrix Q[i, j] = rate j -> i, columns sum to 0 (gtr.py:131-143)."""
        q = (self.W * self.Pi).T
        np.fill_diagonal(q, 0.0)
        np.fill_diagonal(q, -q.sum(axis=0))
        return q

    def assign_rates(self, mu=1.0, pi=None, W=None):
        """gtr.py:235-283.  As in the reference, W's diagonal is zeroed, W is
        divided by the average rate and mu multiplied by it, so the effective
        rate matrix is mu_in * pi_i * W_in[i, j]."""
        n = len(self.alphabet)
        self._mu = mu
        self.is_site_specific = False
        if pi is not None and len(pi) == n:
            p = np.array(pi, dtype=float)
        else:
            if pi is not None:
                self.logger("length of equilibrium frequency vector does not match alphabet length", 4, warn=True)
            p = np.ones(n)
        self._Pi = p / p.sum()
        if W is None or np.shape(W) != (n, n):
            if W is not None:
                self.logger("Substitution matrix size does not match alphabet size", 4, warn=True)
            Wm = np.ones((n, n))
        else:
            Wm = np.array(W, dtype=float)
        np.fill_diagonal(Wm, 0.0)
        avg = avg_transition(Wm, self._Pi, gap_index=self.gap_index)
        self._W = Wm / avg
        self._mu = self._mu * avg
        self._eig()

    def _eig(self):
        self.eigenvals, self.v, self.v_inv = self._eig_single_site(self.W, self.Pi)

    @staticmethod
    def _eig_single_site(W, p):
        """Symmetrised eigendecomposition (gtr.py:612-629):
        S = diag(sqrt pi) W diag(sqrt pi) with the exit rates on the diagonal;
        v = (U sqrt(pi)) / |.|_1 columns, v_inv = its inverse."""
        assert np.abs(np.diag(W).sum()) < 1e-10
        sp = np.sqrt(p)
        S = W * np.outer(sp, sp)
        np.fill_diagonal(S, -np.sum(W * p, axis=1))
        lam, U = np.linalg.eigh(S)
        scaled = U.T * sp
        norm1 = np.sum(np.abs(scaled), axis=1)
        return lam, scaled.T / norm1, (U * norm1).T / sp

    # --- constructors -------------------------------------------------------
    @classmethod
    def custom(cls, mu=1.0, pi=None, W=None, **kwargs):
        gtr = cls(**kwargs)
        gtr.assign_rates(mu=mu, pi=pi, W=W)
        return gtr

    @staticmethod
    def standard(model, **kwargs):
        """Named models (gtr.py:318-461; nuc_models.py)."""
        from .nuc_models import F81, HKY85, JC69, K80, T92, TN93

        m = model.lower()
        if m in ["jc", "jc69", "jukes-cantor", "jukes-cantor69", "jukescantor", "jukescantor69"]:
            g = JC69(**kwargs)
        elif m in ["k80", "kimura80", "kimura1980"]:
            g = K80(**kwargs)
        elif m in ["f81", "felsenstein81", "felsenstein1981"]:
            g = F81(**kwargs)
        elif m in ["hky", "hky85", "hky1985"]:
            g = HKY85(**kwargs)
        elif m in ["t92", "tamura92", "tamura1992"]:
            g = T92(**kwargs)
        elif m in ["tn93", "tamura_nei_93", "tamuranei93"]:
            g = TN93(**kwargs)
        elif m in ["jtt", "jtt92"]:
            raise NotImplementedError("JTT92's empirical matrix is not bundled; build it with GTR.custom")
        else:
            raise KeyError("The GTR model '{}' is not in the list of available models.".format(model))
        g.mu = kwargs.get("mu", 1.0)
        return g

    @classmethod
    def infer(cls, nij, Ti, root_state, fixed_pi=None, pc=1.0, gap_limit=0.01, **kwargs):
        """Model from transition counts and state times (gtr.py:491-599):
        the fixed point of n_ij + pc = mu pi_i W_ij (T_j + pc + root_state),
        W symmetric, iterated until pi moves by less than 1e-5 (at most 40
        rounds).  The n x n statistics come from the device
        (TreeAnc.infer_gtr); the iteration is small and runs on the host."""
        gtr = cls(**kwargs)
        gtr.logger("GTR: model inference ", 1)
        nij = np.array(nij, dtype=float)
        Ti = np.asarray(Ti, dtype=float)
        root_state = np.asarray(root_state, dtype=float)
        pc_mat = pc * np.ones_like(nij)
        np.fill_diagonal(pc_mat, 0.0)
        np.fill_diagonal(nij, 0.0)
        pi = np.ones_like(Ti) if fixed_pi is None else np.array(fixed_pi, dtype=float)
        pi /= pi.sum()
        pi_old = np.zeros_like(Ti)
        W = np.ones_like(nij)
        mu = (nij.sum() + pc) / (Ti.sum() + pc)
        rounds = 0
        while np.linalg.norm(pi_old - pi) > 1e-5 and rounds < 40:
            rounds += 1
            pi_old = pi.copy()
            W = (nij + nij.T + 2 * pc_mat) / mu / (np.outer(pi, Ti) + np.outer(Ti, pi) + ttconf.TINY_NUMBER
                                                     + 2 * pc_mat)
            np.fill_diagonal(W, 0)
            W = W / avg_transition(W, pi, gap_index=gtr.gap_index)
            if fixed_pi is None:
                pi = (np.sum(nij + pc_mat, axis=1) + root_state) / (
                    ttconf.TINY_NUMBER + mu * np.dot(W, Ti) + root_state.sum() + np.sum(pc_mat, axis=1))
                pi /= pi.sum()
                mu = (nij.sum() + pc) / (np.sum(pi * (W.dot(Ti))) + pc)
            else:
                mu = (nij.sum() + pc) / (np.sum(pi * (W.dot(pi))) * Ti.sum() + pc)
        if rounds >= 40:
            gtr.logger("WARNING: maximum number of iterations has been reached in GTR inference", 3, warn=True)
        if gtr.gap_index is not None:
            # the gap frequency is always set to gap_limit (gtr.py:585-596);
            # only the warning depends on the estimate
            if pi[gtr.gap_index] < gap_limit:
                gtr.logger("The model allows for gaps which are estimated to occur at a low fraction of %1.3e; gap "
                           "fraction will be set to %1.4f" % (pi[gtr.gap_index], gap_limit), 2, warn=True)
            pi[gtr.gap_index] = gap_limit
            pi /= pi.sum()
        gtr.assign_rates(mu=mu, W=W, pi=pi)
        return gtr

    @classmethod
    def random(cls, mu=1.0, alphabet="nuc", rng=None):
        rng = rng if rng is not None else np.random.default_rng()
        n = len(alphabets[alphabet_synonyms[alphabet]])
        pi = 1.0 * rng.integers(0, 100, size=n) + 1
        W = 1.0 * rng.integers(0, 100, size=(n, n)) + 1
        W = 0.5 * (W + W.T)
        return cls.custom(mu=mu, alphabet=alphabet, pi=pi, W=W)

    @staticmethod
    def from_file(gtr_fname):
        """Read the text format written by __str__ (gtr.py:174-233)."""
        try:
            with open(gtr_fname) as f:
                lines = f.read().splitlines()
            mu, alpha, pi, W = None, [], [], None
            i = 0
            while i < len(lines):
                s = lines[i].strip()
                if s.startswith("Substitution rate (mu):"):
                    mu = float(s.split(":")[1])
                elif s.startswith("Equilibrium frequencies (pi_i):"):
                    i += 1
                    while i < len(lines) and lines[i].strip():
                        a, b = lines[i].split(":")
                        alpha.append(a.strip())
                        pi.append(float(b))
                        i += 1
                    if not any(len(alpha) == len(a) and np.all(np.array(alpha) == a) for a in alphabets.values()):
                        raise ValueError("alphabet not recognized")
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
                    if r != n:
                        raise ValueError("W rows do not match alphabet length")
                i += 1
            return GTR.custom(mu, pi, W, alphabet=np.array(alpha))
        except Exception as e:
            raise MissingDataError("GTR: was unable to read custom GTR model in " + str(gtr_fname)) from e

    def __str__(self):
        s = "Substitution rate (mu): " + str(np.round(self.mu, 6)) + "\n"
        s += "\nEquilibrium frequencies (pi_i):\n"
        for a, p in zip(self.alphabet, self.Pi):
            s += "  " + a + ": " + str(np.round(p, 4)) + "\n"
        s += "\nSymmetrized rates from j->i (W_ij):\n\t" + "\t".join(self.alphabet) + "\n"
        for a, Wi in zip(self.alphabet, self.W):
            s += "  " + a + "\t" + "\t".join(str(np.round(max(0, p), 4)) for p in Wi) + "\n"
        s += "\nActual rates from j->i (Q_ij):\n\t" + "\t".join(self.alphabet) + "\n"
        for a, Qi in zip(self.alphabet, self.Q):
            s += "  " + a + "\t" + "\t".join(str(np.round(max(0, p), 4)) for p in Qi) + "\n"
        return s

    # --- P(t) ---------------------------------------------------------------
    def _exp_lt(self, t):
        lv = self.mu * t * self.eigenvals
        if np.any(lv > 10):
            raise ValueError(
                "Error in computing exp(Q * t): Q has positive eigenvalues or the branch length t is too large. "
                "This is most likely caused by incorrect input data.")
        return np.exp(lv)

    def expQt(self, t):
        """P(t)[i, j] = Pr(child = i | parent = j), FP64, clipped at 0."""
        return np.maximum(0.0, self.v.dot(np.diag(self._exp_lt(t)).dot(self.v_inv)))

    def expQs(self, s):
        return self.expQt(s**2)

    def expQsds(self, s):
        d = np.diag(2.0 * self._exp_lt(s**2) * self.eigenvals * s)
        return self.v.dot(d.dot(self.v_inv))

    # --- profile operations on the device ------------------------------------
    def _apply(self, profile, Q, transpose, return_log):
        torch = _device()
        lib = _native.load()
        prof = np.ascontiguousarray(profile, dtype=np.float64)
        squeeze = prof.ndim == 1
        if squeeze:
            prof = prof[None, :]
        n = prof.shape[1]
        din = torch.from_numpy(prof).to("cuda")
        dout = torch.empty_like(din)
        Qc = np.ascontiguousarray(Q, dtype=np.float64)
        stream = torch.cuda.current_stream().cuda_stream
        _native.check(lib.tt_propagate_profile(din.data_ptr(), prof.shape[0], n, _native.ptr(Qc, C.c_double),
                                               int(transpose), int(bool(return_log)), dout.data_ptr(), stream))
        out = dout.cpu().numpy()
        return out[0] if squeeze else out

    def propagate_profile(self, profile, t, return_log=False):
        """Message to the parent: profile @ expQt(t) (gtr.py:965-995)."""
        return self._apply(profile, self.expQt(t), 0, return_log)

    def evolve(self, profile, t, return_log=False):
        """Child distribution given the parent: profile @ expQt(t).T (gtr.py:997-1025)."""
        return self._apply(profile, self.expQt(t), 1, return_log)

    def prob_t_profiles(self, profile_pair, multiplicity, t, return_log=False, ignore_gaps=True):
        """Multiplicity-weighted log-probability of a (parent, child) profile
        pair at distance t (gtr.py:922-963).  ``t`` may be a scalar or a
        sequence of candidate lengths evaluated in one launch."""
        ts = np.atleast_1d(np.asarray(t, dtype=float))
        res = np.empty(len(ts))
        ok = ts >= 0
        res[~ok] = -ttconf.BIG_NUMBER
        if ok.any():
            torch = _device()
            lib = _native.load()
            pp = np.ascontiguousarray(profile_pair[0], dtype=np.float64)
            pc = np.ascontiguousarray(profile_pair[1], dtype=np.float64)
            mult = np.ascontiguousarray(multiplicity, dtype=np.float64)
            P, n = pp.shape
            Qs = np.ascontiguousarray(np.stack([self.expQt(x) for x in ts[ok]]), dtype=np.float64)
            dpp = torch.from_numpy(pp).to("cuda")
            dpc = torch.from_numpy(pc).to("cuda")
            dm = torch.from_numpy(mult).to("cuda")
            out = np.empty(int(ok.sum()))
            gap = self.gap_index if (ignore_gaps and self.gap_index is not None) else -1
            stream = torch.cuda.current_stream().cuda_stream
            _native.check(lib.tt_prob_t_profiles(dpp.data_ptr(), dpc.data_ptr(), dm.data_ptr(), P, n,
                                                 _native.ptr(Qs, C.c_double), len(Qs), gap,
                                                 _native.ptr(out, C.c_double), stream))
            res[ok] = out
        if np.ndim(t) == 0:
            res = res[0]
        return res if return_log else np.exp(res)

    # --- fixed-sequence branch statistics (host) -----------------------------
    def _state_indices(self, seq, default):
        """Alphabet index of every character (``default`` for the others)."""
        if seq.dtype.kind == "U" and seq.dtype.itemsize == 4 and all(len(str(a)) == 1 for a in self.alphabet):
            codes = seq.view(np.uint32).reshape(seq.shape)
            if codes.size == 0 or codes.max() < 256:
                table = np.full(256, default, dtype=int)
                for i, a in enumerate(self.alphabet):
                    if ord(str(a)) < 256:
                        table[ord(str(a))] = i
                return table[codes]
        lookup = {c: i for i, c in enumerate(self.alphabet)}
        return np.array([lookup.get(c, default) for c in seq.tolist()], dtype=int)

    def state_pair(self, seq_p, seq_ch, pattern_multiplicity=None, ignore_gaps=False):
        """Multiplicity-weighted (parent, child) state-pair counts, cast to int
        as the reference does (gtr.py:631-712)."""
        seq_p = np.asarray(seq_p)
        seq_ch = np.asarray(seq_ch)
        if seq_ch.shape != seq_p.shape:
            raise ValueError("GTR.state_pair: Sequence lengths do not match!")
        w = np.ones(len(seq_p)) if pattern_multiplicity is None else np.asarray(pattern_multiplicity, dtype=float)
        n = self.n_states
        default = 1 if n >= 10 else -1
        ip, ic = self._state_indices(seq_p, default), self._state_indices(seq_ch, default)
        keep = (ip >= 0) & (ic >= 0)
        if ignore_gaps and self.gap_index is not None:
            keep &= (ip != self.gap_index) & (ic != self.gap_index)
        counts = np.zeros((n, n))
        np.add.at(counts, (ip[keep], ic[keep]), w[keep])
        nz = np.argwhere(counts != 0)
        return nz.astype(int), counts[nz[:, 0], nz[:, 1]].astype(int)

    def prob_t_compressed(self, seq_pair, multiplicity, t, return_log=False):
        """gtr.py:717-749."""
        if t < 0:
            logP = -ttconf.BIG_NUMBER
        else:
            lq = np.log(np.maximum(self.expQt(t), ttconf.SUPERTINY_NUMBER))
            logP = np.sum(lq[seq_pair[:, 1], seq_pair[:, 0]] * multiplicity)
        return logP if return_log else np.exp(logP)

    def prob_t(self, seq_p, seq_ch, t, pattern_multiplicity=None, return_log=False, ignore_gaps=True):
        sp, m = self.state_pair(seq_p, seq_ch, pattern_multiplicity=pattern_multiplicity, ignore_gaps=ignore_gaps)
        return self.prob_t_compressed(sp, m, t, return_log=return_log)

    def optimal_t(self, seq_p, seq_ch, pattern_multiplicity=None, ignore_gaps=False):
        sp, m = self.state_pair(seq_p, seq_ch, pattern_multiplicity=pattern_multiplicity, ignore_gaps=ignore_gaps)
        return self.optimal_t_compressed(sp, m)

    def optimal_t_compressed(self, seq_pair, multiplicity, profiles=False, tol=1e-10):
        """Brent minimisation in s = sqrt(t) with the reference's bracket and
        penalty (gtr.py:816-920)."""
        from scipy.optimize import minimize_scalar

        def neg(s):
            if profiles:
                return -self.prob_t_profiles(seq_pair, multiplicity, s**2, return_log=True) + np.exp(s**4 / 10000)
            return -self.prob_t_compressed(seq_pair, multiplicity, s**2, return_log=True)

        if profiles:
            hd = 1 - np.sum(multiplicity * np.sum(seq_pair[0] * seq_pair[1], axis=1)) / np.sum(multiplicity)
        else:
            hd = np.sum(multiplicity[seq_pair[:, 1] != seq_pair[:, 0]]) / np.sum(multiplicity)
        opt = minimize_scalar(neg, bracket=[-np.sqrt(ttconf.MAX_BRANCH_LENGTH), np.sqrt(hd),
                                            np.sqrt(ttconf.MAX_BRANCH_LENGTH)], tol=tol, method="brent")
        new_len = opt["x"] ** 2
        success = opt["success"] if "success" in opt else True
        if new_len > 0.9 * ttconf.MAX_BRANCH_LENGTH:
            self.logger("WARNING: GTR.optimal_t_compressed -- The branch length seems to be very long!", 4, warn=True)
        if success is not True and success != True:  # noqa: E712
            new_len = hd
        return new_len

    def sequence_logLH(self, seq, pattern_multiplicity=None):
        seq = np.asarray(seq)
        w = np.ones(len(seq)) if pattern_multiplicity is None else pattern_multiplicity
        return float(sum(np.sum((seq == s) * w * np.log(self.Pi[i])) for i, s in enumerate(self.alphabet)))

    def average_rate(self):
        return self.mu * avg_transition(self.W, self.Pi, gap_index=self.gap_index)

    def save_to_npz(self, outfile):
        full = self.mu * np.dot(self.Pi, self.W)
        desc = np.array(["GTR matrix description\n", "Substitution rate: " + str(self.mu)])
        np.savez(outfile, description=desc, full_gtr=full, char_dist=self.Pi, flow_matrix=self.W)
