"""Site-specific GTR model: one exchangeability matrix W, per-site
equilibrium frequencies Pi (n, L) and rates mu (L,).

Same surface as treetime/gtr_site_specific.py: assign_rates (43-116), the
per-site eigensystems (_eig, 312-329), expQt with the linear interpolator
over a fixed grid of branch lengths (331-371), propagate_profile / evolve
(376-437), prob_t (439-489) and average_rate (491-495).  The model set-up is
FP64 numpy on the host; propagate_profile and evolve run as a HIP kernel,
and a TreeAnc with this model runs its passes with per-pattern matrices on
the device (tt_tree_set_site_model).
"""
import ctypes as C

import numpy as np

from . import _native
from . import config as ttconf
from .gtr import GTR, _device
from .seq_utils import alphabets


class GTR_site_specific(GTR):
    """GTR model with frequencies and rates that differ between sites."""

    def __init__(self, seq_len=1, approximate=True, **kwargs):
        self.seq_len = seq_len
        self.approximate = approximate
        super().__init__(**kwargs)
        self.is_site_specific = True

    @property
    def Q(self):
        """Rate matrices (n, n, L): Q[i, j, a] = Pi[i, a] W[i, j], diagonal
        minus the column sums."""
        tmp = np.einsum("ia,ij->ija", self.Pi, self.W)
        diag_vals = np.sum(tmp, axis=0)
        for x in range(tmp.shape[-1]):
            np.fill_diagonal(tmp[:, :, x], -diag_vals[:, x])
        return tmp

    def assign_rates(self, mu=1.0, pi=None, W=None):
        """gtr_site_specific.py:43-116."""
        if not np.isscalar(mu) and pi is not None and len(np.shape(pi)) == 2:
            if np.shape(mu)[0] != np.shape(pi)[1]:
                raise ValueError("GTR_site_specific: length of rate vector (got {}) and equilibrium frequency vector "
                                 "(got {}) must match!".format(np.shape(mu)[0], np.shape(pi)[1]))
        n = len(self.alphabet)
        if np.isscalar(mu):
            self._mu = mu * np.ones(self.seq_len)
        else:
            self._mu = np.array(mu, dtype=float)
            self.seq_len = self._mu.shape[0]
        if pi is not None and np.shape(pi)[0] == n and len(np.shape(pi)) == 2:
            self.seq_len = np.shape(pi)[1]
            Pi = np.array(pi, dtype=float)
        elif pi is not None:
            if len(pi) == n:
                Pi = np.repeat([np.asarray(pi, dtype=float)], self.seq_len, axis=0).T
            else:
                raise ValueError("GTR_site_specific: length of equilibrium frequency vector (got {}) does not match "
                                 "alphabet length {}".format(len(pi), n))
        else:
            Pi = np.ones(shape=(n, self.seq_len))
        self._Pi = Pi / np.sum(Pi, axis=0)
        if W is None or np.shape(W) != (n, n):
            if W is not None and np.shape(W) != (n, n):
                raise ValueError("GTR_site_specific: Size of substitution matrix (got {}) does not match alphabet "
                                 "length {}".format(np.shape(W), n))
            W = np.ones((n, n))
        else:
            W = 0.5 * (np.array(W, dtype=float) + np.array(W, dtype=float).T)
        np.fill_diagonal(W, 0)
        average_rate = np.einsum("ia,ij,ja", self._Pi, W, self._Pi) / self.seq_len
        self._W = W / average_rate
        self._mu = self._mu * average_rate
        self.is_site_specific = True
        self._eig()
        self._make_expQt_interpolator()

    def _eig(self):
        """Per-site eigensystems in the reference's layout: eigenvals (n, L),
        v and v_inv (n, n, L) with v[:, :, a] = (site a's eigenvectors).T."""
        ev, vec, vec_inv = [], [], []
        for a in range(self.seq_len):
            e, v, vi = self._eig_single_site(self._W, self._Pi[:, a])
            ev.append(e)
            vec.append(v)
            vec_inv.append(vi)
        self.eigenvals = np.array(ev).T
        self.v = np.swapaxes(vec, 0, -1)
        self.v_inv = np.swapaxes(vec_inv, 0, -1)

    def _make_expQt_interpolator(self):
        """Grid of branch lengths and P(t) on it (gtr_site_specific.py:331-348)."""
        self.rate_scale = self.average_rate().mean()
        self.t_grid = (1.0 / self.rate_scale) * np.concatenate((
            np.linspace(0, 0.1, 11)[:-1], np.linspace(0.1, 1, 21)[:-1],
            np.linspace(1, 5, 21)[:-1], np.linspace(5, 10, 11)))
        self._grid_expQt = np.stack([self._expQt(t) for t in self.t_grid], axis=0)

    def _expQt(self, t):
        """Exact per-site P(t), shape (n, n, L)."""
        eLambdaT = np.exp(t * self.mu * self.eigenvals)
        return np.einsum("jia,ja,kja->ika", self.v, eLambdaT, self.v_inv)

    def expQt(self, t):
        """Linear interpolation between the grid points around t (scipy
        interp1d 'linear': searchsorted, clipped to [1, G-1]) when
        approximate and t * rate_scale < 10, else the exact P(t)."""
        if t * self.rate_scale < 10 and self.approximate:
            g = self.t_grid
            hi = int(np.clip(np.searchsorted(g, t), 1, len(g) - 1))
            lo = hi - 1
            slope = (self._grid_expQt[hi] - self._grid_expQt[lo]) / (g[hi] - g[lo])
            return slope * (t - g[lo]) + self._grid_expQt[lo]
        return self._expQt(t)

    def average_rate(self):
        if self.Pi.shape[1] > 1:
            return np.einsum("a,ia,ij,ja->a", self.mu, self.Pi, self.W, self.Pi)
        return self.mu * np.einsum("ia,ij,ja->a", self.Pi, self.W, self.Pi)

    # --- constructors -------------------------------------------------------
    @classmethod
    def random(cls, L=1, avg_mu=1.0, alphabet="nuc", pi_dirichlet_alpha=1, W_dirichlet_alpha=3.0,
               mu_gamma_alpha=3.0, rng=None):
        """gtr_site_specific.py:117-172."""
        if rng is None:
            rng = np.random.default_rng()
        gtr = cls(alphabet=alphabet, seq_len=L)
        n = gtr.alphabet.shape[0]
        pi = 1.0 * rng.gamma(pi_dirichlet_alpha, size=(n, L)) if pi_dirichlet_alpha else np.ones((n, L))
        pi /= pi.sum(axis=0)
        tmp = 1.0 * rng.gamma(W_dirichlet_alpha, size=(n, n)) if W_dirichlet_alpha else np.ones((n, n))
        tmp = np.tril(tmp, k=-1)
        W = tmp + tmp.T
        mu = rng.gamma(mu_gamma_alpha, size=(L,)) if mu_gamma_alpha else np.ones(L)
        gtr.assign_rates(mu=mu, pi=pi, W=W)
        gtr.mu *= avg_mu / np.mean(gtr.average_rate())
        return gtr

    @classmethod
    def custom(cls, mu=1.0, pi=None, W=None, **kwargs):
        gtr = cls(**kwargs)
        gtr.assign_rates(mu=mu, pi=pi, W=W)
        return gtr

    @classmethod
    def infer(cls, sub_ija, T_ia, root_state, pc=1.0, gap_limit=0.01, Nit=30, dp=1e-5, **kwargs):
        """Fixed point of n_ija + pc = pi_ia W_ij (T_ja + pc + root_state)
        with one W and per-site pi and mu (gtr_site_specific.py:208-310):
        pseudocounts off the diagonal, W symmetrised from the summed counts,
        pi_ia and mu_a updated per site, stop when |p - p_old| <= dp or after
        Nit rounds; with a gap state its frequency is set to gap_limit at
        every site.  Host FP64 (n x n x L arrays)."""
        gtr = cls(**kwargs)
        gtr.logger("GTR: model inference ", 1)
        q = len(gtr.alphabet)
        L = sub_ija.shape[-1]
        n_iter = 0
        n_ija = np.copy(sub_ija)
        n_ija[range(q), range(q), :] = 0
        n_ij = n_ija.sum(axis=-1)
        m_ia = np.sum(n_ija, axis=1) + root_state + pc
        n_a = n_ija.sum(axis=1).sum(axis=0) + pc
        Lambda = np.sum(root_state, axis=0) + q * pc
        p_ia_old = np.zeros((q, L))
        p_ia = np.ones((q, L)) / q
        mu_a = np.ones(L)
        W_ij = np.ones((q, q)) - np.eye(q)
        while np.linalg.norm(p_ia_old - p_ia) > dp and n_iter < Nit:
            n_iter += 1
            p_ia_old = np.copy(p_ia)
            S_ij = np.einsum("a,ia,ja", mu_a, p_ia, T_ia)
            W_ij = (n_ij + n_ij.T + pc) / (S_ij + S_ij.T + pc)
            avg_pi = p_ia.mean(axis=-1)
            average_rate = W_ij.dot(avg_pi).dot(avg_pi)
            W_ij = W_ij / average_rate
            mu_a *= average_rate
            p_ia = m_ia / (mu_a * np.dot(W_ij, T_ia) + Lambda)
            p_ia = p_ia / p_ia.sum(axis=0)
            mu_a = n_a / (pc + np.einsum("ia,ij,ja->a", p_ia, W_ij, T_ia))
        if n_iter >= Nit:
            gtr.logger("WARNING: maximum number of iterations has been reached in GTR inference", 3, warn=True)
        if gtr.gap_index is not None:
            for p in range(p_ia.shape[-1]):
                if p_ia[gtr.gap_index, p] < gap_limit:
                    gtr.logger("The model allows for gaps which are estimated to occur at a low fraction of %1.3e"
                               % p_ia[gtr.gap_index, p], 4, warn=True)
                p_ia[gtr.gap_index, p] = gap_limit
                p_ia[:, p] /= p_ia[:, p].sum()
        gtr.assign_rates(mu=mu_a, W=W_ij, pi=p_ia)
        return gtr

    # --- device --------------------------------------------------------------
    def device_model(self):
        """Arguments of tt_tree_set_site_model: per-site eigenvalues (L, n),
        eigenvectors and inverse (L, n, n) in row-major per-site layout, mu
        (L,), pi (L, n), and the interpolation grid (empty when not
        approximate) with the rate scale."""
        eig = np.ascontiguousarray(self.eigenvals.T, dtype=np.float64)
        V = np.ascontiguousarray(np.transpose(self.v, (2, 1, 0)), dtype=np.float64)
        Vi = np.ascontiguousarray(np.transpose(self.v_inv, (2, 1, 0)), dtype=np.float64)
        mu = np.ascontiguousarray(self.mu * np.ones(self.seq_len), dtype=np.float64)
        pi = np.ascontiguousarray(self.Pi.T, dtype=np.float64)
        grid = np.ascontiguousarray(self.t_grid if self.approximate else np.zeros(0), dtype=np.float64)
        return eig, V, Vi, mu, pi, grid, float(self.rate_scale)

    def _apply_site(self, profile, t, transpose, return_log):
        torch = _device()
        lib = _native.load()
        prof = np.ascontiguousarray(profile, dtype=np.float64)
        P, n = prof.shape
        Q = np.ascontiguousarray(np.moveaxis(self.expQt(t), -1, 0), dtype=np.float64)   # (L, n, n)
        if Q.shape[0] != P:
            raise ValueError("profile length {} does not match the model's {} sites".format(P, Q.shape[0]))
        din = torch.from_numpy(prof).to("cuda")
        dq = torch.from_numpy(Q).to("cuda")
        dout = torch.empty_like(din)
        stream = torch.cuda.current_stream().cuda_stream
        _native.check(lib.tt_propagate_profile_site(din.data_ptr(), P, n, dq.data_ptr(), int(transpose),
                                                    int(bool(return_log)), dout.data_ptr(), stream))
        return dout.cpu().numpy()

    def propagate_profile(self, profile, t, return_log=False):
        """res[a, j] = sum_i profile[a, i] P[i, j, a]; log(max(TINY, res)) or
        max(0, res) (gtr_site_specific.py:376-406)."""
        return self._apply_site(profile, t, 0, return_log)

    def evolve(self, profile, t, return_log=False):
        """res[a, j] = sum_i profile[a, i] P[j, i, a]; log(res) or res
        (gtr_site_specific.py:408-437)."""
        return self._apply_site(profile, t, 1, return_log)

    def prob_t(self, seq_p, seq_ch, t, pattern_multiplicity=None, return_log=False, ignore_gaps=True):
        """gtr_site_specific.py:439-489 (host)."""
        if t < 0:
            logP = -ttconf.BIG_NUMBER
        else:
            tmp_eQT = self.expQt(t)
            bad = tmp_eQT == 0
            with np.errstate(divide="ignore", invalid="ignore"):
                logQt = np.log(tmp_eQT + ttconf.TINY_NUMBER * bad)
            logQt[np.isnan(logQt) | np.isinf(logQt) | bad] = -ttconf.BIG_NUMBER
            ip = np.zeros(len(seq_p), dtype=int)
            ic = np.zeros(len(seq_ch), dtype=int)
            for ai, a in enumerate(self.alphabet):
                ip[np.asarray(seq_p) == a] = ai
                ic[np.asarray(seq_ch) == a] = ai
            mult = np.ones(len(seq_ch)) if pattern_multiplicity is None else pattern_multiplicity
            logP = np.sum(logQt[ip, ic, np.arange(len(seq_ch))] * mult)
        return logP if return_log else np.exp(logP)


__all__ = ["GTR_site_specific", "alphabets"]
