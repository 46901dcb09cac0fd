"""Standard nucleotide models (treetime/nuc_models.py:18-253)."""
import numpy as np

from .gtr import GTR
from .seq_utils import alphabets


def _alphabet_size(alphabet):
    return len(alphabets[alphabet]) if isinstance(alphabet, str) else len(alphabet)


def _transversion_W(kappa):
    """A,C,G,T rates: transitions A<->G, C<->T at kappa, the rest at 1."""
    W = np.ones((4, 4))
    for i, j in ((0, 2), (1, 3)):
        W[i, j] = W[j, i] = kappa
    return W


def JC69(mu=1.0, alphabet="nuc", **kwargs):
    n = _alphabet_size(alphabet)
    g = GTR(alphabet=alphabet)
    g.assign_rates(mu=mu, pi=np.ones(n), W=np.ones((n, n)))
    return g


def K80(mu=1.0, kappa=0.1, **kwargs):
    g = GTR(alphabet="nuc_nogap")
    g.assign_rates(mu=mu, pi=np.ones(4) / 4, W=_transversion_W(kappa))
    return g


def F81(mu=1.0, pi=None, alphabet="nuc", **kwargs):
    n = _alphabet_size(alphabet)
    pi = 0.25 * np.ones(4) if pi is None else np.array(pi, dtype=float)
    if len(pi) != n:
        pi = np.ones(n)
    g = GTR(alphabet=alphabet)
    g.assign_rates(mu=mu, pi=pi / pi.sum(), W=np.ones((n, n)))
    return g


def HKY85(mu=1.0, pi=None, kappa=0.1, **kwargs):
    pi = 0.25 * np.ones(4) if pi is None else np.array(pi, dtype=float)
    if pi.shape[0] != 4:
        pi = np.ones(4)
    g = GTR(alphabet="nuc_nogap")
    g.assign_rates(mu=mu, pi=pi / pi.sum(), W=_transversion_W(kappa))
    return g


def T92(mu=1.0, pi_GC=0.5, kappa=0.1, **kwargs):
    if pi_GC >= 1.0:
        raise ValueError("The relative GC content specified is larger than 1.0!")
    pi = np.array([(1.0 - pi_GC) * 0.5, pi_GC * 0.5, pi_GC * 0.5, (1 - pi_GC) * 0.5])
    g = GTR(alphabet="nuc_nogap")
    g.assign_rates(mu=mu, pi=pi, W=_transversion_W(kappa))
    return g


def TN93(mu=1.0, kappa1=1.0, kappa2=1.0, pi=None, **kwargs):
    pi = 0.25 * np.ones(4) if pi is None else np.array(pi, dtype=float)
    k1, k2 = kappa1, kappa2
    W = np.array([[1, k1, 1, k1], [k1, 1, k1, k2], [1, k1, 1, k1], [k1, k2, k1, 1]], dtype=float)
    pi = pi / pi.sum()
    # the reference builds this model on the 5-state 'nuc' alphabet, where a
    # 4-state pi and W are replaced by the defaults (nuc_models.py:240-246)
    g = GTR(alphabet="nuc")
    g.assign_rates(mu=mu, pi=pi, W=W)
    return g
