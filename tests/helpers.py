"""Shared helpers for the parity tests: the same network built for the engine
and for the CPU oracle, and small statistics utilities."""
import numpy as np
from scipy import stats

import tl_oracle as O
from tau_leaping_amd import models


def pair(net):
    """(engine Model, OracleModel) for one network."""
    import tau_leaping_amd as T

    csc = net.to_csc()
    return T.Model.from_csc(net.n_species, *csc), O.OracleModel(net.n_species, net.n_channels, *csc)


def oracle_of(net):
    return O.OracleModel(net.n_species, net.n_channels, *net.to_csc())


def random_states(N, R, seed, lo=0, hi=5000, small_frac=0.25):
    """[N][R] populations with a share of tiny counts so criticality and the
    order-3 GetG branch are exercised."""
    rng = np.random.default_rng(seed)
    x = rng.integers(lo, hi + 1, size=(N, R))
    small = rng.random((N, R)) < small_frac
    x[small] = rng.integers(0, 8, size=int(small.sum()))
    return np.ascontiguousarray(x.astype(np.int32))


def mean_ci_ok(a, b, alpha=0.01):
    """Welch two-sample t-test at level alpha (means consistent)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    if a.std() == 0 and b.std() == 0:
        return a.mean() == b.mean()
    return stats.ttest_ind(a, b, equal_var=False).pvalue > alpha


def ks_ok(a, b, alpha=0.01):
    return stats.ks_2samp(np.asarray(a, float), np.asarray(b, float)).pvalue > alpha


SMALL_NETS = {
    "dimer": lambda: models.dimer_decay(100000),
    "toy": lambda: models.toy_network(32, 64, seed=7),
    "cyclic": lambda: models.cyclic_chain(64, seed=3),
    "colloid": lambda: models.colloidal(20, seed=5),
}


def cyclic_generator(c):
    """Generator of one molecule's walk on the ring S_i -> S_(i+1) (rates c)."""
    N = len(c)
    L = np.zeros((N, N))
    for i in range(N):
        L[i, i] -= c[i]
        L[i, (i + 1) % N] += c[i]
    return L


def cyclic_net(N, c, x0):
    rx = [models.Reaction({i: 1}, {(i + 1) % N: 1}, float(c[i])) for i in range(N)]
    return models.Network(N, rx, np.asarray(x0, np.int32))


def cyclic_mean_jacobian(c, x0, T):
    """d E[X_k(T)] / d c_j = x0 . d exp(T L) / d c_j (Frechet derivative), [k][j]."""
    from scipy import linalg

    N = len(c)
    Lg = cyclic_generator(c)
    J = np.zeros((N, N))
    for j in range(N):
        E = np.zeros((N, N))
        E[j, j] = -1.0
        E[j, (j + 1) % N] = 1.0
        _, dP = linalg.expm_frechet(T * Lg, T * E)
        J[:, j] = np.asarray(x0, float) @ dP
    return J
