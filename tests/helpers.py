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
