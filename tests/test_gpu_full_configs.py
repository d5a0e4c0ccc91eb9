"""The benchmark configurations at their full sizes on the GPU.

Realization r's random streams depend only on (seed, model id, r), so a
full-size GPU run and an oracle run over the first realizations of the same
configuration simulate the same trajectories: those realizations must agree
(identical final state for nearly all of them, the rest statistically
consistent), and size-independent invariants (conservation laws) must hold in
every realization of the full run.

  * dimerisation-decay, S1 = 100,000, t_final 10, 10,000 realizations, FP64
    (all 10,000 against the oracle);
  * toy network 128 x 256, t_final 0.1, 100,000 realizations, FP64 and FP32
    (the first 1,000 against the oracle);
  * cyclic chain 1024, t_final 1, 16,384 realizations (the first 4,096; the
    population total is conserved in all 16,384);
  * colloid 250 (31,250 channels), t_final 0.05, 4,096 realizations (the first
    16, whose oracle run takes seconds per realization; the mass sum_i i x_i is
    conserved in all 4,096).
"""
import numpy as np
import pytest

import tau_leaping_amd as T
from helpers import ks_ok, mean_ci_ok, oracle_of
from tau_leaping_amd import models

pytestmark = pytest.mark.gpu


def _agree(gpu_x, ora_x, min_same):
    same = np.all(gpu_x == ora_x, axis=1)
    assert same.mean() >= min_same, f"identical realizations {same.mean():.4f}"
    if gpu_x.shape[0] >= 200:
        for q in range(gpu_x.shape[1]):
            assert mean_ci_ok(gpu_x[:, q], ora_x[:, q])
            assert ks_ok(gpu_x[:, q], ora_x[:, q])
    return same


def test_dimer_decay_full(cuda):
    net = models.dimer_decay(100000)
    m = T.Model.from_network(net)
    res = m.simulate(10.0, 10000, seed=1234, track=[0, 2])
    ora = oracle_of(net).simulate(10000, 10.0, seed=1234)
    assert res.n_failed == 0 and (ora["status"] == 0).all()
    same = _agree(res.final_x, ora["x_final"][:, [0, 2]], 0.97)
    assert np.array_equal(res.steps[same], ora["steps"][same])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_toy_network_full(cuda, prec):
    net = models.toy_network(128, 256, seed=42)
    m = T.Model.from_network(net)
    res = m.simulate(0.1, 100000, seed=1234, precision=prec)
    assert res.n_failed == 0
    ora = oracle_of(net).simulate(1000, 0.1, seed=1234, precision=prec)
    assert (ora["status"] == 0).all()
    same = _agree(res.final_x[:1000], ora["x_final"], 0.97)
    assert np.array_equal(res.steps[:1000][same], ora["steps"][same])


def test_cyclic_chain_full(cuda):
    net = models.cyclic_chain(1024, seed=43)
    m = T.Model.from_network(net)
    res = m.simulate(1.0, 16384, seed=1234)
    assert res.n_failed == 0
    assert np.all(res.final_x.astype(np.int64).sum(axis=1) == int(net.x0.astype(np.int64).sum()))
    ora = oracle_of(net).simulate(4096, 1.0, seed=1234)
    _agree(res.final_x[:4096][:, :8], ora["x_final"][:, :8], 0.97)
    assert np.mean(np.all(res.final_x[:4096] == ora["x_final"], axis=1)) >= 0.97


def test_colloid_full(cuda):
    net = models.colloidal(250, seed=44)
    m = T.Model.from_network(net)
    res = m.simulate(0.05, 4096, seed=1234)
    assert res.n_failed == 0
    w = np.arange(1, 251, dtype=np.int64)
    mass = res.final_x.astype(np.int64) @ w
    assert np.all(mass == int(net.x0.astype(np.int64) @ w))
    ora = oracle_of(net).simulate(16, 0.05, seed=1234)
    assert (ora["status"] == 0).all()
    same = np.all(res.final_x[:16] == ora["x_final"], axis=1)
    assert same.mean() >= 0.8, same.mean()
    assert np.array_equal(res.steps[:16][same], ora["steps"][same])
