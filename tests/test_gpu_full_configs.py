"""The benchmark configurations at their full sizes on the GPU, against the
oracle fixtures of tests/golden/ (written in the build container by
tests/golden/make_config_fixtures.py).

Realization r's random streams depend only on (seed, model id, r), so the
first realizations of a full-size GPU run simulate the same trajectories as
the fixtures' oracle runs: nearly all must agree exactly (state, steps), the
rest statistically (Welch t and two-sample KS at 1 %); size-independent
invariants (conservation laws) hold in every realization of the full run.

  * dimerisation-decay, S1 = 100,000, t_final 10, all 10,000 realizations;
  * toy network 128 x 256, t_final 0.1, 100,000 realizations, FP64 and FP32
    (the first 4,096: species 0-7 and a checksum of all 128);
  * cyclic chain 1024, t_final 1, 16,384 realizations (the first 4,096; the
    population total is conserved in all);
  * colloid 250 (31,250 channels), t_final 0.05, 4,096 realizations, FP64 and
    FP32, and the plain form on 256 (the first 256 against the fixture, all
    250 species; the mass sum_i i x_i is conserved in all);
  * the 64-model heterogeneous batch, FP32, scores on, in one persistent
    launch (each model's first 256 realizations: species 0-3, steps and the
    score projected on 4 fixed directions).
"""
import os

import numpy as np
import pytest

import tau_leaping_amd as T
from helpers import ks_ok, mean_ci_ok
from tau_leaping_amd import models

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def checksum(x):
    return x.astype(np.int64) @ np.arange(1, x.shape[1] + 1, dtype=np.int64)


def _agree(gpu_cols, fix_cols, min_same):
    """gpu_cols / fix_cols: lists of [R] or [R][k] arrays compared together."""
    def flat(cols):
        return np.concatenate([np.asarray(c).reshape(len(c), -1).astype(np.float64) for c in cols], axis=1)

    g, f = flat(gpu_cols), flat(fix_cols)
    same = np.all(g == f, axis=1)
    assert same.mean() >= min_same, f"identical realizations {same.mean():.4f}"
    for q in range(g.shape[1]):
        assert mean_ci_ok(g[:, q], f[:, q])
        assert ks_ok(g[:, q], f[:, q])
    return same


def test_dimer_decay_full(cuda):
    fx = fixture("c1_dimer_decay.npz")
    m = T.Model.from_network(models.dimer_decay(100000))
    res = m.simulate(10.0, 10000, seed=1234, track=[0, 2])
    assert res.n_failed == 0
    _agree([res.final_x, res.steps], [fx["x"], fx["steps"]], 0.97)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_toy_network_full(cuda, prec):
    fx = fixture(f"c2_toy128x256_{prec}.npz")
    m = T.Model.from_network(models.toy_network(128, 256, seed=42))
    res = m.simulate(0.1, 100000, seed=1234, precision=prec)
    assert res.n_failed == 0
    n = fx["steps"].size
    _agree([res.final_x[:n, :8], checksum(res.final_x[:n]), res.steps[:n]],
           [fx["x"], fx["checksum"], fx["steps"]], 0.97)


def test_cyclic_chain_full(cuda):
    fx = fixture("c3_cyclic1024.npz")
    net = models.cyclic_chain(1024, seed=43)
    m = T.Model.from_network(net)
    res = m.simulate(1.0, 16384, seed=1234)
    assert res.n_failed == 0
    assert np.all(res.final_x.astype(np.int64).sum(axis=1) == int(net.x0.astype(np.int64).sum()))
    n = fx["steps"].size
    _agree([res.final_x[:n][:, [0, 256, 512, 768]], checksum(res.final_x[:n]), res.steps[:n]],
           [fx["x"], fx["checksum"], fx["steps"]], 0.97)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_colloid_full(cuda, prec):
    fx = fixture(f"c4_colloid250_{prec}.npz")
    net = models.colloidal(250, seed=44)
    m = T.Model.from_network(net)
    res = m.simulate(0.05, 4096, seed=1234, precision=prec)
    assert res.n_failed == 0 and res.form == "persistent"
    w = np.arange(1, 251, dtype=np.int64)
    assert np.all(res.final_x.astype(np.int64) @ w == int(net.x0.astype(np.int64) @ w))
    n = fx["steps"].size
    _agree([checksum(res.final_x[:n]), res.steps[:n]], [checksum(fx["x"]), fx["steps"]], 0.9)
    same = np.all(res.final_x[:n] == fx["x"], axis=1) & (res.steps[:n] == fx["steps"])
    assert same.mean() >= 0.9, same.mean()
    assert np.array_equal(res.rejections[:n][same], fx["rejections"][same])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_colloid_plain_form(cuda, prec):
    # the plain per-model launch loop on the same realizations
    fx = fixture(f"c4_colloid250_{prec}.npz")
    m = T.Model.from_network(models.colloidal(250, seed=44))
    n = fx["steps"].size
    res = m.simulate(0.05, n, seed=1234, precision=prec, mode="plain")
    assert res.n_failed == 0 and res.form == "plain"
    same = np.all(res.final_x == fx["x"], axis=1) & (res.steps == fx["steps"])
    assert same.mean() >= 0.9, same.mean()
    _agree([checksum(res.final_x), res.steps], [checksum(fx["x"]), fx["steps"]], 0.9)


def test_hetero64_batch(cuda):
    # all 64 models in one persistent batch at FP32 with scores; up to 512
    # realizations each (the full counts run in the benchmark)
    fx = fixture("c5_hetero64_fp32.npz")
    nets, reals = models.hetero_batch(64)
    ms = [T.Model.from_network(n) for n in nets]
    kws = [dict(t_final=0.1, n_realizations=min(r, 512), seed=1234 + i, model_id=i, precision="fp32",
                sensitivities=True, return_scores=True, track=np.arange(4, dtype=np.int32)) for i, r in enumerate(reals)]
    out = T.batch_simulate(ms, kws, mode="persistent")
    ok_all = []
    for i, (net, res) in enumerate(zip(nets, out)):
        assert res.n_failed == 0
        x, steps, proj = fx[f"x_{i}"], fx[f"steps_{i}"], fx[f"proj_{i}"]
        n = steps.size
        # the fixture tracks species 0-3, so the score projections are what
        # shows a path that differs in the other species
        W = np.random.default_rng(7000 + i).standard_normal((net.n_channels, 4))
        got = res.score[:n] @ W
        scale = np.abs(proj).max(axis=1) + 1e-300
        ok = (np.all(res.final_x[:n] == x, axis=1) & (res.steps[:n] == steps) &
              (np.abs(got - proj).max(axis=1) < 1e-8 * scale))
        ok_all.append(ok)
        assert ok.mean() >= 0.9, (i, ok.mean())
    assert np.concatenate(ok_all).mean() >= 0.97, np.concatenate(ok_all).mean()
