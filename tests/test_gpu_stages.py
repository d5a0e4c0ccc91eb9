"""Stage-level parity of the HIP kernels against the CPU oracle on identical
inputs (tl_stage_* entry points of the C ABI).

Deterministic stages use the same operation order as the oracle, so the
bar is: propensities, criticality, tau1 / argmin, leap selection and the
integer update exact; a0 within 1e-12 relative (FP64) / 1e-6 (FP32 values
summed in FP64); firing counts identical except where a last-ulp difference
of the device exp/log flips an inversion comparison (at most 1e-4 of entries).
"""
import numpy as np
import pytest

import tau_leaping_amd as T
from helpers import SMALL_NETS, pair, random_states

pytestmark = pytest.mark.gpu

PRECS = ["fp64", "fp32"]


def _t(x, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _n(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("name", list(SMALL_NETS))
@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("R", [777, 778])  # one / two realizations per thread
def test_propensities_and_tau(cuda, name, prec, R):
    net = SMALL_NETS[name]()
    m, om = pair(net)
    x = random_states(net.n_species, R, seed=11)
    a_o, c_o, a0_o, a0c_o = om.propensities(x, prec)
    a, c, a0, a0c = T.stage_propensities(m, _t(x, cuda), prec)
    assert np.array_equal(_n(a), a_o)
    assert np.array_equal(_n(c), c_o)
    tol = 1e-12 if prec == "fp64" else 1e-6
    np.testing.assert_allclose(_n(a0), a0_o, rtol=tol, atol=0)
    np.testing.assert_allclose(_n(a0c), a0c_o, rtol=tol, atol=0)

    tau_o, am_o, wh_o = om.tau(x, a_o, c_o, 0.03, prec)
    tau, am, wh = T.stage_tau(m, _t(x, cuda), _t(a_o, cuda), _t(c_o, cuda), 0.03, prec)
    np.testing.assert_allclose(_n(tau), tau_o, rtol=1e-12 if prec == "fp64" else 1e-5, atol=0)
    assert np.array_equal(_n(am), am_o)
    assert np.array_equal(_n(wh), wh_o)


@pytest.mark.parametrize("name", list(SMALL_NETS))
@pytest.mark.parametrize("prec", PRECS)
def test_select_fire_update(cuda, name, prec):
    net = SMALL_NETS[name]()
    m, om = pair(net)
    R = 1000
    rng = np.random.default_rng(5)
    x = random_states(net.n_species, R, seed=12)
    a_o, c_o, a0_o, a0c_o = om.propensities(x, prec)
    tau1_o, _, _ = om.tau(x, a_o, c_o, 0.03, prec)
    epoch = rng.integers(0, 1 << 20, size=R).astype(np.uint32)
    t_rem = np.where(rng.random(R) < 0.1, tau1_o * 0.5, 10.0)  # some leaps clamped
    seed = 1234

    tau_o, jc_o, cl_o = om.select(seed, epoch, a_o, c_o, a0c_o, tau1_o, t_rem, prec, model_id=3, rbase=100)
    tau, jc, cl = T.stage_select(m, seed, _t(epoch, cuda), _t(a_o, cuda), _t(c_o, cuda), _t(a0c_o, cuda),
                                 _t(tau1_o, cuda), _t(t_rem, cuda), prec, model_id=3, rbase=100)
    np.testing.assert_allclose(_n(tau), tau_o, rtol=1e-12, atol=0)
    assert np.array_equal(_n(jc), jc_o)
    assert np.array_equal(_n(cl), cl_o)

    k_o = om.fire(seed, epoch, a_o, c_o, tau_o, jc_o, prec, model_id=3, rbase=100)
    k = _n(T.stage_fire(m, seed, _t(epoch, cuda), _t(a_o, cuda), _t(c_o, cuda), _t(tau_o, cuda), _t(jc_o, cuda),
                        prec, model_id=3, rbase=100))
    mism = np.mean(k != k_o)
    assert mism <= 1e-4, f"firing-count mismatch fraction {mism}"

    xo_o, neg_o = om.update(x, k_o)
    xo, neg = T.stage_update(m, _t(x, cuda), _t(k_o, cuda))
    assert np.array_equal(_n(xo), xo_o)
    assert np.array_equal(_n(neg), neg_o)


@pytest.mark.parametrize("R", [999, 1000])
def test_update_odd_and_even_realizations(cuda, R):
    # even R runs two realizations per thread, odd R one; both exact, with
    # some updates driving a species negative
    net = SMALL_NETS["toy"]()
    m, om = pair(net)
    rng = np.random.default_rng(R)
    x = random_states(net.n_species, R, seed=13)
    k = rng.poisson(3.0, size=(m.n_channels, R)).astype(np.int32)
    xo_o, neg_o = om.update(x, k)
    assert neg_o.any() and not neg_o.all()
    xo, neg = T.stage_update(m, _t(x, cuda), _t(k, cuda))
    assert np.array_equal(_n(xo), xo_o)
    assert np.array_equal(_n(neg), neg_o)


@pytest.mark.parametrize("name", ["dimer", "toy", "colloid"])
@pytest.mark.parametrize("prec", PRECS)
def test_direct_events(cuda, name, prec):
    net = SMALL_NETS[name]()
    m, om = pair(net)
    R = 300
    x = random_states(net.n_species, R, seed=13, small_frac=0.05)
    t = np.zeros(R)
    epoch = np.arange(R, dtype=np.uint32) * 7
    xo, to, eo, evo = om.ssa(99, x, t, epoch, 1e9, 50, prec)
    xd, td, ed = _t(x, cuda), _t(t, cuda), _t(epoch, cuda)
    ev = T.stage_ssa(m, 99, xd, td, ed, 1e9, 50, prec)
    assert np.array_equal(_n(ev), evo)
    assert np.array_equal(_n(ed), eo)
    same = np.all(_n(xd) == xo, axis=0)
    assert same.mean() >= 0.99, f"identical direct-event paths: {same.mean()}"
    np.testing.assert_allclose(_n(td)[same], to[same], rtol=1e-12)


def test_poisson_sampler_stream(cuda):
    """Fire on a single-channel model exercises the inversion (< 80) and
    normal (>= 80) branches of one stream against the oracle's draws."""
    from tau_leaping_amd.models import Network, Reaction

    net = Network(1, [Reaction({0: 1}, {}, 1.0)], np.array([100], np.int32))
    m, om = pair(net)
    means = np.concatenate([np.linspace(0, 79.99, 400), np.linspace(80, 400, 200)])
    R = means.size
    a = np.ascontiguousarray(means.reshape(1, R))
    crit = np.zeros((1, R), np.uint8)
    tau = np.ones(R)
    jc = -np.ones(R, np.int32)
    epoch = np.arange(R, dtype=np.uint32)
    k = _n(T.stage_fire(m, 5, _t(epoch, cuda), _t(a, cuda), _t(crit, cuda), _t(tau, cuda), _t(jc, cuda)))
    k_o = om.fire(5, epoch, a, crit, tau, jc)
    assert np.array_equal(k, k_o)


def test_inversion_draws_many_means(cuda):
    """16,384 FP64 inversion searches (means 0.5 to 80, one draw each, fresh
    streams): the device's step divides through a reciprocal table with an
    FMA correction, the oracle divides; the counts agree draw for draw."""
    from tau_leaping_amd.models import Network, Reaction

    net = Network(1, [Reaction({0: 1}, {}, 1.0)], np.array([100], np.int32))
    m, om = pair(net)
    R = 16384
    means = np.random.default_rng(11).uniform(0.5, 79.999, R)
    a = np.ascontiguousarray(means.reshape(1, R))
    crit = np.zeros((1, R), np.uint8)
    tau = np.ones(R)
    jc = -np.ones(R, np.int32)
    epoch = np.arange(R, dtype=np.uint32) * 7 + 3
    k = _n(T.stage_fire(m, 9, _t(epoch, cuda), _t(a, cuda), _t(crit, cuda), _t(tau, cuda), _t(jc, cuda)))
    k_o = om.fire(9, epoch, a, crit, tau, jc)
    assert np.array_equal(k, k_o)
    assert k.max() > 80  # searches well past the mean


def test_stage_waits_for_callers_stream(cuda):
    # the input is produced by torch kernels queued just before the stage
    # call (a long matrix product first, so they are still running when the
    # stage is called); the stage must see the finished input
    import torch

    net = models_toy()
    m, om = pair(net)
    R = 4096
    x = random_states(net.n_species, R, seed=5)
    a_o, c_o, _, _ = om.propensities(x, "fp64")
    big = torch.randn(4096, 4096, device=cuda)
    for _ in range(3):
        big = big @ big / 64.0
    xs = torch.zeros((net.n_species, R), dtype=torch.int32, device=cuda)
    xs += (big[0, 0] * 0).to(torch.int32)  # depends on the product
    xs += torch.from_numpy(x).to(cuda, non_blocking=True)
    a, c, _, _ = T.stage_propensities(m, xs, "fp64")
    assert np.array_equal(_n(a), a_o) and np.array_equal(_n(c), c_o)


def models_toy():
    from tau_leaping_amd import models

    return models.toy_network(32, 64, seed=7)
