"""Whole-trajectory parity of the HIP engine (persistent and plain forms)
against the CPU oracle.

Realization r draws from a counter stream keyed on (seed, r), independent of
launch shape, so the GPU and the oracle follow the same path until a last-ulp
difference of a device math function flips a discrete decision.  Bars:
  * per realization: identical final state for >= 97% of realizations, and
    the rest statistically consistent (Welch t at 1%, two-sample KS at 1%);
  * conservation laws exact in every realization;
  * a batch run is identical to single-model runs, a run repeats bit for
    bit, and wavefront / block launch shapes agree to the oracle bar.
"""
import numpy as np
import pytest

import tau_leaping_amd as T
from helpers import SMALL_NETS, ks_ok, mean_ci_ok, oracle_of, pair
from tau_leaping_amd import models

pytestmark = pytest.mark.gpu


def _compare(gpu_x, ora_x, min_same=0.97):
    same = np.all(gpu_x == ora_x, axis=1)
    assert same.mean() >= min_same, f"identical realizations {same.mean():.4f}"
    for q in range(gpu_x.shape[1]):
        assert mean_ci_ok(gpu_x[:, q], ora_x[:, q])
        assert ks_ok(gpu_x[:, q], ora_x[:, q])
    return same


@pytest.mark.parametrize("name,tf", [("dimer", 10.0), ("toy", 0.1), ("cyclic", 1.0), ("colloid", 0.05)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("mode", ["persistent", "plain"])
def test_trajectories_match_oracle(cuda, name, tf, prec, mode):
    net = SMALL_NETS[name]()
    m, om = pair(net)
    R = 1024
    res = m.simulate(tf, R, seed=1234, precision=prec, mode=mode)
    ora = om.simulate(R, tf, seed=1234, precision=prec)
    assert res.n_failed == 0
    assert (ora["status"] == 0).all()
    same = _compare(res.final_x, ora["x_final"])
    assert np.array_equal(res.steps[same], ora["steps"][same])
    assert res.total_steps == int(res.steps.sum())


@pytest.mark.parametrize("bs", [256, 1024])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_block_groups_match_oracle(cuda, bs, prec):
    # one realization per block: tiled species sums, word-level critical
    # search (the colloid keeps ~40 % of its channels critical), direct
    # events with block barriers (the larger toy network)
    for net, tf, R in [(models.colloidal(60, seed=8), 0.05, 512), (models.toy_network(300, 1100, seed=12), 0.004, 256)]:
        m, om = pair(net)
        res = m.simulate(tf, R, seed=4321, precision=prec, block_size=bs)
        ora = om.simulate(R, tf, seed=4321, precision=prec)
        assert res.n_failed == 0 and (ora["status"] == 0).all()
        assert res.block_size == bs
        same = _compare(res.final_x, ora["x_final"])
        assert np.array_equal(res.steps[same], ora["steps"][same])


@pytest.mark.parametrize("bs", [256, 512, 1024])
def test_launch_shapes_agree(cuda, bs):
    # one realization per wavefront (64) against one per block: the species
    # sums of the tau bound round in another order, so the shapes are held to
    # the oracle bar
    net = models.toy_network(48, 120, seed=9)
    m = T.Model.from_network(net)
    ref = m.simulate(0.05, 600, seed=77, block_size=64)
    res = m.simulate(0.05, 600, seed=77, block_size=bs)
    assert res.block_size == bs and ref.block_size == 64
    _compare(res.final_x, ref.final_x)


@pytest.mark.parametrize("bs", [64, 1024])
def test_repeatable(cuda, bs):
    m = T.Model.from_network(models.toy_network(48, 120, seed=9))
    a = m.simulate(0.05, 600, seed=78, block_size=bs)
    b = m.simulate(0.05, 600, seed=78, block_size=bs)
    assert np.array_equal(a.final_x, b.final_x)
    assert np.array_equal(a.steps, b.steps)


def test_batch_equals_single_runs(cuda):
    nets = [models.toy_network(16 + 8 * i, 40 + 20 * i, seed=100 + i) for i in range(5)]
    ms = [T.Model.from_network(n) for n in nets]
    kws = [dict(t_final=0.05, n_realizations=200 + 37 * i, seed=1234 + i, model_id=i) for i in range(5)]
    batch = T.batch_simulate(ms, kws, mode="persistent")
    for m, kw, b in zip(ms, kws, batch):
        single = m.simulate(**kw)
        assert np.array_equal(b.final_x, single.final_x)
        assert np.array_equal(b.steps, single.steps)
    plain = T.batch_simulate(ms, kws, mode="plain")
    for b, p in zip(batch, plain):
        assert np.mean(np.all(b.final_x == p.final_x, axis=1)) >= 0.97


def test_default_shape_is_deterministic(cuda):
    # without block_size the launch shape is a pure function of the model
    # sizes: two runs give the same trajectories bit for bit; the timed
    # autotuner (block_size -1) may pick another shape, whose trajectories
    # agree to the oracle bar
    m = T.Model.from_network(models.toy_network(33, 70, seed=31))
    a = m.simulate(0.02, 20000, seed=91)
    b = m.simulate(0.02, 20000, seed=91)
    assert a.n_failed == 0 and a.block_size == b.block_size == 64
    assert np.array_equal(a.final_x, b.final_x) and np.array_equal(a.steps, b.steps)
    tuned = m.simulate(0.02, 20000, seed=91, block_size=-1)
    assert tuned.n_failed == 0
    if tuned.block_size == 64:
        assert np.array_equal(tuned.final_x, a.final_x)
    else:
        _compare(tuned.final_x, a.final_x)


def test_mixed_size_batch(cuda):
    # models of very different sizes land in different size classes (one
    # persistent launch each, one graph); each model's results agree with its
    # single run to the oracle bar (a class may pick another launch shape)
    nets = [models.toy_network(12, 30, seed=201), models.toy_network(400, 1400, seed=202),
            models.toy_network(40, 100, seed=203), models.colloidal(30, seed=204)]
    ms = [T.Model.from_network(n) for n in nets]
    kws = [dict(t_final=0.02, n_realizations=400, seed=77 + i, model_id=i) for i in range(len(nets))]
    batch = T.batch_simulate(ms, kws)
    for m, kw, b in zip(ms, kws, batch):
        single = m.simulate(**kw)
        assert b.n_failed == 0
        _compare(b.final_x, single.final_x)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_model_larger_than_lds(cuda, prec):
    # 6000 species x 12000 channels: a realization's populations, snapshots
    # and species sums (28 B per species) exceed the 160 KiB of LDS.  The
    # persistent form runs it on 1024-lane groups whose state lives in HBM
    # (AUTO picks that), the plain form with its state in HBM too; both follow
    # the oracle's trajectories.
    net = models.toy_network(6000, 12000, seed=41)
    m, om = pair(net)
    R, tf = 256, 1e-4
    track = np.arange(0, 6000, 150)
    res = m.simulate(tf, R, seed=5, track=track, precision=prec)
    pers = m.simulate(tf, R, seed=5, track=track, precision=prec, mode="persistent")
    plain = m.simulate(tf, R, seed=5, track=track, precision=prec, mode="plain")
    assert res.block_size == 1024
    assert np.array_equal(res.final_x, pers.final_x) and np.array_equal(res.steps, pers.steps)
    ora = om.simulate(R, tf, seed=5, precision=prec)
    assert (ora["status"] == 0).all()
    for r in (res, plain):
        assert r.n_failed == 0
        same = _compare(r.final_x, ora["x_final"][:, track])
        assert np.array_equal(r.steps[same], ora["steps"][same])


@pytest.mark.parametrize("prec,R", [("fp64", 4096), ("fp32", 4096), ("fp64", 4095)])
def test_plain_form_one_pass_species_minimum(cuda, prec, R):
    # 6000 species, ~4096 realizations: the plain form sums the species rows
    # and takes each row chunk's tau minimum in one pass (93 chunks of 65
    # rows, no per-species bound stored; two realizations per thread when R
    # is even, one when odd); its realizations follow the persistent kernel's
    # and the oracle's
    net = models.toy_network(6000, 12000, seed=41)
    m, om = pair(net)
    tf = 1e-4
    track = np.arange(0, 6000, 150)
    plain = m.simulate(tf, R, seed=5, track=track, precision=prec, mode="plain")
    pers = m.simulate(tf, R, seed=5, track=track, precision=prec, mode="persistent")
    assert plain.n_failed == 0 and pers.n_failed == 0
    same = _compare(plain.final_x, pers.final_x)
    assert np.array_equal(plain.steps[same], pers.steps[same])
    ora = om.simulate(64, tf, seed=5, precision=prec)
    assert (ora["status"] == 0).all()
    same = _compare(plain.final_x[:64], ora["x_final"][:, track], min_same=0.95)
    assert np.array_equal(plain.steps[:64][same], ora["steps"][same])


def test_model_with_many_tiles_stores_a_in_hbm(cuda):
    # 6000 species x 300,000 channels: 19 propensity tiles, so the HBM-state
    # kernel keeps a[M] in global memory and sums the species rows in one
    # pass; trajectories follow the oracle's (AUTO would run the plain form)
    import tl_oracle as O
    from helpers import toy_csc

    csc = toy_csc(6000, 300000, 7)
    m = T.Model.from_csc(6000, *csc)
    om = O.OracleModel(6000, 300000, *csc)
    R, tf = 32, 1e-6
    track = np.arange(0, 6000, 100)
    res = m.simulate(tf, R, seed=5, track=track, mode="persistent")
    ora = om.simulate(R, tf, seed=5)
    assert res.block_size == 1024 and res.n_failed == 0 and (ora["status"] == 0).all()
    same = _compare(res.final_x, ora["x_final"][:, track], min_same=0.9)
    assert np.array_equal(res.steps[same], ora["steps"][same])


def test_large_model_with_extended_forms(cuda):
    # a model too large for LDS that also has sources and three-reactant
    # channels: the HBM-state kernels of the extended build, against the oracle
    import tl_oracle as O
    from helpers import toy_csc

    N, M = 6000, 12000
    col_ptr, row_ind, stoich, rates, x0 = toy_csc(N, M - 2, 11)
    # append a source (-> S0) and S1 + S2 + S3 -> S4
    col_ptr = np.append(col_ptr, [col_ptr[-1] + 1, col_ptr[-1] + 5]).astype(np.int32)
    row_ind = np.append(row_ind, [0, 1, 2, 3, 4]).astype(np.int32)
    stoich = np.append(stoich, [1, -1, -1, -1, 1]).astype(np.int8)
    rates = np.append(rates, [50.0, 1e-9])
    m = T.Model.from_csc(N, col_ptr, row_ind, stoich, rates, x0)
    om = O.OracleModel(N, M, col_ptr, row_ind, stoich, rates, x0)
    assert list(m.structure()[0][-2:]) == [0, 4]
    R, tf = 128, 1e-4
    track = np.arange(0, N, 150)
    res = m.simulate(tf, R, seed=21, track=track)
    ora = om.simulate(R, tf, seed=21)
    assert res.block_size == 1024 and res.n_failed == 0 and (ora["status"] == 0).all()
    same = _compare(res.final_x, ora["x_final"][:, track], min_same=0.95)
    assert np.array_equal(res.steps[same], ora["steps"][same])


@pytest.mark.parametrize("M,form", [(12000, "persistent"), (24000, "plain"), (100000, "plain"), (300000, "plain")])
def test_auto_form_for_large_models(cuda, M, form):
    # models too large for LDS: AUTO keeps the HBM-state persistent kernel
    # while one propensity tile holds the channels (12,000) and runs the plain
    # form beyond (24,000: two tiles; 1e5: 7; 3e5: a[M] in HBM), a pure function
    # of the sizes; the result is that form's, bit for bit, and its first
    # realizations are the oracle's
    import tl_oracle as O
    from helpers import toy_csc

    N, R, tf = 20000, 2048, 1e-6 if M > 100000 else 1e-5
    csc = toy_csc(N, M, 17)
    m = T.Model.from_csc(N, *csc)
    track = np.arange(0, N, 500)
    res = m.simulate(tf, R, seed=8, track=track)
    assert res.form == form
    ref = m.simulate(tf, R, seed=8, track=track, mode=form)
    assert res.n_failed == 0
    assert np.array_equal(res.final_x, ref.final_x) and np.array_equal(res.steps, ref.steps)
    ora = O.OracleModel(N, M, *csc).simulate(16, tf, seed=8)
    same = np.all(res.final_x[:16] == ora["x_final"][:, track], axis=1)
    assert same.mean() >= 0.8, same.mean()


def _source_network():
    # sources, decay, dimerisation (2A -> B), a reversible association and a
    # catalytic step (B + D -> B + E)
    from tau_leaping_amd.models import Network, Reaction
    rx = [Reaction({}, {0: 1}, 50.0), Reaction({0: 1}, {}, 0.1), Reaction({0: 2}, {1: 1}, 0.001),
          Reaction({1: 1}, {}, 0.05), Reaction({}, {2: 1}, 20.0), Reaction({0: 1, 2: 1}, {3: 1}, 0.002),
          Reaction({3: 1}, {0: 1, 2: 1}, 0.5), Reaction({1: 1, 3: 1}, {1: 1, 4: 1}, 0.001),
          Reaction({4: 1}, {}, 0.2), Reaction({}, {5: 2}, 5.0)]
    return Network(6, rx, np.array([100, 10, 50, 0, 0, 3], np.int32))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("mode,bs", [("persistent", 64), ("persistent", 256), ("plain", 0)])
def test_sources_match_oracle(cuda, prec, mode, bs):
    # channels without reactants (a = c, never critical) in every form
    m, om = pair(_source_network())
    R, tf = 2048, 5.0
    res = m.simulate(tf, R, seed=314, precision=prec, mode=mode, block_size=bs)
    ora = om.simulate(R, tf, seed=314, precision=prec)
    assert res.n_failed == 0 and (ora["status"] == 0).all()
    same = _compare(res.final_x, ora["x_final"])
    assert np.array_equal(res.steps[same], ora["steps"][same])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("mode,bs", [("persistent", 64), ("persistent", 1024), ("plain", 0)])
def test_exact_sampler_matches_oracle(cuda, prec, mode, bs):
    # dimerisation-decay from S1 = 100,000: firing means far above 80, so
    # every leap draws through PTRS in the exact option
    m, om = pair(models.dimer_decay(100000))
    R, tf = 1024, 2.0
    res = m.simulate(tf, R, seed=77, precision=prec, mode=mode, block_size=bs, poisson="exact")
    ora = om.simulate(R, tf, seed=77, precision=prec, poisson="exact")
    assert res.n_failed == 0 and (ora["status"] == 0).all()
    same = _compare(res.final_x, ora["x_final"])
    assert np.array_equal(res.steps[same], ora["steps"][same])
    normal = m.simulate(tf, R, seed=77, precision=prec, mode=mode, block_size=bs)
    assert not np.array_equal(normal.final_x, res.final_x)  # the option changes the draws


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("mode,bs", [("persistent", 64), ("persistent", 256), ("persistent", 1024), ("plain", 0)])
def test_fast_sampler_matches_oracle(cuda, prec, mode, bs):
    # the cyclic chain draws means of 30-80 per leap: the fast option samples
    # them by PTRS in every form, as the oracle does
    m, om = pair(models.cyclic_chain(256, seed=9))
    R, tf = 1024, 0.5
    res = m.simulate(tf, R, seed=66, precision=prec, mode=mode, block_size=bs, poisson="fast")
    ora = om.simulate(R, tf, seed=66, precision=prec, poisson="fast")
    assert res.n_failed == 0 and (ora["status"] == 0).all()
    same = _compare(res.final_x, ora["x_final"])
    assert np.array_equal(res.steps[same], ora["steps"][same])
    ref = m.simulate(tf, R, seed=66, precision=prec, mode=mode, block_size=bs)
    assert not np.array_equal(ref.final_x, res.final_x)  # the option changes the draws


def test_exact_sampler_law(cuda):
    # a source at rate 1e4 from x0 = 1e6: tau1 = eps*x0/c = 3 > t_final, so the
    # run is one leap clamped to t_final and x - x0 is one firing count of
    # mean 500.  Exact option: Poisson(500) by chi-square; the default keeps
    # the reference's floor(|N(500, 500)|).
    from scipy import stats

    from tau_leaping_amd.models import Network, Reaction
    m = T.Model.from_network(Network(1, [Reaction({}, {0: 1}, 1e4)], np.array([10 ** 6], np.int32)))
    R = 40000
    res = m.simulate(0.05, R, seed=5, poisson="exact")
    assert (res.steps == 1).all()
    k = res.final_x[:, 0].astype(np.int64) - 10 ** 6
    support = np.arange(380, 630)
    pmf = stats.poisson.pmf(support, 500.0)
    obs = np.array([(k == v).sum() for v in support], float)
    exp_ = np.append(pmf * R, R - (pmf * R).sum())
    obs = np.append(obs, R - obs.sum())
    assert stats.chisquare(obs, exp_).pvalue > 1e-3
    assert abs(k.mean() - 500) < 4 * np.sqrt(500 / R)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("mode,bs", [("persistent", 64), ("persistent", 256), ("persistent", 1024), ("plain", 0)])
def test_three_reactant_channels_match_oracle(cuda, prec, mode, bs):
    # Sa+Sb+Sc, 2Sa+Sb and 3S channels (the extended propensity formulas
    # and their third reactant) in every form
    from helpers import trimolecular_net
    m, om = pair(trimolecular_net())
    assert list(m.structure()[0][:6]) == [4, 1, 5, 1, 4, 1] and m.structure()[0][-1] == 6
    R, tf = 1024, 2.0
    res = m.simulate(tf, R, seed=808, precision=prec, mode=mode, block_size=bs)
    ora = om.simulate(R, tf, seed=808, precision=prec)
    assert res.n_failed == 0 and (ora["status"] == 0).all()
    same = _compare(res.final_x, ora["x_final"])
    assert np.array_equal(res.steps[same], ora["steps"][same])


def test_conservation_laws(cuda):
    # cyclic chain: sum_i x_i constant; colloid: sum_i i * x_i constant
    cyc = models.cyclic_chain(256, seed=1)
    res = T.Model.from_network(cyc).simulate(1.0, 2000, seed=5)
    assert (res.final_x.sum(axis=1) == cyc.x0.sum()).all()
    col = models.colloidal(40, seed=2)
    res = T.Model.from_network(col).simulate(0.05, 500, seed=6)
    w = np.arange(1, 41)
    assert (res.final_x @ w == col.x0 @ w).all()


def test_status_codes(cuda):
    net = models.toy_network(16, 32, seed=2)
    m = T.Model.from_network(net)
    res = m.simulate(0.1, 64, seed=1, max_iter=3)
    assert res.n_failed == 64 and res.first_failure == 8  # TL_ERR_ITER_CAP
    assert (res.steps == 3).all()
    om = oracle_of(net)
    ora = om.simulate(64, 0.1, seed=1, max_iter=3)
    assert (ora["status"] == 8).all()


def test_batch_graph_replays(cuda, monkeypatch):
    # a batch of several size classes runs as one graph whose branches (one
    # per class, up to four streams) run concurrently; the executable graph
    # is kept per batch shape and replayed with new arguments: every replay,
    # with other device work and synchronizations between them, equals the
    # same launches run without a graph
    import torch

    nets = [models.toy_network(12, 30, seed=301), models.toy_network(400, 1400, seed=302),
            models.toy_network(60, 150, seed=303), models.colloidal(30, seed=304), models.toy_network(150, 500, seed=305)]
    ms = [T.Model.from_network(n) for n in nets]
    kws = [dict(t_final=0.02, n_realizations=300 + 50 * i, seed=900 + i, model_id=i) for i in range(len(nets))]
    monkeypatch.setenv("TL_GRAPH", "0")
    eager = T.batch_simulate(ms, kws, mode="persistent")
    monkeypatch.delenv("TL_GRAPH")
    for rep in range(3):
        junk = torch.randn(2048, 2048, device=cuda)
        (junk @ junk).sum().item()
        torch.cuda.synchronize()
        got = T.batch_simulate(ms, kws, mode="persistent")
        for e, g in zip(eager, got):
            assert g.n_failed == 0
            assert np.array_equal(e.final_x, g.final_x), rep
            assert np.array_equal(e.steps, g.steps), rep
    # another realization count (new arguments, same shape): still the eager result
    kws2 = [dict(k, n_realizations=k["n_realizations"] // 2) for k in kws]
    got = T.batch_simulate(ms, kws2, mode="persistent")
    monkeypatch.setenv("TL_GRAPH", "0")
    eager2 = T.batch_simulate(ms, kws2, mode="persistent")
    for e, g in zip(eager2, got):
        assert np.array_equal(e.final_x, g.final_x) and np.array_equal(e.steps, g.steps)
