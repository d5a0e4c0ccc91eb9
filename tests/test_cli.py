"""The command-line tools: toy_system_gen writes StochKit models the reader
accepts, with the structure of each family (random, -w cyclic chain, -s
colloid, --toy benchmark recipe); tau_leaping parses the reference's flags and
fails loudly without a GPU."""
import os
import subprocess

import numpy as np
import pytest

import tau_leaping_amd as T
from tau_leaping_amd import models

BIN = os.path.join(os.path.dirname(T.__file__), "bin")
GEN = os.path.join(BIN, "toy_system_gen")
SIM = os.path.join(BIN, "tau_leaping")

pytestmark = pytest.mark.skipif(not (os.path.exists(GEN) and os.path.exists(SIM)), reason="CLI tools not built")


def gen(tmp_path, name, *args):
    out = tmp_path / name
    with open(out, "w") as f:
        r = subprocess.run([GEN, *args], stdout=f, stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return str(out)


def test_colloid_family_structure(tmp_path):
    m = T.Model.from_stochkit(gen(tmp_path, "s.xml", "-r", "20", "-s", "--seed", "3"))
    ref = T.Model.from_network(models.colloidal(20, seed=5))
    assert (m.n_species, m.n_channels, m.nnz) == (20, 200, 580)
    # same reactions as the generator family in models.py (rates and
    # populations come from different random streams)
    assert [tuple(a) for a in m.structure()[:3]] == [tuple(a) for a in ref.structure()[:3]]


def test_cyclic_family_structure(tmp_path):
    m = T.Model.from_stochkit(gen(tmp_path, "w.xml", "-r", "8", "-w", "--seed", "3"))
    order, s0, s1, rp, csr = m.structure()
    assert (m.n_species, m.n_channels) == (8, 8)
    assert (order == 1).all() and (s0 == np.arange(8)).all()


def test_toy_recipe(tmp_path):
    path = gen(tmp_path, "t.xml", "-r", "16", "-c", "40", "--toy", "--seed", "5", "--log-rates", "0.001", "1")
    m = T.Model.from_stochkit(path)
    assert (m.n_species, m.n_channels) == (16, 40)
    order, s0, s1, _, _ = m.structure()
    assert set(order.tolist()) <= {1, 2}
    r = m.rates
    assert (r >= 1e-3 * (1 - 1e-12)).all() and (r <= 1.0).all()
    # the same seed writes the same file
    assert open(path).read() == open(gen(tmp_path, "t2.xml", "-r", "16", "-c", "40", "--toy", "--seed", "5",
                                         "--log-rates", "0.001", "1")).read()


def test_simulator_flags(tmp_path):
    path = gen(tmp_path, "w.xml", "-r", "8", "-w", "--seed", "3")
    bad = subprocess.run([SIM, "-f", path, "-t", "-1"], capture_output=True, text=True, timeout=60)
    assert bad.returncode != 0 and "-t" in bad.stderr
    nofile = subprocess.run([SIM, "-t", "1"], capture_output=True, text=True, timeout=60)
    assert nofile.returncode != 0 and "-f" in nofile.stderr
    unknown = subprocess.run([SIM, "-f", path, "-t", "1", "--bogus"], capture_output=True, text=True, timeout=60)
    assert unknown.returncode != 0 and "unknown option" in unknown.stderr
    missing = subprocess.run([SIM, "-f", str(tmp_path / "none.xml"), "-t", "1"], capture_output=True, text=True,
                             timeout=60)
    assert missing.returncode != 0


def test_simulator_without_gpu_fails_loudly(tmp_path):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    path = gen(tmp_path, "w.xml", "-r", "8", "-w", "--seed", "3")
    r = subprocess.run([SIM, "-f", path, "-t", "1", "-n", "10"], capture_output=True, text=True, timeout=60,
                       cwd=str(tmp_path))
    assert r.returncode != 0
    assert "HIP" in (r.stdout + r.stderr)


@pytest.mark.gpu
def test_simulator_end_to_end(tmp_path, cuda):
    """tau_leaping -f -t -n -r -i -b on the dimerisation-decay model
    (S1 = 100,000): the mean_/stddev_ files hold the grid moments the Python
    binding returns for the same seed, and the moments at t_final agree with
    the oracle's within their standard errors."""
    import json

    from helpers import oracle_of

    net = models.dimer_decay(100000)
    path = tmp_path / "dimer.xml"
    path.write_text(net.to_stochkit_xml())
    R, tf = 4000, 10.0
    r = subprocess.run([SIM, "-f", str(path), "-t", str(tf), "-n", str(R), "-r", "S1", "S3", "-i", "5",
                        "--seed", "1234", "-b", "run", "--json"], capture_output=True, text=True, timeout=300,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    assert rec["failed"] == 0 and rec["realizations"] == R and rec["mode"] == "auto"
    mean = np.loadtxt(tmp_path / "mean_run.txt")
    sd = np.loadtxt(tmp_path / "stddev_run.txt")
    assert mean.shape == (6, 3) and sd.shape == (6, 3)
    np.testing.assert_allclose(mean[:, 0], np.linspace(0, tf, 6))
    m = T.Model.from_stochkit(str(path))
    res = m.simulate(tf, R, seed=1234, track=[0, 2], n_intervals=5)
    np.testing.assert_allclose(mean[:, 1:], res.mean, rtol=1e-5, atol=1e-3)  # six significant digits in the files
    np.testing.assert_allclose(sd[:, 1:], res.stddev, rtol=1e-5, atol=1e-3)
    ora = oracle_of(net).simulate(R, tf, seed=1234)["x_final"][:, [0, 2]].astype(float)
    se = np.sqrt(sd[-1, 1:] ** 2 / R + ora.var(0, ddof=1) / R)
    assert np.all(np.abs(mean[-1, 1:] - ora.mean(0)) < 3.3 * se)


@pytest.mark.gpu
def test_simulator_net_stoichiometry(tmp_path, cuda):
    """--net-stoichiometry runs the model with catalysts counted once in the
    tau bound: the same realization-steps as the binding with the switch on,
    fewer than without it."""
    import json

    net = models.gene_expression()
    path = tmp_path / "gene.xml"
    path.write_text(net.to_stochkit_xml())
    R, tf = 2000, 1.0
    recs = {}
    for flag in ([], ["--net-stoichiometry"]):
        r = subprocess.run([SIM, "-f", str(path), "-t", str(tf), "-n", str(R), "-r", "all", "--seed", "5",
                            "-b", "g", "--json"] + flag, capture_output=True, text=True, timeout=300,
                           cwd=str(tmp_path))
        assert r.returncode == 0, r.stderr
        recs[bool(flag)] = json.loads(r.stdout.strip().splitlines()[-1])
    m = T.Model.from_stochkit(str(path))
    m.net_stoichiometry = True
    res = m.simulate(tf, R, seed=5)
    assert recs[True]["total_steps"] == res.total_steps
    assert recs[False]["total_steps"] > 1.3 * recs[True]["total_steps"]
