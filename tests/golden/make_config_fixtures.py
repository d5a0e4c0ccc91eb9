"""Write the oracle fixtures of the benchmark configurations (run in the build
container, where the oracle has time to finish; the GPU tests compare the
engine's first realizations with them).

Realization r's random streams depend only on (seed, model id, r), so the
first realizations of a full-size GPU run simulate the same trajectories as
these oracle runs.  Files (numpy .npz, no pickles):

  c1_dimer_decay.npz         dimerisation-decay (S1 = 100,000), t_final 10,
                             all 10,000 realizations, FP64: S1, S3, steps
  c2_toy128x256_{fp64,fp32}.npz  toy network 128 x 256 (generator seed 42),
                             t_final 0.1, the first 4,096 of 100,000
                             realizations: species 0-7, a checksum of all 128
                             (sum_i (i + 1) x_i), steps, rejections
  c3_cyclic1024.npz          cyclic chain 1024 (seed 43), t_final 1, the first
                             4,096 of 16,384: species 0, 256, 512, 768,
                             checksum, steps
  c4_colloid250_{fp64,fp32}.npz  colloid 250 (seed 44, 31,250 channels),
                             t_final 0.05, the first 256 of 4,096: all 250
                             species, steps, rejections
  c5_hetero64_fp32.npz       the 64-model heterogeneous batch (models seeds
                             1000 + m, simulation seed 1234 + m, model id m),
                             t_final 0.1, FP32, scores on, the first 256
                             realizations of each model: species 0-3, steps,
                             and the score projected on 4 fixed directions
                             (score @ W_m, W_m = standard normal [M][4] from
                             numpy seed 7000 + m)

Usage: python tests/golden/make_config_fixtures.py [c1 c2 c3 c4 c5]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import tl_oracle as O  # noqa: E402
from tau_leaping_amd import models  # noqa: E402


def oracle(net):
    return O.OracleModel(net.n_species, net.n_channels, *net.to_csc())


def checksum(x):
    return x.astype(np.int64) @ np.arange(1, x.shape[1] + 1, dtype=np.int64)


def projections(M, m):
    return np.random.default_rng(7000 + m).standard_normal((M, 4))


def c1():
    net = models.dimer_decay(100000)
    r = oracle(net).simulate(10000, 10.0, seed=1234)
    assert (r["status"] == 0).all()
    np.savez_compressed(os.path.join(HERE, "c1_dimer_decay.npz"), x=r["x_final"][:, [0, 2]].astype(np.int32),
                        steps=r["steps"])


def c2():
    net = models.toy_network(128, 256, seed=42)
    for prec in ("fp64", "fp32"):
        r = oracle(net).simulate(4096, 0.1, seed=1234, precision=prec)
        assert (r["status"] == 0).all()
        np.savez_compressed(os.path.join(HERE, f"c2_toy128x256_{prec}.npz"), x=r["x_final"][:, :8].astype(np.int32),
                            checksum=checksum(r["x_final"]), steps=r["steps"], rejections=r["rejections"])


def c3():
    net = models.cyclic_chain(1024, seed=43)
    r = oracle(net).simulate(4096, 1.0, seed=1234)
    assert (r["status"] == 0).all()
    np.savez_compressed(os.path.join(HERE, "c3_cyclic1024.npz"), x=r["x_final"][:, [0, 256, 512, 768]].astype(np.int32),
                        checksum=checksum(r["x_final"]), steps=r["steps"])


def c4():
    net = models.colloidal(250, seed=44)
    for prec in ("fp64", "fp32"):
        r = oracle(net).simulate(256, 0.05, seed=1234, precision=prec)
        assert (r["status"] == 0).all()
        np.savez_compressed(os.path.join(HERE, f"c4_colloid250_{prec}.npz"), x=r["x_final"].astype(np.int32),
                            steps=r["steps"], rejections=r["rejections"])


def c5():
    nets, reals = models.hetero_batch(64)
    out = {}
    for m, net in enumerate(nets):
        r = oracle(net).simulate(256, 0.1, seed=1234 + m, model_id=m, precision="fp32", sensitivities=True)
        assert (r["status"] == 0).all()
        out[f"x_{m}"] = r["x_final"][:, :4].astype(np.int32)
        out[f"steps_{m}"] = r["steps"]
        out[f"proj_{m}"] = r["score"] @ projections(net.n_channels, m)
        print(f"c5 model {m}: {net.n_species} x {net.n_channels}", flush=True)
    np.savez_compressed(os.path.join(HERE, "c5_hetero64_fp32.npz"), **out)


if __name__ == "__main__":
    which = sys.argv[1:] or ["c1", "c2", "c3", "c4", "c5"]
    for w in which:
        t0 = time.time()
        globals()[w]()
        print(f"{w}: {time.time() - t0:.1f} s", flush=True)
