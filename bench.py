"""Benchmark of the tau-leaping engine on one MI355X.

Workload (one "step" = one full simulation of the batch to t_final, one
persistent-kernel launch): random toy network, 128 species x 256 channels
(1-2 reactants with coefficient 1, 1-2 products, rates log-uniform in
[1e-3, 1], populations uniform in [0, 10000], generator seed 42), 100,000
realizations, t_final 0.1, eps 0.03, FP64 propensities, int32 populations,
simulation seed 1234.

value = accepted realization-steps (leaps + direct-method blocks) per second
of wall time over the timed steps, model constants resident on the device.
The roofline entry prices the persistent kernel at the algorithmic bytes of
the staged dataflow, B = M*(3*s_a + 8 + 3/8) + 16*N + 8 bytes per
realization-leap (s_a = 8 for FP64), over its HIP-event-timed duration (the
events are recorded on the library's stream around the launch).  `traffic`
is the HBM bytes per launch measured with rocprofv3 counters (FETCH_SIZE,
WRITE_SIZE), read from the committed profile summary when it matches the
workload.  The CPU baseline is the C oracle (same algorithm and random
streams) on the host cores over a bounded sample of the same workload.

Extra fields (on by default, --no-extras to skip): the FP32 variant; the
plain per-model launch loop against the persistent launch on the same
realizations; the cyclic chain with score-function sensitivities (grad/fwd
cost ratio); the 31,250-channel colloid model (persistent FP64 / FP32 and a
plain sample); the heterogeneous 64-model batch in one persistent launch
against the plain per-model loop; large-model ingestion (generator and
reader at 1e6 channels, the HBM-state persistent kernel on 256 realizations,
and the default mode on 4096, which runs it in the plain form).
`--workload colloid` and `--workload large` time the 31,250-channel colloid
and that 1e6-channel model instead of the toy network.
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
TRAFFIC_FILES = [os.path.join(ROOT, "profiles", n) for n in
                 ("k_persistent_toy128x256_fp64_traffic.json", "k_persistent_hbm_toy200000x1000000_fp64_traffic.json")]
ISSUE_FILE = os.path.join(ROOT, "profiles", "k_persistent_toy128x256_fp64_issue.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bytes_per_leap(N, M, fp64=True, sens=False):
    sa = 8 if fp64 else 4
    return M * (3 * sa + 8 + 3 / 8) + 16 * N + 8 + (16 * M if sens else 0)


def measured_traffic(workload):
    """HBM bytes per launch from the committed counter summaries (one JSON
    object per workload), if one was taken on this workload."""
    for path in TRAFFIC_FILES:
        try:
            with open(path) as f:
                t = json.load(f)
        except (OSError, ValueError):
            continue
        if t.get("workload") == workload:
            return (float(t["fetch_kib"]) * 1024 * float(t["fetch_scale"]) +
                    float(t["write_kib"]) * 1024 * float(t["write_scale"]))
    return None


def measured_issue(kernel_name):
    """VALU issue and LDS bandwidth of the kernel from the committed counter
    summary (SQ counters of one dispatch of the same kernel and model at
    fewer realizations), if it covers this kernel."""
    try:
        with open(ISSUE_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("kernel") != kernel_name:
        return None
    return {"valu_busy": t["valu_busy"], "valu_active_lanes": t["valu_active_lanes"], "lds_TBs": t["lds_tbs"],
            "lds_peak_TBs": t["lds_peak_tbs"], "lds_frac": t["lds_frac"], "counter_workload": t["workload"],
            "summary": t["summary"]}


def cpu_baseline(net, R_sample, t_final, seed):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import tl_oracle

    om = tl_oracle.OracleModel(net.n_species, net.n_channels, *net.to_csc())
    t0 = time.perf_counter()
    out = om.simulate(R_sample, t_final, seed=seed, n_threads=0)
    dt = time.perf_counter() - t0
    steps = int(out["steps"].sum())
    # one core on a smaller sample (the first realizations of the same run)
    R1 = max(1, R_sample // 20)
    t0 = time.perf_counter()
    one = om.simulate(R1, t_final, seed=seed, n_threads=1)
    dt1 = time.perf_counter() - t0
    steps1 = int(one["steps"].sum())
    return dict(value=steps / dt, unit="realization-leaps/s", cores=int(out["threads"]), kind="port",
                sample=f"{R_sample} realizations of the same workload (realization ids 0..{R_sample - 1}), "
                       f"{steps} realization-steps in {dt:.1f} s on {out['threads']} threads",
                single_core={"value": steps1 / dt1, "unit": "realization-leaps/s", "cores": 1,
                             "sample": f"realization ids 0..{R1 - 1}, {steps1} realization-steps in {dt1:.1f} s"})


def rate(res):
    return res.total_steps / (res.kernel_ms * 1e-3) if res.kernel_ms > 0 else float("nan")


def hetero_run(T, models, np, tf=0.1):
    """The 64-model heterogeneous batch: FP32, FP64 time, scores on for every
    channel, per-model seeds 1234 + m and model ids m.  The full
    configuration (all 161,120 realizations) in one persistent launch, then
    the batched launch and the plain per-model loop on the same subset (every
    4th model at 1/16 of its realizations, at least 16; same seeds, scores on
    in both): the plain loop takes minutes per model at full size."""
    nets, reals = models.hetero_batch(64)
    ms = [T.Model.from_network(n) for n in nets]
    kws = [dict(t_final=tf, n_realizations=r, seed=1234 + i, model_id=i, precision="fp32", sensitivities=True,
                track=np.arange(4, dtype=np.int32)) for i, r in enumerate(reals)]
    T.batch_simulate(ms, [dict(k, n_realizations=64) for k in kws], want_final=False)  # uploads, warm caches
    t0 = time.perf_counter()
    hb = T.batch_simulate(ms, kws, want_final=False)
    dtb = time.perf_counter() - t0
    steps_b = sum(r.total_steps for r in hb)
    out = {"models": 64, "realizations": int(sum(reals)), "t_final": tf, "precision": "fp32", "sensitivities": True,
           "species_range": [min(n.n_species for n in nets), max(n.n_species for n in nets)],
           "channels_range": [min(n.n_channels for n in nets), max(n.n_channels for n in nets)],
           "leaps": int(steps_b), "ssa_events": int(sum(r.total_ssa_events for r in hb)),
           "persistent_leaps_per_s": steps_b / dtb, "persistent_seconds": dtb,
           "persistent_kernel_ms": hb[0].kernel_ms, "failed": int(sum(r.n_failed for r in hb))}
    pick = list(range(0, 64, 4))
    sub = [dict(kws[i], n_realizations=max(16, reals[i] // 16)) for i in pick]
    sub_ms = [ms[i] for i in pick]
    log("hetero matched subset: batched")
    t0 = time.perf_counter()
    sb = T.batch_simulate(sub_ms, sub, mode="persistent", want_final=False)
    dsb = time.perf_counter() - t0
    log("hetero matched subset: plain loop")
    # the plain form runs the models one after another; one call per model
    # (the same work) so that each model's time is logged as it ends
    sp, dsp = [], 0.0
    for i, mm, kk in zip(pick, sub_ms, sub):
        t0 = time.perf_counter()
        sp += T.batch_simulate([mm], [kk], mode="plain", want_final=False)
        dt = time.perf_counter() - t0
        dsp += dt
        log(f"  plain model {i}: {dt:.1f} s, {sp[-1].total_steps} leaps")
    lb, lp = sum(r.total_steps for r in sb), sum(r.total_steps for r in sp)
    out["matched_subset"] = {
        "models": pick, "realizations": int(sum(k["n_realizations"] for k in sub)), "sensitivities": True,
        "batched_leaps": int(lb), "plain_leaps": int(lp), "batched_seconds": dsb, "plain_seconds": dsp,
        "batched_leaps_per_s": lb / dsb, "plain_leaps_per_s": lp / dsp, "batched_over_plain": (lb / dsb) / (lp / dsp),
        "note": "same models, realizations, seeds and score setting in both forms; wall time of each call"}
    return out


def cyclic_lr_vs_fd(T, models, np, R=262144, n_fd=8):
    """Likelihood-ratio dE[X_k]/dc_j of the 1024-species cyclic chain against
    central finite differences (relative step 1e-3, common random numbers) on
    the n_fd channels feeding each of the 4 tracked species, both over R
    realizations (at 16,384 the finite differences alone carry ~10 % noise)."""
    net = models.cyclic_chain(1024, seed=43)
    m = T.Model.from_network(net)
    track = np.array([0, 256, 512, 768], np.int32)
    res = m.simulate(1.0, R, seed=1234, track=track, sensitivities=True)
    c = np.ones(1024)
    lr, lrse, fd, fdse = [], [], [], []
    for q, k in enumerate(track):
        for d in range(n_fd):
            j = (int(k) - d) % 1024
            outs = []
            for sgn in (+1, -1):
                cc = c.copy()
                cc[j] += sgn * 1e-3
                m.rates = cc
                outs.append(m.simulate(1.0, R, seed=1234, track=[int(k)]).final_x[:, 0].astype(float))
            diff = (outs[0] - outs[1]) / 2e-3
            fd.append(diff.mean())
            fdse.append(diff.std(ddof=1) / np.sqrt(R))
            lr.append(res.dmean_dc[q, j])
            lrse.append(res.dmean_dc_se[q, j])
    lr, lrse, fd, fdse = map(np.array, (lr, lrse, fd, fdse))
    return {"realizations": R, "entries": int(fd.size), "rel_err": float(np.linalg.norm(lr - fd) / np.linalg.norm(fd)),
            "rel_noise": float(np.linalg.norm(np.hypot(lrse, fdse)) / np.linalg.norm(fd)), "bar": 0.05,
            "max_abs_z": float((np.abs(lr - fd) / np.hypot(lrse, fdse)).max()),
            "mean_z2": float(np.mean(((lr - fd) / np.hypot(lrse, fdse)) ** 2))}


def extras_run(T, models, np, m_toy, kw_toy):
    ex = {}
    tf = 0.1
    # FP32 variant of the same workload
    log("fp32 toy")
    r32 = m_toy.simulate(tf, kw_toy["n_realizations"], **{**kw_toy["kw"], "precision": "fp32"})
    ex["toy_fp32"] = {"leaps_per_s": rate(r32), "kernel_ms": r32.kernel_ms, "block_size": r32.block_size}

    # plain per-model launch loop vs persistent on the same realizations
    log("plain toy sample")
    Rp = 500
    rp = m_toy.simulate(tf, Rp, mode="plain", **kw_toy["kw"])
    rq = m_toy.simulate(tf, Rp, **kw_toy["kw"])
    assert rp.total_steps > 0 and rq.total_steps > 0
    ex["toy_plain_vs_persistent"] = {
        "realizations": Rp, "plain_leaps_per_s": rp.total_steps / rp.sim_seconds,
        "persistent_leaps_per_s": rq.total_steps / rq.sim_seconds,
        "ratio": (rq.total_steps / rq.sim_seconds) / (rp.total_steps / rp.sim_seconds),
        "note": "wall time of each call; the persistent launch is far from filling the GPU at this size"}

    # cyclic chain, 1024 species, 16384 realizations, t_final 1: grad / fwd cost
    log("cyclic grad/fwd")
    cyc = T.Model.from_network(models.cyclic_chain(1024, seed=43))
    track = np.array([0, 256, 512, 768], np.int32)
    ckw = dict(seed=1234, track=track)
    cyc.simulate(1.0, 16384, **ckw)
    cyc.simulate(1.0, 16384, sensitivities=True, **ckw)
    f = [cyc.simulate(1.0, 16384, **ckw) for _ in range(3)]
    g = [cyc.simulate(1.0, 16384, sensitivities=True, **ckw) for _ in range(3)]
    fwd = min(r.sim_seconds for r in f)
    grd = min(r.sim_seconds for r in g)
    ex["cyclic1024_grad"] = {
        "realizations": 16384, "fwd_seconds": fwd, "grad_seconds": grd, "grad_fwd_ratio": grd / fwd,
        "fwd_kernel_ms": min(r.kernel_ms for r in f), "grad_kernel_ms": min(r.kernel_ms for r in g),
        "fwd_leaps_per_s": f[0].total_steps / fwd, "grad_fwd_bar": 1.6,
        "lr_vs_fd": cyclic_lr_vs_fd(T, models, np)}

    # colloid, 250 species, 31250 channels, 4096 realizations, t_final 0.05
    log("colloid")
    col_net = models.colloidal(250, seed=44)
    col = T.Model.from_network(col_net)
    ckw = dict(seed=1234, track=np.arange(4, dtype=np.int32))
    cx = {}
    for prec in ("fp64", "fp32"):
        col.simulate(0.05, 256, precision=prec, **ckw)  # tunes the launch shape
        rc = col.simulate(0.05, 4096, precision=prec, **ckw)
        B = bytes_per_leap(250, col_net.n_channels, prec == "fp64")
        cx[prec] = {"leaps_per_s": rate(rc), "kernel_ms": rc.kernel_ms, "block_size": rc.block_size,
                    "bytes_per_leap": B, "algorithmic_GBs": B * rate(rc) / 1e9,
                    "hbm_frac": B * rate(rc) / 1e9 / HBM_PEAK_GBS}
    log("colloid plain sample")
    rcp = col.simulate(0.05, 256, mode="plain", **ckw)
    cx["plain_fp64_256_realizations"] = {"leaps_per_s": rcp.total_steps / rcp.sim_seconds,
                                          "kernel_ms": rcp.kernel_ms}
    ex["colloid250"] = cx

    # heterogeneous batch: 64 toy-recipe models, FP32, sensitivities on; a
    # quarter of each model's realizations (the full batch takes minutes)
    log("hetero batch (full configuration)")
    ex["hetero64"] = hetero_run(T, models, np)

    # large-model ingestion: toy_system_gen at 1e6 channels over 2e5 species
    # (about 500 MB of StochKit XML), the streaming reader, and the
    # persistent engine with each realization's state in HBM
    log("large model")
    import subprocess
    import tempfile
    gen = os.path.join(ROOT, "tau_leaping_amd", "bin", "toy_system_gen")
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "toy200000x1000000.xml")
        t0 = time.perf_counter()
        with open(path, "w") as f:
            subprocess.run([gen, "-r", "200000", "-c", "1000000", "--toy", "--seed", "3", "--log-rates", "0.001", "1"],
                           stdout=f, stderr=subprocess.DEVNULL, check=True)
        t_gen = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        big = T.Model.from_stochkit(path)
        t_read = time.perf_counter() - t0
    rb = big.simulate(1e-6, 256, seed=1, track=np.arange(4, dtype=np.int32), max_iter=3000, mode="persistent")
    lb = bytes_per_leap(big.n_species, big.n_channels)
    # the default mode on more realizations: a model this large runs in the plain form
    t0 = time.perf_counter()
    ra = big.simulate(1e-6, 4096, seed=1, track=np.arange(4, dtype=np.int32), max_iter=3000)
    t_auto = time.perf_counter() - t0
    ex["large1m"] = {
        "species": big.n_species, "channels": big.n_channels, "nnz": big.nnz, "xml_bytes": size,
        "generate_seconds": t_gen, "read_seconds": t_read, "realizations": 256, "t_final": 1e-6,
        "leaps_per_s": rate(rb), "kernel_ms": rb.kernel_ms, "block_size": rb.block_size, "failed": rb.n_failed,
        "bytes_per_leap": lb, "hbm_frac": lb * rb.total_steps / (rb.kernel_ms * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "auto_4096": {"form": ra.form, "seconds": t_auto,
                      "leaps_per_s": rate(ra), "kernel_ms": ra.kernel_ms, "failed": ra.n_failed,
                      "hbm_frac": lb * ra.total_steps / (ra.kernel_ms * 1e-3) / (HBM_PEAK_GBS * 1e9)}}
    return ex


def hetero_main(args, T, models, np, torch):
    """The 64-model heterogeneous batch as the timed step: every model's
    realizations in one batched call (size classes as branches of one cached
    hipGraph), FP32 propensities, FP64 time, scores on for every channel."""
    nets, reals = models.hetero_batch(64)
    ms = [T.Model.from_network(n) for n in nets]
    mode = "persistent" if args.mode == "auto" else args.mode
    kws = [dict(t_final=0.1, n_realizations=r, seed=1234 + i, model_id=i, precision="fp32", sensitivities=True,
                track=np.arange(4, dtype=np.int32)) for i, r in enumerate(reals)]
    log(f"warmup x{args.warmup}")
    for _ in range(args.warmup):
        T.batch_simulate(ms, kws, mode=mode, want_final=False)
    torch.cuda.synchronize()
    log(f"timed x{args.steps}")
    t0 = time.perf_counter()
    res = []
    for _ in range(args.steps):
        res.append(T.batch_simulate(ms, kws, mode=mode, want_final=False))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    steps_total = sum(r.total_steps for rr in res for r in rr)
    # algorithmic bytes of each model's leaps (FP32, scores on)
    B_total = sum(bytes_per_leap(n.n_species, n.n_channels, False, True) * r.total_steps
                  for rr in res for n, r in zip(nets, rr))
    kernel_ms = float(np.mean([rr[0].kernel_ms for rr in res]))
    achieved = B_total / len(res) / (kernel_ms * 1e-3) / 1e9
    workload = (f"64-model heterogeneous batch (toy recipe, model seeds 1000+m), {int(sum(reals))} realizations, "
                f"t_final 0.1, eps 0.03, fp32 propensities, fp64 time, scores on, {mode}")
    line = {
        "metric": METRIC, "value": steps_total / wall, "unit": "realization-leaps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload, "models": 64, "realizations": int(sum(reals)), "t_final": 0.1,
                   "leaps_per_step": steps_total // len(res),
                   "ssa_events_per_step": int(sum(r.total_ssa_events for r in res[-1]))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_persistent (size classes)",
                     "kernel_ms": kernel_ms},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--realizations", type=int, default=100000)
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--workload", default="toy", choices=["toy", "colloid", "large", "hetero"],
                    help="toy: 128 x 256 toy network, 100,000 realizations, t_final 0.1 (the default line); "
                         "colloid: 250-species colloid (31,250 channels), 4,096 realizations, t_final 0.05; "
                         "hetero: the 64-model heterogeneous batch (161,120 realizations, FP32, scores on) in "
                         "one batched call; "
                         "large: toy_system_gen --toy network of 200,000 species x 1,000,000 channels, 4,096 "
                         "realizations, t_final 1e-6, default mode (AUTO: the plain form for this model)")
    ap.add_argument("--mode", default="auto", choices=["auto", "persistent", "plain"],
                    help="engine form (auto: the library's choice; persistent with --workload large runs the "
                         "kernel that keeps each realization's state in HBM)")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=4000)
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
    if args.workload == "hetero":
        hetero_main(args, T, models, np, torch)
        return
    if args.workload == "colloid":
        net = models.colloidal(250, seed=44)
        R, tf = (args.realizations if args.realizations != 100000 else 4096), 0.05
    elif args.workload == "large":
        import subprocess
        import tempfile
        gen = os.path.join(ROOT, "tau_leaping_amd", "bin", "toy_system_gen")
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "toy200000x1000000.xml")
            with open(path, "w") as f:
                subprocess.run([gen, "-r", "200000", "-c", "1000000", "--toy", "--seed", "3", "--log-rates", "0.001",
                                "1"], stdout=f, stderr=subprocess.DEVNULL, check=True)
            m = T.Model.from_stochkit(path)
        net = None
        R, tf = (args.realizations if args.realizations != 100000 else 4096), 1e-6
    else:
        net = models.toy_network(128, 256, seed=42)
        R, tf = args.realizations, 0.1
    if net is not None:
        m = T.Model.from_network(net)
    fp64 = args.precision == "fp64"
    kw = dict(seed=1234, precision=args.precision, track=np.arange(4, dtype=np.int32), mode=args.mode)

    log(f"warmup x{args.warmup}")
    for _ in range(args.warmup):
        m.simulate(tf, R, **kw)
    torch.cuda.synchronize()
    log(f"timed x{args.steps}")
    t0 = time.perf_counter()
    res = []
    for _ in range(args.steps):
        res.append(m.simulate(tf, R, **kw))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    steps_total = sum(r.total_steps for r in res)
    kernel_ms = [r.kernel_ms for r in res]
    value = steps_total / wall
    B = bytes_per_leap(m.n_species, m.n_channels, fp64)
    avg_ms = float(np.mean(kernel_ms))
    achieved = B * (steps_total / len(res)) / (avg_ms * 1e-3) / 1e9
    plain_form = res[-1].form == "plain"
    form = "plain form" if plain_form else "persistent engine"
    mode = "default mode" if args.mode == "auto" else f"mode {args.mode}"
    if args.workload == "colloid":
        workload = (f"colloid 250 species x {net.n_channels} channels, {R} realizations, t_final 0.05, eps 0.03, "
                    f"{args.precision}, {form}")
    elif args.workload == "large":
        workload = (f"toy network 200000 species x 1000000 channels (toy_system_gen --toy --seed 3), {R} "
                    f"realizations, t_final 1e-06, eps 0.03, {args.precision}, {mode} ({form})")
    else:
        workload = (f"toy network 128 species x 256 channels, {R} realizations, t_final 0.1, eps 0.03, "
                    f"{args.precision}, {form}")
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
        "config": {"workload": workload, "realizations": R, "species": m.n_species, "channels": m.n_channels,
                   "t_final": tf, "leaps_per_step": steps_total // len(res),
                   "ssa_events_per_step": res[-1].total_ssa_events, "block_size": res[-1].block_size},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(workload),
                     "kernel": "plain stage kernels" if plain_form else "k_persistent", "kernel_ms": avg_ms,
                     "bytes_per_leap": B,
                     "issue": measured_issue("k_persistent<double, 64, 64, true, false, 6, false, false>")
                     if (args.workload == "toy" and fp64 and res[-1].block_size == 64) else None},
    }
    if not args.no_extras and args.workload == "toy":
        line["extras"] = extras_run(T, models, np, m,
                                    dict(n_realizations=R, kw={k: v for k, v in kw.items() if k != "mode"}))
    if not args.no_cpu and args.workload == "toy":
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(net, args.cpu_sample, tf, 1234)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
