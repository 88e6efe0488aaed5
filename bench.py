#!/usr/bin/env python3
"""Benchmark of the SGRLD a-MMSB hot path on one MI355X.

Metric: update_phi node*K updates/s -- nodes whose phi/pi rows are updated,
times K, per second of whole SGRLD iterations -- together with the
time-to-target held-out perplexity of the FP32 precision mode against the
FP64 path.  Graph and state are resident in HBM before any timed region.

Default workload (one GPU): N = 4,000,000, K = 1024, community graph with
mean degree 40 (about 80M edges), held-out ratio 0.1, alpha = 1/K,
eta = (1, 1), epsilon = 1e-7, a = 0.01, b = 1024, c = 0.55, stratified
random node minibatches of 4,096 nodes (m = 4,095 non-link pairs; link
minibatches gather sources until the next would exceed m,
sampler.max-source = 1024), 128 neighbour samples (129 drawn per node),
perplexity every 50 steps.  `value` is the FP32 precision mode; the FP64
path runs the same measurement beside it (`fp64`).

Per precision, from the same initial state and Philox seed:
  * time to target: the reference's loop (perplexity before the first step
    and every 50 steps, mcmc_sampler_stochastic.cc:112-130) for
    --ttt-steps steps; the target is the FP64 run's last perplexity, reached
    at the first checkpoint within 1 % of it;
  * throughput: --warmup untimed steps, then exactly --steps timed steps.
    The sample-ahead ring draws `ring_slots` iterations per refill and the
    perplexity runs every 50 steps, so a short window holds more or fewer
    of them than a long run.  The window starts with a refill; the refill
    and one perplexity evaluation are timed in the same run, and
        ms_per_step = (window - refills_in_window * refill) / steps
                      + refill / ring_slots + perplexity / 50,
    the per-step cost of a long run.  The raw window figures are reported
    too (`window`).
The JSON line carries
  roofline      the update_phi kernel's algorithmic bytes per launch / its
                average launch time (HIP events on the library's stream over
                the timed window), against the 8 TB/s HBM peak
  time_to_target, errors_vs_fp64   FP32 against FP64 from identical seeds
  cpu_baseline  the oracle (C++ restatement of the reference loop) on the
                host cores, on the same workload and state: median over 20
                iterations after 5 warm-up, threaded as the distributed
                reference's OpenMP loops, FP32 (`value`) and FP64; and one
                thread (the reference's sequential build) on a shorter
                sample
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
INTERVAL = 50  # perplexity interval (mcmc_sampler_stochastic.cc:17-26)

WORKLOADS = {
    "n4m-k1024": dict(N=4_000_000, K=1024, deg=40.0, n=128, m=4095, primary="fp32",
                      desc="N=4M K=1024 mean-degree-40 community graph, 4096-node minibatches, "
                           "128 neighbour samples; FP32 precision mode (value) and FP64 beside it"),
    "n1m-k256": dict(N=1_000_000, K=256, deg=50.0, n=128, m=4095, primary="fp64",
                     desc="N=1M K=256 mean-degree-50 community graph, 4096-node minibatches, "
                          "128 neighbour samples; FP64 (value) and FP32 beside it"),
    "n5000-k32": dict(N=5000, K=32, deg=None, n=32, m=500, primary="fp64",
                      desc="N=5000 K=32 planted graph, m=500, 32 neighbour samples; FP64 (value) "
                           "and FP32 beside it"),
}
# one-precision names: (workload, primary precision)
ALIASES = {"n4m-k1024-fp32": ("n4m-k1024", "fp32"), "n4m-k1024-fp64": ("n4m-k1024", "fp64"),
           "n1m-k256-fp64": ("n1m-k256", "fp64"), "n1m-k256-fp32": ("n1m-k256", "fp32"),
           "n5000-k32-fp64": ("n5000-k32", "fp64")}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def resolve_workload(name):
    if name in ALIASES:
        wname, prec = ALIASES[name]
        return dict(WORKLOADS[wname], name=wname), prec
    return dict(WORKLOADS[name], name=name), WORKLOADS[name]["primary"]


def build_graph(w, seed):
    import mcmc_for_ammsb_amd as M
    t = time.time()
    if w["deg"] is None:
        g = M.generate("planted", N=w["N"], K=w["K"], seed=seed)
    else:
        g = M.generate("community", N=w["N"], K=w["K"], mean_degree=w["deg"], seed=seed)
    s = M.make_split(g, 0.1, 42)
    log(f"graph N={g.N} E={g.num_edges} held-out={len(s.held_out_labels)} "
        f"({time.time() - t:.1f}s)")
    return g, s


def make_options(w, precision):
    import mcmc_for_ammsb_amd as M
    return M.Options(K=w["K"], num_node_sample=w["n"], mini_batch_size=w["m"],
                     sampler_max_source=1024 if w["deg"] is not None else 1,
                     a=0.01, b=1024.0, c=0.55, epsilon=1e-7, eta0=1.0, eta1=1.0,
                     interval=INTERVAL, random_seed=42, precision=precision)


def host_info():
    """Host cores and CPU model the CPU baseline ran on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = None
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model}


def trajectory(smp, steps, interval):
    """The reference loop's perplexity schedule (mcmc_sampler_stochastic.cc:
    112-130): perplexity before the first step and after every `interval`
    steps.  Checkpoints (step, wall s including perplexity, wall s excluding
    it, perplexity)."""
    out = []
    t_all, t_ppx = 0.0, 0.0
    done = 0
    while True:
        a = time.perf_counter()
        p = smp.perplexity()
        d = time.perf_counter() - a
        t_ppx += d
        t_all += d
        out.append((done + 1, t_all, t_all - t_ppx, p))
        if done >= steps:
            break
        a = time.perf_counter()
        smp.step(interval, sync=True)
        t_all += time.perf_counter() - a
        done += interval
    return out


def tuning(smp, name):
    import ctypes
    from mcmc_for_ammsb_amd import _lib
    v = ctypes.c_int64()
    _lib.check(_lib.lib.ammsb_get_tuning(smp._h, name.encode(), ctypes.byref(v)))
    return v.value


def get_beta(smp):
    import ctypes
    from mcmc_for_ammsb_amd import _lib
    beta = np.empty(smp.info["K"])
    _lib.check(_lib.lib.ammsb_get_state(smp._h, None, None, _lib.ptr(beta, ctypes.c_double)))
    return beta


def measure(smp, args, sample_nodes):
    """Time to target and throughput of one context (see the module
    docstring).  Returns a dict; the context is left after the last step."""
    inf = smp.info
    K, nnb = inf["K"], inf["num_node_sample"] + 1
    fsize = 8 if inf["precision"] == 0 else 4
    seed = smp.options.random_seed
    res = {}
    # device warm-up on the run's own draws, then the initial state again
    smp.init(seed)
    smp.step(2)
    smp.init(seed)
    smp.set_counters(1, 1)
    if args.ttt_steps > 0:
        res["trajectory"] = trajectory(smp, args.ttt_steps, INTERVAL)
        res["rows"] = smp.get_rows(sample_nodes).astype(np.float64)
        res["beta"] = get_beta(smp)
    # ---- throughput
    smp.step(args.warmup, sync=True)
    slots = tuning(smp, "ring_slots")
    step0, avg0 = smp.info["step_count"], smp.info["average_count"]
    smp.enable_timers(2)  # HIP events around update_phi only
    tb = smp.timers()
    n0 = smp.totals()[0]
    smp.set_counters(step0, avg0)  # the window starts with a ring refill
    t0 = time.perf_counter()
    smp.step(args.steps, sync=True)  # ends with a stream synchronisation
    t_window = time.perf_counter() - t0
    n_window = smp.totals()[0] - n0
    ta = smp.timers()
    smp.enable_timers(False)
    phi_s = ta["update_phi"][0] - tb["update_phi"][0]
    phi_n = ta["update_phi"][1] - tb["update_phi"][1]
    refills_in_window = math.ceil(args.steps / slots)
    # the refill: two ring periods with every phase timed, starting at a refill
    smp.enable_timers(True)
    tb = smp.timers()
    st = smp.info
    smp.set_counters(st["step_count"], st["average_count"])
    smp.step(2 * slots, sync=True)
    ta = smp.timers()
    smp.enable_timers(False)
    phases = {k: (ta[k][0] - tb[k][0], ta[k][1] - tb[k][1]) for k in ta}
    fills = max(phases["sample_mini_batch"][1], 1)
    t_fill = (phases["sample_mini_batch"][0] + phases["sample_neighbor_nodes"][0]) / fills
    # the held-out perplexity: median of three evaluations
    ppx_times = []
    for _ in range(3):
        a = time.perf_counter()
        smp.perplexity()
        ppx_times.append(time.perf_counter() - a)
    t_ppx = float(np.median(ppx_times))
    # a contiguous read of the resident pi array: the box's streaming read
    # rate, and a read of known size for calibrating HBM counters
    probe_s, probe_bytes = smp.probe_read(5)
    # update_phi's own row gather without its arithmetic: the practical
    # ceiling of this access pattern on the box
    gather_s, gather_bytes = smp.probe_gather(5)
    t_iter = (t_window - refills_in_window * t_fill) / args.steps
    sec_per_step = t_iter + t_fill / slots + t_ppx / INTERVAL
    nodes_per_step = n_window / args.steps
    avg_launch = phi_s / max(phi_n, 1)
    nodes_per_launch = n_window / max(phi_n, 1)
    # algorithmic bytes of update_phi per node: n_nb neighbour rows + own row
    # read + own row written (K each), the n_nb neighbour ids and phi_sum
    bytes_per_node = (nnb + 2) * K * fsize + nnb * 4 + 2 * fsize
    achieved = nodes_per_launch * bytes_per_node / avg_launch / 1e9
    res.update({
        "value": nodes_per_step * K / sec_per_step,
        "ms_per_step": 1e3 * sec_per_step,
        "nodes_per_step": nodes_per_step,
        "components_ms": {"iteration": 1e3 * t_iter, "ring_refill": 1e3 * t_fill,
                          "ring_slots": slots, "perplexity": 1e3 * t_ppx,
                          "perplexity_interval": INTERVAL},
        "window": {"steps": args.steps, "ms_per_step": 1e3 * t_window / args.steps,
                   "ring_refills": refills_in_window, "perplexity_evaluations": 0},
        "update_phi": {"avg_launch_us": 1e6 * avg_launch, "launches": phi_n,
                       "nodes_per_launch": nodes_per_launch,
                       "algorithmic_bytes_per_launch": nodes_per_launch * bytes_per_node,
                       "achieved_GBs": achieved, "frac": achieved / HBM_PEAK_GBS,
                       "kernel_node_k_per_s": n_window * K / phi_s if phi_s > 0 else None},
        "stream_read_GBs": probe_bytes / probe_s / 1e9,
        "gather_probe": {"us": 1e6 * gather_s, "bytes": gather_bytes,
                         "GBs": gather_bytes / gather_s / 1e9,
                         "frac": gather_bytes / gather_s / 1e9 / HBM_PEAK_GBS},
        "phase_us_per_step": {k: 1e6 * sec / (2 * slots) for k, (sec, n) in phases.items()},
        "phase_ticks": {k: n for k, (sec, n) in phases.items()},
    })
    return res


def reach(traj, target):
    for t, a, b, p in traj:
        if abs(p - target) <= 0.01 * target or p <= target:
            return t, a, b
    return None, None, None


def ttt_summary(r64, r32, K):
    """FP32 against FP64 from the same state and seed: time to the FP64 run's
    final perplexity (within 1 %), and the error of FP32 per quantity."""
    t64, t32 = r64["trajectory"], r32["trajectory"]
    target = t64[-1][3]
    a64, a32 = reach(t64, target), reach(t32, target)
    dpi = np.abs(r32["rows"][:, :K] - r64["rows"][:, :K])
    rb = np.abs(r32["beta"] / r64["beta"] - 1)
    rp = [abs(a[3] - b[3]) / b[3] for a, b in zip(t32, t64)]
    return {
        "target_perplexity": target, "steps": t64[-1][0] - 1, "interval": INTERVAL,
        "fp64_s": a64[1], "fp64_s_excl_perplexity": a64[2], "fp64_steps": a64[0],
        "fp32_s": a32[1], "fp32_s_excl_perplexity": a32[2], "fp32_steps": a32[0],
        "speedup": a64[1] / a32[1] if a32[1] else None,
        "speedup_excl_perplexity": a64[2] / a32[2] if a32[2] else None,
    }, {
        "against": "FP64 GPU run from the same state and Philox seed, after the same steps",
        "rows_compared": int(len(dpi)),
        "pi_max_abs": float(dpi.max()), "pi_mean_abs": float(dpi.mean()),
        "phi_sum_max_rel": float(np.abs(r32["rows"][:, K] / r64["rows"][:, K] - 1).max()),
        "beta_max_rel": float(rb.max()), "beta_mean_rel": float(rb.mean()),
        "perplexity_max_rel": float(max(rp)), "perplexity_mean_rel": float(np.mean(rp)),
        "perplexity_final_rel": float(rp[-1]),
    }


def cpu_leg(orc, opts, start_step, threads, warmup, timed, with_ppx):
    """Oracle iterations on the host: `warmup` untimed, then `timed` timed
    one by one; with_ppx: the first warm-up iteration falls on a perplexity
    step and its evaluation is timed.  Median iteration, amortised
    perplexity."""
    K = opts.K
    step = start_step
    if with_ppx:
        step = (step - 1) // INTERVAL * INTERVAL + 1 + INTERVAL  # (step - 1) % 50 == 0
    else:
        while (step - 1) % INTERVAL == 0 or (step - 1) // INTERVAL != (step + warmup + timed - 2) // INTERVAL:
            step += 1
    orc.set_counters(step, 2)
    t_ppx, it_s, phi_s, nodes = None, [], [], []
    for i in range(warmup + timed):
        _, ph = orc.run(1, 0, opts.mini_batch_size, opts.sampler_max_source, INTERVAL, 42,
                        threads=threads)
        if i == 0 and with_ppx:
            t_ppx = float(ph[0])
        if i >= warmup:
            it_s.append(float(ph[1] + ph[2] + ph[3] + ph[4]))
            phi_s.append(float(ph[3]))
            nodes.append(float(ph[5]))
    t_it = float(np.median(it_s))
    sec = t_it + (t_ppx / INTERVAL if t_ppx is not None else 0.0)
    return {
        "value": float(np.median(nodes)) * K / sec, "unit": "node*K/s", "cores": threads,
        "median_iteration_s": t_it, "perplexity_s": t_ppx,
        "update_phi_node_k_per_s": float(np.median(nodes)) * K / float(np.median(phi_s)),
        "sample": f"{warmup} warm-up + {timed} timed full iterations (samplers, update_phi+pi, "
                  f"update_beta) of the same workload from the GPU run's final state, from step "
                  f"{step}; median iteration"
                  + (", plus one perplexity evaluation amortised over 50 steps" if with_ppx
                     else "; perplexity not included"),
    }


def cpu_baseline(g, s, w, pi32, theta, start_step, args, primary):
    """The oracle loop on the host cores, FP32 and FP64, threaded (the
    distributed reference's OpenMP loops) and on one thread; `value` in the
    workload's primary precision, the other beside it."""
    from tests.oracle_lib import Oracle
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    out = {}
    for prec in ("fp32", "fp64"):
        opts = make_options(w, prec)
        t = time.time()
        orc = Oracle(g, s, opts.K, precision=prec, num_node_sample=opts.num_node_sample,
                     a=opts.a, b=opts.b, c=opts.c, epsilon=opts.epsilon, eta0=opts.eta0,
                     eta1=opts.eta1)
        orc.set_state(pi32 if prec == "fp32" else pi32.astype(np.float64), theta)
        log(f"cpu baseline {prec}: oracle ready ({time.time() - t:.1f}s)")
        t = time.time()
        omp = cpu_leg(orc, opts, start_step, threads, args.cpu_warmup, args.cpu_iters, True)
        log(f"  {threads} threads: {omp['value']:.3e} node*K/s ({time.time() - t:.1f}s)")
        t = time.time()
        seq = cpu_leg(orc, opts, start_step, 1, 1, args.cpu_seq_iters, False)
        log(f"  1 thread: {seq['value']:.3e} node*K/s ({time.time() - t:.1f}s)")
        out[prec] = {"threads": omp, "one_thread": seq}
        del orc
    other = "fp64" if primary == "fp32" else "fp32"
    cb = dict(out[primary]["threads"])
    cb.update({"kind": "port", "precision": primary,
               other: out[other]["threads"],
               "one_thread": {"fp32": out["fp32"]["one_thread"],
                              "fp64": out["fp64"]["one_thread"]},
               "host": host_info()})
    return cb


def run_throughput(args, ttt_only=False):
    import mcmc_for_ammsb_amd as M
    w, primary = resolve_workload(args.workload)
    g, s = build_graph(w, args.seed)
    sample_nodes = np.sort(np.random.default_rng(0).choice(g.N, size=min(g.N, 65536),
                                                           replace=False))
    order = ("fp64", "fp32")
    res, keep = {}, None
    for prec in order:
        opts = make_options(w, prec)
        smp = M.MCMCSamplerStochastic(opts, g, s)
        t = time.time()
        res[prec] = measure(smp, args, sample_nodes)
        r = res[prec]
        log(f"{prec}: {r['ms_per_step']:.4f} ms/step amortised ({r['window']['ms_per_step']:.4f} "
            f"in the window), {r['value']:.3e} node*K/s, update_phi "
            f"{r['update_phi']['avg_launch_us']:.1f} us = {r['update_phi']['achieved_GBs']:.0f} GB/s "
            f"({time.time() - t:.1f}s)")
        if "trajectory" in r:
            log(f"  trajectory: " + " ".join(f"{c[0]}:{c[3]:.5f}@{c[1]:.2f}s"
                                            for c in r["trajectory"][::10]))
        if prec == primary and not args.no_cpu_baseline and not ttt_only:
            pi, theta, _ = smp.get_state()
            keep = (pi if pi.dtype == np.float32 else pi.astype(np.float32), theta,
                    smp.info["step_count"])
            del pi
        smp.close()
    ttt, errs = (ttt_summary(res["fp64"], res["fp32"], w["K"])
                 if args.ttt_steps > 0 else (None, None))
    if ttt:
        log(f"time to target {ttt['target_perplexity']:.6f}: FP64 {ttt['fp64_s']:.3f}s "
            f"({ttt['fp64_steps']} steps), FP32 {ttt['fp32_s']:.3f}s ({ttt['fp32_steps']} steps), "
            f"speedup {ttt['speedup']:.2f}")
    if ttt_only:
        print(json.dumps({"metric": "time-to-target held-out perplexity, FP32 vs FP64",
                          "workload": f"{w['name']}: {w['desc']}", "time_to_target": ttt,
                          "errors_vs_fp64": errs,
                          "fp64_checkpoints": res["fp64"].get("trajectory"),
                          "fp32_checkpoints": res["fp32"].get("trajectory")}), flush=True)
        return 0
    p = res[primary]
    other = "fp64" if primary == "fp32" else "fp32"
    up = p["update_phi"]
    traffic = None
    try:
        with open(args.traffic) as f:
            tr = json.load(f)
        entry = tr.get(f"{w['name']}-{primary}")
        if isinstance(entry, dict):
            traffic = entry.get("bytes_per_launch")
    except (OSError, ValueError):
        traffic = None
    out = {
        "metric": "update_phi node*K updates/s",
        "value": p["value"],
        "unit": "node*K/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": p["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if primary == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{w['name']}: {w['desc']}", "N": g.N, "K": w["K"],
                   "edges": g.num_edges, "num_node_sample": w["n"], "mini_batch_size": w["m"],
                   "sampler_max_source": 1024 if w["deg"] is not None else 1,
                   "perplexity_interval": INTERVAL, "held_out": int(len(s.held_out_labels)),
                   "nodes_per_step": p["nodes_per_step"], "precision": primary,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": up["achieved_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": up["frac"], "traffic": traffic,
                     "kernel": "k_update_phi", "avg_launch_us": up["avg_launch_us"],
                     "algorithmic_bytes_per_launch": up["algorithmic_bytes_per_launch"],
                     "nodes_per_launch": up["nodes_per_launch"]},
        "components_ms": p["components_ms"],
        "window": p["window"],
        "stream_read_GBs": p["stream_read_GBs"],
        "gather_probe": p["gather_probe"],
        other: {k: res[other][k] for k in ("value", "ms_per_step", "components_ms", "window",
                                           "update_phi", "stream_read_GBs", "gather_probe")},
        "time_to_target": ttt,
        "errors_vs_fp64": errs,
        "phase_us_per_step": p["phase_us_per_step"],
        "perplexity_last": p["trajectory"][-1][3] if "trajectory" in p else None,
        "cpu_baseline": None,
    }
    rank = int(os.environ.get("RANK", "0"))
    if rank == 0 and keep is not None:
        pi32, theta, step = keep
        out["cpu_baseline"] = cpu_baseline(g, s, w, pi32, theta, step, args, primary)
        del pi32
    if rank == 0:
        print(json.dumps(out), flush=True)
    return 0


def run_diff(args):
    """Differentiation workload: forward minibatch log-likelihood and its
    adjoint over 8,192 source nodes' training edges + 256 non-edges each, on
    an N=1M, K=2048 community graph (mean degree 50), FP64 and FP32; central
    finite differences (h=1e-6, FP64) on 1,000 random coordinates (800 of
    phi, 200 of theta)."""
    import torch
    import mcmc_for_ammsb_amd as M
    import mcmc_for_ammsb_amd.autograd as A
    from mcmc_for_ammsb_amd import _lib
    import ctypes
    N, K, deg, MS, NNL = 1_000_000, 2048, 50.0, 8192, 256
    t = time.time()
    g = M.generate("community", N=N, K=K, mean_degree=deg, seed=args.seed)
    s = M.make_split(g, 0.1, 42)
    rng = np.random.default_rng(7)
    nodes = np.sort(rng.choice(N, size=MS, replace=False)).astype(np.int32)
    ptr, nbr, y = A.make_pairs(g, s, nodes, NNL, seed=1)
    log(f"graph N={g.N} E={g.num_edges}, {MS} sources, {len(nbr)} pairs ({time.time() - t:.1f}s)")
    dev = torch.device("cuda:0")
    n_t = torch.from_numpy(nodes).to(dev)
    p_t = torch.from_numpy(ptr).to(dev)
    b_t = torch.from_numpy(nbr).to(dev)
    y_t = torch.from_numpy(y).to(dev)
    out = {"metric": "minibatch log-likelihood forward + adjoint, pair*K/s",
           "workload": f"N={N} K={K} community graph mean degree {deg:g}, {MS} sources, "
                       f"all training edges + {NNL} non-edges each ({len(nbr)} pairs)",
           "pairs": int(len(nbr)), "sources": MS, "results": {}}
    for prec in ("fp64", "fp32"):
        F = 8 if prec == "fp64" else 4
        opts = M.Options(K=K, num_node_sample=128, precision=prec, random_seed=42)
        smp = M.MCMCSamplerStochastic(opts, g, s)
        smp.init(42)
        if args.ll_fwd_rows >= 0:
            smp.set_tuning(ll_fwd_rows=args.ll_fwd_rows)
        if args.ll_bwd_rows >= 0:
            smp.set_tuning(ll_bwd_rows=args.ll_bwd_rows)
        rows = smp.get_rows(nodes)
        theta = np.empty((K, 2))
        _lib.check(_lib.lib.ammsb_get_state(smp._h, None, _lib.ptr(theta, ctypes.c_double), None))
        dt = torch.float64 if prec == "fp64" else torch.float32
        phi = torch.from_numpy(np.ascontiguousarray(rows[:, :K] * rows[:, K:K + 1])).to(dev, dt)
        th = torch.from_numpy(theta).to(dev)
        for _ in range(2):
            A.loglik_per_source(smp, phi, th, n_t, p_t, b_t, y_t)
            A.loglik_grad(smp, phi, th, n_t, p_t, b_t, y_t)
        torch.cuda.synchronize()
        reps = args.diff_reps
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record()
        for _ in range(reps):
            per = A.loglik_per_source(smp, phi, th, n_t, p_t, b_t, y_t)
        ev[1].record()
        ev[2].record()
        for _ in range(reps):
            gphi, gth = A.loglik_grad(smp, phi, th, n_t, p_t, b_t, y_t)
        ev[3].record()
        torch.cuda.synchronize()
        t_f = ev[0].elapsed_time(ev[1]) / 1e3 / reps
        t_b = ev[2].elapsed_time(ev[3]) / 1e3 / reps
        P = len(nbr)
        fwd_bytes = P * K * F + MS * K * F + P * 5 + (MS + 1) * 8 + MS * 8
        bwd_bytes = fwd_bytes - MS * 8 + MS * K * F + min(MS, 1024) * K * 8
        r = {"fwd_ms": 1e3 * t_f, "bwd_ms": 1e3 * t_b,
             "fwd_pair_k_per_s": P * K / t_f, "bwd_pair_k_per_s": P * K / t_b,
             "fwd_GBs": fwd_bytes / t_f / 1e9, "bwd_GBs": bwd_bytes / t_b / 1e9,
             "fwd_frac_of_hbm_peak": fwd_bytes / t_f / 1e9 / HBM_PEAK_GBS,
             "bwd_frac_of_hbm_peak": bwd_bytes / t_b / 1e9 / HBM_PEAK_GBS,
             "loglik": float(per.sum().item())}
        if prec == "fp64":
            r.update(finite_differences(A, smp, phi, th, n_t, p_t, b_t, y_t, gphi, gth, ptr))
            ref64 = (per.clone(), gphi.clone(), gth.clone())
        else:
            r["fp32_vs_fp64"] = {
                "loglik_rel": float(abs(per.sum().item() / ref64[0].sum().item() - 1)),
                "grad_phi_max_abs_over_max": float((gphi.double() - ref64[1]).abs().max()
                                                   / ref64[1].abs().max()),
                "grad_theta_max_abs_over_max": float((gth - ref64[2]).abs().max()
                                                     / ref64[2].abs().max())}
        out["results"][prec] = r
        log(f"{prec}: fwd {1e3 * t_f:.3f} ms ({r['fwd_GBs']:.0f} GB/s), "
            f"bwd {1e3 * t_b:.3f} ms ({r['bwd_GBs']:.0f} GB/s)")
        del phi, gphi, gth, per
        smp.close()
    print(json.dumps(out), flush=True)
    return 0


def finite_differences(A, smp, phi, th, n_t, p_t, b_t, y_t, gphi, gth, ptr, h=1e-6):
    """Central differences on 800 phi and 200 theta coordinates (FP64)."""
    import torch
    rng = np.random.default_rng(11)
    MS, K = phi.shape
    errs_phi, errs_th = [], []
    gmax = float(gphi.abs().max())
    for _ in range(800):
        q, k = int(rng.integers(MS)), int(rng.integers(K))
        lo, hi = int(ptr[q]), int(ptr[q + 1])
        sub_p = torch.tensor([0, hi - lo], dtype=torch.int64, device=phi.device)
        vals = []
        for sg in (1, -1):
            ph = phi[q:q + 1].clone()
            ph[0, k] += sg * h
            vals.append(A.loglik_per_source(smp, ph, th, n_t[q:q + 1], sub_p, b_t[lo:hi],
                                            y_t[lo:hi])[0].item())
        fd = (vals[0] - vals[1]) / (2 * h)
        errs_phi.append(abs(fd - gphi[q, k].item()) / (abs(gphi[q, k].item()) + 1e-3 * gmax))
    tmax = float(gth.abs().max())
    for _ in range(200):
        k, i = int(rng.integers(K)), int(rng.integers(2))
        vals = []
        for sg in (1, -1):
            t2 = th.clone()
            t2[k, i] += sg * h
            vals.append(A.loglik_per_source(smp, phi, t2, n_t, p_t, b_t, y_t).sum().item())
        fd = (vals[0] - vals[1]) / (2 * h)
        errs_th.append(abs(fd - gth[k, i].item()) / (abs(gth[k, i].item()) + 1e-3 * tmax))
    return {"fd_coordinates": 1000, "fd_h": h,
            "fd_phi_max_rel_err": float(max(errs_phi)),
            "fd_phi_median_rel_err": float(np.median(errs_phi)),
            "fd_theta_max_rel_err": float(max(errs_th)),
            "fd_theta_median_rel_err": float(np.median(errs_th))}


def batch_workload(n_chains, seed):
    """Chains of the batching workload: N log-uniform in [10k, 500k], K
    cycling through {16, 64, 256, 512}, mean degree uniform in [20, 80],
    per-chain seeds; minibatches of up to 1,024 nodes (m = 1023 non-link
    pairs, link minibatches gathering sources up to m), 32 neighbour
    samples, FP32, a=0.01, b=1024, c=0.55."""
    import mcmc_for_ammsb_amd as M
    rng = np.random.default_rng(seed)
    ks = (16, 64, 256, 512)
    chains = []
    for i in range(n_chains):
        N = int(round(np.exp(rng.uniform(np.log(10_000), np.log(500_000)))))
        chains.append(dict(N=N, K=ks[i % len(ks)], deg=float(rng.uniform(20.0, 80.0)),
                           seed=int(1000 + i)))
    graphs, splits, opts = [], [], []
    t = time.time()
    for c in chains:
        g = M.generate("community", N=c["N"], K=c["K"], mean_degree=c["deg"], seed=c["seed"])
        graphs.append(g)
        splits.append(M.make_split(g, 0.1, c["seed"]))
        opts.append(M.Options(K=c["K"], num_node_sample=32, mini_batch_size=1023,
                              sampler_max_source=1024, a=0.01, b=1024.0, c=0.55,
                              random_seed=c["seed"], precision="fp32"))
    log(f"{n_chains} chains, N {min(c['N'] for c in chains)}..{max(c['N'] for c in chains)}, "
        f"{sum(g.num_edges for g in graphs)} edges ({time.time() - t:.1f}s)")
    return chains, graphs, splits, opts


def run_operator(args):
    """The injected-draw operator ammsb_update_phi (update_phi + update_pi of
    given nodes with given neighbour lists and noise) called with host
    buffers: the ids and the noise cross PCIe on every call, and the call
    checks its ids on the host and allocates its device copies.  Reported
    next to the resident kernel rate; never the headline value."""
    import mcmc_for_ammsb_amd as M
    w, prec = resolve_workload(args.workload)
    g, s = build_graph(w, args.seed)
    opts = make_options(w, prec)
    smp = M.MCMCSamplerStochastic(opts, g, s)
    smp.init(opts.random_seed)
    inf = smp.info
    N, K, nnb = inf["N"], inf["K"], inf["num_node_sample"] + 1
    dt = np.float64 if opts.precision == "fp64" else np.float32
    n_nodes, sets = 4096, 8
    rng = np.random.default_rng(args.seed)
    calls = []
    for _ in range(sets):
        nodes = rng.choice(N, n_nodes, replace=False).astype(np.int32)
        nbrs = rng.integers(0, N, size=(n_nodes, nnb), dtype=np.int32)
        noise = rng.standard_normal((n_nodes, K)).astype(dt)
        calls.append((nodes, nbrs, noise))
    out = {"metric": "injected update_phi node*K updates/s, host buffers (PCIe-inclusive)",
           "unit": "node*K/s", "workload": w["desc"], "nodes_per_call": n_nodes,
           "neighbours_per_node": nnb, "dtype": "f64" if dt is np.float64 else "f32"}
    for label, with_noise in (("host_noise", True), ("device_noise", False)):
        for i in range(args.warmup):
            nodes, nbrs, noise = calls[i % sets]
            smp.update_phi(nodes, nbrs, noise if with_noise else None, 0.01)
        times = []
        for i in range(args.steps):
            nodes, nbrs, noise = calls[i % sets]
            t = time.perf_counter()
            smp.update_phi(nodes, nbrs, noise if with_noise else None, 0.01)
            times.append(time.perf_counter() - t)
        sec = float(np.median(times))
        h2d = n_nodes * 4 + n_nodes * nnb * 4 + (n_nodes * K * dt().itemsize if with_noise else 0)
        out[label] = {"ms_per_call": sec * 1e3, "value": n_nodes * K / sec,
                      "host_to_device_bytes": h2d}
        log(f"{label}: {sec * 1e3:.3f} ms per call, {n_nodes * K / sec:.3e} node*K/s, "
            f"{h2d / 1e6:.1f} MB host to device")
    smp.close()
    print(json.dumps(out), flush=True)
    return 0


def run_batch(args):
    """Batching workload: aggregate update_phi node*K/s of all chains, for
    the batched path (hipGraph replay, autotuned shapes) and for the same
    chains as separate contexts on separate streams."""
    import mcmc_for_ammsb_amd as M
    chains, graphs, splits, opts = batch_workload(args.chains, args.seed)
    Ks = np.array([c["K"] for c in chains], dtype=np.float64)
    out = {"metric": "batched chains: update_phi node*K updates/s, all chains",
           "unit": "node*K/s", "n_chains": args.chains, "steps": args.steps,
           "warmup": args.warmup, "dtype": "f32", "data": "synthetic",
           "workload": "N log-uniform [10k,500k], K in {16,64,256,512}, mean degree U[20,80], "
                       "m=1023, 32 neighbour samples, FP32",
           "chains": [{"N": c["N"], "K": c["K"], "mean_degree": round(c["deg"], 2)}
                      for c in chains]}

    def totals(views):
        return np.array([v.totals()[0] for v in views], dtype=np.float64)

    b = M.MCMCBatch(opts, graphs, splits)
    b.init()
    b.step(args.warmup)
    if not args.no_autotune:
        t = time.time()
        shapes = b.autotune(reps=3)
        out["autotune_s"] = time.time() - t
        out["shapes"] = [{k: (round(v, 2) if k == "us" else v) for k, v in sh.items()}
                         for sh in shapes]
    views = [b.chain(i) for i in range(len(b))]
    res = {}
    # "graph" is the default configuration: hipGraph replay, refills
    # overlapped on a second stream
    for mode, graph, in_order in (("graph", True, False), ("graph_refill_in_order", True, True),
                                  ("direct", False, False)):
        b.use_graph(graph)
        b.refill_in_order(in_order)
        b.step(2)
        n0 = totals(views)
        t0 = time.perf_counter()
        b.step(args.steps)  # ends with the error check (a stream sync)
        el = time.perf_counter() - t0
        nk = float(((totals(views) - n0) * Ks).sum())
        res[mode] = {"value": nk / el, "ms_per_step": 1e3 * el / args.steps}
        log(f"batched ({mode}): {1e3 * el / args.steps:.3f} ms/step, {nk / el:.3e} node*K/s")
    out["ppx_after"] = [round(float(x), 5) for x in b.perplexity()]
    b.close()
    if not args.no_separate:
        ctxs = []
        for o, g, s in zip(opts, graphs, splits):
            smp = M.MCMCSamplerStochastic(o, g, s)
            smp.init()
            ctxs.append(smp)
        for smp in ctxs:
            smp.step(args.warmup, sync=False)
        for smp in ctxs:
            smp.step(0, sync=True)
        n0 = totals(ctxs)
        t0 = time.perf_counter()
        for smp in ctxs:
            smp.step(args.steps, sync=False)
        for smp in ctxs:
            smp.step(0, sync=True)
        el = time.perf_counter() - t0
        nk = float(((totals(ctxs) - n0) * Ks).sum())
        res["separate_contexts"] = {"value": nk / el, "ms_per_step": 1e3 * el / args.steps}
        log(f"separate contexts: {1e3 * el / args.steps:.3f} ms/step, {nk / el:.3e} node*K/s")
        for smp in ctxs:
            smp.close()
    out["results"] = res
    out["value"] = res["graph"]["value"]
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="n4m-k1024", choices=sorted(WORKLOADS) + sorted(ALIASES))
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--ttt-steps", type=int, default=3000,
                    help="steps of the time-to-target trajectory (0: none)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-warmup", type=int, default=5)
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--cpu-seq-iters", type=int, default=3)
    ap.add_argument("--traffic", default=os.path.join(HERE, "profiles", "update_phi_traffic.json"))
    ap.add_argument("--mode", default="throughput",
                    choices=["throughput", "ttt", "diff", "batch", "operator"])
    ap.add_argument("--chains", type=int, default=64)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-separate", action="store_true")
    ap.add_argument("--diff-reps", type=int, default=10)
    ap.add_argument("--ll-fwd-rows", type=int, default=-1,
                    help="diff mode: 1 whole rows per wavefront, 0 the K-chunked forward")
    ap.add_argument("--ll-bwd-rows", type=int, default=-1,
                    help="diff mode: 1 whole rows per wavefront, 0 the K-chunked adjoint")
    args = ap.parse_args()
    if args.gpus != 1:
        log("bench.py measures one GPU: the library has no multi-GPU path "
            "(one MI355X per process); run with --gpus 1")
        return 2
    if args.steps < 1 or args.warmup < 0:
        log("--steps must be >= 1 and --warmup >= 0")
        return 2
    if args.mode == "ttt":
        return run_throughput(args, ttt_only=True)
    if args.mode == "diff":
        return run_diff(args)
    if args.mode == "batch":
        return run_batch(args)
    if args.mode == "operator":
        return run_operator(args)
    return run_throughput(args)


if __name__ == "__main__":
    sys.exit(main())
