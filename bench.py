#!/usr/bin/env python3
"""Benchmark of the SGRLD a-MMSB hot path on one MI355X.

Metric: update_phi node*K updates/s -- nodes whose phi/pi rows are updated,
times K, per second of whole SGRLD iterations (minibatch and neighbour
sampling, update_phi, update_pi, update_beta, and the held-out perplexity
at the reference's interval of 50 steps), with graph and state resident in
HBM before the timed region starts.

Default workload (one GPU): N=1,000,000, K=256, FP64, community graph with
mean degree 50 (~25M edges), held-out ratio 0.1, alpha=1/K, eta=(1,1),
epsilon=1e-7, a=0.01, b=1024, c=0.55, stratified random node minibatches of
4,096 nodes (m=4095 non-link pairs; link minibatches gather sources until
the next one would exceed m), 128 neighbour samples (129 drawn per node).

The JSON line also carries
  roofline      achieved HBM GB/s of the update_phi kernel (algorithmic
                bytes per launch / its average launch time measured with HIP
                events on the library's stream over the timed region)
  cpu_baseline  the CPU oracle (C++ restatement of the reference loop) on
                the host cores, on a bounded sample of the same workload
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (N, K, mean_degree, n, m, precision)
    "n1m-k256-fp64": dict(N=1_000_000, K=256, deg=50.0, n=128, m=4095, precision="fp64",
               desc="N=1M K=256 FP64 mean-degree-50 community graph, 4096-node minibatch, "
                    "128 neighbour samples"),
    "n4m-k1024-fp32": dict(N=4_000_000, K=1024, deg=40.0, n=128, m=4095, precision="fp32",
               desc="N=4M K=1024 FP32 precision mode, mean-degree-40 community graph, "
                    "4096-node minibatch, 128 neighbour samples"),
    "n4m-k1024-fp64": dict(N=4_000_000, K=1024, deg=40.0, n=128, m=4095, precision="fp64",
                    desc="N=4M K=1024 FP64, mean-degree-40 community graph, 4096-node "
                         "minibatch, 128 neighbour samples"),
    "n5000-k32-fp64": dict(N=5000, K=32, deg=None, n=32, m=500, precision="fp64",
               desc="N=5000 K=32 FP64 planted graph, m=500, 32 neighbour samples"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(w, seed):
    import mcmc_for_ammsb_amd as M
    t = time.time()
    if w["deg"] is None:
        g = M.generate("planted", N=w["N"], K=w["K"], seed=seed)
    else:
        g = M.generate("community", N=w["N"], K=w["K"], mean_degree=w["deg"], seed=seed)
    s = M.make_split(g, 0.1, 42)
    log(f"graph N={g.N} E={g.num_edges} held-out={len(s.held_out_labels)} "
        f"({time.time() - t:.1f}s)")
    opts = M.Options(K=w["K"], num_node_sample=w["n"], mini_batch_size=w["m"],
                     sampler_max_source=1024 if w["deg"] is not None else 1,
                     a=0.01, b=1024.0, c=0.55, epsilon=1e-7, eta0=1.0, eta1=1.0,
                     interval=50, random_seed=42, precision=w["precision"])
    return g, s, opts


def cpu_baseline(g, s, opts, pi, theta, start_step, budget_s, threads):
    """The oracle loop on the host cores (bounded sample, no perplexity)."""
    from tests.oracle_lib import Oracle
    K = opts.K
    orc = Oracle(g, s, K, precision=opts.precision, num_node_sample=opts.num_node_sample,
                 a=opts.a, b=opts.b, c=opts.c, epsilon=opts.epsilon, eta0=opts.eta0,
                 eta1=opts.eta1)
    orc.set_state(pi, theta)
    step = start_step if (start_step - 1) % 50 != 0 else start_step + 1
    orc.set_counters(step, 2)
    done, nodes, secs, phi_secs = 0, 0.0, 0.0, 0.0
    t0 = time.time()
    while time.time() - t0 < budget_s and done < 200:
        ta = time.time()
        _, ph = orc.run(1, 0, opts.mini_batch_size, opts.sampler_max_source,
                        10 ** 9, 42, threads=threads)
        secs += time.time() - ta
        phi_secs += ph[3]
        nodes += ph[5]
        done += 1
    return {
        "value": nodes * K / secs,
        "unit": "node*K/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{done} full iterations (samplers, update_phi+pi, update_beta) of the same "
                  f"workload from the GPU run's state at step {step}, C++ oracle with "
                  f"{threads} threads for the node loop",
        "update_phi_node_k_per_s": nodes * K / phi_secs if phi_secs > 0 else None,
        "seconds": secs,
    }


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
    """Checkpoints (step, wall s including perplexity, wall s excluding it, perplexity),
    perplexity before the first step and then every `interval` steps
    (mcmc_sampler_stochastic.cc:112-130)."""
    out = []
    t_all, t_ppx = 0.0, 0.0
    done = 0
    while True:
        a = time.perf_counter()
        p = smp.perplexity()
        t_ppx += time.perf_counter() - a
        t_all += time.perf_counter() - a
        out.append((done + 1, t_all, t_all - t_ppx, p))
        if done >= steps:
            break
        a = time.perf_counter()
        smp.step(interval, sync=True)
        t_all += time.perf_counter() - a
        done += interval
    return out


def run_ttt(args):
    """Time-to-target of the FP32 precision mode against the FP64 path on the
    same workload, initial state and Philox seed; error per quantity."""
    import mcmc_for_ammsb_amd as M
    w = dict(WORKLOADS[args.workload])
    w["precision"] = "fp64"
    g, s, opts = build_workload(w, args.seed)
    res = {}
    states = {}
    sample_nodes = np.sort(np.random.default_rng(0).choice(g.N, size=min(g.N, 65536), replace=False))
    for prec in ("fp64", "fp32"):
        o = M.Options(**{**opts.__dict__, "precision": prec, "remains": []})
        smp = M.MCMCSamplerStochastic(o, g, s)
        smp.init(o.random_seed)
        smp.step(2)  # device warm-up on the same draws as the run (state reset below)
        smp.init(o.random_seed)
        smp.set_counters(1, 1)
        traj = trajectory(smp, args.ttt_steps, args.ttt_interval)
        rows = smp.get_rows(sample_nodes).astype(np.float64)  # 65,536 random rows
        _, _, beta = smp.get_state() if g.N * (opts.K + 1) <= 50_000_000 else (None, None, None)
        if beta is None:
            beta = np.empty(opts.K)
            import ctypes
            from mcmc_for_ammsb_amd import _lib
            _lib.check(_lib.lib.ammsb_get_state(smp._h, None, None, _lib.ptr(beta, ctypes.c_double)))
        states[prec] = (rows, beta)
        res[prec] = traj
        log(f"{prec}: " + " ".join(f"{t}:{p:.5f}@{a:.2f}s" for t, a, _, p in traj))
        smp.close()
    target = res["fp64"][-1][3]
    def reach(traj):
        for t, a, b, p in traj:
            if abs(p - target) <= 0.01 * target or p <= target:
                return t, a, b
        return None, None, None
    t64 = reach(res["fp64"])
    t32 = reach(res["fp32"])
    pi64, b64 = states["fp64"]
    pi32, b32 = states["fp32"]
    K = pi64.shape[1] - 1
    dpi = np.abs(pi32[:, :K] - pi64[:, :K])
    rel_ppx = [abs(a[3] - b[3]) / b[3] for a, b in zip(res["fp32"], res["fp64"])]
    out = {
        "metric": "time-to-target held-out perplexity, FP32 precision mode vs FP64",
        "workload": f"{args.workload}: {w['desc']} (FP32 and FP64 from the same state and seed)",
        "steps": args.ttt_steps, "interval": args.ttt_interval,
        "target_perplexity": target,
        "fp64_time_to_target_s": t64[1], "fp64_time_to_target_s_excl_perplexity": t64[2],
        "fp64_steps_to_target": t64[0],
        "fp32_time_to_target_s": t32[1], "fp32_time_to_target_s_excl_perplexity": t32[2],
        "fp32_steps_to_target": t32[0],
        "speedup": (t64[1] / t32[1]) if t32[1] else None,
        "speedup_excl_perplexity": (t64[2] / t32[2]) if t32[2] else None,
        "errors_vs_fp64": {
            "rows_compared": int(len(sample_nodes)),
            "pi_max_abs": float(dpi.max()), "pi_mean_abs": float(dpi.mean()),
            "pi_max_rel": float((dpi / np.maximum(pi64[:, :K], 1e-30)).max()),
            "phi_sum_max_rel": float(np.abs(pi32[:, K] / pi64[:, K] - 1).max()),
            "beta_max_rel": float(np.abs(b32 / b64 - 1).max()),
            "beta_mean_rel": float(np.abs(b32 / b64 - 1).mean()),
            "perplexity_max_rel": float(max(rel_ppx)), "perplexity_mean_rel": float(np.mean(rel_ppx)),
            "perplexity_final_rel": rel_ppx[-1]},
        "fp64_checkpoints": res["fp64"], "fp32_checkpoints": res["fp32"],
    }
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
    w = WORKLOADS[args.workload]
    g, s, opts = build_workload(w, args.seed)
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
    for mode, graph, in_order in (("graph", True, True), ("graph_refill_overlapped", True, False),
                                  ("direct", False, True)):
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
    ap.add_argument("--workload", default="n1m-k256-fp64", choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU sample")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--phi-waves", type=int, default=0, help="update_phi wavefronts per node")
    ap.add_argument("--phi-unroll", type=int, default=0, help="update_phi rows in flight")
    ap.add_argument("--timer-mode", default="phi", choices=["phi", "all", "off"],
                    help="HIP events in the timed region: update_phi only, every phase, none")
    ap.add_argument("--async-ring", type=int, default=-1,
                    help="1: samplers draw ahead on a second stream; 0: on the main stream")
    ap.add_argument("--phi-split", type=int, default=-1,
                    help="1: update_phi gather + finish kernels; 0: fused step + scatter")
    ap.add_argument("--traffic", default=os.path.join(HERE, "profiles", "update_phi_traffic.json"))
    ap.add_argument("--mode", default="throughput", choices=["throughput", "ttt", "diff", "batch", "operator"])
    ap.add_argument("--chains", type=int, default=64)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-separate", action="store_true")
    ap.add_argument("--diff-reps", type=int, default=10)
    ap.add_argument("--ll-fwd-rows", type=int, default=-1,
                    help="diff mode: 1 whole rows per wavefront, 0 the K-chunked forward")
    ap.add_argument("--ll-bwd-rows", type=int, default=-1,
                    help="diff mode: 1 whole rows per wavefront, 0 the K-chunked adjoint")
    ap.add_argument("--ttt-steps", type=int, default=3000)
    ap.add_argument("--ttt-interval", type=int, default=100)
    args = ap.parse_args()
    if args.gpus != 1:
        log("bench.py measures one GPU: the library has no multi-GPU path "
            "(one MI355X per process); run with --gpus 1")
        return 2
    rank = int(os.environ.get("RANK", "0"))
    if args.mode == "ttt":
        return run_ttt(args)
    if args.mode == "diff":
        return run_diff(args)
    if args.mode == "batch":
        return run_batch(args)
    if args.mode == "operator":
        return run_operator(args)

    import mcmc_for_ammsb_amd as M
    w = WORKLOADS[args.workload]
    g, s, opts = build_workload(w, args.seed)
    smp = M.MCMCSamplerStochastic(opts, g, s)
    if args.phi_waves or args.phi_unroll:
        cur = smp.get_tuning()
        smp.set_tuning(phi_waves=args.phi_waves or cur["phi_waves"],
                       phi_unroll=args.phi_unroll or cur["phi_unroll"])
    if args.phi_split >= 0:
        smp.set_tuning(phi_split=args.phi_split)
    if args.async_ring >= 0:
        smp.set_tuning(async_ring=args.async_ring)
    inf = smp.info
    smp.init(opts.random_seed)
    K, nnb = inf["K"], inf["num_node_sample"] + 1
    fsize = 8 if opts.precision == "fp64" else 4

    smp.run(max_steps=args.warmup, print_log=False)
    # timed region: HIP events around update_phi only (for the roofline);
    # the per-phase breakdown comes from a separate pass below, because an
    # event pair around every phase adds gaps to the stream
    smp.enable_timers(2 if args.timer_mode == "phi" else (1 if args.timer_mode == "all" else 0))
    t_before = smp.timers()
    n0, e0 = smp.totals()
    t0 = time.perf_counter()
    ckpt = smp.run(max_steps=args.steps, print_log=False)  # ends with a stream sync
    t1 = time.perf_counter()
    n1, e1 = smp.totals()
    t_after = smp.timers()
    smp.enable_timers(False)
    elapsed = t1 - t0
    nodes = n1 - n0
    value = nodes * K / elapsed

    ph = {k: (t_after[k][0] - t_before[k][0], t_after[k][1] - t_before[k][1]) for k in t_after}
    bd_steps = min(args.steps, 100)
    smp.enable_timers(True)
    tb0 = smp.timers()
    nb0 = smp.totals()[0]
    smp.run(max_steps=bd_steps, print_log=False)
    nb1 = smp.totals()[0]
    tb1 = smp.timers()
    smp.enable_timers(False)
    ph_all = {k: (tb1[k][0] - tb0[k][0], tb1[k][1] - tb0[k][1]) for k in tb1}
    phi_nodes = nodes
    if args.timer_mode == "off":
        ph, phi_nodes = ph_all, nb1 - nb0  # roofline from the breakdown pass
    phi_s, phi_n = ph["update_phi"]
    avg_launch = phi_s / max(phi_n, 1)
    nodes_per_launch = phi_nodes / max(phi_n, 1)
    # algorithmic bytes of one (node, k) update: n_nb neighbour rows + own row
    # read + own row written; per node the n_nb neighbour ids and phi_sum
    bytes_per_node = (nnb + 2) * K * fsize + nnb * 4 + 2 * fsize
    bytes_per_launch = nodes_per_launch * bytes_per_node
    achieved = bytes_per_launch / avg_launch / 1e9
    traffic = None
    if os.path.exists(args.traffic):
        try:
            with open(args.traffic) as f:
                tr = json.load(f)
            entry = tr.get(args.workload)
            if isinstance(entry, dict):
                traffic = entry.get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    probe_s, probe_bytes = smp.probe_read(5)  # contiguous read of the resident state
    log(f"contiguous read of {probe_bytes / 1e9:.2f} GB: {probe_bytes / probe_s / 1e9:.0f} GB/s")
    log(f"phase breakdown ({bd_steps} further steps, every phase timed)")
    log("phase            total_s   ticks   per_tick_us")
    for k, (sec, n) in ph_all.items():
        log(f"{k:<16} {sec:9.4f} {n:7d} {1e6 * sec / max(n, 1):12.2f}")
    log(f"checkpoints (step, perplexity): {ckpt}")

    out = {
        "metric": "update_phi node*K updates/s",
        "value": value,
        "unit": "node*K/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if opts.precision == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {w['desc']}", "N": g.N, "K": K,
                   "edges": g.num_edges, "num_node_sample": inf["num_node_sample"],
                   "mini_batch_size": inf["mini_batch_size"],
                   "sampler_max_source": opts.sampler_max_source, "perplexity_interval": 50,
                   "held_out": inf["held_out_count"], "nodes_per_step": nodes / args.steps,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_update_phi", "avg_launch_us": 1e6 * avg_launch,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "nodes_per_launch": nodes_per_launch, "launch_shape": smp.get_tuning()},
        "update_phi_kernel_node_k_per_s": phi_nodes * K / phi_s if phi_s > 0 else None,
        "timed_region_events": args.timer_mode,
        "phase_us_per_step": {k: 1e6 * sec / bd_steps for k, (sec, n) in ph_all.items()},
        "perplexity_last": ckpt[-1][1] if ckpt else None,
        "stream_read_GBs": probe_bytes / probe_s / 1e9,
        "cpu_baseline": None,
    }
    if rank == 0 and not args.no_cpu_baseline:
        pi, theta, _ = smp.get_state()
        step = smp.info["step_count"]
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cb = cpu_baseline(g, s, opts, pi, theta, step, args.cpu_budget, threads)
        # the reference's SEQ build runs the loop on one thread
        seq = cpu_baseline(g, s, opts, pi, theta, step, args.cpu_budget / 2, 1)
        cb["single_thread"] = {k: seq[k] for k in ("value", "unit", "cores", "sample",
                                                   "update_phi_node_k_per_s", "seconds")}
        cb["host"] = host_info()
        out["cpu_baseline"] = cb
        del pi
    smp.close()
    if rank == 0:
        print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
