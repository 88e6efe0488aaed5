"""Parity of the HIP path against the CPU oracle (GPU tests).

Each test calls the device kernels through the C-ABI (libammsb.so) and the
oracle (oracle/liboracle.so) on the same seeded inputs.  Tolerances:
  * sampled indices (minibatch edges, flags, neighbours): bit-exact
  * FP64 operators with injected draws: 1e-12 relative, with an absolute
    floor of 1e-12 x the row's largest |value| (summation order differs)
  * perplexity trajectory: within 1% of the oracle at every checkpoint
"""
import numpy as np
import pytest

import mcmc_for_ammsb_amd as M
from tests import oracle_lib
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def _state(N, K, seed, dtype=np.float64):
    rng = np.random.default_rng(seed)
    phi = rng.gamma(1.0, 1.0, size=(N, K))
    s = phi.sum(1)
    pi = np.concatenate([phi / s[:, None], s[:, None]], 1).astype(dtype)
    theta = rng.gamma(2.0, 0.5, size=(K, 2))
    return pi, theta


def _close(gpu, ref, tol):
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.max(np.abs(ref), axis=-1, keepdims=True) if ref.ndim > 1 else np.max(np.abs(ref))
    err = np.abs(gpu - ref)
    bound = tol * np.abs(ref) + tol * scale
    return bool(np.all(err <= bound)), float(np.max(err / (np.abs(ref) + 1e-300)))


GRAPHS = {}


def graph(kind, N, K, deg=20.0, seed=3):
    key = (kind, N, K, deg, seed)
    if key not in GRAPHS:
        if kind == "planted":
            g = M.generate("planted", N=N, K=K, seed=seed)
        else:
            g = M.generate("community", N=N, K=K, mean_degree=deg, seed=seed)
        GRAPHS[key] = (g, M.make_split(g, 0.1, 42))
    return GRAPHS[key]


def make_pair(kind, N, K, precision="fp64", n=32, m=None, strategy="stratified-random-node",
              max_source=1, seed=42, deg=20.0, **hyp):
    g, s = graph(kind, N, K, deg)
    m = m if m is not None else max(1, N // 10)
    opts = M.Options(K=K, num_node_sample=n, mini_batch_size=m, strategy=strategy,
                     sampler_max_source=max_source, random_seed=seed, precision=precision, **hyp)
    smp = M.MCMCSamplerStochastic(opts, g, s)
    orc = Oracle(g, s, K, precision=precision, num_node_sample=n, alpha=hyp.get("alpha"),
                 a=hyp.get("a"), b=hyp.get("b", 1024.0), c=hyp.get("c", 0.5),
                 epsilon=hyp.get("epsilon", 1e-7), eta0=hyp.get("eta0", 1.0),
                 eta1=hyp.get("eta1", 1.0))
    return g, s, smp, orc


@pytest.mark.parametrize("kind,N,K", [("planted", 2000, 32), ("community", 20000, 256),
                                      ("community", 8000, 300), ("community", 4000, 7),
                                      ("community", 3000, 1024), ("community", 3000, 1),
                                      ("community", 1500, 2048)])
def test_update_phi_fp64_injected(kind, N, K):
    print(f"update_phi fp64 {kind} N={N} K={K}", flush=True)
    g, s, smp, orc = make_pair(kind, N, K, n=24)
    pi, theta = _state(N, K, 7)
    smp.set_state(pi, theta)
    orc.set_state(pi, theta)
    rng = np.random.default_rng(11)
    nodes = rng.choice(N, size=min(300, N // 4), replace=False).astype(np.int32)
    nnb = 25
    nbrs = np.stack([orc.sample_neighbors(int(v), nnb, 42, 5) & 0x7FFFFFFF for v in nodes])
    nbrs[0, 3] = nodes[0]  # a self entry is skipped (mcmc_sampler_stochastic.cc:275-277)
    noise = rng.standard_normal((len(nodes), K))
    eps_t = 0.0123
    smp.update_phi(nodes, nbrs, noise, eps_t)
    orc.update_phi(nodes, nbrs, noise, eps_t)
    gpi, _, _ = smp.get_state()
    opi, _, _ = orc.get_state()
    ok, rel = _close(gpi[nodes, :K], opi[nodes, :K], TOL64)
    assert ok, rel
    ok, rel = _close(gpi[nodes, K], opi[nodes, K], TOL64)
    assert ok, rel
    untouched = np.setdiff1d(np.arange(N), nodes)
    assert np.array_equal(gpi[untouched], pi[untouched])


@pytest.mark.parametrize("nnb", [25, 301])
@pytest.mark.parametrize("waves,unroll", [(4, 1), (2, 2), (2, 4), (4, 2), (4, 4), (8, 1), (8, 2),
                                          (8, 4)])
def test_update_phi_launch_shapes(waves, unroll, nnb):
    """Every update_phi launch shape (wavefronts per node, rows in flight)
    against the oracle, with neighbour lists that fit the wavefronts' id
    registers (25) and that are staged in LDS (301 > 64 x wavefronts for 2
    and 4 wavefronts)."""
    print(f"update_phi shape waves={waves} unroll={unroll} nnb={nnb}", flush=True)
    N, K = 6000, 256
    g, s, smp, orc = make_pair("community", N, K, n=nnb - 1)
    smp.set_tuning(phi_waves=waves, phi_unroll=unroll)
    pi, theta = _state(N, K, 17)
    smp.set_state(pi, theta)
    orc.set_state(pi, theta)
    rng = np.random.default_rng(19)
    nodes = rng.choice(N, size=150, replace=False).astype(np.int32)
    nbrs = np.stack([orc.sample_neighbors(int(v), nnb, 42, 7) & 0x7FFFFFFF for v in nodes])
    noise = rng.standard_normal((len(nodes), K))
    smp.update_phi(nodes, nbrs, noise, 0.011)
    orc.update_phi(nodes, nbrs, noise, 0.011)
    gpi, _, _ = smp.get_state()
    opi, _, _ = orc.get_state()
    ok, rel = _close(gpi[nodes], opi[nodes], TOL64)
    assert ok, rel


@pytest.mark.parametrize("K", [32, 1024, 4096])
def test_update_phi_fp32_injected(K):
    print(f"update_phi fp32 K={K}", flush=True)
    N = 4000
    g, s, smp, orc = make_pair("community", N, K, precision="fp32", n=24)
    pi, theta = _state(N, K, 8, np.float32)
    smp.set_state(pi, theta)
    orc.set_state(pi, theta)
    rng = np.random.default_rng(12)
    nodes = rng.choice(N, size=200, replace=False).astype(np.int32)
    nbrs = np.stack([orc.sample_neighbors(int(v), 25, 42, 9) & 0x7FFFFFFF for v in nodes])
    noise = rng.standard_normal((len(nodes), K)).astype(np.float32)
    smp.update_phi(nodes, nbrs, noise, 0.01)
    orc.update_phi(nodes, nbrs, noise, 0.01)
    gpi, _, _ = smp.get_state()
    opi, _, _ = orc.get_state()
    # FP32 bulk arithmetic, different summation order: 1e-4 relative
    ok, rel = _close(gpi[nodes], opi[nodes], 1e-4)
    assert ok, rel


@pytest.mark.parametrize("strategy,max_source", [("stratified-random-node-nonlinks", 1),
                                                 ("stratified-random-node-links", 1),
                                                 ("stratified-random-node-links", 8)])
@pytest.mark.parametrize("K", [32, 256, 2048])
def test_update_beta_fp64_injected(strategy, max_source, K):
    print(f"update_beta fp64 {strategy} {max_source} K={K}", flush=True)
    N = 6000
    strat = {"stratified-random-node-nonlinks": 2, "stratified-random-node-links": 1}[strategy]
    g, s, smp, orc = make_pair("community", N, K, n=16, m=600, max_source=max_source)
    pi, theta = _state(N, K, 9)
    smp.set_state(pi, theta)
    orc.set_state(pi, theta)
    mb = orc.sample_minibatch(strat, 600, max_source, 42, 3)
    assert len(mb["y"]) > 0
    noise = np.random.default_rng(3).standard_normal((K, 2))
    smp.update_beta(mb["edges"], mb["y"], mb["scale"], noise, 0.02)
    orc.update_beta(mb["edges"], mb["y"], mb["scale"], noise, 0.02)
    _, gth, gbe = smp.get_state()
    _, oth, obe = orc.get_state()
    ok, rel = _close(gth, oth, TOL64)
    assert ok, rel
    ok, rel = _close(gbe, obe, TOL64)
    assert ok, rel


@pytest.mark.parametrize("K,prec", [(32, "fp64"), (256, "fp64"), (300, "fp64"), (512, "fp64"),
                                    (2048, "fp64"), (64, "fp32"), (1024, "fp32"), (4096, "fp32")])
def test_perplexity_running_average(K, prec):
    print(f"perplexity K={K} {prec}", flush=True)
    N = 5000
    dt = np.float64 if prec == "fp64" else np.float32
    g, s, smp, orc = make_pair("community", N, K, precision=prec)
    tol = TOL64 if prec == "fp64" else 1e-5
    for it in range(3):
        pi, theta = _state(N, K, 20 + it, dt)
        smp.set_state(pi, theta)
        orc.set_state(pi, theta)
        a = smp.perplexity()
        b = orc.perplexity()
        assert abs(a - b) <= tol * abs(b), (it, a, b)
    assert smp.info["average_count"] == 4


@pytest.mark.parametrize("strategy,max_source", [("stratified-random-node", 1),
                                                 ("stratified-random-node-links", 16),
                                                 ("stratified-random-node-nonlinks", 1)])
def test_samplers_bit_exact(strategy, max_source):
    print(f"samplers {strategy} {max_source}", flush=True)
    N, K, n = 20000, 16, 40
    strat = {"stratified-random-node": 0, "stratified-random-node-links": 1,
             "stratified-random-node-nonlinks": 2}[strategy]
    g, s, smp, orc = make_pair("community", N, K, n=n, m=1500, strategy=strategy,
                               max_source=max_source, seed=77)
    flags = set()
    for step in range(1, 13):
        d = smp.debug_sample(step)
        r = orc.sample_minibatch(strat, 1500, max_source, 77, step)
        flags.add(d["flag"])
        assert d["flag"] == r["flag"], step
        assert np.array_equal(d["edges"], r["edges"]), step
        assert d["scale"] == pytest.approx(r["scale"], rel=1e-15)
        ref_nodes = np.unique(r["edges"])
        assert np.array_equal(np.sort(d["nodes"]), ref_nodes), step
        for q in range(0, len(d["nodes"]), max(1, len(d["nodes"]) // 25)):
            v = int(d["nodes"][q])
            assert np.array_equal(d["nbrs"][q], orc.sample_neighbors(v, n + 1, 77, step)), (step, v)
    if strat == 0:
        assert flags == {0, 1}


@pytest.mark.parametrize("strategy,max_source", [("stratified-random-node", 4),
                                                 ("random-edge", 1)])
def test_tiny_graph(strategy, max_source):
    """N=64 (185 edges): the node list fills most of the graph, the
    neighbour draws 9 of 63 nodes.  Draws bit-exact against the oracle, then
    40 iterations keep every row normalised and the perplexity finite."""
    print(f"tiny graph {strategy}", flush=True)
    N, K, n, m = 64, 4, 8, 20
    strat = {"stratified-random-node": 0, "random-edge": 3}[strategy]
    g, s, smp, orc = make_pair("community", N, K, n=n, m=m, strategy=strategy,
                               max_source=max_source, seed=5, deg=6.0)
    mm = smp.info["mini_batch_size"]
    for step in range(1, 9):
        d = smp.debug_sample(step)
        r = orc.sample_minibatch(strat, mm, max_source, 5, step)
        assert d["flag"] == r["flag"], step
        assert np.array_equal(d["edges"], r["edges"]), step
        assert np.array_equal(np.sort(d["nodes"]), np.unique(r["edges"])), step
        for q, v in enumerate(d["nodes"]):
            assert np.array_equal(d["nbrs"][q], orc.sample_neighbors(int(v), n + 1, 5, step))
    smp.init(3)
    smp.step(40)
    pi, theta, beta = smp.get_state()
    assert np.all(np.isfinite(pi)) and np.all(pi[:, K] > 0)
    assert np.allclose(pi[:, :K].sum(1), 1.0, atol=1e-12)
    assert np.all((beta > 0) & (beta < 1))
    assert np.isfinite(smp.perplexity())


def test_run_stops_at_convergence():
    """run() stops once the relative change of the last two held-out
    perplexities is at most convergence_threshold (learner.cc:130-136): with
    a threshold of 1 the loop ends after the second check, one iteration
    past it; with 0 it runs to max_iteration."""
    print("convergence stop", flush=True)
    N, K = 2000, 16
    g, s = graph("community", N, K)
    opts = M.Options(K=K, num_node_sample=8, interval=10, max_iteration=200,
                     convergence_threshold=1.0)
    smp = M.MCMCSamplerStochastic(opts, g, s)
    smp.init(2)
    ck = smp.run(print_log=False)
    assert [t for t, _ in ck] == [1, 11]
    assert smp.info["step_count"] == 12
    smp.close()
    opts = M.Options(K=K, num_node_sample=8, interval=10, max_iteration=45,
                     convergence_threshold=0.0)
    smp = M.MCMCSamplerStochastic(opts, g, s)
    smp.init(2)
    ck = smp.run(print_log=False)
    assert [t for t, _ in ck] == [1, 11, 21, 31, 41]
    assert smp.info["step_count"] == 45
    smp.close()


def test_driver_matches_python_run(tmp_path):
    """The command-line driver (the reference's main: edge-list input,
    split, init, run with its log line) prints the same held-out
    perplexities as the Python front end on the same file and options."""
    import os
    import re
    import subprocess
    print("command-line driver", flush=True)
    driver = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "mcmc_for_ammsb_amd",
                          "ammsb_mcmc")
    g0 = M.generate("community", N=3000, K=8, mean_degree=10.0, seed=4)
    src = np.repeat(np.arange(g0.N), np.diff(g0.row_ptr))
    keep = src < g0.cols
    path = tmp_path / "graph.txt"
    with open(path, "w") as f:
        f.write("# synthetic\n# community graph\n# Nodes: 3000\n# FromNodeId ToNodeId\n")
        for a, b in zip(src[keep], g0.cols[keep]):
            f.write(f"{a}\t{b}\n")
    argv = ["-f", str(path), "--mcmc.input.contiguous", "-K", "8", "-n", "16", "-m", "300",
            "-x", "61", "-i", "20", "--mcmc.seed", "3"]
    r = subprocess.run([driver] + argv, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    pat = r"step count: (\d+) time: [0-9.]+ perplexity for hold out set: ([0-9.eE+-]+)"
    got = [(int(m.group(1)), float(m.group(2))) for m in re.finditer(pat, r.stdout)]
    opts = M.Options.parse(argv)
    g = M.Graph.from_edge_list_file(str(path), contiguous=True)
    smp = M.MCMCSamplerStochastic(opts, g)
    smp.init()
    ref = smp.run(print_log=False)
    smp.close()
    assert [t for t, _ in got] == [t for t, _ in ref] == [1, 21, 41]
    for (_, a), (_, b) in zip(got, ref):
        assert abs(a - b) <= 1e-11, (a, b)
    # the timer table printed at the end of run() (PrintStats,
    # mcmc_sampler_stochastic.cc:185-198): timer.cc's tabular layout with
    # timer.h's widths (name 36, total 12, ticks 8, per tick 14), fixed
    # 3-decimal numbers, the eight phases in the reference's order
    lines = r.stdout.splitlines()
    head = "timer".ljust(37) + "total (s)".rjust(13) + "ticks".rjust(9) + "per tick (us)".rjust(15)
    assert head in lines
    i = lines.index(head)
    names = ["outer", "perplexity", "sample_mini_batch", "nodes_in_mini_batch",
             "sample_neighbor_nodes", "update_phi", "update_pi", "update_beta"]
    row = re.compile(r"^(  [a-z_]+) {1,}( {0,12}\d+\.\d{3}) ( {0,8}\d+) ( {0,14}\d+\.\d{3})$")
    for k, name in enumerate(names):
        text = lines[i + 1 + k]
        assert text[:36] == ("  " + name).ljust(36), text
        if text[36:] == "<unused>":
            continue
        m = row.match(text)
        assert m and len(text) == 36 + 1 + 12 + 1 + 8 + 1 + 14, text
    ticks = {n: lines[i + 1 + k][36:] for k, n in enumerate(names)}
    assert ticks["outer"] != "<unused>" and ticks["update_phi"] != "<unused>"


def test_trajectory_n5000_within_one_percent():
    """The reference's CPU-sized workload (N=5000, K=32, n=32, a=0.01,
    b=1024, c=0.55) from the same initial state and Philox seed: held-out
    perplexity within 1% at every checkpoint of 1000 iterations."""
    print("trajectory N=5000 K=32", flush=True)
    N, K = 5000, 32
    g, s, smp, orc = make_pair("planted", N, K, n=32, m=500, a=0.01, b=1024.0, c=0.55,
                               interval=50, seed=42)
    pi, theta = _state(N, K, 1)
    smp.set_state(pi, theta)
    orc.set_state(pi, theta)
    gres = smp.run(max_steps=1000, print_log=False)
    ores, _ = orc.run(1000, 0, 500, 1, 50, 42)
    assert len(gres) == len(ores) == 20
    for (gs, gp), (os_, op) in zip(gres, ores):
        assert gs == os_
        assert abs(gp - op) <= 0.01 * abs(op), (gs, gp, op)
    print("checkpoints", [(a, round(b, 6)) for a, b in gres[::5]], flush=True)


def test_trajectory_n5000_fp32_within_one_percent():
    """Precision mode on the same workload: the FP32 device run keeps the
    held-out perplexity within 1% of the FP32 oracle (same operation order in
    float, FP64 renormalisation and perplexity sums) and within 1% of the
    FP64 oracle at every checkpoint of 1000 iterations."""
    print("trajectory N=5000 K=32 fp32", flush=True)
    N, K = 5000, 32
    hyp = dict(n=32, m=500, a=0.01, b=1024.0, c=0.55, interval=50, seed=42)
    g, s, smp, orc32 = make_pair("planted", N, K, precision="fp32", **hyp)
    _, _, _, orc64 = make_pair("planted", N, K, precision="fp64", **hyp)
    pi, theta = _state(N, K, 1)
    smp.set_state(pi.astype(np.float32), theta)
    orc32.set_state(pi.astype(np.float32), theta)
    orc64.set_state(pi, theta)
    gres = smp.run(max_steps=1000, print_log=False)
    r32, _ = orc32.run(1000, 0, 500, 1, 50, 42)
    r64, _ = orc64.run(1000, 0, 500, 1, 50, 42)
    assert len(gres) == len(r32) == len(r64) == 20
    for (gs, gp), (s32, p32), (s64, p64) in zip(gres, r32, r64):
        assert gs == s32 == s64
        assert abs(gp - p32) <= 0.01 * abs(p32), (gs, gp, p32)
        assert abs(gp - p64) <= 0.01 * abs(p64), (gs, gp, p64)
    print("checkpoints", [(a, round(b, 6)) for a, b in gres[::5]], flush=True)
    # error of the precision mode against the FP64 oracle after 1000
    # iterations, per quantity (pi, beta, perplexity over all checkpoints)
    gpi, _, gbe = smp.get_state()
    opi, _, obe = orc64.get_state()
    dpi = np.abs(gpi[:, :K].astype(np.float64) - opi[:, :K])
    rb = np.abs(gbe / obe - 1)
    rp = np.array([abs(gp - p64) / p64 for (_, gp), (_, p64) in zip(gres, r64)])
    print(f"fp32 vs fp64 oracle after 1000 iterations: pi max abs {dpi.max():.3e} mean abs "
          f"{dpi.mean():.3e}; beta max rel {rb.max():.3e} mean rel {rb.mean():.3e}; "
          f"perplexity max rel {rp.max():.3e} mean rel {rp.mean():.3e}", flush=True)
    assert rp.max() <= 0.01
    assert dpi.mean() <= 1e-2 and rb.mean() <= 0.1


def test_init_state_properties():
    N, K = 3000, 64
    g, s = graph("community", N, K)
    opts = M.Options(K=K, num_node_sample=10)
    a = M.MCMCSamplerStochastic(opts, g, s)
    b = M.MCMCSamplerStochastic(opts, g, s)
    a.init(5)
    b.init(5)
    pa, ta, ba = a.get_state()
    pb, tb, bb = b.get_state()
    assert np.array_equal(pa, pb) and np.array_equal(ta, tb)
    assert np.allclose(pa[:, :K].sum(1), 1.0, atol=1e-12)
    assert np.all(pa[:, K] > 0) and np.all(ta > 0)
    assert np.all((ba > 0) & (ba < 1))
    assert np.allclose(ba, ta[:, 1] / ta.sum(1), rtol=1e-15)
    # Gamma(1,1) rows: E[phi_sum] = K
    assert abs(pa[:, K].mean() - K) < 0.05 * K


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("eta0,eta1", [(1.0, 1.0), (2.0, 0.5), (0.5, 3.0)])
def test_init_state_matches_oracle(eta0, eta1, prec):
    """init() (mcmc_sampler_stochastic.cc:54-76): theta = gamma(eta0, eta1,
    K, 2) bit-exact against the oracle's exp(log Gamma(eta0)) * eta1 on the
    same counters (shape eta0, scale eta1: random.cc:472-511; a swap changes
    every value once eta0 != eta1), beta = theta1 / (theta0 + theta1), and
    the pi ++ phi_sum rows against the oracle's -log U draws row-normalised
    (FP64 1e-13: the device's log and summation order differ; FP32 rows are
    the FP64 values rounded, 1e-6)."""
    print(f"init eta=({eta0},{eta1}) {prec}", flush=True)
    N, K, seed = 3000, 64, 77
    g, s = graph("community", N, K)
    smp = M.MCMCSamplerStochastic(M.Options(K=K, num_node_sample=10, eta0=eta0, eta1=eta1,
                                            precision=prec), g, s)
    smp.init(seed)
    _, theta, beta = smp.get_state()
    ref = oracle_lib.init_theta(seed, eta0, eta1, K)
    np.testing.assert_array_equal(theta, ref)
    np.testing.assert_array_equal(beta, ref[:, 1] / (ref[:, 0] + ref[:, 1]))
    nodes = np.unique(np.concatenate([[0, N - 1], np.random.default_rng(1).choice(N, 200)]))
    rows = smp.get_rows(nodes).astype(np.float64)
    want = oracle_lib.init_phi_rows(seed, nodes, K)
    ok, rel = _close(rows, want, 1e-13 if prec == "fp64" else 1e-6)
    assert ok, rel
    smp.close()


@pytest.mark.parametrize("strategy,max_source", [("stratified-random-node-nonlinks", 1),
                                                 ("stratified-random-node-links", 8)])
@pytest.mark.parametrize("K", [32, 1024])
def test_update_beta_fp32_injected(strategy, max_source, K):
    """Precision mode: update_beta with FP32 rows and FP64 gradient
    accumulation against the FP32 oracle (the reference's FP32 build,
    accumulating in float) and the FP64 oracle on the same float rows, to
    1e-4."""
    print(f"update_beta fp32 {strategy} K={K}", flush=True)
    N = 6000
    strat = {"stratified-random-node-nonlinks": 2, "stratified-random-node-links": 1}[strategy]
    g, s, smp, orc32 = make_pair("community", N, K, precision="fp32", n=16, m=600,
                                 max_source=max_source)
    orc64 = Oracle(g, s, K, precision="fp64", num_node_sample=16)
    pi, theta = _state(N, K, 9, np.float32)
    smp.set_state(pi, theta)
    mb = orc64.sample_minibatch(strat, 600, max_source, 42, 3)
    assert len(mb["y"]) > 0
    noise = np.random.default_rng(3).standard_normal((K, 2))
    smp.update_beta(mb["edges"], mb["y"], mb["scale"], noise, 0.02)
    _, gth, gbe = smp.get_state()
    for orc, state_pi in ((orc32, pi), (orc64, pi.astype(np.float64))):
        orc.set_state(state_pi, theta)
        orc.update_beta(mb["edges"], mb["y"], mb["scale"], noise, 0.02)
        _, oth, obe = orc.get_state()
        ok, rel = _close(gth, oth, 1e-4)
        assert ok, (orc.prec, rel)
        ok, rel = _close(gbe, obe, 1e-4)
        assert ok, (orc.prec, rel)


@pytest.mark.parametrize("flag", [0, 1], ids=["nonlinks", "links"])
@pytest.mark.parametrize("max_source", [1, 16])
def test_stratified_iteration_fp64(flag, max_source):
    """One full FP64 iteration of the default strategy (stratified random
    node) through ammsb_step -- samplers, update_phi + update_pi,
    update_beta + theta step -- against the oracle's run() from the same
    state, step and seed: 1e-12."""
    N, K, seed = 6000, 32, 13
    t = 2
    while (int(oracle_lib.philox([0, 0, t, 1], [seed, 0])[0]) & 1) != flag:
        t += 1
    print(f"stratified iteration flag={flag} max_source={max_source} step={t}", flush=True)
    g, s, smp, orc = make_pair("community", N, K, n=20, m=600, max_source=max_source, seed=seed,
                               a=0.01, c=0.55)
    pi, theta = _state(N, K, 6)
    smp.set_state(pi, theta)
    orc.set_state(pi, theta)
    smp.set_counters(t, 1)
    orc.set_counters(t, 1)
    smp.step(1)
    orc.run(1, 0, 600, max_source, 10 ** 9, seed)
    gpi, gth, gbe = smp.get_state()
    rpi, rth, rbe = orc.get_state()
    assert np.sum(np.any(gpi != pi, axis=1)) > 1
    for gv, rv in ((gpi, rpi), (gth, rth), (gbe, rbe)):
        ok, err = _close(gv, rv, TOL64)
        assert ok, err
    assert smp.info["step_count"] == t + 1


def test_empty_inputs_and_operator_reuse():
    """Empty operator inputs: update_phi of no nodes changes nothing;
    update_beta of no edges is the theta step with a zero gradient, which
    has a closed form (mcmc_sampler_stochastic.cc:245-266 with G = 0).  Then
    one context serves operator calls of growing and shrinking sizes (its
    device workspace reused), each to the oracle at 1e-12."""
    print("empty inputs / operator reuse", flush=True)
    N, K = 4000, 32
    g, s, smp, orc = make_pair("community", N, K, n=16)
    pi, theta = _state(N, K, 5)
    smp.set_state(pi, theta)
    orc.set_state(pi, theta)
    smp.update_phi(np.zeros(0, np.int32), np.zeros((0, 17), np.int32), None, 0.01)
    gpi, gth, _ = smp.get_state()
    assert np.array_equal(gpi, pi) and np.array_equal(gth, theta)
    xi = np.random.default_rng(1).standard_normal((K, 2))
    eps_t = 0.03
    smp.update_beta(np.zeros((0, 2), np.int32), np.zeros(0, np.uint8), 1.0, xi, eps_t)
    eta = np.array([1.0, 1.0])  # eta0, eta1 defaults
    th = np.abs(theta + eps_t / 2.0 * (eta - theta) + np.sqrt(eps_t * theta) * xi)
    th = np.maximum(th, 1e-24)
    _, gth, gbe = smp.get_state()
    ok, rel = _close(gth, th, TOL64)
    assert ok, rel
    ok, rel = _close(gbe, th[:, 1] / th.sum(1), TOL64)
    assert ok, rel
    orc.set_state(pi, gth)
    rng = np.random.default_rng(2)
    for n_nodes, nnb in ((300, 25), (20, 25), (40, 301), (300, 25)):
        nodes = rng.choice(N, size=n_nodes, replace=False).astype(np.int32)
        nbrs = np.stack([orc.sample_neighbors(int(v), nnb, 42, 7) & 0x7FFFFFFF for v in nodes])
        noise = rng.standard_normal((n_nodes, K))
        smp.update_phi(nodes, nbrs, noise, 0.011)
        orc.update_phi(nodes, nbrs, noise, 0.011)
        gpi, _, _ = smp.get_state()
        opi, _, _ = orc.get_state()
        ok, rel = _close(gpi[:, :K], opi[:, :K], TOL64)
        assert ok, (n_nodes, nnb, rel)
        ok, rel = _close(gpi[:, K], opi[:, K], TOL64)
        assert ok, (n_nodes, nnb, rel)


def test_bandwidth_probes_leave_the_state_alone():
    """The read probe and update_phi's gather probe (no counterpart in the
    reference) report the bytes they read -- the gather: the step's nodes'
    own and n+1 neighbour rows plus the ids -- change no state, and the
    iterations after them equal those of a context that never probed."""
    print("bandwidth probes", flush=True)
    N, K, n = 4000, 64, 16
    g, s, a, _ = make_pair("community", N, K, n=n, m=400, seed=9)
    _, _, b, _ = make_pair("community", N, K, n=n, m=400, seed=9)
    pi, theta = _state(N, K, 3)
    a.set_state(pi, theta)
    b.set_state(pi, theta)
    a.step(3)
    b.step(3)
    t = a.info["step_count"]
    nodes = a.debug_sample(t)["nodes"]
    sec, nb = a.probe_gather(2)
    assert sec > 0 and nb == len(nodes) * ((n + 2) * K * 8 + 4 * (n + 1))
    sec, nb = a.probe_read(2)
    assert sec > 0 and nb == N * a.info["ld"] * 8
    for x, y in zip(a.get_state(), b.get_state()):
        np.testing.assert_array_equal(x, y)
    a.step(4)
    b.step(4)
    for x, y in zip(a.get_state(), b.get_state()):
        np.testing.assert_array_equal(x, y)


def test_error_paths():
    N, K = 2000, 16
    g, s = graph("community", N, K)
    smp = M.MCMCSamplerStochastic(M.Options(K=K, num_node_sample=8), g, s)
    with pytest.raises(M.AmmsbError):
        smp.step(1)  # before init
    smp.init(1)
    with pytest.raises(M.AmmsbError):
        smp.update_phi([1, 1], np.zeros((2, 4), np.int32), None, 0.01)  # duplicate nodes
    with pytest.raises(M.AmmsbError):
        smp.update_phi([N + 5], np.zeros((1, 4), np.int32), None, 0.01)
    with pytest.raises(M.AmmsbError):
        M.MCMCSamplerStochastic(M.Options(K=K, num_node_sample=N), g, s)
    # a caller's split with a node id outside [0, N)
    for bad in (N, -1):
        hoe = np.array(s.held_out_edges, dtype=np.int32, copy=True)
        hoe[len(hoe) // 2, 1] = bad
        bad_split = M.Split(hoe, s.held_out_labels, s.test_edges, s.test_labels)
        with pytest.raises(M.AmmsbError):
            M.MCMCSamplerStochastic(M.Options(K=K, num_node_sample=8), g, bad_split)
    # the largest rows the device kernels take: K 2048 (FP64), 4096 (FP32)
    for k_max, prec in ((2048, "fp64"), (4096, "fp32")):
        M.MCMCSamplerStochastic(M.Options(K=k_max, num_node_sample=8, precision=prec), g, s).close()
        with pytest.raises(M.AmmsbError):
            M.MCMCSamplerStochastic(M.Options(K=k_max + 1, num_node_sample=8, precision=prec), g, s)
    smp.step(3)
    assert smp.info["step_count"] == 4


@pytest.mark.parametrize("N,m,deg", [(20000, 1500, 20.0), (300, 299, 2.0)])
def test_random_edge_sampler_bit_exact(N, m, deg):
    """random-edge minibatches (network.cc:237-268): the device draws equal
    the oracle's sequential definition; m >= E - |held-out| - |test| takes
    the whole training set."""
    print(f"random-edge N={N} m={m}", flush=True)
    K, n = 16, 20
    g, s, smp, orc = make_pair("community", N, K, n=n, m=m, strategy="random-edge", seed=31,
                               deg=deg)
    mm = smp.info["mini_batch_size"]
    full = mm >= g.num_edges - len(s.held_out_labels) - len(s.test_labels)
    assert full == (deg < 5)  # the second case takes the whole training set
    for step in range(1, 6):
        d = smp.debug_sample(step)
        r = orc.sample_minibatch(3, mm, 1, 31, step)
        assert d["flag"] == r["flag"] == 2
        assert np.array_equal(d["edges"], r["edges"]), step
        assert d["scale"] == pytest.approx(r["scale"], rel=1e-15)
        assert np.array_equal(np.sort(d["nodes"]), np.unique(r["edges"]))
        for q in range(0, len(d["nodes"]), max(1, len(d["nodes"]) // 10)):
            v = int(d["nodes"][q])
            assert np.array_equal(d["nbrs"][q], orc.sample_neighbors(v, n + 1, 31, step))


def test_random_edge_iteration_fp64():
    """One full FP64 iteration with random-edge minibatches against the
    oracle from the same state and seed (1e-12)."""
    print("random-edge iteration", flush=True)
    N, K = 6000, 32
    g, s, smp, orc = make_pair("community", N, K, n=16, m=700, strategy="random-edge", seed=8)
    pi, theta = _state(N, K, 4)
    smp.set_state(pi, theta)
    orc.set_state(pi, theta)
    smp.step(1)
    orc.run(1, 3, 700, 1, 10 ** 9, 8)
    gpi, gth, gbe = smp.get_state()
    rpi, rth, rbe = orc.get_state()
    for gv, rv in ((gpi, rpi), (gth, rth), (gbe, rbe)):
        ok, err = _close(gv, rv, TOL64)
        assert ok, err
