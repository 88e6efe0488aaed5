"""Parity at the reference's default sizes and at the benchmark sizes (GPU).

The reference derives num_node_sample = N/50 and mini_batch_size = N/10 when
they are not given (mcmc_sampler_stochastic.cc:35-44).  From N of a few ten
thousand on, these sizes move the device samplers' de-dup tables from LDS to
HBM and make update_phi stage its neighbour ids in LDS chunk by chunk; from
N of about 370,000 on, one iteration's draws exceed the sample-ahead ring's
1 GiB budget and the ring holds a single iteration.  These tests run those
paths and the benchmark sizes against the oracle:

  * samplers bit-exact at N = 250,000 with the derived sizes (n = 5,000,
    m = 25,000): minibatch and neighbour de-dup tables in HBM
  * one full FP64 iteration at N = 400,000 with the derived sizes
    (n = 8,000, m = 40,000: HBM tables, four chunks of neighbour ids, a
    one-slot ring) to 1e-12, for a non-link and a link minibatch
  * update_phi with more than 2,048 neighbours (several LDS chunks), FP64
    to 1e-12 and FP32 to 1e-4, with injected draws
  * one full FP64 iteration at N = 1M, K = 256 (the FP64 benchmark size,
    4,096-node minibatches, sampler.max-source 1024) to 1e-12, for both
    minibatch kinds
  * N = 4M, K = 1024, FP32 (the precision-mode benchmark size, row offsets
    beyond 2^31 elements): one iteration leaves every row outside the
    minibatch bit-identical and renormalises every minibatch row; the
    update_phi of the highest-id minibatch nodes and the update_beta of the
    iteration equal the oracle; the injected update_phi of high-id nodes
    equals the oracle.
Tolerances: draws bit-exact; FP64 1e-12 relative with an absolute floor of
1e-12 x the row maximum; FP32 1e-4.
"""
import ctypes
import os

import numpy as np
import pytest

import mcmc_for_ammsb_amd as M
from mcmc_for_ammsb_amd import _lib
from tests import oracle_lib
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
THREADS = min(16, os.cpu_count() or 1)
_GRAPHS = {}


def _graph(N, K, deg, seed=3):
    key = (N, K, deg, seed)
    if key not in _GRAPHS:
        g = M.generate("community", N=N, K=K, mean_degree=deg, seed=seed)
        _GRAPHS[key] = (g, M.make_split(g, 0.1, 42))
    return _GRAPHS[key]


def _tuning(smp, name):
    v = ctypes.c_int64()
    _lib.check(_lib.lib.ammsb_get_tuning(smp._h, name.encode(), ctypes.byref(v)))
    return v.value


def _flag_step(seed, flag, start=2):
    """First step >= start whose stratified coin is `flag` (0 non-links, 1
    links): u64(FLAG; 0, 0, t) % 2 of the seed's Philox stream."""
    t = start
    while True:
        b = oracle_lib.philox([0, 0, t, 1], [seed & 0xFFFFFFFF, seed >> 32])
        if int(b[0]) & 1 == flag:
            return t
        t += 1


def _close_rows(gpu, ref, tol):
    """(ok, max relative error) with an absolute floor of tol x row max."""
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.max(np.abs(ref), axis=-1, keepdims=True)
    err = np.abs(gpu - ref)
    ok = bool(np.all(err <= tol * np.abs(ref) + tol * scale))
    return ok, float(np.max(err / (np.abs(ref) + 1e-300)))


def _compare_states(gpu, ref, tol, chunk=65536):
    """Full pi ++ phi_sum states, chunk by chunk (no full-size temporaries)."""
    worst = 0.0
    for lo in range(0, gpu.shape[0], chunk):
        ok, rel = _close_rows(gpu[lo:lo + chunk], ref[lo:lo + chunk], tol)
        assert ok, (lo, rel)
        worst = max(worst, rel)
    return worst


def test_samplers_reference_defaults_bit_exact():
    """N = 250,000, num_node_sample and mini_batch_size derived as the
    reference does (5,000 and 25,000): both de-dup tables live in HBM.  The
    stratified minibatch, its node set and the 5,001 neighbours of sampled
    nodes equal the oracle's sequential definitions bit for bit, for
    non-link and link minibatches."""
    N, K, seed = 250_000, 32, 23
    print(f"samplers at the reference defaults N={N}", flush=True)
    g, s = _graph(N, K, 20.0)
    smp = M.MCMCSamplerStochastic(M.Options(K=K, random_seed=seed), g, s)
    inf = smp.info
    n, m = inf["num_node_sample"], inf["mini_batch_size"]
    assert (n, m) == (N // 50, N // 10)
    assert _tuning(smp, "sampler_lds") == 0  # both tables in HBM
    orc = Oracle(g, s, K, num_node_sample=n)
    for flag in (0, 1):
        t = _flag_step(seed, flag)
        d = smp.debug_sample(t)
        r = orc.sample_minibatch(0, m, 1, seed, t)
        assert d["flag"] == r["flag"] == flag, t
        assert np.array_equal(d["edges"], r["edges"]), t
        assert d["scale"] == pytest.approx(r["scale"], rel=1e-15)
        assert np.array_equal(np.sort(d["nodes"]), np.unique(r["edges"])), t
        nn = len(d["nodes"])
        picks = sorted(set([0, nn - 1] + list(range(0, nn, max(1, nn // 30)))))
        for q in picks:
            v = int(d["nodes"][q])
            assert np.array_equal(d["nbrs"][q], orc.sample_neighbors(v, n + 1, seed, t)), (t, v)
        print(f"  flag {flag} step {t}: {len(r['y'])} edges, {nn} nodes, "
              f"{len(picks)} neighbour lists checked", flush=True)
    smp.close()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("flag", [0, 1], ids=["nonlinks", "links"])
def test_iteration_reference_defaults_fp64(flag):
    """N = 400,000 with the derived sizes (n = 8,000, m = 40,000): one
    iteration's draws (about 1.3 GB of neighbour ids) exceed the ring's
    1 GiB budget, so the ring holds one iteration; the de-dup tables live
    in HBM and update_phi stages 8,001 neighbour ids in four LDS chunks.
    One full FP64 iteration through ammsb_step equals the oracle's run() to
    1e-12 from the same state, step and seed."""
    N, K, seed = 400_000, 16, 29
    g, s = _graph(N, K, 20.0)
    smp = M.MCMCSamplerStochastic(M.Options(K=K, random_seed=seed, a=0.01, b=1024.0, c=0.55),
                                  g, s)
    inf = smp.info
    n, m = inf["num_node_sample"], inf["mini_batch_size"]
    assert _tuning(smp, "ring_slots") == 1
    assert _tuning(smp, "sampler_lds") == 0
    t = _flag_step(seed, flag)
    print(f"iteration at the reference defaults N={N} n={n} m={m} flag={flag} step={t}",
          flush=True)
    smp.init(seed)
    pi, theta, _ = smp.get_state()
    smp.set_counters(t, 1)
    orc = Oracle(g, s, K, num_node_sample=n, a=0.01, b=1024.0, c=0.55)
    orc.set_state(pi, theta)
    orc.set_counters(t, 1)
    smp.step(1)
    orc.run(1, 0, m, 1, 10 ** 9, seed, threads=THREADS)
    gpi, gth, gbe = smp.get_state()
    rpi, rth, rbe = orc.get_state()
    changed = np.flatnonzero(np.any(gpi != pi, axis=1))
    print(f"  {len(changed)} rows updated", flush=True)
    assert len(changed) > (1000 if flag == 0 else 1)
    worst = _compare_states(gpi, rpi, TOL64)
    for gv, rv in ((gth, rth), (gbe[None, :], rbe[None, :])):
        ok, rel = _close_rows(gv, rv, TOL64)
        assert ok, rel
    print(f"  max relative error of pi {worst:.2e}", flush=True)
    smp.close()


@pytest.mark.parametrize("prec,K", [("fp64", 256), ("fp32", 1024)])
def test_update_phi_many_neighbours_injected(prec, K):
    """update_phi with 4,097 and 2,049 neighbours per node: the ids are
    staged in LDS in chunks of 2,048 (the last chunk of one id), with a self
    entry in a later chunk skipped as the reference does
    (mcmc_sampler_stochastic.cc:275-277)."""
    N = 6000
    dt = np.float64 if prec == "fp64" else np.float32
    tol = TOL64 if prec == "fp64" else 1e-4
    g, s = _graph(N, K, 20.0)
    smp = M.MCMCSamplerStochastic(M.Options(K=K, num_node_sample=16, precision=prec), g, s)
    orc = Oracle(g, s, K, precision=prec, num_node_sample=16)
    rng = np.random.default_rng(5)
    phi = rng.gamma(1.0, 1.0, size=(N, K))
    pi = np.concatenate([phi / phi.sum(1, keepdims=True), phi.sum(1, keepdims=True)], 1).astype(dt)
    theta = rng.gamma(2.0, 0.5, size=(K, 2))
    for nnb in (4097, 2049):
        print(f"update_phi {prec} K={K} with {nnb} neighbours", flush=True)
        smp.set_state(pi, theta)
        orc.set_state(pi, theta)
        nodes = rng.choice(N, size=64, replace=False).astype(np.int32)
        nbrs = np.stack([orc.sample_neighbors(int(v), nnb, 42, 3) & 0x7FFFFFFF for v in nodes])
        nbrs[0, 3000 % nnb] = nodes[0]
        noise = rng.standard_normal((len(nodes), K)).astype(dt)
        smp.update_phi(nodes, nbrs, noise, 0.0125)
        orc.update_phi(nodes, nbrs, noise, 0.0125)
        gpi, _, _ = smp.get_state()
        opi, _, _ = orc.get_state()
        ok, rel = _close_rows(gpi[nodes], opi[nodes], tol)
        assert ok, rel
        rest = np.setdiff1d(np.arange(N), nodes)
        assert np.array_equal(gpi[rest], pi[rest])
    smp.close()


_C2 = {}


@pytest.mark.timeout(600)
@pytest.mark.parametrize("flag", [0, 1], ids=["nonlinks", "links"])
def test_iteration_n1m_k256_fp64(flag):
    """The FP64 benchmark size: N = 1M, K = 256, mean degree 50 (25M
    edges), 4,096-node minibatches (m = 4,095, sampler.max-source = 1024),
    128 neighbour samples.  One full FP64 iteration through ammsb_step
    equals the oracle's run() to 1e-12 from the same state, step and seed,
    for a non-link and a link minibatch."""
    N, K, seed, m, n, ms = 1_000_000, 256, 42, 4095, 128, 1024
    g, s = _graph(N, K, 50.0, seed=2)
    t = _flag_step(seed, flag)
    print(f"iteration N={N} K={K} fp64 flag={flag} step={t}", flush=True)
    opts = M.Options(K=K, num_node_sample=n, mini_batch_size=m, sampler_max_source=ms,
                     a=0.01, b=1024.0, c=0.55, random_seed=seed)
    smp = M.MCMCSamplerStochastic(opts, g, s)
    smp.init(seed)
    pi, theta, _ = smp.get_state()
    smp.set_counters(t, 1)
    if "orc" not in _C2:
        _C2["orc"] = Oracle(g, s, K, num_node_sample=n, a=0.01, b=1024.0, c=0.55)
    orc = _C2["orc"]
    orc.set_state(pi, theta)
    orc.set_counters(t, 1)
    smp.step(1)
    print("  oracle iteration", flush=True)
    orc.run(1, 0, m, ms, 10 ** 9, seed, threads=THREADS)
    gpi, gth, gbe = smp.get_state()
    smp.close()
    del pi
    rpi, rth, rbe = orc.get_state()
    worst = _compare_states(gpi, rpi, TOL64)
    for gv, rv in ((gth, rth), (gbe[None, :], rbe[None, :])):
        ok, rel = _close_rows(gv, rv, TOL64)
        assert ok, rel
    print(f"  max relative error of pi {worst:.2e}", flush=True)
    if flag == 1:
        del _C2["orc"]


class RowsOracle:
    """The oracle over a subset U of a large graph's nodes, renumbered
    0..|U|-1 inside a graph of N_v = N/n virtual nodes with one neighbour
    sample, so that the update_phi factor N/num_node_sample
    (mcmc_sampler_stochastic.cc:300) is the large graph's.  Its edges are the
    large graph's links among the given pairs; its rows are copies of the
    large state's rows of U.  update_phi and update_beta read only these
    rows, the links and theta, so the results equal those the oracle would
    compute on the whole state."""

    def __init__(self, g, K, precision, n, U, pairs, hyper):
        assert g.N % n == 0, "N / num_node_sample must be an integer"
        self.Nv = g.N // n
        self.U = np.unique(np.asarray(U, dtype=np.int64))
        assert len(self.U) <= self.Nv
        loc = self.local
        links = [(a, b) for a, b in pairs if g.linked(int(a), int(b))]
        le = (np.array([[loc(a), loc(b)] for a, b in links], dtype=np.int64)
              if links else np.zeros((0, 2), np.int64))
        mg = M.Graph.from_edges(self.Nv, le)
        z2, z1 = np.zeros((0, 2), np.int32), np.zeros(0, np.uint8)
        ms = M.Split(z2, z1, z2.copy(), z1.copy())
        self.orc = Oracle(mg, ms, K, precision=precision, num_node_sample=1, **hyper)
        self.K = K
        self.dtype = np.float64 if precision == "fp64" else np.float32

    def local(self, v):
        i = np.searchsorted(self.U, np.asarray(v, dtype=np.int64))
        assert np.all(self.U[i] == v)
        return i

    def set_rows(self, rows_of_U, theta):
        st = np.zeros((self.Nv, self.K + 1), dtype=self.dtype)
        st[:, :self.K] = 1.0 / self.K
        st[:, self.K] = self.K
        st[:len(self.U)] = rows_of_U
        self.orc.set_state(st, theta)

    def rows(self, v):
        pi, theta, beta = self.orc.get_state()
        return pi[self.local(v)], theta, beta


_C3 = {}


def _c3():
    if "g" not in _C3:
        g = M.generate("community", N=4_000_000, K=1024, mean_degree=40.0, seed=2)
        _C3["g"], _C3["s"] = g, M.make_split(g, 0.1, 42)
    return _C3["g"], _C3["s"]


@pytest.mark.timeout(900)
def test_iteration_n4m_k1024_fp32():
    """The precision-mode benchmark size: N = 4M, K = 1024, FP32 rows of
    1,056 floats, so rows of nodes above 2,033,601 start beyond element 2^31.
    One iteration of a non-link minibatch through ammsb_step:
      * leaves every row outside the minibatch's node set bit-identical and
        changes every row inside it;
      * renormalises every updated row (sum of pi = 1 to 1e-6, phi_sum > 0);
      * the new rows of the 64 highest-id minibatch nodes equal the FP32
        oracle's update_phi (its Philox noise, the iteration's neighbours) to
        1e-4;
      * theta and beta equal the oracle's update_beta over the iteration's
        minibatch on the updated rows to 1e-4 (FP64 oracle on the same float
        rows; FP32 oracle to 1e-3).
    Then ammsb_update_phi on 64 of the highest node ids, with neighbours from
    the top of the id range, equals the oracle to 1e-4."""
    K, n, m, ms, seed = 1024, 128, 4095, 1024, 42
    hyp = dict(a=0.01, b=1024.0, c=0.55)
    print("iteration N=4M K=1024 fp32: graph", flush=True)
    g, s = _c3()
    N = g.N
    opts = M.Options(K=K, num_node_sample=n, mini_batch_size=m, sampler_max_source=ms,
                     random_seed=seed, precision="fp32", **hyp)
    smp = M.MCMCSamplerStochastic(opts, g, s)
    ld = smp.info["ld"]
    smp.init(seed)
    t = _flag_step(seed, 0)
    smp.set_counters(t, 1)
    print(f"  step {t}: state download", flush=True)
    pre, theta0, _ = smp.get_state()
    d = smp.debug_sample(t)
    assert d["flag"] == 0
    nodes = np.sort(d["nodes"])
    smp.step(1)
    post, theta1, beta1 = smp.get_state()
    print(f"  {len(nodes)} minibatch nodes; comparing {N} rows", flush=True)
    changed = []
    for lo in range(0, N, 1 << 18):
        diff = np.any(pre[lo:lo + (1 << 18)] != post[lo:lo + (1 << 18)], axis=1)
        changed.append(lo + np.flatnonzero(diff))
    changed = np.concatenate(changed)
    assert np.array_equal(changed, nodes)
    rows = post[nodes].astype(np.float64)
    assert np.all(np.isfinite(rows)) and np.all(rows[:, K] > 0)
    assert np.max(np.abs(rows[:, :K].sum(1) - 1.0)) <= 1e-6

    # the 64 highest-id minibatch nodes: row offsets beyond 2^31 elements
    order = np.argsort(d["nodes"])
    top = order[-64:]
    hi_nodes = d["nodes"][top].astype(np.int64)
    assert int(hi_nodes.min()) * ld > 2 ** 31
    hi_nbrs = d["nbrs"][top] & 0x7FFFFFFF
    pairs = [(int(v), int(b)) for v, row in zip(hi_nodes, hi_nbrs) for b in row]
    U = np.concatenate([hi_nodes, hi_nbrs.ravel(), d["edges"].ravel()])
    ro = RowsOracle(g, K, "fp32", n, U, pairs, hyp)
    ro.set_rows(pre[ro.U], theta0)
    eps_t = ro.orc.eps_t(t)
    noise = np.stack([ro.orc.phi_noise(int(v), t, seed) for v in hi_nodes])
    ro.orc.update_phi(ro.local(hi_nodes).astype(np.int32),
                      ro.local(hi_nbrs.ravel()).reshape(hi_nbrs.shape).astype(np.int32),
                      noise, eps_t)
    got, _, _ = ro.rows(hi_nodes)
    ok, rel = _close_rows(post[hi_nodes], got, 1e-4)
    print(f"  update_phi of the highest-id nodes: max rel err {rel:.2e}", flush=True)
    assert ok, rel

    # update_beta of the iteration on the updated rows
    tn = ro.orc.theta_noise(t, seed)
    le = ro.local(d["edges"].ravel()).reshape(-1, 2).astype(np.int32)
    y = np.zeros(len(le), np.uint8)
    for prec, tol in (("fp64", 1e-4), ("fp32", 1e-3)):
        rb = RowsOracle(g, K, prec, n, U, [], hyp)
        rb.set_rows(post[rb.U].astype(rb.dtype), theta0)
        rb.orc.update_beta(le, y, d["scale"], tn, eps_t)
        _, oth, obe = rb.orc.get_state()
        ok_t, rel_t = _close_rows(theta1, oth, tol)
        ok_b, rel_b = _close_rows(beta1[None, :], obe[None, :], tol)
        print(f"  update_beta vs {prec} oracle: theta {rel_t:.2e}, beta {rel_b:.2e}", flush=True)
        assert ok_t and ok_b, (prec, rel_t, rel_b)
    del pre

    # injected update_phi on high node ids
    rng = np.random.default_rng(3)
    inj = rng.choice(np.arange(N - 200_000, N), size=64, replace=False).astype(np.int32)
    nb = rng.integers(N - 1_000_000, N, size=(64, n + 1)).astype(np.int32)
    nb[nb == inj[:, None]] = 0
    xi = rng.standard_normal((64, K)).astype(np.float32)
    pairs = [(int(v), int(b)) for v, row in zip(inj, nb) for b in row]
    ri = RowsOracle(g, K, "fp32", n, np.concatenate([inj, nb.ravel()]), pairs, hyp)
    ri.set_rows(post[ri.U], theta1)
    smp.update_phi(inj, nb, xi, 0.0117)
    ri.orc.update_phi(ri.local(inj).astype(np.int32),
                      ri.local(nb.ravel()).reshape(nb.shape).astype(np.int32), xi, 0.0117)
    got, _, _ = ri.rows(inj)
    ok, rel = _close_rows(smp.get_rows(inj), got, 1e-4)
    print(f"  injected update_phi of high ids: max rel err {rel:.2e}", flush=True)
    assert ok, rel
    smp.close()
