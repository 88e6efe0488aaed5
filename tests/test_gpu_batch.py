"""Batched chains (GPU): many independent samplers advanced by one launch
per phase, replayed from a hipGraph.

Pins:
  * one FP64 iteration of every chain of an uneven batch against the oracle
    run from the same state and Philox seed: 1e-12 relative (draws are the
    chain's own Philox streams, so they equal the oracle's sequential
    replica bit for bit)
  * every chain of the FP32 batch (K up to 512) after one and three
    iterations against the FP32 oracle: 1e-4 relative
  * FP32 batch against single contexts with the same options and state:
    1e-4 relative after three iterations (summation order differs)
  * hipGraph replay against direct launches: bit-identical
  * autotuned shapes: valid for each chain's K, results within FP32
    tolerance of the default shapes
"""
import numpy as np
import pytest

import mcmc_for_ammsb_amd as M
from tests.oracle_lib import Oracle

pytestmark = pytest.mark.gpu

# (N, K, mean degree, n, m, seed): uneven sizes, K in {16, 64, 256, 512}
CHAINS = [(3000, 16, 20.0, 16, 300, 11), (12000, 64, 40.0, 32, 800, 12),
          (6000, 256, 30.0, 24, 500, 13), (20000, 512, 25.0, 32, 1000, 14),
          (2500, 16, 60.0, 40, 250, 15), (9000, 64, 20.0, 20, 700, 16)]
_GRAPHS = {}


def _graph(N, K, deg, seed):
    key = (N, K, deg, seed)
    if key not in _GRAPHS:
        g = M.generate("community", N=N, K=K, mean_degree=deg, seed=seed)
        _GRAPHS[key] = (g, M.make_split(g, 0.1, 42))
    return _GRAPHS[key]


def _state(N, K, seed, dtype):
    rng = np.random.default_rng(seed)
    phi = rng.gamma(1.0, 1.0, size=(N, K))
    s = phi.sum(1)
    pi = np.concatenate([phi / s[:, None], s[:, None]], 1).astype(dtype)
    return pi, rng.gamma(2.0, 0.5, size=(K, 2))


def _setup(precision, chains):
    opts, graphs, splits, states = [], [], [], []
    dt = np.float64 if precision == "fp64" else np.float32
    for N, K, deg, n, m, seed in chains:
        g, s = _graph(N, K, deg, seed)
        opts.append(M.Options(K=K, num_node_sample=n, mini_batch_size=m, random_seed=seed,
                              precision=precision))
        graphs.append(g)
        splits.append(s)
        states.append(_state(N, K, seed + 100, dt))
    return opts, graphs, splits, states


def _batch(precision, chains, graph=True):
    opts, graphs, splits, states = _setup(precision, chains)
    b = M.MCMCBatch(opts, graphs, splits)
    b.use_graph(graph)
    for i, (pi, th) in enumerate(states):
        b.chain(i).set_state(pi, th)
    return b, opts, graphs, splits, states


def _close(gpu, ref, tol):
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.max(np.abs(ref), axis=-1, keepdims=True) if ref.ndim > 1 else np.max(np.abs(ref))
    return bool(np.all(np.abs(gpu - ref) <= tol * np.abs(ref) + tol * scale))


def test_batch_fp64_one_iteration_matches_oracle():
    chains = [c for c in CHAINS if c[1] <= 256]
    print(f"batch fp64 vs oracle, {len(chains)} chains", flush=True)
    b, opts, graphs, splits, states = _batch("fp64", chains)
    b.step(1)
    for i, (N, K, deg, n, m, seed) in enumerate(chains):
        orc = Oracle(graphs[i], splits[i], K, precision="fp64", num_node_sample=n)
        orc.set_state(*states[i])
        orc.run(1, 0, m, 1, 10 ** 9, seed)
        rpi, rth, rbeta = orc.get_state()
        gpi, gth, gbeta = b.chain(i).get_state()
        assert _close(gpi, rpi, 1e-12), (i, "pi")
        assert _close(gth, rth, 1e-12), (i, "theta")
        assert _close(gbeta, rbeta, 1e-12), (i, "beta")
        assert b.chain(i).info["step_count"] == 2
    b.close()


@pytest.mark.parametrize("iters", [1, 3])
def test_batch_fp32_matches_oracle(iters):
    """The batching configuration's own precision: every chain of the uneven
    FP32 batch (K 16 ... 512) after one and after three iterations equals
    the FP32 oracle's run() (the reference loop per chain,
    mcmc_sampler_stochastic.cc:93-183, in float with FP64 renormalisation)
    from the same state and Philox seed, to 1e-4."""
    print(f"batch fp32 vs oracle, {len(CHAINS)} chains, {iters} iterations", flush=True)
    b, opts, graphs, splits, states = _batch("fp32", CHAINS)
    b.step(iters)
    assert max(c[1] for c in CHAINS) == 512
    for i, (N, K, deg, n, m, seed) in enumerate(CHAINS):
        orc = Oracle(graphs[i], splits[i], K, precision="fp32", num_node_sample=n)
        orc.set_state(*states[i])
        orc.run(iters, 0, m, 1, 10 ** 9, seed)
        rpi, rth, rbeta = orc.get_state()
        gpi, gth, gbeta = b.chain(i).get_state()
        assert _close(gpi, rpi, 1e-4), (i, K, "pi")
        assert _close(gth, rth, 1e-4), (i, K, "theta")
        assert _close(gbeta, rbeta, 1e-4), (i, K, "beta")
        assert b.chain(i).info["step_count"] == 1 + iters
    b.close()


def test_batch_fp32_matches_single_contexts():
    print(f"batch fp32 vs single contexts, {len(CHAINS)} chains", flush=True)
    b, opts, graphs, splits, states = _batch("fp32", CHAINS)
    b.step(3)
    for i in range(len(CHAINS)):
        smp = M.MCMCSamplerStochastic(opts[i], graphs[i], splits[i])
        smp.set_state(*states[i])
        smp.step(3)
        rpi, rth, rbeta = smp.get_state()
        gpi, gth, gbeta = b.chain(i).get_state()
        assert _close(gpi, rpi, 1e-4), (i, "pi")
        assert _close(gth, rth, 1e-4), (i, "theta")
        assert _close(gbeta, rbeta, 1e-4), (i, "beta")
        # the held-out perplexity of a chain is the single-context evaluation
        smp2 = M.MCMCSamplerStochastic(opts[i], graphs[i], splits[i])
        smp2.set_state(gpi, gth)
        assert b.chain(i).perplexity() == smp2.perplexity()
        smp.close()
        smp2.close()
    b.close()


def test_graph_replay_equals_direct_launches():
    print("batch graph replay vs direct launches", flush=True)
    ga = _batch("fp32", CHAINS, graph=True)[0]
    da = _batch("fp32", CHAINS, graph=False)[0]
    # several replays with host work and synchronisation between them; every
    # replay checked against the direct launches (crossing a refill at 16)
    for steps in (2, 7, 9, 3):
        ga.step(steps)
        da.step(steps)
        np.testing.assert_array_equal(ga.perplexity(), da.perplexity())
        for i in range(len(CHAINS)):
            for x, y in zip(ga.chain(i).get_state(), da.chain(i).get_state()):
                np.testing.assert_array_equal(x, y)
    ga.close()
    da.close()


def test_refill_modes_same_trajectory():
    """Sampler refills in stream order and overlapped on a second stream (the
    default) draw the same iterations: bit-identical states across refills."""
    print("batch refill in order vs overlapped", flush=True)
    a = _batch("fp32", CHAINS)[0]
    b = _batch("fp32", CHAINS)[0]
    a.refill_in_order(True)
    for steps in (3, 17, 20):
        a.step(steps)
        b.step(steps)
        for i in range(len(CHAINS)):
            for x, y in zip(a.chain(i).get_state(), b.chain(i).get_state()):
                np.testing.assert_array_equal(x, y)
    a.close()
    b.close()


def test_autotune_shapes():
    print("batch autotune", flush=True)
    b, opts, graphs, splits, states = _batch("fp32", CHAINS)
    ref = _batch("fp32", CHAINS)[0]
    shapes = b.autotune(reps=2)
    assert len(shapes) == len(CHAINS)
    for (N, K, deg, n, m, seed), sh in zip(CHAINS, shapes):
        assert sh["us"] > 0
        assert sh["lanes_per_row"] * sh["vectors_per_lane"] * 4 >= K
        groups = 4 * 64 // sh["lanes_per_row"]
        assert 1 <= sh["nodes_per_block"] <= groups
    b.step(2)
    ref.step(2)
    for i in range(len(CHAINS)):
        assert _close(b.chain(i).get_state()[0], ref.chain(i).get_state()[0], 1e-4), i
    # shapes valid for chain 1 (K=64: 16 vectors of 4 floats) are accepted
    for lpr, ch, npb in ((16, 1, 1), (8, 2, 2), (16, 1, 4), (8, 2, 32)):
        b.set_shape(1, lpr, ch, npb)
    with pytest.raises(M.AmmsbError):
        b.set_shape(0, 64, 2, 1)  # K=16 would leave lanes idle
    with pytest.raises(M.AmmsbError):
        b.set_shape(1, 16, 1, 3)  # not a power of two
    b.close()
    ref.close()


@pytest.mark.parametrize("lpr,ch,npb", [(32, 1, 1), (32, 1, 2), (32, 1, 8), (16, 2, 1),
                                        (16, 2, 4), (16, 2, 16)])
def test_every_shape_same_result(lpr, ch, npb):
    """K=64 chain: each (lanes per row, vectors per lane, nodes per block)
    shape against the default, FP64 to 1e-12 after one iteration."""
    print(f"shape {lpr} {ch} {npb}", flush=True)
    chains = [CHAINS[1]]
    a = _batch("fp64", chains)[0]
    bb = _batch("fp64", chains)[0]
    bb.set_shape(0, lpr, ch, npb)
    a.step(1)
    bb.step(1)
    assert _close(bb.chain(0).get_state()[0], a.chain(0).get_state()[0], 1e-12)
    a.close()
    bb.close()
