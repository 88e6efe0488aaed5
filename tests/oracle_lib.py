"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- test-side only.

The oracle restates the reference's hot path on the CPU; tests use it as the
checker for the device kernels.  Build it with ``make -C oracle``.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

c_i32, c_i64, c_u64, c_dbl, c_vp = (ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64,
                                    ctypes.c_double, ctypes.c_void_p)
P = ctypes.POINTER


def _load():
    if not os.path.exists(LIB):
        raise ImportError(f"{LIB} missing: run `make -C oracle`")
    lib = ctypes.CDLL(LIB)
    sig = {
        "orc_philox": (None, [P(ctypes.c_uint32)] * 3),
        "orc_create": (c_vp, [c_i32, c_i64, P(c_i64), P(c_i32), P(c_i32), P(ctypes.c_uint8), c_i64,
                              P(c_i32), P(ctypes.c_uint8), c_i64, c_i64, c_i32, P(c_dbl)]),
        "orc_destroy": (None, [c_vp]),
        "orc_set_state": (None, [c_vp, c_vp, P(c_dbl)]),
        "orc_get_state": (None, [c_vp, c_vp, P(c_dbl), P(c_dbl)]),
        "orc_set_counters": (None, [c_vp, c_i64, c_i64]),
        "orc_update_phi": (c_i32, [c_vp, P(c_i32), c_i64, P(c_i32), c_i64, c_vp, c_dbl, c_u64, c_i64,
                                   c_i32]),
        "orc_update_beta": (c_i32, [c_vp, P(c_i32), P(ctypes.c_uint8), c_i64, c_dbl, P(c_dbl), c_dbl]),
        "orc_perplexity": (c_dbl, [c_vp]),
        "orc_eps_t": (c_dbl, [P(c_dbl), c_i64, c_i32]),
        "orc_sample_minibatch": (c_i32, [c_vp, c_i32, c_i64, c_i64, c_u64, c_i64, P(c_i32), P(c_dbl),
                                         P(c_i32), P(ctypes.c_uint8), c_i64, P(c_i64)]),
        "orc_sample_neighbors": (c_i32, [c_vp, c_i32, c_i64, c_u64, c_i64, P(c_i32)]),
        "orc_phi_noise": (None, [c_i32, c_u64, c_i32, c_i64, c_i64, c_vp]),
        "orc_theta_noise": (None, [c_u64, c_i64, c_i64, P(c_dbl)]),
        "orc_log_gamma": (c_dbl, [c_u64, c_dbl, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
        "orc_init_theta": (None, [c_u64, c_dbl, c_dbl, c_i64, P(c_dbl)]),
        "orc_init_phi_row": (None, [c_u64, c_i32, c_i64, P(c_dbl)]),
        "orc_phi_grads": (None, [c_vp, c_i32, P(c_i32), c_i64, P(c_dbl)]),
        "orc_beta_grads": (None, [c_vp, P(c_i32), P(ctypes.c_uint8), c_i64, P(c_dbl)]),
        "orc_run": (c_i32, [c_vp, c_i64, c_i32, c_i64, c_i64, c_i64, c_u64, c_i32, P(c_dbl), P(c_i64),
                            c_i64, P(c_i64), P(c_dbl)]),
    }
    for n, (r, a) in sig.items():
        f = getattr(lib, n)
        f.restype = r
        f.argtypes = a
    return lib


lib = _load()


def _p(a, t):
    return None if a is None else a.ctypes.data_as(P(t))


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib.orc_philox(_p(c, ctypes.c_uint32), _p(k, ctypes.c_uint32), _p(o, ctypes.c_uint32))
    return o


def init_theta(seed, eta0, eta1, K):
    """theta [K][2] of init() (mcmc_sampler_stochastic.cc:54): Gamma(shape
    eta0, scale eta1) per entry on the counter stream."""
    out = np.zeros((K, 2))
    lib.orc_init_theta(seed, eta0, eta1, K, _p(out, c_dbl))
    return out


def init_phi_rows(seed, nodes, K):
    """pi ++ phi_sum rows [len(nodes)][K+1] of init() (phi ~ Gamma(1,1),
    mcmc_sampler_stochastic.cc:66-76), in FP64."""
    out = np.zeros((len(nodes), K + 1))
    for i, v in enumerate(nodes):
        lib.orc_init_phi_row(seed, int(v), K, _p(out[i], c_dbl))
    return out


def hyper_array(alpha, eta0, eta1, epsilon, a, b, c, num_node_sample):
    return np.array([alpha, eta0, eta1, epsilon, a, b, c, num_node_sample], dtype=np.float64)


class Oracle:
    """One oracle network + state (pi ++ phi_sum, theta, beta, held-out R)."""

    def __init__(self, graph, split, K, precision="fp64", alpha=None, eta0=1.0, eta1=1.0,
                 epsilon=1e-7, a=None, b=1024.0, c=0.5, num_node_sample=None):
        self.graph, self.split, self.K = graph, split, int(K)
        self.prec = 0 if precision == "fp64" else 1
        self.dtype = np.float64 if self.prec == 0 else np.float32
        alpha = 1.0 / K if alpha is None else alpha
        a = b ** (-c) if a is None else a
        n = graph.N // 50 if num_node_sample is None else num_node_sample
        self.hyper = hyper_array(alpha, eta0, eta1, epsilon, a, b, c, n)
        self.num_node_sample = n
        self._keep = [np.ascontiguousarray(x) for x in (graph.row_ptr, graph.cols,
                                                       split.held_out_edges, split.held_out_labels,
                                                       split.test_edges, split.test_labels)]
        rp, cl, he, hl, te, tl = self._keep
        self.h = lib.orc_create(graph.N, graph.num_edges, _p(rp, c_i64), _p(cl, c_i32),
                                _p(he, c_i32), _p(hl, ctypes.c_uint8), len(hl), _p(te, c_i32),
                                _p(tl, ctypes.c_uint8), len(tl), self.K, self.prec,
                                _p(self.hyper, c_dbl))

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_destroy(self.h)
            self.h = None

    def set_state(self, pi, theta):
        pi = np.ascontiguousarray(pi, dtype=self.dtype)
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        lib.orc_set_state(self.h, pi.ctypes.data, _p(theta, c_dbl))

    def get_state(self):
        N, K = self.graph.N, self.K
        pi = np.empty((N, K + 1), dtype=self.dtype)
        theta = np.empty((K, 2))
        beta = np.empty(K)
        lib.orc_get_state(self.h, pi.ctypes.data, _p(theta, c_dbl), _p(beta, c_dbl))
        return pi, theta, beta

    def set_counters(self, step, average_count):
        lib.orc_set_counters(self.h, step, average_count)

    def update_phi(self, nodes, nbrs, noise, eps_t, seed=0, step=1, threads=1):
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        nbrs = np.ascontiguousarray(nbrs, dtype=np.int32)
        nz = None if noise is None else np.ascontiguousarray(noise, dtype=self.dtype)
        rc = lib.orc_update_phi(self.h, _p(nodes, c_i32), len(nodes), _p(nbrs, c_i32), nbrs.shape[1],
                                None if nz is None else nz.ctypes.data, eps_t, seed, step, threads)
        assert rc == 0

    def update_beta(self, edges, y, scale, noise, eps_t):
        edges = np.ascontiguousarray(edges, dtype=np.int32)
        y = np.ascontiguousarray(y, dtype=np.uint8)
        noise = np.ascontiguousarray(noise, dtype=np.float64)
        assert lib.orc_update_beta(self.h, _p(edges, c_i32), _p(y, ctypes.c_uint8), len(y), scale,
                                   _p(noise, c_dbl), eps_t) == 0

    def perplexity(self):
        return lib.orc_perplexity(self.h)

    def eps_t(self, step):
        return lib.orc_eps_t(_p(self.hyper, c_dbl), step, self.prec)

    def sample_minibatch(self, strategy, m, max_source, seed, step, cap=None):
        # random-edge: the full training set (m >= E - |held-out| - |test|)
        # can hold more than m links
        cap = cap or (max(m, self.graph.num_edges if strategy == 3 else 0)
                      + int(self.graph.degree().max()) + 8)
        edges = np.zeros((cap, 2), dtype=np.int32)
        y = np.zeros(cap, dtype=np.uint8)
        flag, scale, cnt = c_i32(), c_dbl(), c_i64()
        rc = lib.orc_sample_minibatch(self.h, strategy, m, max_source, seed, step, ctypes.byref(flag),
                                      ctypes.byref(scale), _p(edges, c_i32), _p(y, ctypes.c_uint8),
                                      cap, ctypes.byref(cnt))
        assert rc == 0, rc
        n = cnt.value
        return dict(flag=flag.value, scale=scale.value, edges=edges[:n].copy(), y=y[:n].copy())

    def sample_neighbors(self, node, nnb, seed, step):
        out = np.zeros(nnb, dtype=np.int32)
        assert lib.orc_sample_neighbors(self.h, node, nnb, seed, step, _p(out, c_i32)) == 0
        return out

    def phi_noise(self, node, step, seed):
        out = np.zeros(self.K, dtype=self.dtype)
        lib.orc_phi_noise(self.prec, seed, node, step, self.K, out.ctypes.data)
        return out

    def theta_noise(self, step, seed):
        out = np.zeros((self.K, 2))
        lib.orc_theta_noise(seed, step, self.K, _p(out, c_dbl))
        return out

    def phi_grads(self, node, nbrs):
        nbrs = np.ascontiguousarray(nbrs, dtype=np.int32)
        out = np.zeros(self.K)
        lib.orc_phi_grads(self.h, int(node), _p(nbrs, c_i32), len(nbrs), _p(out, c_dbl))
        return out

    def beta_grads(self, edges, y):
        edges = np.ascontiguousarray(edges, dtype=np.int32)
        y = np.ascontiguousarray(y, dtype=np.uint8)
        out = np.zeros((self.K, 2))
        lib.orc_beta_grads(self.h, _p(edges, c_i32), _p(y, ctypes.c_uint8), len(y), _p(out, c_dbl))
        return out

    def run(self, iters, strategy, m, max_source, interval, seed, threads=1, cap=100000):
        ppx = np.zeros(cap)
        steps = np.zeros(cap, dtype=np.int64)
        n = c_i64()
        ph = np.zeros(6)
        rc = lib.orc_run(self.h, iters, strategy, m, max_source, interval, seed, threads,
                         _p(ppx, c_dbl), _p(steps, c_i64), cap, ctypes.byref(n), _p(ph, c_dbl))
        assert rc == 0, rc
        k = min(n.value, cap)
        return list(zip(steps[:k].tolist(), ppx[:k].tolist())), ph
