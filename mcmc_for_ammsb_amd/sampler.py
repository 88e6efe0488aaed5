"""MCMCSamplerStochastic: the Python face of the device sampler.

Mirrors the reference's learner plugin
  ``MCMCSamplerStochastic(const Options&)``, ``init()``, ``run()``
  (c++/src/mcmc/learning/mcmc_sampler_stochastic.h:72-80) on top of the
C-ABI of libammsb.so.  Everything numeric runs in the native library on the
GPU; this class only owns the context handle and moves host arrays.
"""
import ctypes

import numpy as np

from . import _lib
from .graph import Graph, Split, make_split
from .options import Options

PHASES = ("outer", "perplexity", "sample_mini_batch", "nodes_in_mini_batch",
          "sample_neighbor_nodes", "update_phi", "update_pi", "update_beta")


# set in this order: each launch-shape pair is validated on its second key
TUNING_KEYS = ("phi_waves", "phi_unroll", "ll_fwd_rows", "ll_bwd_rows", "ll_fwd_pairs",
               "ll_fwd_prefetch", "ll_bwd_pairs",
               "ll_bwd_prefetch")


class MCMCSamplerStochastic:
    """SGRLD sampler for the assortative MMSB with HBM-resident state.

    ``options``: an :class:`Options` (or an argv list for ``Options.parse``).
    ``graph``: a :class:`Graph`; ``split``: a :class:`Split` or None (drawn
    by the library with the reference's semantics, network.cc:503-614).
    """

    def __init__(self, options, graph, split=None, _borrowed=None):
        if _borrowed is not None:  # a chain of an MCMCBatch (the batch owns it)
            self.options, self.graph, self.split = options, graph, split
            self._h = _borrowed
            self._owned = False
            self.dtype = np.float64 if options.precision == "fp64" else np.float32
            return
        self._owned = True
        if not isinstance(options, Options):
            options = Options.parse(options)
        if not isinstance(graph, Graph):
            raise TypeError("graph must be a mcmc_for_ammsb_amd.Graph")
        self.options = options
        self.graph = graph
        if split is None:
            ratio = options.held_out_ratio if options.held_out_ratio != 0.0 else 0.1
            split = make_split(graph, ratio, options.random_seed)
        self.split = split
        self._opts_c = options.to_c()
        self._g_c = graph.to_c()
        self._s_c = split.to_c()
        h = ctypes.c_void_p()
        _lib.check(_lib.lib.ammsb_create(ctypes.byref(self._opts_c), ctypes.byref(self._g_c),
                                         ctypes.byref(self._s_c), ctypes.byref(h)))
        self._h = h
        self.dtype = np.float64 if options.precision == "fp64" else np.float32

    # ---- lifecycle -------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                _lib.lib.ammsb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- reference interface ----------------------------------------------
    def init(self, seed=None):
        """MCMCSamplerStochastic::init() state draw (mcmc_sampler_stochastic.cc:54-76)."""
        seed = self.options.random_seed if seed is None else seed
        _lib.check(_lib.lib.ammsb_init_state(self._h, int(seed) & 0xFFFFFFFFFFFFFFFF))

    def run(self, max_steps=0, print_log=True, max_checkpoints=100000):
        """MCMCSamplerStochastic::run() (mcmc_sampler_stochastic.cc:93-183).
        Returns [(step_count, perplexity)] of the held-out checks."""
        ppx = np.zeros(max_checkpoints, dtype=np.float64)
        steps = np.zeros(max_checkpoints, dtype=np.int64)
        n = ctypes.c_int64()
        _lib.check(_lib.lib.ammsb_run(self._h, int(max_steps), 1 if print_log else 0,
                                      _lib.ptr(ppx, ctypes.c_double),
                                      _lib.ptr(steps, ctypes.c_int64), max_checkpoints,
                                      ctypes.byref(n)))
        k = min(n.value, max_checkpoints)
        return list(zip(steps[:k].tolist(), ppx[:k].tolist()))

    def step(self, iters=1, sync=True):
        """``iters`` iterations of the loop body without perplexity checks."""
        _lib.check(_lib.lib.ammsb_step(self._h, int(iters), 1 if sync else 0))

    def perplexity(self):
        """Learner::cal_perplexity_held_out (learner.cc:126-183)."""
        out = ctypes.c_double()
        _lib.check(_lib.lib.ammsb_perplexity(self._h, ctypes.byref(out)))
        return out.value

    # ---- state -----------------------------------------------------------
    @property
    def info(self):
        i = _lib.Info()
        _lib.check(_lib.lib.ammsb_get_info(self._h, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in _lib.Info._fields_}

    def get_state(self):
        """(pi [N][K+1] with phi_sum in slot K, theta [K][2], beta [K])."""
        inf = self.info
        N, K = inf["N"], inf["K"]
        pi = np.empty((N, K + 1), dtype=self.dtype)
        theta = np.empty((K, 2), dtype=np.float64)
        beta = np.empty(K, dtype=np.float64)
        _lib.check(_lib.lib.ammsb_get_state(self._h, _lib.vptr(pi), _lib.ptr(theta, ctypes.c_double),
                                            _lib.ptr(beta, ctypes.c_double)))
        return pi, theta, beta

    def set_state(self, pi, theta):
        pi = np.ascontiguousarray(pi, dtype=self.dtype)
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        inf = self.info
        if pi.shape != (inf["N"], inf["K"] + 1) or theta.shape != (inf["K"], 2):
            raise ValueError("pi must be [N][K+1] and theta [K][2]")
        _lib.check(_lib.lib.ammsb_set_state(self._h, _lib.vptr(pi), _lib.ptr(theta, ctypes.c_double)))

    def get_rows(self, nodes):
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        K = self.info["K"]
        out = np.empty((len(nodes), K + 1), dtype=self.dtype)
        _lib.check(_lib.lib.ammsb_get_rows(self._h, _lib.ptr(nodes, ctypes.c_int32), len(nodes),
                                           _lib.vptr(out)))
        return out

    def set_counters(self, step_count, average_count):
        _lib.check(_lib.lib.ammsb_set_counters(self._h, int(step_count), int(average_count)))

    # ---- single operators with injected draws ------------------------------
    def update_phi(self, nodes, nbrs, noise, eps_t):
        """update_phi + update_pi for distinct ``nodes`` with neighbours
        ``nbrs`` [n][n_nb] and noise [n][K] (None: the Philox noise)."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        nbrs = np.ascontiguousarray(nbrs, dtype=np.int32)
        if nbrs.ndim != 2 or nbrs.shape[0] != len(nodes):
            raise ValueError("nbrs must be [len(nodes)][n_nb]")
        nz = None if noise is None else np.ascontiguousarray(noise, dtype=self.dtype)
        _lib.check(_lib.lib.ammsb_update_phi(self._h, _lib.ptr(nodes, ctypes.c_int32), len(nodes),
                                             _lib.ptr(nbrs, ctypes.c_int32), nbrs.shape[1],
                                             _lib.vptr(nz), float(eps_t)))

    def update_beta(self, edges, y, scale, noise, eps_t):
        """update_beta for minibatch ``edges`` [E][2] with labels ``y``."""
        edges = np.ascontiguousarray(edges, dtype=np.int32).reshape(-1, 2)
        y = np.ascontiguousarray(y, dtype=np.uint8)
        nz = None if noise is None else np.ascontiguousarray(noise, dtype=np.float64)
        _lib.check(_lib.lib.ammsb_update_beta(self._h, _lib.ptr(edges, ctypes.c_int32),
                                              _lib.ptr(y, ctypes.c_uint8), len(y), float(scale),
                                              _lib.ptr(nz, ctypes.c_double), float(eps_t)))

    def debug_sample(self, step):
        """Device minibatch + neighbour draws of iteration ``step``:
        dict(flag, scale, edges [E][2], nodes [n] (unordered),
        nbrs [n][n+1] with bit 31 = link)."""
        inf = self.info
        ecap, ncap, nnb = inf["max_minibatch_edges"], inf["max_minibatch_nodes"], inf["num_node_sample"] + 1
        edges = np.zeros((ecap, 2), dtype=np.int32)
        nodes = np.zeros(ncap, dtype=np.int32)
        nbrs = np.zeros((ncap, nnb), dtype=np.int32)
        flag = ctypes.c_int32()
        scale = ctypes.c_double()
        ne = ctypes.c_int64()
        nn = ctypes.c_int64()
        _lib.check(_lib.lib.ammsb_debug_sample(self._h, int(step), ctypes.byref(flag), ctypes.byref(scale),
                                               _lib.ptr(edges, ctypes.c_int32), ctypes.byref(ne),
                                               _lib.ptr(nodes, ctypes.c_int32), ctypes.byref(nn),
                                               _lib.ptr(nbrs, ctypes.c_int32)))
        return dict(flag=flag.value, scale=scale.value, edges=edges[:ne.value].copy(),
                    nodes=nodes[:nn.value].copy(), nbrs=nbrs[:nn.value].copy())

    # ---- timers -----------------------------------------------------------
    def enable_timers(self, on=True):
        """Per-phase HIP-event timers: True/1 every phase, 2 update_phi only,
        False/0 off."""
        mode = 2 if on == 2 else (1 if on else 0)
        _lib.check(_lib.lib.ammsb_enable_timers(self._h, mode))

    def timers(self):
        secs = np.zeros(8, dtype=np.float64)
        ticks = np.zeros(8, dtype=np.int64)
        _lib.check(_lib.lib.ammsb_timers(self._h, _lib.ptr(secs, ctypes.c_double),
                                         _lib.ptr(ticks, ctypes.c_int64)))
        return {p: (float(s), int(t)) for p, s, t in zip(PHASES, secs, ticks)}

    def set_tuning(self, **kw):
        """Launch shapes, e.g. set_tuning(phi_waves=8, phi_unroll=2) for update_phi
        or set_tuning(ll_bwd_pairs=2, ll_bwd_prefetch=1) for the log-likelihood adjoint."""
        for name in TUNING_KEYS:
            if name in kw:
                _lib.check(_lib.lib.ammsb_set_tuning(self._h, name.encode(), int(kw.pop(name))))
        if kw:
            raise ValueError(f"unknown tuning parameters {sorted(kw)}")

    def get_tuning(self):
        out = {}
        for name in TUNING_KEYS:
            v = ctypes.c_int64()
            _lib.check(_lib.lib.ammsb_get_tuning(self._h, name.encode(), ctypes.byref(v)))
            out[name] = v.value
        return out

    def probe_read(self, reps=5):
        """(seconds, bytes) of one contiguous read of the resident pi array."""
        sec, nb = ctypes.c_double(), ctypes.c_int64()
        _lib.check(_lib.lib.ammsb_probe_read(self._h, int(reps), ctypes.byref(sec), ctypes.byref(nb)))
        return sec.value, nb.value

    def probe_gather(self, reps=5):
        """(seconds, bytes) of update_phi's row gather for the current step
        without its arithmetic: the practical ceiling of that access pattern."""
        sec, nb = ctypes.c_double(), ctypes.c_int64()
        _lib.check(_lib.lib.ammsb_probe_gather(self._h, int(reps), ctypes.byref(sec),
                                               ctypes.byref(nb)))
        return sec.value, nb.value

    def totals(self):
        """(nodes updated, minibatch edges) summed over all iterations."""
        n, e = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.lib.ammsb_get_totals(self._h, ctypes.byref(n), ctypes.byref(e)))
        return n.value, e.value

    def device_ptrs(self):
        pi, th, be = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(_lib.lib.ammsb_device_ptrs(self._h, ctypes.byref(pi), ctypes.byref(th), ctypes.byref(be)))
        return pi.value, th.value, be.value


class MCMCBatch:
    """Many independent SGRLD chains advanced together on one GPU.

    ``options``: one :class:`Options` per chain (same precision); ``graphs``
    and ``splits`` likewise (a split of None is drawn with the reference's
    semantics from the chain's held-out ratio and seed).  Each iteration of
    every chain is one launch per phase over all chains, replayed from a
    captured hipGraph.  ``chain(i)`` is a :class:`MCMCSamplerStochastic` view
    of chain i (state, perplexity, info) owned by the batch.
    """

    def __init__(self, options, graphs, splits=None):
        n = len(options)
        if len(graphs) != n or (splits is not None and len(splits) != n):
            raise ValueError("one graph (and split) per chain")
        if splits is None:
            splits = [None] * n
        self.options = list(options)
        self.graphs = list(graphs)
        self.splits = [s if s is not None else
                       make_split(g, o.held_out_ratio if o.held_out_ratio != 0.0 else 0.1,
                                  o.random_seed)
                       for o, g, s in zip(self.options, self.graphs, splits)]
        self._opts_c = (_lib.Options * n)(*[o.to_c() for o in self.options])
        self._g_c = (_lib.Graph * n)(*[g.to_c() for g in self.graphs])
        self._s_c = (_lib.Split * n)(*[s.to_c() for s in self.splits])
        h = ctypes.c_void_p()
        _lib.check(_lib.lib.ammsb_batch_create(n, self._opts_c, self._g_c, self._s_c,
                                               ctypes.byref(h)))
        self._h = h
        self.n = n
        self._chains = []
        for i in range(n):
            c = ctypes.c_void_p()
            _lib.check(_lib.lib.ammsb_batch_chain(self._h, i, ctypes.byref(c)))
            self._chains.append(MCMCSamplerStochastic(self.options[i], self.graphs[i],
                                                      self.splits[i], _borrowed=c))

    def close(self):
        if getattr(self, "_h", None):
            for c in self._chains:
                c._h = None
            _lib.lib.ammsb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return self.n

    def chain(self, i):
        return self._chains[i]

    def init(self, seeds=None):
        """State draw of every chain (seeds: one per chain, default each
        chain's random_seed)."""
        arr = None
        if seeds is not None:
            arr = (ctypes.c_uint64 * self.n)(*[int(x) & 0xFFFFFFFFFFFFFFFF for x in seeds])
        _lib.check(_lib.lib.ammsb_batch_init_state(self._h, arr))

    def step(self, iters=1):
        _lib.check(_lib.lib.ammsb_batch_step(self._h, int(iters)))

    def perplexity(self):
        out = np.empty(self.n, dtype=np.float64)
        _lib.check(_lib.lib.ammsb_batch_perplexity(self._h, _lib.ptr(out, ctypes.c_double)))
        return out

    def use_graph(self, enable=True):
        _lib.check(_lib.lib.ammsb_batch_use_graph(self._h, 1 if enable else 0))

    def refill_in_order(self, enable=True):
        """Draw the ring's next 16 iterations on the batch's own stream (in
        order) instead of overlapping them on a second stream (the default)."""
        _lib.check(_lib.lib.ammsb_batch_refill_in_order(self._h, 1 if enable else 0))

    def get_shape(self, i):
        sh = _lib.BatchShape()
        _lib.check(_lib.lib.ammsb_batch_get_shape(self._h, int(i), ctypes.byref(sh)))
        return {f: getattr(sh, f) for f, _ in _lib.BatchShape._fields_}

    def set_shape(self, i, lanes_per_row, vectors_per_lane, nodes_per_block):
        sh = _lib.BatchShape(int(lanes_per_row), int(vectors_per_lane), int(nodes_per_block), 0.0)
        _lib.check(_lib.lib.ammsb_batch_set_shape(self._h, int(i), ctypes.byref(sh)))

    def autotune(self, reps=5):
        """Times every valid update_phi shape of each chain on its current
        minibatch and keeps the fastest; returns the chosen shapes."""
        out = (_lib.BatchShape * self.n)()
        _lib.check(_lib.lib.ammsb_batch_autotune(self._h, int(reps), out))
        return [{f: getattr(o, f) for f, _ in _lib.BatchShape._fields_} for o in out]
