"""Graphs and held-out/test splits for the sampler.

``Graph`` is the symmetric CSR that replaces the reference's NetworkGraph
(c++/src/mcmc/data.h:135-313); ``Split`` holds the held-out and test EdgeMaps
(network.h:326-327).  Synthetic graphs and the split are produced by the
native library (host C++), not in Python.
"""
import ctypes
import dataclasses

import numpy as np

from . import _lib


@dataclasses.dataclass
class Graph:
    N: int
    row_ptr: np.ndarray  # int64 [N+1]
    cols: np.ndarray     # int32 [2E], rows ascending

    @property
    def num_edges(self):
        return int(self.row_ptr[-1] // 2)

    def degree(self):
        return np.diff(self.row_ptr)

    def to_c(self):
        self.row_ptr = np.ascontiguousarray(self.row_ptr, dtype=np.int64)
        self.cols = np.ascontiguousarray(self.cols, dtype=np.int32)
        g = _lib.Graph()
        g.N = self.N
        g.num_edges = self.num_edges
        g.row_ptr = _lib.ptr(self.row_ptr, ctypes.c_int64)
        g.cols = _lib.ptr(self.cols, ctypes.c_int32)
        return g

    def linked(self, a, b):
        row = self.cols[self.row_ptr[a]:self.row_ptr[a + 1]]
        i = np.searchsorted(row, b)
        return bool(i < len(row) and row[i] == b)

    @classmethod
    def from_edges(cls, N, edges):
        """Symmetric CSR from an undirected edge array [E, 2]; drops self
        loops and duplicates (the reference's relativity loader keeps one
        copy of each undirected edge, relativity.cc:20-193)."""
        e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        e = e[e[:, 0] != e[:, 1]]
        lo = np.minimum(e[:, 0], e[:, 1])
        hi = np.maximum(e[:, 0], e[:, 1])
        key = np.unique(lo * N + hi)
        lo, hi = key // N, key % N
        src = np.concatenate([lo, hi])
        dst = np.concatenate([hi, lo])
        order = np.lexsort((dst, src))
        src, dst = src[order], dst[order]
        row_ptr = np.zeros(N + 1, dtype=np.int64)
        np.add.at(row_ptr, src + 1, 1)
        row_ptr = np.cumsum(row_ptr)
        return cls(N, row_ptr, dst.astype(np.int32))

    @classmethod
    def from_edge_list_file(cls, path, contiguous=False):
        """SNAP edge list read by the native loader with the reference's
        "relativity" rules (relativity.cc:20-193): the first 4 lines are the
        header, then two integer ids per line, plain or gzip; contiguous ids
        are kept (N = max id + 1), others are renumbered in order of first
        appearance; self-links and repeated pairs are dropped."""
        N = ctypes.c_int32()
        E = ctypes.c_int64()
        rp = ctypes.POINTER(ctypes.c_int64)()
        cl = ctypes.POINTER(ctypes.c_int32)()
        self_links, dups = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.lib.ammsb_graph_load_edge_list(
            str(path).encode(), 1 if contiguous else 0, ctypes.byref(N), ctypes.byref(E),
            ctypes.byref(rp), ctypes.byref(cl), ctypes.byref(self_links), ctypes.byref(dups)))
        g = cls(int(N.value), _take(rp, N.value + 1, np.int64), _take(cl, 2 * E.value, np.int32))
        g.self_links, g.duplicates = int(self_links.value), int(dups.value)
        return g

    @classmethod
    def from_options(cls, options):
        """The reference's DataFactory (data_factory.cc:7-60) for its edge-list
        classes: input_class relativity / rc / rz / rcz (contiguous and
        compressed switches), reading options.input_filename.  The XML
        (netscience) and sparsehash classes are dataset parsers outside the
        sampler and are refused."""
        klass = options.input_class
        contiguous = options.input_contiguous or klass in ("rc", "rcz")
        if klass in ("relativity", "rc", "rz", "rcz"):
            return cls.from_edge_list_file(options.input_filename, contiguous=contiguous)
        if klass == "netscience":
            raise NotImplementedError(
                "input class 'netscience' is an XML dataset parser, outside this library; "
                "convert the graph to an edge list")
        if klass in ("sparsehash", "gz", "preprocessed"):
            raise NotImplementedError(
                f"input class {klass!r} is google sparsehash's serialized table format, which "
                "only sparsehash itself reads; convert the graph to an edge list")
        raise ValueError(f'Unknown dataset name "{klass}"')


@dataclasses.dataclass
class Split:
    held_out_edges: np.ndarray   # int32 [H, 2], a < b
    held_out_labels: np.ndarray  # uint8 [H]
    test_edges: np.ndarray       # int32 [T, 2]
    test_labels: np.ndarray      # uint8 [T]

    def to_c(self):
        self.held_out_edges = np.ascontiguousarray(self.held_out_edges, dtype=np.int32)
        self.held_out_labels = np.ascontiguousarray(self.held_out_labels, dtype=np.uint8)
        self.test_edges = np.ascontiguousarray(self.test_edges, dtype=np.int32)
        self.test_labels = np.ascontiguousarray(self.test_labels, dtype=np.uint8)
        s = _lib.Split()
        s.held_out_count = len(self.held_out_labels)
        s.held_out_edges = _lib.ptr(self.held_out_edges, ctypes.c_int32)
        s.held_out_labels = _lib.ptr(self.held_out_labels, ctypes.c_uint8)
        s.test_count = len(self.test_labels)
        s.test_edges = _lib.ptr(self.test_edges, ctypes.c_int32)
        s.test_labels = _lib.ptr(self.test_labels, ctypes.c_uint8)
        return s


def _take(p, n, dtype):
    """Copy n elements from a library-owned pointer and free it."""
    out = np.ctypeslib.as_array(p, shape=(max(n, 1),))[:n].astype(dtype, copy=True)
    _lib.lib.ammsb_free(ctypes.cast(p, ctypes.c_void_p))
    return out


def generate(kind="community", N=1000, K=16, gen_alpha=None, beta_lo=0.02, beta_hi=0.08,
             delta=1e-7, mean_degree=20.0, max_comms=3, seed=0):
    """Synthetic graph from the native generators (see include/ammsb.h):
    'planted' -- pi_a ~ Dir(gen_alpha), per-pair link probability
                 sum_k pi_ak pi_bk beta_k + delta (1 - sum_k pi_ak pi_bk);
    'community' -- 1..max_comms communities per node, Erdos-Renyi inside
                 each with p for mean_degree, plus delta background pairs."""
    p = _lib.GenParams()
    p.kind = _lib.GEN_PLANTED if kind == "planted" else _lib.GEN_COMMUNITY
    p.N, p.K = int(N), int(K)
    p.gen_alpha = (0.05 if kind == "planted" else 0.1) if gen_alpha is None else gen_alpha
    p.beta_lo, p.beta_hi, p.delta = beta_lo, beta_hi, delta
    p.mean_degree, p.max_comms, p.seed = mean_degree, max_comms, seed
    rp = ctypes.POINTER(ctypes.c_int64)()
    cl = ctypes.POINTER(ctypes.c_int32)()
    E = ctypes.c_int64()
    _lib.check(_lib.lib.ammsb_graph_generate(ctypes.byref(p), ctypes.byref(rp), ctypes.byref(cl),
                                             ctypes.byref(E)))
    row_ptr = _take(rp, N + 1, np.int64)
    cols = _take(cl, 2 * E.value, np.int32)
    return Graph(int(N), row_ptr, cols)


def make_split(graph, held_out_ratio=0.1, seed=42):
    """Held-out/test split with the reference's semantics (network.cc:503-614)."""
    g = graph.to_c()
    he = ctypes.POINTER(ctypes.c_int32)()
    hl = ctypes.POINTER(ctypes.c_uint8)()
    te = ctypes.POINTER(ctypes.c_int32)()
    tl = ctypes.POINTER(ctypes.c_uint8)()
    hc = ctypes.c_int64()
    tc = ctypes.c_int64()
    _lib.check(_lib.lib.ammsb_split_generate(ctypes.byref(g), held_out_ratio, seed, ctypes.byref(he),
                                             ctypes.byref(hl), ctypes.byref(hc), ctypes.byref(te),
                                             ctypes.byref(tl), ctypes.byref(tc)))
    H, T = hc.value, tc.value
    return Split(_take(he, 2 * H, np.int32).reshape(H, 2), _take(hl, H, np.uint8),
                 _take(te, 2 * T, np.int32).reshape(T, 2), _take(tl, T, np.uint8))


def philox(ctr, key):
    """One Philox4x32-10 block from the library (the samplers' stream)."""
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    _lib.lib.ammsb_philox(_lib.ptr(c, ctypes.c_uint32), _lib.ptr(k, ctypes.c_uint32),
                          _lib.ptr(out, ctypes.c_uint32))
    return out
