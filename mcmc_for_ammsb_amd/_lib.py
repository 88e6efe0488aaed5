"""ctypes binding of libammsb.so (the C-ABI declared in include/ammsb.h).

The shared library is built in-tree by ``mcmc_for_ammsb_amd.build`` and is
required: there is no CPU fallback, and importing the package without it
raises immediately.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libammsb.so")

AMMSB_OK = 0
AMMSB_ERR_ARG = -1
AMMSB_ERR_OOM = -2
AMMSB_ERR_HIP = -3
AMMSB_ERR_INVARIANT = -4
AMMSB_ERR_STATE = -5

STRATIFIED_RANDOM_NODE = 0
STRATIFIED_RANDOM_NODE_LINKS = 1
STRATIFIED_RANDOM_NODE_NONLINKS = 2
RANDOM_EDGE = 3

FP64 = 0
FP32 = 1

GEN_PLANTED = 0
GEN_COMMUNITY = 1

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
P_i32 = ctypes.POINTER(ctypes.c_int32)
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_u8 = ctypes.POINTER(ctypes.c_uint8)
P_dbl = ctypes.POINTER(ctypes.c_double)


class Options(ctypes.Structure):
    _fields_ = [
        ("alpha", c_dbl), ("eta0", c_dbl), ("eta1", c_dbl), ("epsilon", c_dbl),
        ("a", c_dbl), ("b", c_dbl), ("c", c_dbl),
        ("K", c_i64), ("mini_batch_size", c_i64), ("num_node_sample", c_i64),
        ("strategy", c_i32), ("sampler_max_source", c_i64), ("max_iteration", c_i64),
        ("interval", c_i64), ("stats_print_interval", c_i64), ("num_updates", c_i64),
        ("held_out_ratio", c_dbl), ("random_seed", c_i32), ("convergence_threshold", c_dbl),
        ("precision", c_i32), ("device", c_i32), ("verbose", c_i32),
    ]


class Input(ctypes.Structure):
    _fields_ = [("filename", ctypes.c_char * 4096), ("input_class", ctypes.c_char * 64),
                ("contiguous", c_i32), ("compressed", c_i32)]


class Graph(ctypes.Structure):
    _fields_ = [("N", c_i32), ("num_edges", c_i64), ("row_ptr", P_i64), ("cols", P_i32)]


class Split(ctypes.Structure):
    _fields_ = [
        ("held_out_count", c_i64), ("held_out_edges", P_i32), ("held_out_labels", P_u8),
        ("test_count", c_i64), ("test_edges", P_i32), ("test_labels", P_u8),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("N", c_i32), ("K", c_i64), ("num_edges", c_i64), ("ld", c_i64),
        ("mini_batch_size", c_i64), ("num_node_sample", c_i64), ("interval", c_i64),
        ("stats_print_interval", c_i64), ("max_minibatch_edges", c_i64),
        ("max_minibatch_nodes", c_i64), ("held_out_count", c_i64), ("test_count", c_i64),
        ("step_count", c_i64), ("average_count", c_i64),
        ("alpha", c_dbl), ("a", c_dbl), ("b", c_dbl), ("c", c_dbl), ("epsilon", c_dbl),
        ("eta0", c_dbl), ("eta1", c_dbl),
        ("precision", c_i32), ("strategy", c_i32), ("sampler_max_source", c_i64),
        ("max_iteration", c_i64),
    ]


class GenParams(ctypes.Structure):
    _fields_ = [
        ("kind", c_i32), ("N", c_i32), ("K", c_i32), ("gen_alpha", c_dbl),
        ("beta_lo", c_dbl), ("beta_hi", c_dbl), ("delta", c_dbl), ("mean_degree", c_dbl),
        ("max_comms", c_i32), ("seed", c_u64),
    ]


class BatchShape(ctypes.Structure):
    _fields_ = [("lanes_per_row", c_i32), ("vectors_per_lane", c_i32),
                ("nodes_per_block", c_i32), ("us", c_dbl)]


class AmmsbError(RuntimeError):
    """Raised for every non-zero status of the C-ABI."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _share_torch_hip_runtime():
    """PyTorch-ROCm wheels bundle their own libamdhip64 (soname
    libamdhip64.so.7, the same as /opt/rocm's).  Two HIP runtimes in one
    process each claim the GPU and the second one finds none, and device
    pointers and streams are only meaningful inside one runtime.  Mapping
    torch's copy before libammsb.so makes the loader satisfy our
    libamdhip64.so.7 dependency with it (match by soname), so the sampler and
    torch tensors share one runtime whichever is imported first.  torch itself
    is not imported.  AMMSB_HIP_RUNTIME=system keeps /opt/rocm's runtime."""
    if os.environ.get("AMMSB_HIP_RUNTIME", "torch") == "system":
        return None
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return None
    if spec is None or spec.origin is None:
        return None
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if not os.path.exists(path):
        return None
    ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    return path


HIP_RUNTIME = None


def _load():
    global HIP_RUNTIME
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python mcmc_for_ammsb_amd/build.py` "
            "(the sampler has no CPU fallback)")
    HIP_RUNTIME = _share_torch_hip_runtime()
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "ammsb_last_error": (ctypes.c_char_p, []),
        "ammsb_version": (ctypes.c_char_p, []),
        "ammsb_options_default": (c_i32, [ctypes.POINTER(Options)]),
        "ammsb_options_parse": (c_i32, [c_i32, ctypes.POINTER(ctypes.c_char_p),
                                        ctypes.POINTER(Options), ctypes.POINTER(Input),
                                        ctypes.POINTER(c_i32)]),
        "ammsb_graph_generate": (c_i32, [ctypes.POINTER(GenParams), ctypes.POINTER(P_i64),
                                         ctypes.POINTER(P_i32), P_i64]),
        "ammsb_split_generate": (c_i32, [ctypes.POINTER(Graph), c_dbl, c_u64,
                                         ctypes.POINTER(P_i32), ctypes.POINTER(P_u8), P_i64,
                                         ctypes.POINTER(P_i32), ctypes.POINTER(P_u8), P_i64]),
        "ammsb_free": (None, [c_vp]),
        "ammsb_graph_load_edge_list": (c_i32, [ctypes.c_char_p, c_i32, ctypes.POINTER(c_i32),
                                               P_i64, ctypes.POINTER(P_i64),
                                               ctypes.POINTER(P_i32), P_i64, P_i64]),
        "ammsb_philox": (None, [ctypes.POINTER(ctypes.c_uint32)] * 3),
        "ammsb_create": (c_i32, [ctypes.POINTER(Options), ctypes.POINTER(Graph),
                                 ctypes.POINTER(Split), ctypes.POINTER(c_vp)]),
        "ammsb_destroy": (c_i32, [c_vp]),
        "ammsb_get_info": (c_i32, [c_vp, ctypes.POINTER(Info)]),
        "ammsb_init_state": (c_i32, [c_vp, c_u64]),
        "ammsb_set_state": (c_i32, [c_vp, c_vp, P_dbl]),
        "ammsb_get_state": (c_i32, [c_vp, c_vp, P_dbl, P_dbl]),
        "ammsb_get_rows": (c_i32, [c_vp, P_i32, c_i64, c_vp]),
        "ammsb_set_counters": (c_i32, [c_vp, c_i64, c_i64]),
        "ammsb_step": (c_i32, [c_vp, c_i64, c_i32]),
        "ammsb_perplexity": (c_i32, [c_vp, P_dbl]),
        "ammsb_run": (c_i32, [c_vp, c_i64, c_i32, P_dbl, P_i64, c_i64, P_i64]),
        "ammsb_timers": (c_i32, [c_vp, P_dbl, P_i64]),
        "ammsb_enable_timers": (c_i32, [c_vp, c_i32]),
        "ammsb_update_phi": (c_i32, [c_vp, P_i32, c_i64, P_i32, c_i64, c_vp, c_dbl]),
        "ammsb_update_beta": (c_i32, [c_vp, P_i32, P_u8, c_i64, c_dbl, P_dbl, c_dbl]),
        "ammsb_debug_sample": (c_i32, [c_vp, c_i64, P_i32, P_dbl, P_i32, P_i64, P_i32, P_i64,
                                       P_i32]),
        "ammsb_get_totals": (c_i32, [c_vp, P_i64, P_i64]),
        "ammsb_loglik_fwd": (c_i32, [c_vp, c_i64, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64,
                                     c_vp, c_vp, c_vp]),
        "ammsb_loglik_bwd": (c_i32, [c_vp, c_i64, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64,
                                     c_vp, c_vp, c_vp, c_vp]),
        "ammsb_loglik_bwd_noalloc": (c_i32, [c_vp, c_i64, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp,
                                             c_i64, c_vp, c_vp, c_vp, c_vp]),
        "ammsb_probe_read": (c_i32, [c_vp, c_i32, P_dbl, P_i64]),
        "ammsb_probe_gather": (c_i32, [c_vp, c_i32, P_dbl, P_i64]),
        "ammsb_set_tuning": (c_i32, [c_vp, ctypes.c_char_p, c_i64]),
        "ammsb_get_tuning": (c_i32, [c_vp, ctypes.c_char_p, P_i64]),
        "ammsb_device_ptrs": (c_i32, [c_vp, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp),
                                      ctypes.POINTER(c_vp)]),
        "ammsb_batch_create": (c_i32, [c_i32, ctypes.POINTER(Options), ctypes.POINTER(Graph),
                                       ctypes.POINTER(Split), ctypes.POINTER(c_vp)]),
        "ammsb_batch_destroy": (c_i32, [c_vp]),
        "ammsb_batch_chain": (c_i32, [c_vp, c_i32, ctypes.POINTER(c_vp)]),
        "ammsb_batch_init_state": (c_i32, [c_vp, ctypes.POINTER(c_u64)]),
        "ammsb_batch_step": (c_i32, [c_vp, c_i64]),
        "ammsb_batch_perplexity": (c_i32, [c_vp, P_dbl]),
        "ammsb_batch_use_graph": (c_i32, [c_vp, c_i32]),
        "ammsb_batch_refill_in_order": (c_i32, [c_vp, c_i32]),
        "ammsb_batch_set_shape": (c_i32, [c_vp, c_i32, ctypes.POINTER(BatchShape)]),
        "ammsb_batch_get_shape": (c_i32, [c_vp, c_i32, ctypes.POINTER(BatchShape)]),
        "ammsb_batch_autotune": (c_i32, [c_vp, c_i32, ctypes.POINTER(BatchShape)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def last_error():
    m = lib.ammsb_last_error()
    return m.decode() if m else ""


def check(rc):
    if rc != AMMSB_OK:
        raise AmmsbError(rc, last_error())
    return rc


def ptr(arr, ctype):
    """ctypes pointer into a contiguous numpy array (or None)."""
    if arr is None:
        return None
    return arr.ctypes.data_as(ctypes.POINTER(ctype))


def vptr(arr):
    if arr is None:
        return None
    return ctypes.c_void_p(arr.ctypes.data)
