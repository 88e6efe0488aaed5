"""Host-side checks of libammsb.so (no GPU): exported symbols, defaults,
options parsing, synthetic generators, CSR construction and the split."""
import ctypes
import os
import re

import numpy as np
import pytest

import mcmc_for_ammsb_amd as M
from mcmc_for_ammsb_amd import _lib
from mcmc_for_ammsb_amd.options import Options, parse_size

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "ammsb.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ammsb_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 25
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_options_default_matches_reference():
    # c++/src/mcmc/options.h:139-189
    o = _lib.Options()
    assert _lib.lib.ammsb_options_default(ctypes.byref(o)) == 0
    assert (o.alpha, o.eta0, o.eta1, o.epsilon) == (0.0, 1.0, 1.0, 1e-7)
    assert (o.a, o.b, o.c, o.K) == (0.0, 1024.0, 0.5, 300)
    assert (o.mini_batch_size, o.num_node_sample, o.strategy, o.sampler_max_source) == (0, 0, 0, 1)
    assert (o.max_iteration, o.interval, o.stats_print_interval, o.num_updates) == (10000000, 0, 0, 1000)
    assert (o.held_out_ratio, o.random_seed, o.convergence_threshold) == (0.0, 42, 1e-12)


def test_options_parse_flags_aliases_config(tmp_path):
    o = Options.parse(["--mcmc.K", "32", "-m", "500", "--mcmc.num-node-sample=32", "--mcmc.a", "0.01",
                       "--mcmc.c=0.55", "--mcmc.strategy", "stratified-random-node-links",
                       "--mcmc.sampler.max-source", "8", "-x", "1k", "--unknown", "7", "pos"])
    assert (o.K, o.mini_batch_size, o.num_node_sample, o.a, o.c) == (32, 500, 32, 0.01, 0.55)
    assert o.strategy == "stratified-random-node-links" and o.sampler_max_source == 8
    assert o.max_iteration == 1024
    assert o.remains == ["--unknown", "7", "pos"]
    cfg = tmp_path / "c.cfg"
    cfg.write_text("[mcmc]\nK = 16\nepsilon = 1e-6\ninterval=10\n")
    o = Options.parse(["--config", str(cfg), "--mcmc.K", "64"])
    assert o.K == 64 and o.epsilon == 1e-6 and o.interval == 10  # command line wins
    a, alpha = Options(K=64, b=1024.0, c=0.5).derived()
    assert a == pytest.approx(1024 ** -0.5) and alpha == 1 / 64
    with pytest.raises(ValueError):
        Options.parse(["--mcmc.strategy", "nope"])
    # the reference's parse_size_t: powers of 1024 for k/kb..., 0x hex, 0 octal
    assert parse_size("2M") == 2 * 1024 ** 2 and parse_size("3kB") == 3 * 1024
    assert parse_size("0x1b") == 27 and parse_size("010") == 8 and parse_size("1g") == 1 << 30
    for bad in ("-5", "1q", "2kk"):
        with pytest.raises(ValueError):
            parse_size(bad)


def _native_parse(argv):
    """ammsb_options_parse (the C++ restatement) on argv."""
    from mcmc_for_ammsb_amd import _lib
    import ctypes
    arr = (ctypes.c_char_p * max(1, len(argv)))(*[a.encode() for a in argv])
    o, inp, n = _lib.Options(), _lib.Input(), ctypes.c_int32()
    rc = _lib.lib.ammsb_options_parse(len(argv), arr, ctypes.byref(o), ctypes.byref(inp),
                                      ctypes.byref(n))
    return rc, o, inp, n.value


_INT_TEXT = "0123456789xXabcdefkKmMgGbB-"
_FLOAT_TEXT = "0123456789.eE-+"


def _py_parse(argv):
    try:
        return Options.parse(argv)
    except ValueError:
        return None


@pytest.mark.parametrize("flag,field", [("--mcmc.max-iteration", "max_iteration"),
                                        ("--mcmc.seed", "random_seed"),
                                        ("--mcmc.epsilon", "epsilon")])
def test_native_and_python_value_parsing_agree(flag, field):
    """Random value strings (the reference's integer forms and their
    mistakes, floats): the C++ and the Python parser accept the same ones
    and agree on the value.  Integer options follow parse<T>
    (options.h:39-57, options.cc:5-67), float options lexical_cast."""
    hypothesis = pytest.importorskip("hypothesis")
    from hypothesis import strategies as st

    alphabet = _FLOAT_TEXT if field == "epsilon" else _INT_TEXT

    @hypothesis.settings(max_examples=400, deadline=None, derandomize=True)
    @hypothesis.given(st.text(alphabet=alphabet, min_size=1, max_size=12))
    def check(text):
        py = _py_parse([flag, text])
        rc, o, _, _ = _native_parse([flag, text])
        assert (rc == 0) == (py is not None), text
        if py is not None:
            assert getattr(o, field) == getattr(py, field), text

    check()
    for text in ("0x10", "010", "019", "1k", "2KB", "0x1gb", "-0", "-", "0x", "1m1",
                 "18446744073709551615", "18446744073709551616", "-2147483648",
                 "-2147483649", "2147483647", "2147483648", "17179869184g"):
        py = _py_parse([flag, text])
        rc, o, _, _ = _native_parse([flag, text])
        assert (rc == 0) == (py is not None), text
        if py is not None:
            assert getattr(o, field) == getattr(py, field), text


def test_native_options_parser_matches_python(tmp_path):
    """The C++ parser (for C/C++ hosts replacing the reference's main) and
    Options.parse agree field by field, including the config file, aliases,
    the reference's integer forms and the count of unrecognised arguments."""
    cfg = tmp_path / "c.cfg"
    cfg.write_text("# comment\n[mcmc]\nK = 16\nepsilon = 1e-6\ninterval=10\n"
                   "input.compressed = true\nseed = 7\n")
    cases = [
        ["--mcmc.K", "32", "-m", "500", "--mcmc.num-node-sample=32", "--mcmc.a", "0.01",
         "--mcmc.c=0.55", "--mcmc.strategy", "stratified-random-node-links",
         "--mcmc.sampler.max-source", "8", "-x", "1k", "--unknown", "7", "pos"],
        ["-K32", "-n0x10", "-i", "010", "--mcmc.seed=-3", "--mcmc.held-out-ratio", "0.2",
         "-f", "/data/x.txt.gz", "-c", "rz", "--mcmc.input.contiguous", "--mcmc.precision",
         "fp32", "--verbose", "--mcmc.stats", "2k", "--mcmc.strategy", "random-edge"],
        ["--config", str(cfg), "-K", "64", "--mcmc.eta0=2.5"],
        [],
    ]
    fields = ("alpha", "eta0", "eta1", "epsilon", "a", "b", "c", "K", "mini_batch_size",
              "num_node_sample", "sampler_max_source", "max_iteration", "interval",
              "stats_print_interval", "num_updates", "held_out_ratio", "random_seed",
              "convergence_threshold", "device")
    for argv in cases:
        py = Options.parse(argv)
        rc, o, inp, nun = _native_parse(argv)
        assert rc == 0, argv
        for f in fields:
            assert getattr(o, f) == getattr(py, f), (argv, f)
        assert o.strategy == M.options.STRATEGIES[py.strategy]
        assert o.precision == (0 if py.precision == "fp64" else 1)
        assert o.verbose == int(py.verbose)
        assert inp.filename.decode() == py.input_filename
        assert inp.input_class.decode() == py.input_class
        assert (inp.contiguous, inp.compressed) == (int(py.input_contiguous),
                                                    int(py.input_compressed))
        assert nun == len(py.remains)
    assert _native_parse(cases[2])[1].K == 64  # the command line (alias) wins over the file
    for bad in (["--mcmc.K", "-5"], ["--mcmc.strategy", "nope"], ["-K", "1q"], ["--mcmc.K"],
                ["--mcmc.precision", "fp16"]):
        assert _native_parse(bad)[0] != 0, bad
        with pytest.raises(ValueError):
            Options.parse(bad)


def _check_csr(g):
    rp, c = g.row_ptr, g.cols
    assert rp[0] == 0 and rp[-1] == len(c) and len(rp) == g.N + 1
    src = np.repeat(np.arange(g.N), np.diff(rp))
    assert np.all(src != c)
    for i in range(0, g.N, max(1, g.N // 200)):
        row = c[rp[i]:rp[i + 1]]
        assert np.all(np.diff(row) > 0)
    # symmetric: multiset of (src, dst) equals (dst, src)
    a = np.sort(src.astype(np.int64) * g.N + c)
    b = np.sort(c.astype(np.int64) * g.N + src)
    assert np.array_equal(a, b)


def test_planted_generator():
    g = M.generate("planted", N=2000, K=16, seed=3)
    _check_csr(g)
    g2 = M.generate("planted", N=2000, K=16, seed=3)
    assert np.array_equal(g.cols, g2.cols)
    # expected edges ~ N^2/2 * E[sum_k pi_ak pi_bk beta_k] with Dir(0.05): ~ 2e6 * 0.05/16*... > 0
    assert g.num_edges > 500


def test_community_generator_mean_degree():
    g = M.generate("community", N=50000, K=64, mean_degree=30.0, seed=1)
    _check_csr(g)
    assert abs(2 * g.num_edges / g.N - 30.0) < 1.5


def test_split_semantics():
    g = M.generate("community", N=20000, K=32, mean_degree=20.0, seed=2)
    s = M.make_split(g, 0.1, 42)
    p = int(0.1 * g.num_edges) // 2  # network.cc:72,504,562
    # held-out: p links + p non-links; test: p + 1 links (init_test_set adds
    # one more after p reaches 0, network.cc:566-580) + p non-links
    assert len(s.held_out_labels) == 2 * p and len(s.test_labels) == 2 * p + 1
    assert s.held_out_labels.sum() == p and s.test_labels.sum() == p + 1
    ho = {tuple(e) for e in s.held_out_edges}
    te = {tuple(e) for e in s.test_edges}
    assert len(ho) == 2 * p and len(te) == 2 * p + 1 and not (ho & te)
    for (a, b), y in zip(s.held_out_edges, s.held_out_labels):
        assert a < b and g.linked(a, b) == bool(y)
    for (a, b), y in zip(s.test_edges, s.test_labels):
        assert a < b and g.linked(a, b) == bool(y)
    s2 = M.make_split(g, 0.1, 42)
    assert np.array_equal(s.held_out_edges, s2.held_out_edges)
    with pytest.raises(M.AmmsbError):
        M.make_split(M.generate("community", N=100, K=4, mean_degree=2.0, seed=1), 3.0, 1)


def test_split_draws_are_uniform():
    """The split's draws over many seeds: every training edge is equally
    likely to be a held-out link and a test link (init_held_out_set /
    init_test_set sample links uniformly, network.cc:503-580), and every
    node about equally likely as a held-out non-link endpoint (non-links are
    uniform pairs, :533-557).  Chi-square statistics within five standard
    deviations of their degrees of freedom; a split that favoured low ids,
    CSR order or high-degree nodes would land far outside."""
    g = M.generate("community", N=200, K=4, mean_degree=6.0, seed=5)
    E = g.num_edges
    src = np.repeat(np.arange(g.N), np.diff(g.row_ptr))
    keep = src < g.cols
    eid = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(src[keep], g.cols[keep]))}
    assert len(eid) == E
    ho = np.zeros(E)
    te = np.zeros(E)
    ends = np.zeros(g.N)
    seeds = 400
    for seed in range(seeds):
        s = M.make_split(g, 0.1, 1000 + seed)
        for (a, b), y in zip(s.held_out_edges, s.held_out_labels):
            if y:
                ho[eid[(int(a), int(b))]] += 1
            else:
                ends[a] += 1
                ends[b] += 1
        for (a, b), y in zip(s.test_edges, s.test_labels):
            if y:
                te[eid[(int(a), int(b))]] += 1

    def chi2_ok(counts, expected):
        x = float((((counts - expected) ** 2) / expected).sum())
        df = len(counts) - 1
        return abs(x - df) < 5 * np.sqrt(2 * df), x

    for counts in (ho, te):
        ok, x = chi2_ok(counts, counts.sum() / E)
        assert ok, x
    # a non-link endpoint: node v is in N-1-deg(v) non-link pairs
    deg = np.diff(g.row_ptr).astype(np.float64)
    w = (g.N - 1 - deg) / (g.N - 1 - deg).sum()
    ok, x = chi2_ok(ends, ends.sum() * w)
    assert ok, x


def test_graph_from_edges_and_file(tmp_path):
    """relativity.cc:20-193: 4 header lines, dense renumbering in order of
    first appearance, self-links and repeated pairs dropped, gzip input,
    contiguous ids (N = max + 1)."""
    import gzip
    body = "10 20\n20 10\n10 30\n30 30\n40 10\n"
    header = "# Directed graph\n# toy\n# Nodes: 4 Edges: 4\n# FromNodeId\tToNodeId\n"
    f = tmp_path / "g.txt"
    f.write_text(header + body)
    g = M.Graph.from_edge_list_file(str(f))
    assert g.N == 4 and g.num_edges == 3 and g.self_links == 1 and g.duplicates == 1
    _check_csr(g)
    assert g.linked(0, 1) and g.linked(0, 2) and g.linked(0, 3) and not g.linked(1, 2)
    fz = tmp_path / "g.txt.gz"
    with gzip.open(fz, "wt") as h:
        h.write(header + body)
    gz = M.Graph.from_edge_list_file(str(fz))
    assert np.array_equal(gz.row_ptr, g.row_ptr) and np.array_equal(gz.cols, g.cols)
    gc = M.Graph.from_edge_list_file(str(f), contiguous=True)
    assert gc.N == 41 and gc.num_edges == 3 and gc.linked(10, 20) and gc.linked(10, 40)
    bad = tmp_path / "bad.txt"
    bad.write_text(header + "1 2\nx y\n")
    with pytest.raises(M.AmmsbError, match="Fail to parse"):
        M.Graph.from_edge_list_file(str(bad))
    with pytest.raises(M.AmmsbError, match="Cannot open"):
        M.Graph.from_edge_list_file(str(tmp_path / "missing.txt"))


def test_graph_from_options(tmp_path):
    with pytest.raises(NotImplementedError):
        M.Graph.from_options(Options.parse(["--mcmc.input.class", "netscience", "-f", "x.xml"]))
    f = tmp_path / "e.txt"
    f.write_text("#\n#\n#\n#\n0 1\n1 2\n")
    g = M.Graph.from_options(Options.parse(["-c", "rc", "-f", str(f)]))
    assert g.N == 3 and g.num_edges == 2
    with pytest.raises(NotImplementedError):
        M.Graph.from_options(Options.parse(["-c", "preprocessed", "-f", str(tmp_path)]))


def test_sampler_requires_graph_type():
    with pytest.raises(TypeError):
        M.MCMCSamplerStochastic(Options(K=4), "not a graph")


def test_make_pairs_invariants():
    """Pairs of the differentiation workload: every training edge of each
    source exactly once (held-out/test excluded), then distinct non-edges
    that avoid the source, its neighbours and the held-out/test pairs."""
    pytest.importorskip("torch")
    from mcmc_for_ammsb_amd.autograd import make_pairs
    g = M.generate("community", N=20000, K=48, mean_degree=20.0, seed=9)
    s = M.make_split(g, 0.1, 42)
    nodes = np.sort(np.random.default_rng(3).choice(g.N, 40, replace=False))
    ptr, nbr, y = make_pairs(g, s, nodes, 24, seed=1)
    assert ptr[0] == 0 and ptr[-1] == len(nbr) == len(y)
    ex = np.sort(np.concatenate([s.held_out_edges, s.test_edges]), 1)
    ht = set(map(tuple, ex.tolist()))
    for q, sn in enumerate(nodes):
        row = set(g.cols[g.row_ptr[sn]:g.row_ptr[sn + 1]].tolist())
        nb, yy = nbr[ptr[q]:ptr[q + 1]], y[ptr[q]:ptr[q + 1]]
        assert len(set(nb.tolist())) == len(nb)
        assert (yy == 0).sum() == 24
        for v, lab in zip(nb.tolist(), yy.tolist()):
            assert v != sn and (min(sn, v), max(sn, v)) not in ht
            assert (v in row) == bool(lab)
        train = {v for v in row if (min(sn, v), max(sn, v)) not in ht}
        assert train == set(nb[yy == 1].tolist())


def test_batch_create_rejects_unsupported_chains():
    """Argument checks of ammsb_batch_create run before any device call."""
    g = M.generate("community", N=3000, K=16, mean_degree=20.0, seed=1)
    s = M.make_split(g, 0.1, 42)
    with pytest.raises(M.AmmsbError, match="exceeds the batched kernels"):
        M.MCMCBatch([M.Options(K=1024, num_node_sample=10, precision="fp32")], [g], [s])
    with pytest.raises(M.AmmsbError, match="share precision"):
        M.MCMCBatch([M.Options(K=16, num_node_sample=10, precision="fp32"),
                     M.Options(K=16, num_node_sample=10, precision="fp64")], [g, g], [s, s])
    with pytest.raises(ValueError):
        M.MCMCBatch([M.Options(K=16)], [g, g])


REF_DATASETS = "/root/reference/datasets"


@pytest.mark.skipif(not os.path.isdir(REF_DATASETS), reason="reference datasets not present")
@pytest.mark.parametrize("name", ["CA-GrQc-tiny.txt", "network.txt", "netscience.txt",
                                  "CA-HepTh.txt"])
def test_reference_datasets_load(name):
    """The reference's own SNAP files through the native reader against an
    independent Python parse of the same rules (4 header lines, undirected
    pairs, no self-links)."""
    path = os.path.join(REF_DATASETS, name)
    g = M.Graph.from_edge_list_file(path)
    ids, pairs = {}, set()
    with open(path) as f:
        for i, line in enumerate(f):
            if i < 4:
                continue
            a, b = (int(x) for x in line.split()[:2])
            for x in (a, b):
                ids.setdefault(x, len(ids))
            if a != b:
                pairs.add((min(ids[a], ids[b]), max(ids[a], ids[b])))
    assert g.N == len(ids) and g.num_edges == len(pairs)
    _check_csr(g)
    for a, b in list(pairs)[:200]:
        assert g.linked(a, b) and g.linked(b, a)


DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "mcmc_for_ammsb_amd",
                      "ammsb_mcmc")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_option_and_input_errors(tmp_path):
    """The command-line driver (the reference's main): --help, and exit code
    33 for option and input errors (main.cc:79-86), all before any GPU use."""
    import subprocess
    r = subprocess.run([DRIVER, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--mcmc.K" in r.stdout and r.stdout.startswith("Command line:")
    for argv in (["--mcmc.K", "-5"], ["--mcmc.strategy", "nope"], ["-K", "8"],
                 ["-f", str(tmp_path / "missing.txt")], ["-c", "netscience", "-f", "x.xml"]):
        r = subprocess.run([DRIVER] + argv, capture_output=True, text=True)
        assert r.returncode == 33, (argv, r.stdout, r.stderr)
        assert "error" in r.stderr.lower()
