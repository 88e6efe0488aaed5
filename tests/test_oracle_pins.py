"""Pins of the CPU oracle (no GPU).

The reference holds no golden vectors for this path (its tests print and do
not assert), so every oracle function is pinned here by something the
mathematics or an independent implementation fixes:
  * Philox4x32-10: the Random123 known-answer vectors, and agreement with the
    library's own (separately written) Philox on random counters
  * edge likelihood / perplexity: closed forms (one-hot memberships), the
    invariant L(1)+L(0) = 1 - eps (1 - <pi_a,pi_b>), the 1e-30 floor, and the
    running average over evaluations
  * update_phi / update_beta gradients: central finite differences of the
    a-MMSB log-likelihood computed by brute force over all (z_ab, z_ba)
    community pairs (the model's definition, not the reference's O(K) form)
  * the SGRLD steps: closed forms with injected noise (K = 1, no edges),
    reflection and the 1e-24 floor
  * samplers: a pure-Python restatement of the documented draw definitions
    on tiny graphs, plus invariants and uniformity
"""
import numpy as np
import pytest

import mcmc_for_ammsb_amd as M
from tests import oracle_lib
from tests.oracle_lib import Oracle

KAT = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
       ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
       ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
        (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]


@pytest.mark.parametrize("ctr,key,exp", KAT)
def test_philox_known_answers(ctr, key, exp):
    assert tuple(int(x) for x in oracle_lib.philox(ctr, key)) == exp
    assert tuple(int(x) for x in M.philox(ctr, key)) == exp


def test_philox_two_implementations_agree():
    rng = np.random.default_rng(0)
    for _ in range(500):
        c = rng.integers(0, 2 ** 32, 4, dtype=np.uint64).astype(np.uint32)
        k = rng.integers(0, 2 ** 32, 2, dtype=np.uint64).astype(np.uint32)
        assert np.array_equal(oracle_lib.philox(c, k), M.philox(c, k))


# ---------------------------------------------------------------- fixtures
def tiny_graph(N=40, seed=1):
    rng = np.random.default_rng(seed)
    e = []
    for a in range(N):
        for b in range(a + 1, N):
            if rng.random() < 0.15:
                e.append((a, b))
    return M.Graph.from_edges(N, np.array(e))


def tiny_split(g, seed=42):
    return M.make_split(g, 0.2, seed)


def state(N, K, seed):
    rng = np.random.default_rng(seed)
    phi = rng.gamma(1.0, 1.0, size=(N, K))
    s = phi.sum(1)
    return np.concatenate([phi / s[:, None], s[:, None]], 1), rng.gamma(2.0, 0.5, size=(K, 2))


def brute_p(pa, pb, y, beta, eps):
    """p(y | pi_a, pi_b, beta): sum over z_ab ~ pi_a, z_ba ~ pi_b; link with
    prob beta_k if z_ab == z_ba == k, else eps."""
    K = len(pa)
    p1 = 0.0
    for k in range(K):
        for l in range(K):
            p1 += pa[k] * pb[l] * (beta[k] if k == l else eps)
    return p1 if y else 1.0 - p1


# ----------------------------------------------------------- eps_t, likelihood
def test_eps_t_closed_form():
    h = oracle_lib.hyper_array(0.1, 1, 1, 1e-7, 0.01, 1024.0, 0.55, 10)
    import ctypes
    for t in (1, 2, 50, 1000, 123456):
        got = oracle_lib.lib.orc_eps_t(h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), t, 0)
        assert got == pytest.approx(0.01 * (1 + t / 1024.0) ** -0.55, rel=1e-15)


def _ppx_setup(K=4):
    g = tiny_graph(30, 3)
    s = tiny_split(g)
    return g, s, Oracle(g, s, K, num_node_sample=5)


def test_likelihood_one_hot_and_running_average():
    K = 4
    g, s, o = _ppx_setup(K)
    N = g.N
    theta = np.array([[1.0, 3.0], [2.0, 2.0], [1.0, 1.0], [3.0, 1.0]])
    beta = theta[:, 1] / theta.sum(1)
    pi = np.zeros((N, K + 1))
    pi[:, 0] = 1.0
    pi[:, K] = 2.0
    o.set_state(pi, theta)
    y = s.held_out_labels.astype(bool)
    L1 = np.where(y, beta[0], 1 - beta[0])
    p1 = o.perplexity()
    assert p1 == pytest.approx(-np.log(L1).mean(), rel=1e-14)
    theta2 = theta.copy()
    theta2[0] = [3.0, 1.0]
    b2 = theta2[:, 1] / theta2.sum(1)
    o.set_state(pi, theta2)
    L2 = np.where(y, b2[0], 1 - b2[0])
    p2 = o.perplexity()  # running mean of the two evaluations
    assert p2 == pytest.approx(-np.log((L1 + L2) / 2).mean(), rel=1e-14)
    # orthogonal one-hot memberships: link floored at 1e-30, non-link 1-eps
    pi3 = np.zeros((N, K + 1))
    pi3[:, K] = 1.0
    pi3[np.arange(N), np.arange(N) % 2] = 1.0
    o3 = Oracle(g, s, K, num_node_sample=5)
    o3.set_state(pi3, theta)
    e = s.held_out_edges
    same = (e[:, 0] % 2) == (e[:, 1] % 2)
    bk = beta[e[:, 0] % 2]
    L = np.where(y, np.where(same, bk, 1e-30), np.where(same, 1 - bk, 1 - 1e-7))
    assert o3.perplexity() == pytest.approx(-np.log(L).mean(), rel=1e-13)


def test_likelihood_invariant_and_brute_force():
    K = 5
    g, s, _ = _ppx_setup(K)
    pi, theta = state(g.N, K, 4)
    beta = theta[:, 1] / theta.sum(1)
    eps = 1e-7
    # flip every held-out label to get L(1) and L(0) for the same pairs
    for lab in (0, 1):
        s2 = M.Split(s.held_out_edges, np.full_like(s.held_out_labels, lab), s.test_edges,
                     s.test_labels)
        o = Oracle(g, s2, K, num_node_sample=5, epsilon=eps)
        o.set_state(pi, theta)
        ppx = o.perplexity()
        ref = []
        for a, b in s.held_out_edges:
            p = brute_p(pi[a, :K], pi[b, :K], lab, beta, eps)
            if lab == 1:  # the reference's link likelihood omits eps (learner.h:106-109)
                p -= eps * (1 - np.dot(pi[a, :K], pi[b, :K]))
            ref.append(max(p, 1e-30))
        assert ppx == pytest.approx(-np.mean(np.log(ref)), rel=1e-12)
    # invariant L(1) + L(0) = 1 - eps (1 - <pi_a, pi_b>)
    a, b = s.held_out_edges[0]
    L1 = brute_p(pi[a, :K], pi[b, :K], 1, beta, eps) - eps * (1 - pi[a, :K] @ pi[b, :K])
    L0 = brute_p(pi[a, :K], pi[b, :K], 0, beta, eps)
    assert L1 + L0 == pytest.approx(1 - eps * (1 - pi[a, :K] @ pi[b, :K]), rel=1e-14)


# ----------------------------------------------------------- update_phi
def _phi_objective(pi_row_i, phi_i, nbr_rows, labels, beta, eps):
    pi_i = phi_i / phi_i.sum()
    return sum(np.log(brute_p(pi_i, pb, y, beta, eps)) for pb, y in zip(nbr_rows, labels))


def test_update_phi_gradient_matches_finite_differences():
    K = 4
    g, s, _ = _ppx_setup(K)
    N = g.N
    pi, theta = state(N, K, 5)
    beta = theta[:, 1] / theta.sum(1)
    eps_t, alpha, eps = 1e-4, 0.25, 1e-7
    o = Oracle(g, s, K, num_node_sample=6, alpha=alpha, epsilon=eps)
    o.set_state(pi, theta)
    node = 3
    nbrs = np.array([[1, 5, 8, 13, 21, 2, 9]], dtype=np.int32)
    o.update_phi([node], nbrs, np.zeros((1, K)), eps_t)
    new, _, _ = o.get_state()
    phi = pi[node, :K] * pi[node, K]
    phi_new = new[node, :K] * new[node, K]
    Nn = N / 6
    g_oracle = ((phi_new - phi) * 2 / eps_t - alpha + phi) / Nn
    labels = [g.linked(node, int(b)) for b in nbrs[0]]
    rows = [pi[b, :K] for b in nbrs[0]]
    h = 1e-6
    fd = np.zeros(K)
    for k in range(K):
        d = np.zeros(K)
        d[k] = h
        fd[k] = (_phi_objective(None, phi + d, rows, labels, beta, eps) -
                 _phi_objective(None, phi - d, rows, labels, beta, eps)) / (2 * h)
    assert np.allclose(g_oracle, fd, rtol=1e-5, atol=1e-7), (g_oracle, fd)


def test_update_phi_K1_closed_form_reflection_and_floor():
    K = 1
    g, s, _ = _ppx_setup(K)
    N = g.N
    pi = np.ones((N, 2))
    pi[:, 1] = np.linspace(0.5, 2.0, N)
    theta = np.ones((1, 2))
    alpha, eps_t = 0.3, 0.05
    o = Oracle(g, s, K, num_node_sample=4, alpha=alpha)
    o.set_state(pi, theta)
    nodes = np.array([0, 7, 11], dtype=np.int32)
    nbrs = np.array([[1, 2, 3, 4]] * 3, dtype=np.int32)
    xi = np.array([[0.7], [-40.0], [0.0]])
    o.update_phi(nodes, nbrs, xi, eps_t)
    new, _, _ = o.get_state()
    for q, v in enumerate(nodes):
        phi = pi[v, 1]
        ref = abs(phi + eps_t / 2 * (alpha - phi) + np.sqrt(eps_t * phi) * xi[q, 0])
        assert new[v, 1] == pytest.approx(max(ref, 1e-24), rel=1e-15)
        assert new[v, 0] == 1.0
    # floor: a vanishing phi with no drift stays at 1e-24
    pi2 = pi.copy()
    pi2[0, 1] = 1e-40
    o2 = Oracle(g, s, K, num_node_sample=4, alpha=0.0)
    o2.set_state(pi2, theta)
    o2.update_phi([0], [[1, 2, 3, 4]], np.zeros((1, 1)), 1e-3)
    assert o2.get_state()[0][0, 1] == 1e-24


def test_update_pi_normalisation():
    K = 7
    g, s, _ = _ppx_setup(K)
    pi, theta = state(g.N, K, 6)
    o = Oracle(g, s, K, num_node_sample=5)
    o.set_state(pi, theta)
    nodes = np.arange(10, dtype=np.int32)
    nbrs = np.tile(np.arange(12, 18, dtype=np.int32), (10, 1))
    o.update_phi(nodes, nbrs, np.random.default_rng(1).standard_normal((10, K)), 0.01)
    new, _, _ = o.get_state()
    assert np.allclose(new[nodes, :K].sum(1), 1.0, atol=4e-16)
    assert np.all(new[nodes, K] > 0)


# ----------------------------------------------------------- update_beta
def test_update_beta_gradient_matches_finite_differences():
    K = 3
    g, s, _ = _ppx_setup(K)
    pi, theta = state(g.N, K, 7)
    eps, eps_t, scale = 1e-7, 1e-5, 37.0
    o = Oracle(g, s, K, num_node_sample=5, epsilon=eps, eta0=1.0, eta1=1.0)
    o.set_state(pi, theta)
    edges = np.array([[0, 3], [2, 9], [4, 5], [7, 11], [1, 20]], dtype=np.int32)
    y = np.array([g.linked(a, b) for a, b in edges], dtype=np.uint8)
    o.update_beta(edges, y, scale, np.zeros((K, 2)), eps_t)
    _, th_new, be_new = o.get_state()
    eta = np.array([1.0, 1.0])
    G = ((th_new - theta) * 2 / eps_t - eta + theta) / scale

    def F(th):
        b = th[:, 1] / th.sum(1)
        return sum(np.log(brute_p(pi[a, :K], pi[c, :K], yy, b, eps)) for (a, c), yy in zip(edges, y))

    h = 1e-6
    fd = np.zeros((K, 2))
    for k in range(K):
        for i in range(2):
            d = np.zeros((K, 2))
            d[k, i] = h
            fd[k, i] = (F(theta + d) - F(theta - d)) / (2 * h)
    assert np.allclose(G, fd, rtol=1e-5, atol=1e-6), (G, fd)
    assert np.allclose(be_new, th_new[:, 1] / th_new.sum(1), rtol=1e-15)


def test_update_beta_no_edges_closed_form():
    K = 4
    g, s, _ = _ppx_setup(K)
    pi, theta = state(g.N, K, 8)
    o = Oracle(g, s, K, num_node_sample=5, eta0=0.7, eta1=1.3)
    o.set_state(pi, theta)
    xi = np.random.default_rng(2).standard_normal((K, 2))
    eps_t = 0.03
    o.update_beta(np.zeros((0, 2), np.int32), np.zeros(0, np.uint8), 5.0, xi, eps_t)
    _, th, be = o.get_state()
    eta = np.array([0.7, 1.3])
    ref = np.maximum(np.abs(theta + eps_t / 2 * (eta - theta) + np.sqrt(eps_t * theta) * xi), 1e-24)
    assert np.allclose(th, ref, rtol=1e-15)
    assert np.allclose(be, ref[:, 1] / ref.sum(1), rtol=1e-15)


# ----------------------------------------------------------- samplers
def _u64(ctr, seed):
    b = M.philox(ctr, (seed & 0xFFFFFFFF, seed >> 32))
    return int(b[0]) | (int(b[1]) << 32)


def _heldtest(s):
    st = set()
    for e in list(s.held_out_edges) + list(s.test_edges):
        st.add((int(min(e)), int(max(e))))
    return st


def py_nonlinks(g, s, m, seed, t):
    ht = _heldtest(s)
    node = _u64((0, 0, t, 2), seed) % g.N
    out, taken, j = [], set(), 0
    while len(out) < m:
        v = _u64((j, 0, t, 3), seed) % g.N
        j += 1
        e = (min(node, v), max(node, v))
        if v == node or g.linked(node, v) or e in ht or v in taken:
            continue
        taken.add(v)
        out.append(e)
    N, E = g.N, g.num_edges
    return out, N * N / m * (E / (N * (N - 1) / 2 - E))


def py_links(g, s, m, max_source, seed, t):
    ht = _heldtest(s)
    r = 0
    while True:
        out, seen = [], set()
        for src in range(max_source):
            node = _u64((r, src, t, 2), seed) % g.N
            deg = g.row_ptr[node + 1] - g.row_ptr[node]
            if src > 0 and len(out) + deg > m:
                break
            for v in g.cols[g.row_ptr[node]:g.row_ptr[node + 1]]:
                e = (min(node, int(v)), max(node, int(v)))
                if e in ht or e in seen:
                    continue
                seen.add(e)
                out.append(e)
        if out:
            return out, (g.N if max_source == 1 else g.num_edges / len(out))
        r += 1


def py_neighbors(g, s, node, nnb, seed, t):
    ht = _heldtest(s)
    out, taken, j = [], set(), 0
    while len(out) < nnb:
        v = _u64((j, node, t, 4), seed) % g.N
        j += 1
        if v == node or (min(node, v), max(node, v)) in ht or v in taken:
            continue
        taken.add(v)
        out.append(v | (0x80000000 if g.linked(node, v) else 0))
    return np.array(out, dtype=np.uint32).astype(np.int32)


def py_random_edge(g, s, m, seed, t):
    """random_edge_sampling (network.cc:237-268): uniform directed entries of
    the CSR, distinct training links, in attempt order; the full training set
    in CSR order when m >= E - |held-out| - |test|."""
    ht = _heldtest(s)
    N = g.N
    if m >= g.num_edges - len(s.held_out_labels) - len(s.test_labels):
        out = [(a, int(b)) for a in range(N) for b in g.cols[g.row_ptr[a]:g.row_ptr[a + 1]]
               if a < b and (a, int(b)) not in ht]
    else:
        out, taken, j = [], set(), 0
        while len(out) < m:
            q = _u64((j, 0, t, 11), seed) % (2 * g.num_edges)
            j += 1
            a = int(np.searchsorted(g.row_ptr, q, side="right") - 1)
            b = int(g.cols[q])
            e = (min(a, b), max(a, b))
            if e in ht or e in taken:
                continue
            taken.add(e)
            out.append(e)
    return out, (N - 1) * N / 2.0 / len(out)


@pytest.mark.parametrize("seed", [1, 99])
def test_random_edge_matches_python_restatement(seed):
    g = tiny_graph(60, 5)
    s = tiny_split(g)
    o = Oracle(g, s, 3, num_node_sample=6)
    n_train = g.num_edges - len(s.held_out_labels) - len(s.test_labels)
    for t in range(1, 10):
        for m in (10, n_train // 2, n_train + 5):
            d = o.sample_minibatch(3, m, 1, seed, t)
            ref, scale = py_random_edge(g, s, m, seed, t)
            assert d["flag"] == 2
            assert [tuple(e) for e in d["edges"]] == ref
            assert np.all(d["y"] == 1)
            assert d["scale"] == pytest.approx(scale, rel=1e-15)


def test_random_edge_invariants_and_uniformity():
    """Distinct training links only; every training link equally likely."""
    g = tiny_graph(30, 2)
    s = tiny_split(g)
    ht = _heldtest(s)
    o = Oracle(g, s, 3, num_node_sample=6)
    train = [(a, int(b)) for a in range(g.N) for b in g.cols[g.row_ptr[a]:g.row_ptr[a + 1]]
             if a < b and (a, int(b)) not in ht]
    counts = dict.fromkeys(train, 0)
    for t in range(1, 2001):
        d = o.sample_minibatch(3, 3, 1, 7, t)
        e = [tuple(x) for x in d["edges"]]
        assert len(e) == len(set(e)) == 3
        for x in e:
            assert g.linked(*x) and x not in ht
            counts[x] += 1
    c = np.array(list(counts.values()), dtype=float)
    exp = c.sum() / len(c)
    chi2 = ((c - exp) ** 2 / exp).sum()
    assert chi2 < len(c) + 6 * np.sqrt(2 * len(c))


@pytest.mark.parametrize("seed", [1, 99])
def test_samplers_match_python_restatement(seed):
    g = tiny_graph(60, 5)
    s = tiny_split(g)
    o = Oracle(g, s, 3, num_node_sample=6)
    for t in range(1, 15):
        d = o.sample_minibatch(0, 12, 1, seed, t)
        flag = _u64((0, 0, t, 1), seed) % 2
        assert d["flag"] == flag
        ref, scale = (py_links(g, s, 12, 1, seed, t) if flag else py_nonlinks(g, s, 12, seed, t))
        assert [tuple(e) for e in d["edges"]] == ref
        assert d["scale"] == pytest.approx(scale, rel=1e-15)
        assert np.all(d["y"] == flag)
        d = o.sample_minibatch(1, 25, 4, seed, t)
        ref, scale = py_links(g, s, 25, 4, seed, t)
        assert [tuple(e) for e in d["edges"]] == ref
        assert d["scale"] == pytest.approx(scale, rel=1e-15)
        for v in (0, 17, 59):
            assert np.array_equal(o.sample_neighbors(v, 7, seed, t), py_neighbors(g, s, v, 7, seed, t))


def test_sampler_invariants():
    g = M.generate("community", N=3000, K=8, mean_degree=15.0, seed=4)
    s = M.make_split(g, 0.1, 42)
    ht = _heldtest(s)
    o = Oracle(g, s, 8, num_node_sample=20)
    for t in range(1, 30):
        d = o.sample_minibatch(2, 300, 1, 5, t)  # non-links
        e = [tuple(x) for x in d["edges"]]
        assert len(e) == 300 == len(set(e))
        src = set(e[0]) & set(e[1])
        assert len(src) == 1
        node = src.pop()
        assert all(node in x and not g.linked(*x) and x not in ht and x[0] < x[1] for x in e)
        d = o.sample_minibatch(1, 300, 1, 5, t)  # links of one node
        e = [tuple(x) for x in d["edges"]]
        assert all(g.linked(*x) and x not in ht for x in e)
        assert d["scale"] == g.N
        nb = o.sample_neighbors(node, 21, 5, t)
        v = nb & 0x7FFFFFFF
        assert len(set(v.tolist())) == 21 and node not in v
        assert all(((x >> 31) & 1) == g.linked(node, int(y)) for x, y in zip(nb.astype(np.int64) & 0xFFFFFFFF, v))


def test_neighbor_draws_uniform():
    g = tiny_graph(12, 9)
    s = M.Split(np.zeros((0, 2), np.int32), np.zeros(0, np.uint8), np.zeros((0, 2), np.int32),
                np.zeros(0, np.uint8))
    o = Oracle(g, s, 2, num_node_sample=1)
    counts = np.zeros(12)
    for t in range(1, 3001):
        for v in o.sample_neighbors(0, 2, 3, t) & 0x7FFFFFFF:
            counts[v] += 1
    assert counts[0] == 0
    exp = counts.sum() / 11
    chi2 = ((counts[1:] - exp) ** 2 / exp).sum()
    assert chi2 < 35.0  # 10 dof, p ~ 1e-4


def test_gamma_and_normal_moments():
    # Marsaglia-Tsang Gamma(shape) and Box-Muller normals of the stream
    import ctypes
    for shape in (0.3, 1.0, 2.5):
        x = np.exp([oracle_lib.lib.orc_log_gamma(7, shape, i, 0, 8) for i in range(20000)])
        assert abs(x.mean() - shape) < 4 * np.sqrt(shape / 20000)
        assert abs(x.var() - shape) < 0.1 * shape
    z = np.zeros((20000, 2))
    oracle_lib.lib.orc_theta_noise(3, 1, 20000, z.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1) < 0.03
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 0.03


@pytest.mark.parametrize("eta0,eta1", [(1.0, 1.0), (2.0, 0.5), (0.5, 3.0)])
def test_init_theta_is_gamma_shape_scale(eta0, eta1):
    """init()'s theta = gamma(eta0, eta1, K, 2) (mcmc_sampler_stochastic.cc:54)
    is Gamma(shape eta0, scale eta1) in the GSL order of Random::gamma
    (random.cc:472-511): a Kolmogorov-Smirnov test against that law passes
    and one against the swapped law (shape eta1, scale eta0) fails, so a swap
    cannot go unnoticed.  Also the moments E = eta0*eta1, Var = eta0*eta1^2."""
    from scipy import stats
    th = oracle_lib.init_theta(7, eta0, eta1, 10000).ravel()
    assert stats.kstest(th, stats.gamma(a=eta0, scale=eta1).cdf).pvalue > 1e-3
    if eta0 != eta1:
        assert stats.kstest(th, stats.gamma(a=eta1, scale=eta0).cdf).pvalue < 1e-6
    n = len(th)
    assert abs(th.mean() - eta0 * eta1) < 5 * np.sqrt(eta0 * eta1 ** 2 / n)
    assert abs(th.var() / (eta0 * eta1 ** 2) - 1) < 0.1
    # entry (k, i) is draw id 2k+i of the INIT_THETA stream
    assert th[5] == np.exp(oracle_lib.lib.orc_log_gamma(7, eta0, 5, 0, 8)) * eta1


def test_init_phi_rows_are_normalised_exponentials():
    """init()'s phi = gamma(1, 1, N, K), pi = phi row-normalised
    (mcmc_sampler_stochastic.cc:66-76): phi = pi * phi_sum is Exp(1)
    (Kolmogorov-Smirnov), rows sum to 1, and the rows of different nodes are
    different draws."""
    from scipy import stats
    K = 64
    rows = oracle_lib.init_phi_rows(9, np.arange(300), K)
    phi = rows[:, :K] * rows[:, K:]
    assert stats.kstest(phi.ravel(), stats.expon().cdf).pvalue > 1e-3
    assert np.allclose(rows[:, :K].sum(1), 1.0, atol=1e-14)
    assert abs(rows[:, K].mean() - K) < 5 * np.sqrt(K / 300)
    assert not np.array_equal(rows[0], rows[1])
    # element k of node v is -log U from Philox block (k>>1, v, 0, INIT_PHI),
    # words (0, 1) for even k and (2, 3) for odd k, U = (top 53 bits + 1/2) 2^-53
    b = oracle_lib.philox([3 >> 1, 17, 0, 7], [9, 0])
    u = (((int(b[2]) | (int(b[3]) << 32)) >> 11) + 0.5) / 2.0 ** 53
    assert phi[17, 3] == pytest.approx(-np.log(u), rel=1e-14)


def test_oracle_threads_match_sequential():
    """orc_run with several threads (the distributed reference's OpenMP
    loops: neighbour draws, update_phi, beta gradients, perplexity, with
    per-thread accumulators) equals the one-thread loop up to summation
    order."""
    g = M.generate("community", N=3000, K=16, mean_degree=15.0, seed=5)
    s = M.make_split(g, 0.1, 42)
    res, states = [], []
    for threads in (1, 4):
        o = Oracle(g, s, 16, num_node_sample=30, a=0.01, b=1024.0, c=0.55)
        pi, theta = state(g.N, 16, 3)
        o.set_state(pi, theta)
        r, _ = o.run(25, 0, 300, 1, 10, 42, threads=threads)
        res.append(r)
        states.append(o.get_state())
    assert [t for t, _ in res[0]] == [t for t, _ in res[1]] == [1, 11, 21]
    for (_, a), (_, b) in zip(res[0], res[1]):
        assert abs(a - b) <= 1e-12 * abs(a)
    for x, y in zip(states[0], states[1]):
        np.testing.assert_allclose(x, y, rtol=1e-11, atol=1e-14)


def test_oracle_loop_decreases_perplexity():
    g = M.generate("planted", N=2000, K=8, seed=11)
    s = M.make_split(g, 0.1, 42)
    o = Oracle(g, s, 8, num_node_sample=20, a=0.01, b=1024.0, c=0.55)
    pi, theta = state(g.N, 8, 1)
    o.set_state(pi, theta)
    res, _ = o.run(400, 0, 200, 1, 100, 42)
    assert [t for t, _ in res] == [1, 101, 201, 301]
    assert res[-1][1] < res[0][1]
