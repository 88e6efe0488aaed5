"""Differentiation: forward minibatch log-likelihood and its adjoint (GPU).

Pins:
  * forward value: brute force over all (z_ab, z_ba) community pairs of the
    a-MMSB likelihood (the model's definition), on a few sources
  * dL/dphi: the reference's closed-form update_phi gradient (oracle,
    update_phi_node's grads) to 1e-12; central finite differences
    (h = 1e-6, FP64) on random coordinates
  * dL/dtheta: the reference's update_beta gradient (oracle) to 1e-12;
    central finite differences on every coordinate
  * FP32 precision mode against FP64 to 1e-4
K = 48 runs the one-wavefront-per-row kernel; K = 600 (ragged last chunk)
and K = 2048 run the whole-row kernels (one neighbour row per wavefront),
and test_kchunk_launch_shapes_agree runs the K-chunked ones (several
wavefronts per row, Z exchanged through LDS) against them.
"""
import numpy as np
import pytest

import mcmc_for_ammsb_amd as M
from tests.oracle_lib import Oracle

torch = pytest.importorskip("torch")
import mcmc_for_ammsb_amd.autograd  # noqa: E402,F401  (M.autograd)

pytestmark = pytest.mark.gpu

MSRC, NNL = 40, 24
SHAPES = [(20000, 48), (8000, 600), (4000, 2048)]


@pytest.fixture(scope="module", params=SHAPES, ids=[f"K{k}" for _, k in SHAPES])
def setup(request):
    N, K = request.param
    g = M.generate("community", N=N, K=K, mean_degree=20.0, seed=9)
    s = M.make_split(g, 0.1, 42)
    rng = np.random.default_rng(3)
    phi = rng.gamma(1.0, 1.0, size=(N, K))
    pi = np.concatenate([phi / phi.sum(1, keepdims=True), phi.sum(1, keepdims=True)], 1)
    theta = rng.gamma(2.0, 0.5, size=(K, 2))
    nodes = np.sort(rng.choice(N, size=MSRC, replace=False)).astype(np.int32)
    ptr, nbr, y = M.autograd.make_pairs(g, s, nodes, NNL, seed=1)
    return g, s, pi, theta, nodes, ptr, nbr, y


def _dev(x, dt=None):
    t = torch.from_numpy(np.ascontiguousarray(x))
    return (t.to(dt) if dt is not None else t).cuda()


def _sampler(g, s, pi, theta, precision="fp64"):
    K = theta.shape[0]
    smp = M.MCMCSamplerStochastic(M.Options(K=K, num_node_sample=10, precision=precision), g, s)
    smp.set_state(pi.astype(np.float64 if precision == "fp64" else np.float32), theta)
    return smp


def _inputs(pi, theta, nodes, ptr, nbr, y, dt=torch.float64):
    K = theta.shape[0]
    phi_src = pi[nodes, :K] * pi[nodes, K:K + 1]
    return (_dev(phi_src, dt).requires_grad_(True), _dev(theta).requires_grad_(True), _dev(nodes),
            _dev(ptr), _dev(nbr), _dev(y))


def test_forward_matches_brute_force(setup):
    g, s, pi, theta, nodes, ptr, nbr, y = setup
    N, K = g.N, theta.shape[0]
    smp = _sampler(g, s, pi, theta)
    phi_t, th_t, n_t, p_t, b_t, y_t = _inputs(pi, theta, nodes, ptr, nbr, y)
    per = M.autograd.loglik_per_source(smp, phi_t, th_t, n_t, p_t, b_t, y_t).cpu().numpy()
    beta = theta[:, 1] / theta.sum(1)
    eps = 1e-7
    for q in range(4 if K <= 600 else 2):
        ref = 0.0
        pa = pi[nodes[q], :K]
        for p in range(ptr[q], ptr[q + 1]):
            pb = pi[nbr[p], :K]
            outer = np.outer(pa, pb)  # z_ab = k, z_ba = l
            p1 = (np.diag(outer) * beta).sum() + eps * (outer.sum() - np.trace(outer))
            ref += np.log(p1 if y[p] else 1.0 - p1)
        assert per[q] == pytest.approx(ref, rel=1e-12)
    total = M.autograd.loglik(smp, phi_t, th_t, n_t, p_t, b_t, y_t)
    assert total.item() == pytest.approx(per.sum(), rel=1e-13)


def test_adjoint_matches_reference_closed_form(setup):
    g, s, pi, theta, nodes, ptr, nbr, y = setup
    N, K = g.N, theta.shape[0]
    smp = _sampler(g, s, pi, theta)
    orc = Oracle(g, s, K, num_node_sample=10)
    orc.set_state(pi, theta)
    phi_t, th_t, n_t, p_t, b_t, y_t = _inputs(pi, theta, nodes, ptr, nbr, y)
    L = M.autograd.loglik(smp, phi_t, th_t, n_t, p_t, b_t, y_t)
    L.backward()
    gphi = phi_t.grad.cpu().numpy()
    gth = th_t.grad.cpu().numpy()
    for q in range(MSRC):
        ref = orc.phi_grads(nodes[q], nbr[ptr[q]:ptr[q + 1]])
        assert np.allclose(gphi[q], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), q
    src = np.repeat(nodes, np.diff(ptr))
    edges = np.stack([src, nbr], 1)
    ref = orc.beta_grads(edges, y)
    assert np.allclose(gth, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_adjoint_matches_finite_differences(setup):
    g, s, pi, theta, nodes, ptr, nbr, y = setup
    N, K = g.N, theta.shape[0]
    smp = _sampler(g, s, pi, theta)
    phi_t, th_t, n_t, p_t, b_t, y_t = _inputs(pi, theta, nodes, ptr, nbr, y)
    L = M.autograd.loglik(smp, phi_t, th_t, n_t, p_t, b_t, y_t)
    L.backward()
    gphi = phi_t.grad.cpu().numpy()
    gth = th_t.grad.cpu().numpy()
    h = 1e-6
    rng = np.random.default_rng(5)
    base_phi = phi_t.detach().clone()
    for _ in range(200):
        q, k = int(rng.integers(MSRC)), int(rng.integers(K))
        vals = []
        for sgn in (1, -1):
            ph = base_phi.clone()
            ph[q, k] += sgn * h
            per = M.autograd.loglik_per_source(smp, ph, th_t.detach(), n_t, p_t, b_t, y_t)
            vals.append(per[q].item())  # only source q depends on phi_q
        fd = (vals[0] - vals[1]) / (2 * h)
        assert fd == pytest.approx(gphi[q, k], rel=1e-5, abs=1e-6), (q, k)
    base_th = th_t.detach().clone()
    ks = range(K) if K <= 64 else rng.choice(K, size=48, replace=False)
    for k in ks:
        for i in range(2):
            vals = []
            for sgn in (1, -1):
                th = base_th.clone()
                th[k, i] += sgn * h
                vals.append(M.autograd.loglik_per_source(smp, base_phi, th, n_t, p_t, b_t, y_t).sum().item())
            fd = (vals[0] - vals[1]) / (2 * h)
            assert fd == pytest.approx(gth[k, i], rel=1e-5, abs=1e-4), (k, i)


def test_fp32_against_fp64(setup):
    g, s, pi, theta, nodes, ptr, nbr, y = setup
    N, K = g.N, theta.shape[0]
    out = {}
    for prec, dt in (("fp64", torch.float64), ("fp32", torch.float32)):
        smp = _sampler(g, s, pi, theta, prec)
        phi_t, th_t, n_t, p_t, b_t, y_t = _inputs(pi, theta, nodes, ptr, nbr, y, dt)
        L = M.autograd.loglik(smp, phi_t, th_t, n_t, p_t, b_t, y_t)
        L.backward()
        out[prec] = (L.item(), phi_t.grad.double().cpu().numpy(), th_t.grad.cpu().numpy())
    assert out["fp32"][0] == pytest.approx(out["fp64"][0], rel=1e-5)
    assert np.allclose(out["fp32"][1], out["fp64"][1], rtol=1e-3, atol=1e-4 * np.abs(out["fp64"][1]).max())
    assert np.allclose(out["fp32"][2], out["fp64"][2], rtol=1e-3, atol=1e-4 * np.abs(out["fp64"][2]).max())


def test_bad_neighbour_id_is_an_error(setup):
    g, s, pi, theta, nodes, ptr, nbr, y = setup
    N, K = g.N, theta.shape[0]
    smp = _sampler(g, s, pi, theta)
    bad = nbr.copy()
    bad[3] = N + 7
    phi_t, th_t, n_t, p_t, b_t, y_t = _inputs(pi, theta, nodes, ptr, bad, y)
    with pytest.raises(M.AmmsbError):
        M.autograd.loglik(smp, phi_t, th_t, n_t, p_t, b_t, y_t)


def test_argument_checks_before_native_calls(setup):
    """Every entry of the torch front end (the autograd function and the
    forward-only / adjoint-only helpers) checks dtype, shape, contiguity and
    device before any native call: a phi of the other precision or of the
    wrong width, or a non-contiguous or 2-d input, is an error and nothing
    is launched."""
    g, s, pi, theta, nodes, ptr, nbr, y = setup
    smp = _sampler(g, s, pi, theta)
    phi_t, th_t, n_t, p_t, b_t, y_t = _inputs(pi, theta, nodes, ptr, nbr, y)
    phi_t, th_t = phi_t.detach(), th_t.detach()
    K = theta.shape[0]
    bad_args = [
        ((phi_t.float(), th_t, n_t, p_t, b_t, y_t), TypeError),           # FP32 phi, FP64 sampler
        ((phi_t[:, :K - 1].contiguous(), th_t, n_t, p_t, b_t, y_t), ValueError),  # narrow rows
        ((phi_t[:-1], th_t, n_t, p_t, b_t, y_t), ValueError),            # rows != sources
        ((phi_t, th_t.float(), n_t, p_t, b_t, y_t), TypeError),
        ((phi_t, th_t, n_t, p_t[:-1], b_t, y_t), ValueError),
        ((phi_t, th_t, n_t, p_t, b_t, y_t[:-1]), ValueError),
        ((phi_t, th_t, n_t, p_t, b_t.reshape(1, -1), y_t.reshape(1, -1)), ValueError),
        ((phi_t.t().contiguous().t(), th_t, n_t, p_t, b_t, y_t), ValueError),  # not contiguous
        ((phi_t.cpu(), th_t, n_t, p_t, b_t, y_t), TypeError),
    ]
    for args, exc in bad_args:
        for fn in (M.autograd.loglik_per_source, M.autograd.loglik_grad, M.autograd.loglik):
            with pytest.raises(exc):
                fn(smp, *args)
    with pytest.raises(TypeError):
        M.autograd.loglik_grad(smp, phi_t, th_t, n_t, p_t, b_t, y_t,
                               gout=torch.ones((), dtype=torch.float32, device=phi_t.device))
    # valid arguments still work after the rejected calls
    per = M.autograd.loglik_per_source(smp, phi_t, th_t, n_t, p_t, b_t, y_t)
    assert torch.isfinite(per).all()
    smp.close()


def test_kchunk_launch_shapes_agree(setup):
    """Every (ll_pairs, ll_prefetch) shape of the K-chunked kernels sums in
    the same order, so values and gradients are bit-identical; the whole-row
    kernels (the default for wide rows) agree with the K-chunked ones to
    1e-12 (FP64) or 1e-5 (FP32) of the largest value: they form Z in another
    order (the regrouped sums), so a source whose log-likelihood is near 0
    differs by the rounding of the larger terms."""
    g, s, pi, theta, nodes, ptr, nbr, y = setup
    if theta.shape[0] < 600:
        pytest.skip("one wavefront per row at this K")
    for prec, dt in (("fp64", torch.float64), ("fp32", torch.float32)):
        smp = _sampler(g, s, pi, theta, prec)
        phi_t, th_t, n_t, p_t, b_t, y_t = _inputs(pi, theta, nodes, ptr, nbr, y, dt)
        assert smp.get_tuning()["ll_fwd_rows"] == 1
        smp.set_tuning(ll_bwd_rows=1)
        rows = M.autograd.loglik_per_source(smp, phi_t.detach(), th_t.detach(), n_t, p_t, b_t,
                                            y_t).cpu().numpy()
        rgp, rgt = M.autograd.loglik_grad(smp, phi_t.detach(), th_t.detach(), n_t, p_t, b_t, y_t)
        rows_g = (rgp.cpu().numpy(), rgt.cpu().numpy())
        smp.set_tuning(ll_fwd_rows=0, ll_bwd_rows=0)
        ref = None
        for pairs, pf in ((4, 1), (2, 0), (2, 1), (4, 0), (8, 0), (8, 1)):
            smp.set_tuning(ll_fwd_pairs=pairs, ll_fwd_prefetch=pf, ll_bwd_pairs=pairs,
                           ll_bwd_prefetch=pf)
            per = M.autograd.loglik_per_source(smp, phi_t.detach(), th_t.detach(), n_t, p_t, b_t, y_t)
            gp, gt = M.autograd.loglik_grad(smp, phi_t.detach(), th_t.detach(), n_t, p_t, b_t, y_t)
            got = (per.cpu().numpy(), gp.cpu().numpy(), gt.cpu().numpy())
            if ref is None:
                ref = got
                tol = 1e-12 if prec == "fp64" else 1e-5
                np.testing.assert_allclose(rows, got[0], rtol=tol, atol=tol * np.abs(got[0]).max())
                for r, o in zip(rows_g, got[1:]):
                    np.testing.assert_allclose(r, o, rtol=tol * 10, atol=tol * np.abs(o).max())
            else:
                for r, o in zip(ref, got):
                    np.testing.assert_array_equal(o, r, err_msg=f"{prec} shape ({pairs},{pf})")
        with pytest.raises(M.AmmsbError):
            smp.set_tuning(ll_bwd_pairs=3)
