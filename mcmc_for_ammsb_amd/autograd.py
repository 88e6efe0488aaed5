"""Differentiable minibatch log-likelihood (torch front end).

``loglik(sampler, phi, theta, nodes, pair_ptr, pair_nbr, pair_y)`` returns

    L = sum_s sum_{p in pairs(s)} log p(y_p | pi_s, pi_{b_p}, beta(theta))

as a 0-d float64 tensor; its backward pass is the hand-written adjoint kernel
of libammsb.so (``ammsb_loglik_bwd``), which recomputes each pair's
likelihood instead of storing it.  ``phi`` ([M][K], the sources' unnormalised
memberships, pi_s = phi_s / sum(phi_s)) and ``theta`` ([K][2], beta =
theta1/(theta0+theta1)) receive gradients; the neighbours' rows come from the
sampler's resident state and are constants, as in update_phi
(c++/src/mcmc/learning/mcmc_sampler_stochastic.cc:269-296).  dL/dphi equals
update_phi's ``grads`` and dL/dtheta equals update_beta's ``grads``
(:234-239).  Both kernels run on torch's current stream.
"""
import numpy as np
import torch

from . import _lib


def _check(t, name, dtype, shape=None):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a GPU tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def _check_args(sampler, phi, theta, nodes, pair_ptr, pair_nbr, pair_y, gout=None):
    """Dtype, shape, contiguity and device of every argument, checked before
    any native call (the kernels read and write these buffers by pointer).
    Returns the number of sources M."""
    K = sampler.info["K"]
    fdt = torch.float64 if sampler.dtype == np.float64 else torch.float32
    _check(nodes, "nodes", torch.int32)
    if nodes.dim() != 1:
        raise ValueError(f"nodes must be 1-d, got shape {tuple(nodes.shape)}")
    M = nodes.shape[0]
    _check(phi, "phi", fdt, (M, K))
    _check(theta, "theta", torch.float64, (K, 2))
    _check(pair_ptr, "pair_ptr", torch.int64, (M + 1,))
    _check(pair_nbr, "pair_nbr", torch.int32)
    if pair_nbr.dim() != 1:
        raise ValueError(f"pair_nbr must be 1-d, got shape {tuple(pair_nbr.shape)}")
    _check(pair_y, "pair_y", torch.uint8, pair_nbr.shape)
    if gout is not None:
        _check(gout, "gout", torch.float64, ())
    devs = {t.device for t in (phi, theta, nodes, pair_ptr, pair_nbr, pair_y)
            + ((gout,) if gout is not None else ())}
    if len(devs) != 1:
        raise ValueError(f"all tensors must be on one device, got {sorted(map(str, devs))}")
    return M


class LogLikelihood(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi, theta, sampler, nodes, pair_ptr, pair_nbr, pair_y):
        M = _check_args(sampler, phi, theta, nodes, pair_ptr, pair_nbr, pair_y)
        per = torch.empty(M, dtype=torch.float64, device=phi.device)
        total = torch.empty((), dtype=torch.float64, device=phi.device)
        stream = torch.cuda.current_stream(phi.device).cuda_stream
        _lib.check(_lib.lib.ammsb_loglik_fwd(
            sampler._h, M, nodes.data_ptr(), phi.data_ptr(), phi.stride(0), theta.data_ptr(),
            pair_ptr.data_ptr(), pair_nbr.data_ptr(), pair_y.data_ptr(), pair_nbr.shape[0],
            per.data_ptr(), total.data_ptr(), stream))
        ctx.save_for_backward(phi, theta, nodes, pair_ptr, pair_nbr, pair_y)
        ctx.sampler = sampler
        ctx.per_source = per
        return total

    @staticmethod
    def backward(ctx, gout):
        phi, theta, nodes, pair_ptr, pair_nbr, pair_y = ctx.saved_tensors
        gout = gout.detach().to(torch.float64).contiguous().reshape(())
        _check_args(ctx.sampler, phi, theta, nodes, pair_ptr, pair_nbr, pair_y, gout)
        grad_phi = torch.empty_like(phi)
        grad_theta = torch.empty_like(theta)
        stream = torch.cuda.current_stream(phi.device).cuda_stream
        _lib.check(_lib.lib.ammsb_loglik_bwd(
            ctx.sampler._h, nodes.shape[0], nodes.data_ptr(), phi.data_ptr(), phi.stride(0),
            theta.data_ptr(), pair_ptr.data_ptr(), pair_nbr.data_ptr(), pair_y.data_ptr(),
            pair_nbr.shape[0], gout.data_ptr(), grad_phi.data_ptr(), grad_theta.data_ptr(), stream))
        return (grad_phi if ctx.needs_input_grad[0] else None,
                grad_theta if ctx.needs_input_grad[1] else None,
                None, None, None, None, None)


def loglik(sampler, phi, theta, nodes, pair_ptr, pair_nbr, pair_y):
    """Differentiable L(phi, theta); see the module docstring."""
    return LogLikelihood.apply(phi, theta, sampler, nodes, pair_ptr, pair_nbr, pair_y)


def loglik_per_source(sampler, phi, theta, nodes, pair_ptr, pair_nbr, pair_y):
    """The forward pass only, one FP64 value per source node (no autograd)."""
    M = _check_args(sampler, phi, theta, nodes, pair_ptr, pair_nbr, pair_y)
    per = torch.empty(M, dtype=torch.float64, device=phi.device)
    stream = torch.cuda.current_stream(phi.device).cuda_stream
    _lib.check(_lib.lib.ammsb_loglik_fwd(
        sampler._h, M, nodes.data_ptr(), phi.data_ptr(), phi.stride(0), theta.data_ptr(),
        pair_ptr.data_ptr(), pair_nbr.data_ptr(), pair_y.data_ptr(), pair_nbr.shape[0],
        per.data_ptr(), None, stream))
    return per


def loglik_grad(sampler, phi, theta, nodes, pair_ptr, pair_nbr, pair_y, gout=None):
    """The adjoint only (no autograd graph): returns (dL/dphi, dL/dtheta),
    scaled by ``gout`` (a 0-d float64 GPU tensor, default 1)."""
    if gout is None:
        gout = torch.ones((), dtype=torch.float64, device=phi.device)
    _check_args(sampler, phi, theta, nodes, pair_ptr, pair_nbr, pair_y, gout)
    grad_phi = torch.empty_like(phi)
    grad_theta = torch.empty_like(theta)
    stream = torch.cuda.current_stream(phi.device).cuda_stream
    _lib.check(_lib.lib.ammsb_loglik_bwd(
        sampler._h, nodes.shape[0], nodes.data_ptr(), phi.data_ptr(), phi.stride(0),
        theta.data_ptr(), pair_ptr.data_ptr(), pair_nbr.data_ptr(), pair_y.data_ptr(),
        pair_nbr.shape[0], gout.data_ptr(), grad_phi.data_ptr(), grad_theta.data_ptr(), stream))
    return grad_phi, grad_theta


def make_pairs(graph, split, nodes, n_nonlinks, seed=0):
    """Pair lists of the differentiation workload: for each source node, all
    its training edges (y = 1; held-out and test edges excluded) in CSR order,
    followed by ``n_nonlinks`` distinct random non-edges (y = 0, not
    held-out/test, not the node itself) in draw order.  Returns numpy
    (pair_ptr int64 [M+1], pair_nbr int32, pair_y uint8)."""
    rng = np.random.default_rng(seed)
    N = np.int64(graph.N)
    ex = np.concatenate([split.held_out_edges, split.test_edges]).astype(np.int64)
    exkeys = np.unique(np.minimum(ex[:, 0], ex[:, 1]) * N + np.maximum(ex[:, 0], ex[:, 1]))

    def excluded(s, v):
        key = np.minimum(s, v) * N + np.maximum(s, v)
        pos = np.minimum(np.searchsorted(exkeys, key), max(len(exkeys) - 1, 0))
        return (exkeys[pos] == key) if len(exkeys) else np.zeros(len(v), dtype=bool)

    ptr = [0]
    nbr, ys = [], []
    for s in np.asarray(nodes, dtype=np.int64):
        row = graph.cols[graph.row_ptr[s]:graph.row_ptr[s + 1]].astype(np.int64)
        links = row[~excluded(s, row)]
        got = np.empty(0, dtype=np.int64)
        while len(got) < n_nonlinks:
            v = rng.integers(0, graph.N, size=2 * n_nonlinks + 16)
            pos = np.minimum(np.searchsorted(row, v), max(len(row) - 1, 0))
            in_row = (row[pos] == v) if len(row) else np.zeros(len(v), dtype=bool)
            v = v[(v != s) & ~in_row & ~excluded(s, v)]
            v = np.concatenate([got, v])
            _, first = np.unique(v, return_index=True)
            got = v[np.sort(first)][:n_nonlinks]
        nbr += [links, got]
        ys += [np.ones(len(links), np.uint8), np.zeros(len(got), np.uint8)]
        ptr.append(ptr[-1] + len(links) + len(got))
    cat = (lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.empty(0, dt))
    return np.asarray(ptr, dtype=np.int64), cat(nbr, np.int32), cat(ys, np.uint8)
