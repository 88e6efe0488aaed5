"""The no-allocation adjoint (GPU): ammsb_loglik_bwd_noalloc, whose scratch
is a file-scope device array instead of device memory the library
allocates, against ammsb_loglik_bwd through the C-ABI.

Both entry points run the same kernels with the same launch shapes, so
dL/dphi and dL/dtheta must be bit-identical: on two calls in a row, the
second with fewer sources and pairs (the array keeps the first call's
contents), for the one-wavefront-per-row kernel (K 48), the K-chunked kernel
(FP32 K 1024, selected with ll_bwd_rows = 0) and the whole-row kernel (FP64
and FP32 K 2048, the default).  A call with a
neighbour id outside [0, N) is an error and the next call succeeds (the
error word in the array is cleared by every call).
"""
import numpy as np
import pytest

import mcmc_for_ammsb_amd as M
from mcmc_for_ammsb_amd import _lib

torch = pytest.importorskip("torch")
import mcmc_for_ammsb_amd.autograd  # noqa: E402,F401

pytestmark = pytest.mark.gpu


def _bwd(fn, smp, phi, theta, nodes, ptr, nbr, y):
    gout = torch.ones((), dtype=torch.float64, device=phi.device)
    gphi = torch.full_like(phi, float("nan"))
    gth = torch.full_like(theta, float("nan"))
    stream = torch.cuda.current_stream(phi.device).cuda_stream
    _lib.check(fn(smp._h, nodes.shape[0], nodes.data_ptr(), phi.data_ptr(), phi.stride(0),
                  theta.data_ptr(), ptr.data_ptr(), nbr.data_ptr(), y.data_ptr(), nbr.shape[0],
                  gout.data_ptr(), gphi.data_ptr(), gth.data_ptr(), stream))
    torch.cuda.synchronize()
    return gphi.cpu().numpy(), gth.cpu().numpy()


@pytest.mark.parametrize("prec,N,K", [("fp64", 20000, 48), ("fp32", 8000, 1024),
                                      ("fp64", 4000, 2048), ("fp32", 4000, 2048)])
def test_noalloc_adjoint_equals_allocating_adjoint(prec, N, K):
    print(f"no-allocation adjoint {prec} K={K}", flush=True)
    dt = torch.float64 if prec == "fp64" else torch.float32
    g = M.generate("community", N=N, K=K, mean_degree=20.0, seed=9)
    s = M.make_split(g, 0.1, 42)
    smp = M.MCMCSamplerStochastic(M.Options(K=K, num_node_sample=10, precision=prec), g, s)
    smp.init(5)
    assert smp.get_tuning()["ll_bwd_rows"] == 1  # whole-row adjoint (one partial row per block)
    if K == 1024:
        smp.set_tuning(ll_bwd_rows=0)  # the K-chunked adjoint
    rng = np.random.default_rng(4)
    theta = torch.from_numpy(rng.gamma(2.0, 0.5, size=(K, 2))).cuda()
    for M_src, nnl in ((48, 40), (9, 7)):  # the second call is smaller
        nodes = np.sort(rng.choice(N, size=M_src, replace=False)).astype(np.int32)
        ptr, nbr, y = M.autograd.make_pairs(g, s, nodes, nnl, seed=M_src)
        phi = torch.from_numpy(rng.gamma(1.0, 1.0, size=(M_src, K))).to(dt).cuda()
        args = (phi, theta, torch.from_numpy(nodes).cuda(), torch.from_numpy(ptr).cuda(),
                torch.from_numpy(nbr).cuda(), torch.from_numpy(y).cuda())
        ref_phi, ref_th = _bwd(_lib.lib.ammsb_loglik_bwd, smp, *args)
        got_phi, got_th = _bwd(_lib.lib.ammsb_loglik_bwd_noalloc, smp, *args)
        assert np.all(np.isfinite(ref_phi)) and np.all(np.isfinite(ref_th))
        np.testing.assert_array_equal(got_phi, ref_phi)
        np.testing.assert_array_equal(got_th, ref_th)
    # an out-of-range neighbour is reported, and the next call starts clean
    bad = args[4].clone()
    bad[1] = N + 3
    with pytest.raises(M.AmmsbError):
        _bwd(_lib.lib.ammsb_loglik_bwd_noalloc, smp, *args[:4], bad, args[5])
    got_phi, got_th = _bwd(_lib.lib.ammsb_loglik_bwd_noalloc, smp, *args)
    np.testing.assert_array_equal(got_phi, ref_phi)
    np.testing.assert_array_equal(got_th, ref_th)
    smp.close()
