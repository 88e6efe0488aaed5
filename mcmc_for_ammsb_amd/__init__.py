"""MI355X-native SGRLD sampler for the assortative MMSB.

A drop-in for the reference's ``mcmc_sampler_stochastic`` loop
(56592aya/MCMC_FOR_AMMSB, c++/src/mcmc/learning/mcmc_sampler_stochastic.*):
HIP kernels for gfx950 behind the C-ABI of ``libammsb.so``
(include/ammsb.h), driven from C++; this package is the Python front end.
"""
from ._lib import AmmsbError, FP32, FP64, lib  # noqa: F401  (fails loudly if unbuilt)
from .graph import Graph, Split, generate, make_split, philox  # noqa: F401
from .options import Options  # noqa: F401
from .sampler import MCMCBatch, MCMCSamplerStochastic  # noqa: F401

__all__ = ["AmmsbError", "Graph", "Split", "generate", "make_split", "philox", "Options",
           "MCMCSamplerStochastic", "MCMCBatch"]
