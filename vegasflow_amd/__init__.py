"""vegasflow_amd: an MI355X-native VEGAS / VEGAS+ / plain Monte Carlo engine
with the Python API of vegasflow (src/vegasflow/__init__.py:3-10).

Python host code drives hand-written gfx950 HIP kernels (libmivegas.so,
include/mivegas.h) through ctypes over PyTorch-ROCm device tensors.
"""
from .configflow import DTYPE, DTYPEINT, float_me, int_me, run_eager
from .integrands import BreitWignerPower, Constant, Gaussian, GaussianMix, Product
from .plain import PlainFlow, plain_sampler, plain_wrapper
from .vflow import VegasFlow, vegas_sampler, vegas_wrapper
from .vflowplus import VegasFlowPlus, vegasflowplus_sampler, vegasflowplus_wrapper

__version__ = "0.1.0"

__all__ = [
    "VegasFlow",
    "VegasFlowPlus",
    "PlainFlow",
    "vegas_wrapper",
    "vegasflowplus_wrapper",
    "plain_wrapper",
    "vegas_sampler",
    "vegasflowplus_sampler",
    "plain_sampler",
    "DTYPE",
    "DTYPEINT",
    "float_me",
    "int_me",
    "run_eager",
    "Gaussian",
    "GaussianMix",
    "Product",
    "BreitWignerPower",
    "Constant",
]
