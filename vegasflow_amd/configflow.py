"""Constants, dtype flags and logging (the contract of
src/vegasflow/configflow.py:13-116).

Environment variables, as in the reference:
- ``VEGASFLOW_LOG_LEVEL`` 0..3 (configflow.py:29-61);
- ``VEGASFLOW_FLOAT`` "64" (default) or "32" (configflow.py:35, 68-78).
  "64" selects the FP64 drop-in kernels. "32" selects the mixed-precision
  kernels: sampling, grid map and integrand in FP32, event sums, histograms,
  sigma and grid refinement in FP64 (never below the reference's float32);
- ``VEGASFLOW_INT`` "32" (default) or "64" (configflow.py:36, 80-86).
"""
import logging
import os

import numpy as np
import torch

# Some global parameters (configflow.py:13-16)
BINS_MAX = 50
ALPHA = 1.5
BETA = 0.75  # Vegas+
TECH_CUT = 1e-8

# Events Limit limits how many events are done in one single step of the
# unfused path (configflow.py:22)
MAX_EVENTS_LIMIT = int(1e6)
# Maximum number of hypercubes for VegasFlowPlus (configflow.py:24)
MAX_NEVAL_HCUBE = int(1e4)

# Devices to look for (configflow.py:27); one HIP device is used
DEFAULT_ACTIVE_DEVICES = ["GPU"]

LOG_DICT = {"0": logging.ERROR, "1": logging.WARNING, "2": logging.INFO, "3": logging.DEBUG}

_log_level_idx = os.environ.get("VEGASFLOW_LOG_LEVEL")
_float_env = os.environ.get("VEGASFLOW_FLOAT", "64")
_int_env = os.environ.get("VEGASFLOW_INT", "32")

_bad_log_warning = None
if _log_level_idx not in LOG_DICT:
    _bad_log_warning = _log_level_idx
    _log_level_idx = None
_log_level = LOG_DICT["2"] if _log_level_idx is None else LOG_DICT[_log_level_idx]

logger = logging.getLogger("vegasflow_amd")
logger.setLevel(_log_level)
if not logger.handlers:
    _console_handler = logging.StreamHandler()
    _console_handler.setLevel(_log_level)
    _console_handler.setFormatter(logging.Formatter("[%(levelname)s] (%(name)s) %(message)s"))
    logger.addHandler(_console_handler)
if _bad_log_warning is not None:
    logger.warning("Accepted log levels are: %s, received: %s", list(LOG_DICT), _bad_log_warning)

if _float_env == "64":
    DTYPE = torch.float64
    FMAX = np.finfo(np.float64).max
elif _float_env == "32":
    DTYPE = torch.float32
    FMAX = np.finfo(np.float32).max
else:
    DTYPE = torch.float64
    FMAX = np.finfo(np.float64).max
    logger.warning(f"VEGASFLOW_FLOAT={_float_env} not understood, defaulting to 64 bits")

if _int_env == "64":
    DTYPEINT = torch.int64
elif _int_env == "32":
    DTYPEINT = torch.int32
else:
    DTYPEINT = torch.int64
    logger.warning(f"VEGASFLOW_INT={_int_env} not understood, defaulting to 64 bits")


def precision_of(dtype):
    """Kernel precision for a float dtype: FP64 drop-in or mixed."""
    return "mixed" if dtype == torch.float32 else "f64"


def default_device():
    """The HIP device the engine runs on (torch's current one); CPU only as
    a place to hold host-side state when no GPU is present."""
    if torch.cuda.is_available():
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device("cpu")


def run_eager(flag=True):
    """Kept for API compatibility (configflow.py:89-96): this engine has no
    graph compilation step, kernels are always launched directly."""
    return flag


def int_me(i):
    """Cast the input to the ``DTYPEINT`` type (configflow.py:103-105)."""
    return torch.as_tensor(i, dtype=DTYPEINT, device=default_device())


def float_me(i):
    """Cast the input to the ``DTYPE`` type (configflow.py:108-110)."""
    return torch.as_tensor(i, dtype=DTYPE, device=default_device())
