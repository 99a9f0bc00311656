"""ctypes binding of libmivegas.so (declared in include/mivegas.h).

Loaded lazily on first use; a missing or stale library raises immediately,
there is no fallback. Pointers are passed as integers taken from
``torch.Tensor.data_ptr()``; the stream is torch's current HIP stream.
"""
import ctypes
import os
import threading

from . import _build

ABI_VERSION = 3
BINS = 50
FUSED_MAX_DIM = 16
MAX_COMP = 4
MAX_OUT = 16
MAX_BLOCKS = 2048
ROWS = 32
DIGEST_DIMS = 16

F64 = 0
MIXED = 1

FN_CONSTANT = 0
FN_GAUSSIAN = 1
FN_PRODUCT = 2
FN_GAUSSMIX = 3
FN_BWPOWER = 4

ERR_ARG = -1
ERR_UNSUPPORTED = -2
ERR_LAUNCH = -3
ERR_WORKSPACE = -4
ERR_DEVICE = -5

STATUS_NONFINITE = 1

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_f64 = ctypes.c_double
_vp = ctypes.c_void_p


class Integrand(ctypes.Structure):
    _fields_ = [
        ("family", _i32),
        ("n_comp", _i32),
        ("reserved", _i32 * 2),
        ("comp", (_f64 * 4) * MAX_COMP),
        ("centre", (_f64 * FUSED_MAX_DIM) * MAX_COMP),
    ]


class Step(ctypes.Structure):
    _fields_ = [
        ("n_dim", _i32),
        ("precision", _i32),
        ("train", _i32),
        ("reserved", _i32),
        ("n_events", _i64),
        ("event_offset", _i64),
        ("seed", _u64),
        ("draw_id", _u32),
        ("reserved2", _u32),
        ("xjac", _f64),
        ("xdeltajac", _f64),
        ("xmin", _vp),
        ("xdelta", _vp),
        ("inject", _vp),
        ("draw_dev", _vp),
    ]


class Partials(ctypes.Structure):
    _fields_ = [
        ("rows", _vp),
        ("stride", _i64),
    ]


class Strata(ctypes.Structure):
    _fields_ = [
        ("n_strat", _i32),
        ("n_cubes", _i32),
        ("n_ev", _vp),
        ("offsets", _vp),
    ]


class Finalize(ctypes.Structure):
    _fields_ = [
        ("n_dim", _i32),
        ("n_out", _i32),
        ("train", _i32),
        ("refine", _i32),
        ("raw", _i32),
        ("reserved", _i32),
        ("n_events", _f64),
        ("draw_dev", _vp),
        ("history", _vp),
        ("history_len", _i32),
        ("reserved2", _i32),
        ("status", _vp),
    ]


class BatchItem(ctypes.Structure):
    _fields_ = [
        ("n_dim", _i32),
        ("train", _i32),
        ("n_events", _i64),
        ("seed", _u64),
        ("grid_offset", _i64),
        ("row_offset", _i64),
        ("fn", Integrand),
    ]


class BatchTile(ctypes.Structure):
    _fields_ = [
        ("item", _i32),
        ("count", _i32),
        ("first", _i64),
    ]


class BatchVplusItem(ctypes.Structure):
    _fields_ = [
        ("n_dim", _i32),
        ("train", _i32),
        ("n_strat", _i32),
        ("n_cubes", _i32),
        ("adaptive", _i32),
        ("min_neval", _i32),
        ("init_calls", _f64),
        ("seed", _u64),
        ("grid_offset", _i64),
        ("row_offset", _i64),
        ("cube_offset", _i64),
        ("off_offset", _i64),
        ("out_offset", _i64),
        ("fn", Integrand),
    ]


STRUCTS = [Integrand, Step, Partials, Strata, Finalize, BatchItem, BatchTile, BatchVplusItem]

_P = ctypes.POINTER
# name -> (restype, argtypes); the names are exactly those of include/mivegas.h
SIGNATURES = {
    "mv_abi_version": (ctypes.c_int, []),
    "mv_last_error": (ctypes.c_char_p, []),
    "mv_struct_size": (_i64, [_i32]),
    "mv_partial_stride": (_i64, [_i32, _i32, _i32]),
    "mv_producer_blocks": (_i32, [_i64]),
    "mv_producer_rows": (_i32, [_i64]),
    "mv_vegas_fused": (ctypes.c_int, [_P(Step), _P(Integrand), _vp, _P(Partials), _vp]),
    "mv_vegas_sample": (ctypes.c_int, [_P(Step), _vp, _vp, _vp, _vp, _vp]),
    "mv_vegas_digest": (ctypes.c_int, [_P(Step), _vp, _i64, _i32, _i32, _vp, _i64, _vp,
                                       _P(Partials), _vp]),
    "mv_vegas_finalize": (ctypes.c_int, [_P(Finalize), _vp, _i64, _i32, _vp, _vp, _vp, _vp]),
    "mv_refine_grid": (ctypes.c_int, [_i32, _vp, _i64, _vp, _vp]),
    "mv_consume_array_into_indices": (ctypes.c_int, [_vp, _vp, _i64, _i32, _vp, _vp]),
    "mv_plain_fused": (ctypes.c_int, [_P(Step), _P(Integrand), _P(Partials), _vp]),
    "mv_plain_sample": (ctypes.c_int, [_P(Step), _vp, _vp]),
    "mv_vplus_fused": (ctypes.c_int, [_P(Step), _P(Strata), _P(Integrand), _vp, _vp, _vp,
                                      _P(Partials), _vp]),
    "mv_vplus_sample": (ctypes.c_int, [_P(Step), _P(Strata), _vp, _vp, _vp, _vp, _vp, _vp]),
    "mv_vplus_digest": (ctypes.c_int, [_P(Step), _P(Strata), _vp, _vp, _vp, _vp, _vp, _vp,
                                       _P(Partials), _vp]),
    "mv_vplus_out_len": (_i64, [_i32]),
    "mv_vplus_finalize": (ctypes.c_int, [_i32, _P(Strata), _vp, _vp, _vp, _vp, _i32, _i32, _f64,
                                         _f64, _vp, _i64, _i32, _i32, _i32, _vp, _vp, _vp, _vp,
                                         _vp, _vp, _i32, _vp, _vp]),
    "mv_vegas_iteration_static": (ctypes.c_int, [_P(Step), _P(Integrand), _vp, _i32, _vp, _vp, _vp, _vp]),
    "mv_batch_fused": (ctypes.c_int, [_vp, _i32, _vp, _i32, _i32, _i32, _i32, _vp, _vp, _vp, _i32, _vp]),
    "mv_batch_fused_pairs": (ctypes.c_int, [_vp, _i32, _vp, _i32, _i32, _i32, _vp, _vp, _vp, _i32, _vp]),
    "mv_batch_pairs_blocks": (_i32, [_i32, _i32]),
    "mv_batch_fused_pairs_f64": (ctypes.c_int, [_vp, _i32, _vp, _i32, _i32, _i32, _vp, _vp, _vp, _i32, _vp]),
    "mv_batch_pairs_blocks_f64": (_i32, [_i32, _i32]),
    "mv_batch_finalize": (ctypes.c_int, [_vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _i32, _vp]),
    "mv_batch_vplus_fused": (ctypes.c_int, [_vp, _vp, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp,
                                            _vp, _vp, _i32, _vp]),
    "mv_batch_vplus_finalize": (ctypes.c_int, [_vp, _vp, _i32, _vp, _i32, _f64, _vp, _vp, _vp, _vp, _vp, _vp,
                                               _vp, _vp, _vp, _vp]),
}

_lib = None
_lock = threading.Lock()


class MivegasError(RuntimeError):
    """A launch or device failure reported by libmivegas.so."""


def lib_path():
    return os.environ.get("MIVEGAS_LIB") or _build.lib_path()


def load():
    """Load and check libmivegas.so (cached)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is not built: run `python -m vegasflow_amd._build` "
                "(or __graft_entry__.build()) first"
            )
        # the HIP runtime must be the one torch uses: import torch first so
        # that its libamdhip64 satisfies the library's dependency
        import torch  # noqa: F401

        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        v = lib.mv_abi_version()
        if v != ABI_VERSION:
            raise RuntimeError(f"{path}: ABI version {v}, expected {ABI_VERSION}; rebuild it")
        for i, st in enumerate(STRUCTS):
            if lib.mv_struct_size(i) != ctypes.sizeof(st):
                raise RuntimeError(f"{path}: layout of {st.__name__} differs from the binding")
        _lib = lib
        return lib


def last_error():
    msg = load().mv_last_error()
    return msg.decode() if msg else ""


def check(rc, what=""):
    """Map a negative status to the exception the reference raises."""
    if rc >= 0:
        return rc
    msg = last_error() or what
    if rc == ERR_ARG:
        raise ValueError(msg)
    if rc == ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise MivegasError(f"{msg} (status {rc})")


def call(name, *args):
    return check(getattr(load(), name)(*args), name)
