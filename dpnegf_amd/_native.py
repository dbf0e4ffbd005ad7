"""ctypes binding of libdpnegf_amd.so (the HIP C ABI in ``include/dpnegf_amd.h``).

There is no CPU implementation behind these calls: if the library is missing
or no HIP device is visible, every entry point raises.
"""
import ctypes
import os

import torch

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libdpnegf_amd.so")

ABI_VERSION = 7  # DPNEGF_ABI_VERSION of include/dpnegf_amd.h this binding matches

DPNEGF_FP64 = 0
DPNEGF_MIXED = 1
DPNEGF_MIXED_REFINED = 2

STATUS_SINGULAR = 1
STATUS_LS_NOCONV = 2
STATUS_LS_RESID_FAIL = 4
STATUS_LS_RESID_WARN = 8
STATUS_REFINE_FALLBACK = 16
STATUS_BAD_INDEX = 32

TUNE_GEMM_TILE = 0
TUNE_INVERSE = 1
TILE_32x32, TILE_32x64, TILE_64x32, TILE_64x64 = 1, 2, 3, 4
INV_REGISTER, INV_LDS, INV_SCHUR = 1, 2, 3

TRANSMISSION_NMAX = 64
LEAD_NMAX = 48
FULL_NMAX = 1024
NOALLOC_SCRATCH_ELEMS = 8388608

_lib = None

_vp = ctypes.c_void_p
_i32 = ctypes.c_int
_i64 = ctypes.c_int64
_f64 = ctypes.c_double

_SIGS = {
    "dpnegf_abi_version": ([], _i32),
    "dpnegf_last_error": ([], ctypes.c_char_p),
    "dpnegf_arena_elems": ([_i32, _vp], _i64),
    "dpnegf_rgf_transmission": ([_vp, _i32, _i32, _vp, _vp, _vp, _i64, _i32, _vp, _i32, _vp, _vp,
                                 _i32, _vp, _i32, _i32, _vp, _vp, _vp], _i32),
    "dpnegf_rgf_transmission_workspace_elems": ([_i32, _vp, _i32], _i64),
    "dpnegf_lead_self_energy": ([_vp, _i32, _i32, _i32, _vp, _i64, _i32, _vp, _i32, _vp, _f64, _f64,
                                 _vp, _vp, _vp, _i32, _vp, _vp, _vp], _i32),
    "dpnegf_lead_self_energy_noalloc": ([_vp, _i32, _i32, _i32, _vp, _i64, _i32, _vp, _i32, _vp,
                                         _f64, _f64, _vp, _vp, _i32, _vp, _vp, _vp], _i32),
    "dpnegf_rgf_full_workspace_elems": ([_i32, _vp, _i32], _i64),
    "dpnegf_lead_self_energy_workspace_elems": ([_i32, _i32, _i32], _i64),
    "dpnegf_rgf_full": ([_vp, _i32, _vp, _vp, _i32, _vp, _vp, _i32, _vp, _i32, _i32] + [_vp] * 23,
                        _i32),
    "dpnegf_rgf_grad_workspace_elems": ([_i32, _vp, _i32], _i64),
    "dpnegf_rgf_transmission_grad": ([_vp, _i32, _vp, _vp, _i32, _vp, _vp, _i32, _vp, _i32, _i32,
                                      _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
                                     _i32),
    "dpnegf_rgf_grad_ckpt_workspace_elems": ([_i32, _vp, _i32, _i32], _i64),
    "dpnegf_rgf_transmission_grad_ckpt": ([_vp, _i32, _vp, _vp, _i32, _vp, _vp, _i32, _vp, _i32,
                                           _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                           _i32, _vp, _vp],
                                          _i32),
    "dpnegf_batch_sum_conjt": ([_vp, _vp, _vp, _i32, _i32, _i32, _i32], _i32),
    "dpnegf_weighted_re_sum": ([_vp, _vp, _vp, _vp, _i64, _i32, _i32], _i32),
    "dpnegf_zgemm_batched": ([_vp, _i32, _i32, _i32, _i32, _vp, _i32, _i64, _i32, _vp, _i32, _i64,
                              _i32, _f64, _f64, _f64, _f64, _vp, _i32, _i64], _i32),
    "dpnegf_diag_product": ([_vp, _vp, _vp, _i64, _vp, _i64, _i32, _i32, _i32, _i32], _i32),
    "dpnegf_zinv_batched": ([_vp, _vp, _i32, _i32, _vp], _i32),
    "dpnegf_rgf_transmission_blocked_workspace_elems": ([_i32, _vp, _i32], _i64),
    "dpnegf_autotune_query": ([_i32, _vp, _i64], _i64),
    "dpnegf_tune_set": ([_i32, _i32, _i32, _i32, _i32], _i32),
    "dpnegf_tune_get": ([_i32, _i32, _i32, _i32], _i32),
    "dpnegf_tune_builtin": ([_i32, _i32, _i32, _i32], _i32),
    "dpnegf_tune_clear": ([], None),
    "dpnegf_tune_record": ([_i32], None),
    "dpnegf_tune_recorded": ([_vp, _i32], _i32),
    "dpnegf_rgf_transmission_blocked": ([_vp, _i32, _vp, _vp, _i32, _vp, _vp, _i32, _vp, _i32, _i32,
                                         _vp, _vp, _vp], _i32),
}

EXPORTED = tuple(_SIGS)


class NativeError(RuntimeError):
    pass


def lib_path():
    return _LIB_PATH


def load():
    """Load the library (no device needed).  Raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise NativeError(f"{_LIB_PATH} not found: run __graft_entry__.build() "
                              "(or dpnegf_amd/csrc/build.sh) first")
        lib = ctypes.CDLL(_LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        if lib.dpnegf_abi_version() != ABI_VERSION:
            raise NativeError(f"{_LIB_PATH} has ABI version {lib.dpnegf_abi_version()}, this "
                              f"binding expects {ABI_VERSION}: rebuild the library")
        _lib = lib
    return _lib


def call(name, *args):
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.dpnegf_last_error().decode()
        if rc == -1:
            raise ValueError(f"{name}: {msg}")
        raise NativeError(f"{name} failed ({rc}): {msg}")
    return rc


def require_device(*tensors):
    """The HIP library only runs on the GPU: reject host tensors loudly."""
    for t in tensors:
        if torch.is_tensor(t) and t.device.type != "cuda":
            raise RuntimeError("dpnegf_amd runs on the GPU only: got a tensor on "
                               f"'{t.device}'; move inputs to a HIP device ('cuda')")
    if not torch.cuda.is_available():
        raise RuntimeError("dpnegf_amd needs a HIP device; none is visible")


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def host_i32(values):
    arr = (ctypes.c_int32 * len(values))(*[int(v) for v in values])
    return arr
