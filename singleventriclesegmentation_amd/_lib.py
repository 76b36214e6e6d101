"""Loads libsvsflow.so (the HIP kernels behind a C ABI) through ctypes.

There is no fallback: if the library is missing or cannot be loaded, every
operator raises.  torch is imported first so that the library binds to the HIP
runtime torch already loaded (both carry the soname libamdhip64.so.7).
"""
import contextlib
import ctypes
import os

import torch  # noqa: F401  (must precede loading the HIP library)

_HERE = os.path.dirname(os.path.abspath(__file__))
# SVS_FLOW_LIB selects another build of the same sources (the host-sanitized
# one of build.py --host-sanitize); either way there is no fallback.
LIB_PATH = os.environ.get("SVS_FLOW_LIB") or os.path.join(_HERE, "libsvsflow.so")

c_float_p = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_i64 = ctypes.c_int64


class TVL1Params(ctypes.Structure):
    """Mirror of svs_tvl1_params (include/svsflow.h)."""

    _fields_ = [
        ("num_scales", _i), ("max_warps", _i), ("max_outer_iterations", _i),
        ("weight_matching", _f), ("weight_tv", _f),
        ("pd_type", _i),
        ("sigma", ctypes.c_double), ("tau", ctypes.c_double), ("theta", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("use_aniso", _i), ("alpha", _f), ("beta", _f),
        ("use_blur", _i), ("blur_K", _i), ("blur_weights", _vp),
        ("interpolation", _i), ("boundary", _i),
        ("length_fixed", _i), ("LZ", _f), ("LY", _f), ("LX", _f),
        ("use_initial_state", _i),
        ("sweep_bx", _i), ("sweep_by", _i), ("sweep_cz", _i),
        ("huber_eps", ctypes.c_double),
        ("relax_theta", ctypes.c_double),
    ]


class AugVolume(ctypes.Structure):
    """Mirror of svs_aug_volume (include/svsflow.h)."""

    _fields_ = [("src", _vp), ("dst", _vp), ("batch", _i), ("channels", _i), ("mode", _i), ("vector", _i)]


_SIGNATURES = {
    "svs_last_error": ([], ctypes.c_char_p),
    "svs_fill": ([_vp, _i64, _f, _vp], _i),
    "svs_nabla3d_forward": ([_vp, _i, _i, _i, _i, _f, _f, _f, _i, _vp, _vp], _i),
    "svs_nabla3d_backward": ([_vp, _i, _i, _i, _i, _f, _f, _f, _i, _vp, _vp], _i),
    "svs_aniso_tangents": ([_vp, _i64, _f, _f, _vp, _vp, _vp, _vp, _vp], _i),
    "svs_aniso_apply": ([_vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp], _i),
    "svs_prox_primal": ([_vp, _f, _f, _vp, _vp, _i64, _vp, _vp], _i),
    "svs_prox_dual": ([_vp, _f, _f, _i64, _vp, _vp], _i),
    "svs_warp3d": ([_vp, _i, _vp, _i, _i, _i, _f, _f, _f, _i, _i, _vp, _vp], _i),
    "svs_warp3d_cnn": ([_vp, _vp, _i, _i, _i, _i, _i, _f, _f, _f, _i, _i, _vp, _vp], _i),
    "svs_warp3d_backward": ([_vp, _vp, _vp, _i, _i, _i, _i, _i, _f, _f, _f, _i, _i, _vp, _vp, _vp], _i),
    "svs_prolongate3d": ([_vp, _i, _i, _i, _i, _f, _f, _f, _i, _i, _i, _f, _f, _f, _i, _i, _vp, _vp], _i),
    "svs_blur3d": ([_vp, _i, _i, _i, _vp, _i, _vp, _vp], _i),
    "svs_median3d": ([_vp, _i, _i, _i, _i, _i, _i, _vp, _vp], _i),
    "svs_mask_counts": ([_vp, _vp, _i, _i, _i, _vp, _vp], _i),
    "svs_mask_hausdorff_workspace_bytes": ([_i, _i, _i], ctypes.c_size_t),
    "svs_mask_hausdorff2": ([_vp, _vp, _i, _i, _i, _vp, ctypes.c_size_t, _vp, _vp], _i),
    "svs_mask_hausdorff2_noalloc": ([_vp, _vp, _i, _i, _i, _vp, _vp, _vp], _i),
    "svs_spatial_augment": ([ctypes.POINTER(AugVolume), _i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _i, _i, _i,
                             _i, _i, _vp], _i),
    "svs_tvl1_workspace_bytes": ([ctypes.POINTER(TVL1Params), _i, ctypes.POINTER(_i)], ctypes.c_size_t),
    "svs_tvl1_plan_create": ([ctypes.POINTER(TVL1Params), _i, ctypes.POINTER(_i), _vp, ctypes.c_size_t, _vp,
                              ctypes.POINTER(_vp)], _i),
    "svs_tvl1_plan_bind": ([_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "svs_tvl1_plan_run": ([_vp, _i, _vp], _i),
    "svs_tvl1_plan_destroy": ([_vp], _i),
    "svs_tvl1_plan_info": ([_vp, _i, ctypes.POINTER(_i)], _i),
    "svs_tvl1_plan_sweeps": ([_vp, _i, _i, _vp], _i),
    "svs_tvl1_plan_set_shape": ([_vp, _i, _i, _i, _i, _vp], _i),
    "svs_tvl1_plan_set_resident": ([_vp, _i, _i, _vp], _i),
    "svs_tvl1_plan_set_stats": ([_vp, _vp, _vp], _i),
    "svs_tvl1_plan_autotune": ([_vp, _i, _vp, ctypes.POINTER(_i)], _i),
    "svs_tvl1_solve": ([ctypes.POINTER(TVL1Params), _vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, ctypes.c_size_t,
                        _vp], _i),
}

_lib = None


def lib():
    """The loaded library (raises if libsvsflow.so was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m singleventriclesegmentation_amd.build` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        l = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(l, name)
            fn.argtypes = args
            fn.restype = res
        _lib = l
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def check(rc):
    """Raises RuntimeError with the library's message for a non-zero status."""
    if rc != 0:
        msg = lib().svs_last_error().decode("utf-8", "replace")
        raise RuntimeError(msg or f"libsvsflow call failed with status {rc}")


# torch's raw current-stream and current-device queries (what its own
# generated code calls); they skip building a Stream object and re-checking
# lazy initialisation per call, which dominated the host time of the
# operator-by-operator path.  Builds without them use the public API.  Every
# caller holds a CUDA tensor already, so torch's CUDA state is initialised.
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_raw_device = getattr(torch._C, "_cuda_getDevice", None) or torch.cuda.current_device
_SAME_DEVICE = contextlib.nullcontext()


def stream_ptr(device=None):
    """torch's current stream on ``device`` (default: the current device) as a hipStream_t
    address for a c_void_p argument."""
    if _raw_stream is not None:
        if device is None or (isinstance(device, torch.device) and device.index is None):
            idx = _raw_device()
        else:
            idx = device.index if isinstance(device, torch.device) else int(device)
        return _raw_stream(idx)
    return torch.cuda.current_stream(device).cuda_stream


def on_device(device):
    """Context that makes ``device`` current for the launches inside it; a
    no-op when it already is (the common case, and a measurable share of an
    operator call's host time otherwise)."""
    cur = _raw_device()
    idx = device.index if device.index is not None else cur
    if idx == cur:
        return _SAME_DEVICE
    return torch.cuda.device(idx)


def ptr(t):
    """A tensor's device address for a c_void_p argument (None passes NULL);
    ctypes converts the plain int itself."""
    return t.data_ptr() if t is not None else None
