"""Helpers of svs/utils/flow_utils.py with the same names and behaviour.

The Gaussian pre-blur runs on the GPU (svs_blur3d); its weights are built on
the host with the same torch calls the reference makes, so the kernel taps
are the reference's (including the 1.5-voxel tap spacing of the linspace at
flow_utils.py:87 for sigma = 1).  The median post-filter (flow_utils.py:115-132,
a SciPy host round trip in the reference) runs on the GPU as well
(svs_median3d), with SciPy's window, rank and 'reflect' extension.
"""
import enum
from typing import Literal, Tuple, Union

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr, stream_ptr
from .opticalFlow import BoundaryType, InterpolationType


class FlowDirection(str, enum.Enum):
    """svs/utils/enums.py:5-7."""

    FORWARD = "forward"
    BACKWARD = "backward"

    def __str__(self):
        return self.value


def get_interpolation_type(inter_type_str: Literal['NEAREST', 'LINEAR', 'CUBIC_HERMITESPLINE'],
                           bound_type_str: Literal['NEAREST', 'MIRROR', 'REFLECT']):
    """String names -> (InterpolationType, BoundaryType) (flow_utils.py:11-45)."""
    interp = {"NEAREST": InterpolationType.INTERPOLATE_NEAREST,
              "LINEAR": InterpolationType.INTERPOLATE_LINEAR,
              "CUBIC_HERMITESPLINE": InterpolationType.INTERPOLATE_CUBIC_HERMITESPLINE}.get(inter_type_str)
    if interp is None:
        raise Exception("Wrong interpolation type")
    bound = {"NEAREST": BoundaryType.BOUNDARY_NEAREST,
             "MIRROR": BoundaryType.BOUNDARY_MIRROR,
             "REFLECT": BoundaryType.BOUNDARY_REFLECT}.get(bound_type_str)
    if bound is None:
        raise Exception("wrong boundary type")
    return interp, bound


def get_mesh_length(length_type: Literal['numDofs', 'fixed'], nz: int, ny: int, nx: int, lz, ly, lx):
    """numDofs: L = N - 1 per axis (h = 1); fixed: the given lengths (flow_utils.py:48-71)."""
    if length_type == "numDofs":
        return nz - 1, ny - 1, nx - 1
    if length_type == "fixed":
        return lz, ly, lx
    raise Exception("Wrong length type")


def _generate_gaussian_kernel3d(sigma: float) -> torch.Tensor:
    """1D taps of the pre-blur (flow_utils.py:74-91), computed with the same torch calls.

    Size int(5 sigma) made odd; positions from linspace(-K//2, K//2 + 1, K),
    i.e. a spacing of (K + 2)/(K - 1) voxels; normalised to sum 1."""
    K = int(sigma * 5)
    if K % 2 == 0:
        K += 1
    pos = torch.linspace(-K // 2, K // 2 + 1, K)
    g = torch.exp(-(pos / sigma) ** 2 / 2)
    return g / g.sum()


def gaussian_weights3d(sigma: float) -> np.ndarray:
    """The normalised K^3 outer-product kernel (flow_utils.py:107-109), as float32 [K,K,K]."""
    k = _generate_gaussian_kernel3d(sigma)
    k3 = torch.einsum('i,j,k->ijk', k, k, k)
    k3 = k3 / k3.sum()
    return np.ascontiguousarray(k3.numpy(), dtype=np.float32)


_weights_cache = {}


def _device_weights(sigma, device):
    key = (float(sigma), str(device))
    w = _weights_cache.get(key)
    if w is None:
        w = torch.from_numpy(gaussian_weights3d(sigma)).to(device)
        _weights_cache[key] = w
    return w


def apply_gaussian_blur3d(vol: torch.Tensor, sigma: float) -> torch.Tensor:
    """Zero-padded 3D Gaussian blur of a [Z,Y,X] CUDA volume (flow_utils.py:94-112)."""
    if vol.device.type != "cuda" or vol.dtype != torch.float32 or vol.dim() != 3:
        raise RuntimeError("apply_gaussian_blur3d expects a float32 [Z,Y,X] CUDA tensor")
    vol = vol.contiguous()
    w = _device_weights(sigma, vol.device)
    with torch.cuda.device(vol.device):
        out = torch.empty_like(vol)
        check(_lib.lib().svs_blur3d(ptr(vol), *vol.shape, ptr(w), w.shape[0], ptr(out), stream_ptr()))
    return out


def median_filter3d(u: torch.Tensor, ks: int) -> torch.Tensor:
    """Per-component median over a ks^3 window of a [Z,Y,X,C] flow or a batch
    [B,Z,Y,X,C] of them, on the tensor's GPU; a new float32 tensor.

    The result is scipy.ndimage.median_filter(u, size=(ks,ks,ks,1)) (per flow):
    the element of rank ks^3 // 2 of each window, with SciPy's default
    'reflect' extension at the faces."""
    if u.device.type != "cuda" or u.dtype != torch.float32 or u.dim() not in (4, 5):
        raise RuntimeError("median filter expects a float32 [Z,Y,X,C] or [B,Z,Y,X,C] CUDA tensor")
    ub = (u if u.dim() == 5 else u.unsqueeze(0)).contiguous()
    B, NZ, NY, NX, C = ub.shape
    with torch.cuda.device(ub.device):
        out = torch.empty_like(ub)
        check(_lib.lib().svs_median3d(ptr(ub), B, NZ, NY, NX, C, int(ks), ptr(out), stream_ptr()))
    return out if u.dim() == 5 else out[0]


def apply_median_filter(u: torch.Tensor, ks: int, device: torch.device) -> torch.Tensor:
    """flow_utils.py:115-132: median over (ks, ks, ks, 1) of a [Z,Y,X,C] flow,
    returned on ``device``.  Computed on the GPU (``device`` must be a CUDA
    device) instead of through SciPy on the host; the values are SciPy's."""
    device = torch.device(device)
    if device.type != "cuda":
        raise RuntimeError("apply_median_filter runs on a CUDA device (got %s)" % device)
    if u.dim() != 4:
        raise RuntimeError("apply_median_filter expects a [Z,Y,X,C] tensor (size (ks, ks, ks, 1))")
    return median_filter3d(u.detach().to(device=device, dtype=torch.float32), ks)


def compute_timepoints(dia_ts: int, sys_ts: int, seg_dia, seg_sys, mode: Union[str, FlowDirection]):
    """Frame order from the earlier to the later phase, reversed for backward flow (flow_utils.py:135-169)."""
    if sys_ts < dia_ts:
        timepoints, segmentations = (sys_ts, dia_ts), (seg_sys, seg_dia)
    else:
        timepoints, segmentations = (dia_ts, sys_ts), (seg_dia, seg_sys)
    times = np.arange(timepoints[0], timepoints[1] + 1, 1)
    if str(mode) == FlowDirection.BACKWARD.value:
        times = np.flip(times)
    return timepoints, segmentations, times.tolist()
