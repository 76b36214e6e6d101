"""Mask-propagation metrics of svs/modules/flow/warping.py:245-270 on the GPU.

The reference calls monai's compute_dice and compute_hausdorff_distance with
include_background=False on one binary channel (which monai then keeps).
Here the counts and the edge-to-edge distances come from HIP kernels
(svs_mask_counts, svs_mask_hausdorff2) and only the final division and square
root run on the host, in the float32 / float64 precision monai uses:

* Dice = 2|A and B| / (|A| + |B|) as a float32 quotient; NaN when the ground
  truth is empty (monai's ignore_empty=True).
* Hausdorff = the larger of the two directed distances between the masks'
  6-connected edge voxels (mask XOR binary erosion), Euclidean in voxel units,
  the float32 of the float64 root; NaN when either mask has no edge voxel.
"""
import math
from typing import Tuple

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr, stream_ptr


def _mask3d(t: torch.Tensor, name: str) -> torch.Tensor:
    if t.device.type != "cuda" or t.dtype != torch.float32:
        raise RuntimeError(f"{name} must be a float32 CUDA tensor")
    while t.dim() > 3 and t.shape[0] == 1:
        t = t[0]
    if t.dim() != 3:
        raise RuntimeError(f"{name} must be one mask [Z,Y,X] (leading singleton dims allowed)")
    return t.contiguous()


def mask_counts(y_pred: torch.Tensor, y: torch.Tensor) -> Tuple[int, int, int]:
    """(|pred|, |gt|, |pred and gt|) of two binary masks (value > 0.5 is set)."""
    a, b = _mask3d(y_pred, "y_pred"), _mask3d(y, "y")
    if a.shape != b.shape or a.device != b.device:
        raise RuntimeError("masks must have the same shape and device")
    with torch.cuda.device(a.device):
        out = torch.empty(3, dtype=torch.int64, device=a.device)
        check(_lib.lib().svs_mask_counts(ptr(a), ptr(b), *a.shape, ptr(out), stream_ptr()))
        c = out.cpu().tolist()
    return c[0], c[1], c[2]


def compute_dice(y_pred: torch.Tensor, y: torch.Tensor) -> float:
    """monai compute_dice(y_pred, y, include_background=False) on one binary channel."""
    na, nb, nab = mask_counts(y_pred, y)
    if nb == 0:
        return float("nan")
    return float(np.float32(2.0 * nab) / np.float32(na + nb))


def directed_hausdorff2(y_pred: torch.Tensor, y: torch.Tensor) -> Tuple[int, int, int, int]:
    """(d2 pred->gt, d2 gt->pred, edge voxels of pred, edge voxels of gt)."""
    a, b = _mask3d(y_pred, "y_pred"), _mask3d(y, "y")
    if a.shape != b.shape or a.device != b.device:
        raise RuntimeError("masks must have the same shape and device")
    lib = _lib.lib()
    nbytes = lib.svs_mask_hausdorff_workspace_bytes(*a.shape)
    with torch.cuda.device(a.device):
        ws = torch.empty((nbytes + 7) // 8, dtype=torch.int64, device=a.device)
        out = torch.empty(2, dtype=torch.int64, device=a.device)
        check(lib.svs_mask_hausdorff2(ptr(a), ptr(b), *a.shape, ptr(ws), ws.numel() * 8, ptr(out), stream_ptr()))
        d = out.cpu().tolist()
        e = ws[:2].cpu().tolist()
    return d[0], d[1], e[0], e[1]


def compute_hausdorff_distance(y_pred: torch.Tensor, y: torch.Tensor) -> float:
    """monai compute_hausdorff_distance(y_pred, y, include_background=False)
    (Euclidean, symmetric, no percentile) on one binary channel."""
    d_ab, d_ba, ea, eb = directed_hausdorff2(y_pred, y)
    if ea == 0 or eb == 0:
        return float("nan")
    return float(np.float32(math.sqrt(float(max(d_ab, d_ba)))))


def propagation_metrics(forward_masks: torch.Tensor, backward_masks: torch.Tensor):
    """warping.py:245-270 (_compute_metrics): masks [T,1,Z,Y,X] from propagate();
    returns (dice_bwd, dice_fwd, hausdorff_bwd, hausdorff_fwd)."""
    initial_gt, initial_pred = forward_masks[0], backward_masks[0]
    final_gt, final_pred = backward_masks[-1], forward_masks[-1]
    return (compute_dice(initial_pred, initial_gt), compute_dice(final_pred, final_gt),
            compute_hausdorff_distance(initial_pred, initial_gt), compute_hausdorff_distance(final_pred, final_gt))
