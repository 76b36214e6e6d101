"""Per-patient resampling of svs/preprocessing/prolongation.py:55-102 on the GPU.

The reference resizes every patient's 4D image [Z,Y,X,T] and its two masks to
one mesh (conf/preprocessing.yaml: 80^3, LINEAR, MIRROR, numDofs) with
opticalFlow.Prolongation3D, rounds the resampled masks and pads the time axis
with zero frames to ``pad_time``.  NIfTI reading and writing stay with the
caller (nibabel is not part of this package).
"""
from typing import Tuple

import torch
import torch.nn.functional as F

from . import opticalFlow
from .flow_utils import get_interpolation_type, get_mesh_length


def time_padding(maxt: int, img: torch.Tensor) -> torch.Tensor:
    """prolongation.py:15-22: zero frames appended along the last (time) axis."""
    return F.pad(img, [0, maxt - img.shape[-1], 0, 0, 0, 0, 0, 0])


def prolongate_patient(img_zyxt: torch.Tensor, seg_dia_zyx: torch.Tensor, seg_sys_zyx: torch.Tensor,
                       shape_zyx: Tuple[int, int, int] = (80, 80, 80), pad_time: int = 35,
                       interpolation: str = "LINEAR", boundary: str = "MIRROR", length_type: str = "numDofs",
                       length_xyz: Tuple[float, float, float] = (1.0, 1.0, 1.0)):
    """prolongation.py:69-85 for one patient (tensors already in the reference's
    z,y,x(,t) axis order, float32 on a CUDA device).

    Returns (image [Z',Y',X',pad_time], end-diastole mask, end-systole mask)."""
    for t, name in ((img_zyxt, "image"), (seg_dia_zyx, "seg_dia"), (seg_sys_zyx, "seg_sys")):
        if t.device.type != "cuda" or t.dtype != torch.float32:
            raise RuntimeError(f"{name} must be a float32 CUDA tensor")
    if img_zyxt.dim() != 4 or seg_dia_zyx.dim() != 3 or seg_sys_zyx.dim() != 3:
        raise RuntimeError("expected an image [Z,Y,X,T] and masks [Z,Y,X]")
    if img_zyxt.shape[:3] != seg_dia_zyx.shape or seg_dia_zyx.shape != seg_sys_zyx.shape:
        raise RuntimeError("image and masks must share the spatial size")
    if img_zyxt.shape[3] > pad_time:
        raise RuntimeError(f"{img_zyxt.shape[3]} frames exceed pad_time={pad_time}")
    ip, bc = get_interpolation_type(interpolation, boundary)
    lx, ly, lz = length_xyz
    nz, ny, nx = img_zyxt.shape[:3]
    Lz, Ly, Lx = get_mesh_length(length_type, nz, ny, nx, lz, ly, lx)
    mz, my, mx = shape_zyx
    Mz, My, Mx = get_mesh_length(length_type, mz, my, mx, lz, ly, lx)
    op = opticalFlow.Prolongation3D(opticalFlow.MeshInfo3D(nz, ny, nx, Lz, Ly, Lx),
                                    opticalFlow.MeshInfo3D(mz, my, mx, Mz, My, Mx), ip, bc)
    seg_dia = op.forward(seg_dia_zyx.contiguous()).round()
    seg_sys = op.forward(seg_sys_zyx.contiguous()).round()
    img = op.forwardVectorField(img_zyxt.contiguous())
    return time_padding(pad_time, img), seg_dia, seg_sys
