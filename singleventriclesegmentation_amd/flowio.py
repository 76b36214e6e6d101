"""Flow files of scripts/flow.py and scripts/warp.py.

processor.py:108-120 stores a patient's flows as one float32 .npy of shape
(T-1, 3, X, Y, Z), ``us[j] = u.permute(3, 2, 1, 0)`` of each [Z,Y,X,3] flow,
named ``{patient}_{mode}_flow.npy``; datasets.py:279-291 reads it back and the
warp harness reorders it to [T-1, Z, Y, X, 3] with t3xyz_to_tzyx3 = (0, 4, 3,
2, 1) (datasets.py:34, warping.py:182-193).  Masks are stored (X, Y, Z) and
reordered with xyz_to_zyx = (2, 1, 0).
"""
import os

import numpy as np
import torch

T3XYZ_TO_TZYX3 = (0, 4, 3, 2, 1)   # datasets.py:34
TZYX3_TO_T3XYZ = (0, 4, 3, 2, 1)   # its inverse (the permutation is an involution)
XYZ_TO_ZYX = (2, 1, 0)             # datasets.py:27


def flow_filename(patient: str, mode: str) -> str:
    """processor.py:118."""
    return f"{patient}_{mode}_flow.npy"


def to_file_layout(flows_tzyx3: torch.Tensor) -> np.ndarray:
    """[T-1, Z, Y, X, 3] flows -> the (T-1, 3, X, Y, Z) float32 array processor.py saves."""
    if flows_tzyx3.dim() != 5 or flows_tzyx3.shape[-1] != 3:
        raise RuntimeError("flows must be [T-1, Z, Y, X, 3]")
    return np.ascontiguousarray(flows_tzyx3.detach().permute(*TZYX3_TO_T3XYZ).float().cpu().numpy())


def save_flows(directory: str, patient: str, mode: str, flows_tzyx3: torch.Tensor) -> str:
    path = os.path.join(directory, flow_filename(patient, mode))
    np.save(path, to_file_layout(flows_tzyx3))
    return path


def load_flows(directory: str, patient: str, mode: str, device=None) -> torch.Tensor:
    """datasets.py:279-291 + warping.py:182-193: the file's (T-1, 3, X, Y, Z) array as
    [T-1, Z, Y, X, 3] float32 on ``device``."""
    arr = np.load(os.path.join(directory, flow_filename(patient, mode)), allow_pickle=False)
    if arr.ndim != 5 or arr.shape[1] != 3:
        raise RuntimeError(f"expected a (T-1, 3, X, Y, Z) flow file, got {arr.shape}")
    t = torch.from_numpy(np.ascontiguousarray(arr.transpose(T3XYZ_TO_TZYX3), dtype=np.float32))
    return t.to(device) if device is not None else t


def mask_to_zyx(mask_xyz: np.ndarray, device=None) -> torch.Tensor:
    """warping.py:150-170: an (X, Y, Z) mask as a float32 [1, 1, Z, Y, X] tensor."""
    t = torch.from_numpy(np.ascontiguousarray(np.asarray(mask_xyz).transpose(XYZ_TO_ZYX), dtype=np.float32))
    t = t[None, None]
    return t.to(device) if device is not None else t
