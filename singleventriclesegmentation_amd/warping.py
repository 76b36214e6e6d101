"""Drop-in for the warp harness of svs/modules/flow/warping.py: ``Warp``
(warping.py:35-65, including the misspelt ``lenghts`` keyword) and mask
propagation along forward/backward flow chains (``propagate``, the
semantics of OpticalFlowWarper._propagate, warping.py:205-243).
"""
from typing import Tuple

import torch

from . import opticalFlow
from .flow_utils import get_interpolation_type, get_mesh_length


class Warp:
    """3D warping of volumes x [BS,C,Z,Y,X] by flows u [BS,Z,Y,X,3] (WarpingCNN3D)."""

    def __init__(self, interpolation_type: str, boundary_type: str, mesh_length_type: str,
                 lenghts: Tuple[int, int, int]):
        self.interpolation, self.boundary = get_interpolation_type(interpolation_type, boundary_type)
        self.mesh_length_type = mesh_length_type
        self.LZ, self.LY, self.LX = lenghts

    def op(self, x: torch.Tensor) -> opticalFlow.WarpingCNN3D:
        NZ, NY, NX = x.shape[2:]
        LZ, LY, LX = get_mesh_length(self.mesh_length_type, NZ, NY, NX, self.LZ, self.LY, self.LX)
        mesh_info = opticalFlow.MeshInfo3D(NZ, NY, NX, LZ, LY, LX)
        return opticalFlow.WarpingCNN3D(mesh_info, self.interpolation, self.boundary)

    def __call__(self, x: torch.Tensor, u: torch.Tensor) -> torch.Tensor:
        return self.op(x).forward(x.contiguous(), u.contiguous())


def propagate(warp: Warp, initial_mask: torch.Tensor, final_mask: torch.Tensor, forward_flow: torch.Tensor,
              backward_flow: torch.Tensor, threshold: float = 0.5):
    """Mask propagation of OpticalFlowWarper._propagate (warping.py:205-243).

    initial_mask/final_mask [1,1,Z,Y,X]; flows [T,Z,Y,X,3].  The forward chain
    starts at the initial mask, the backward chain at the final mask; both
    warp the *continuous* mask and threshold at 0.5 only at the end.  Returns
    (forward_masks, backward_masks), each [T+1,1,Z,Y,X] in {0, 1}, the
    backward chain reversed so index 0 is the estimate of the initial mask."""
    assert forward_flow.shape[0] == backward_flow.shape[0], "Forward and backward flow have different timesteps"
    fwd = [initial_mask]
    bwd = [final_mask]
    for t in range(forward_flow.shape[0]):
        fwd.append(warp(fwd[-1], forward_flow[t].unsqueeze(0)))
        bwd.append(warp(bwd[-1], backward_flow[t].unsqueeze(0)))
    bwd.reverse()
    fwd_m = torch.where(torch.cat(fwd, dim=0) > threshold, 1.0, 0.0)
    bwd_m = torch.where(torch.cat(bwd, dim=0) > threshold, 1.0, 0.0)
    return fwd_m, bwd_m
