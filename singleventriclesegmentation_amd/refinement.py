"""Propagation-refinement training with a differentiable tricubic warp.

One training step on a batch of frame pairs (t, t+1): the U-Net predicts
4-class probabilities at both frames from (image, propagated mask); a
forward-backward consistency loss warps the predictions of frame t forward
and those of frame t+1 backward along the precomputed flows with the
hand-written tricubic warp (WarpingCNN3D) and its adjoint kernels (gradient
with respect to the warped volume and to the flow), next to a supervised
term at frame t and a penalty on the raw head output (the three terms of
svs/modules/loss.py:21-64).  An optional learnable flow residual makes the
flow adjoint part of every step.

The reference trains with a LINEAR warp that has no backward, so gradients
stop at the warp (svs/models/lit_unet.py:114, warpingOpsCNN.h:36-41); here
they flow through it.  This objective extends the reference; its own training
loop (LitUNet, PropagationLoss) is training.py.
"""
from dataclasses import dataclass
from typing import Dict

import torch
import torch.nn.functional as F

from . import opticalFlow
from .unet import UNet3D
from .warpingOps import warp


@dataclass
class RefinementConfig:
    num_classes: int = 4
    features_start: int = 16
    num_layers: int = 4
    lambda_u: float = 1.0     # consistency weight (conf/model/unet.yaml:3)
    gamma_p: float = 0.1      # penalty weight (conf/model/unet.yaml:4)
    lr: float = 1e-3
    betas: tuple = (0.9, 0.999)
    weight_decay: float = 1e-5
    learn_flow_residual: bool = True
    # Let MIOpen search its convolution algorithms once per shape (torch's
    # cudnn.benchmark switch).  Its default heuristics choose naive direct
    # kernels for several FP32 3D convolutions of this U-Net; the searched
    # implicit-GEMM ones train batch 4 x 80^3 about 40x faster.
    conv_search: bool = True


def mesh_for(shape):
    nz, ny, nx = shape
    return opticalFlow.MeshInfo3D(nz, ny, nx, nz - 1, ny - 1, nx - 1)


class Refiner:
    """Model, optimiser and warp operator for one volume size."""

    def __init__(self, shape, batch, device, cfg: RefinementConfig = RefinementConfig()):
        self.cfg = cfg
        self.device = device
        if cfg.conv_search:
            torch.backends.cudnn.benchmark = True
        self.model = UNet3D(cfg.num_layers, cfg.num_classes, 2, cfg.features_start).to(device)
        params = list(self.model.parameters())
        self.flow_res = None
        if cfg.learn_flow_residual:
            self.flow_res = torch.zeros((2, batch) + tuple(shape) + (3,), device=device, requires_grad=True)
            params.append(self.flow_res)
        self.opt = torch.optim.Adam(params, lr=cfg.lr, betas=cfg.betas, weight_decay=cfg.weight_decay)
        self.warp_op = opticalFlow.WarpingCNN3D(mesh_for(shape), opticalFlow.INTERPOLATE_CUBIC_HERMITESPLINE,
                                                opticalFlow.BOUNDARY_MIRROR)

    def losses(self, img_t, img_t1, mask_t, mask_prop_t1, label_t, flow_fwd, flow_bwd) -> Dict[str, torch.Tensor]:
        """img_*, mask_* [B,1,Z,Y,X]; label_t [B,Z,Y,X] class indices; flows [B,Z,Y,X,3]."""
        if self.flow_res is not None:
            flow_fwd = flow_fwd + self.flow_res[0]
            flow_bwd = flow_bwd + self.flow_res[1]
        out_t, head_t = self.model(torch.cat([img_t, mask_t], dim=1))
        out_t1, head_t1 = self.model(torch.cat([img_t1, mask_prop_t1], dim=1))
        p_t = torch.softmax(out_t, dim=1)
        p_t1 = torch.softmax(out_t1, dim=1)
        bs = p_t.shape[0]
        onehot = F.one_hot(label_t.long(), self.cfg.num_classes).permute(0, 4, 1, 2, 3).to(p_t.dtype)
        sup = F.mse_loss(p_t, onehot, reduction="sum") / bs
        fwd = warp(p_t, flow_fwd, self.warp_op)    # frame t carried to t+1
        bwd = warp(p_t1, flow_bwd, self.warp_op)   # frame t+1 carried to t
        cons = (F.mse_loss(fwd, p_t1, reduction="sum") + F.mse_loss(bwd, p_t, reduction="sum")) / bs
        pen = (head_t.pow(2).sum() + head_t1.pow(2).sum()) / bs
        total = sup + self.cfg.lambda_u * cons + self.cfg.gamma_p * pen
        return {"total": total, "supervised": sup, "consistency": cons, "penalty": pen}

    def step(self, *batch) -> Dict[str, torch.Tensor]:
        self.opt.zero_grad(set_to_none=True)
        out = self.losses(*batch)
        out["total"].backward()
        self.opt.step()
        return out


def synthetic_batch(shape, batch, device, seed=0):
    """Frames t, t+1 of ``batch`` phantom subjects, the mask at t, its
    propagation to t+1, a 4-class label at t and small smooth flows."""
    import numpy as np
    from . import phantom

    rng = np.random.default_rng(seed)
    imgs0, imgs1, masks, labels, ff, fb = [], [], [], [], [], []
    for b in range(batch):
        f0, f1, d = phantom.frame_pair(shape, seed0=seed + 3 * b, seed1=seed + 3 * b + 1, disp_seed=seed + 3 * b + 2,
                                       peak=1.5)
        m = phantom.ventricle_mask(shape).astype(np.float32)
        lab = m.astype(np.int64) + (phantom.clean_frame(shape, smooth=0) > 0.4) + (rng.random(shape) > 0.999)
        imgs0.append(f0); imgs1.append(f1); masks.append(m); labels.append(np.clip(lab, 0, 3))
        ff.append(d); fb.append(-d)
    t = lambda a, dt=torch.float32: torch.from_numpy(np.stack(a)).to(device=device, dtype=dt)  # noqa: E731
    img_t, img_t1, mask_t = t(imgs0)[:, None], t(imgs1)[:, None], t(masks)[:, None]
    flow_fwd, flow_bwd = t(ff), t(fb)
    op = opticalFlow.WarpingCNN3D(mesh_for(shape), opticalFlow.INTERPOLATE_CUBIC_HERMITESPLINE,
                                  opticalFlow.BOUNDARY_MIRROR)
    mask_prop = op.forward(mask_t.contiguous(), flow_fwd.contiguous())
    return img_t, img_t1, mask_t, mask_prop, t(labels, torch.int64), flow_fwd, flow_bwd
