"""3D residual U-Net of the propagation-refinement model (architecture of
svs/models/unet.py:9-139): ``num_layers`` resolution levels, two 3x3x3
convolutions with instance norm and LeakyReLU(0.1) per level, max-pool
down-sampling, transposed-convolution up-sampling with skip concatenation,
a 1x1x1 head, and the residual ``output + x[:, 1:]`` (the propagated mask
channel) returned together with the raw head output.

With 4 levels, 16 base channels, 2 input channels and 4 classes it has
1,401,044 parameters.
"""
from typing import Tuple

import torch
from torch import nn
import torch.nn.functional as F


def _conv_block(cin: int, cout: int, k: int, pad: int) -> nn.Sequential:
    return nn.Sequential(
        nn.Conv3d(cin, cout, kernel_size=k, padding=pad), nn.InstanceNorm3d(cout), nn.LeakyReLU(0.1),
        nn.Conv3d(cout, cout, kernel_size=k, padding=pad), nn.InstanceNorm3d(cout), nn.LeakyReLU(0.1))


class _Up(nn.Module):
    def __init__(self, cin: int, cout: int, k: int, pad: int, trilinear: bool):
        super().__init__()
        if trilinear:
            self.up = nn.Sequential(nn.Upsample(scale_factor=2, mode="trilinear", align_corners=True),
                                    nn.Conv3d(cin, cin // 2, kernel_size=1))
        else:
            self.up = nn.ConvTranspose3d(cin, cin // 2, kernel_size=2, stride=2)
        self.conv = _conv_block(cin, cout, k, pad)

    def forward(self, x: torch.Tensor, skip: torch.Tensor) -> torch.Tensor:
        x = self.up(x)
        dz, dy, dx = (s - t for s, t in zip(skip.shape[2:], x.shape[2:]))
        x = F.pad(x, [dx // 2, dx - dx // 2, dy // 2, dy - dy // 2, dz // 2, dz - dz // 2])
        return self.conv(torch.cat([skip, x], dim=1))


class UNet3D(nn.Module):
    def __init__(self, num_layers: int = 4, num_classes: int = 4, input_channels: int = 2,
                 features_start: int = 16, trilinear: bool = False, padding: int = 1, kernel_size: int = 3):
        super().__init__()
        if num_layers < 1:
            raise ValueError(f"num_layers = {num_layers}, expected num_layers > 0")
        self.num_layers = num_layers
        f = features_start
        self.inc = _conv_block(input_channels, f, kernel_size, padding)
        self.down = nn.ModuleList()
        for _ in range(num_layers - 1):
            self.down.append(nn.Sequential(nn.MaxPool3d(2, 2), _conv_block(f, 2 * f, kernel_size, padding)))
            f *= 2
        self.up = nn.ModuleList()
        for _ in range(num_layers - 1):
            self.up.append(_Up(f, f // 2, kernel_size, padding, trilinear))
            f //= 2
        self.head = nn.Conv3d(f, num_classes, kernel_size=1)

    def forward(self, x: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        identity = x[:, 1:]
        feats = [self.inc(x)]
        for layer in self.down:
            feats.append(layer(feats[-1]))
        y = feats.pop()
        for layer in self.up:
            y = layer(y, feats.pop())
        out = self.head(y)
        return out + identity, out


def parameter_count(model: nn.Module) -> int:
    return sum(p.numel() for p in model.parameters())
