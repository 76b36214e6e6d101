"""A plain-PyTorch tricubic Catmull-Rom warp (MIRROR boundary) built from
gathers and elementwise ops, so autograd differentiates it with respect to
the volume and the flow.  Test-only reference for the hand-written adjoint."""
import torch


def _mirror(i, n):
    i = torch.where(i < 0, -i - 1, i)
    i = torch.where(i > n - 1, 2 * n - 1 - i, i)
    return i.clamp(0, n - 1)


def _cr_weights(w):
    w2 = w * w
    w3 = w2 * w
    return (-0.5 * w3 + w2 - 0.5 * w, 1.5 * w3 - 2.5 * w2 + 1.0, -1.5 * w3 + 2.0 * w2 + 0.5 * w, 0.5 * w3 - 0.5 * w2)


def warp_ref(u, phi):
    """u [B,C,Z,Y,X], phi [B,Z,Y,X,3] (x, y, z) -> u(x + phi(x)) with unit spacing."""
    B, C, NZ, NY, NX = u.shape
    dev = u.device
    z, y, x = torch.meshgrid(torch.arange(NZ, device=dev), torch.arange(NY, device=dev),
                             torch.arange(NX, device=dev), indexing="ij")
    cx = x.to(u.dtype) + phi[..., 0]
    cy = y.to(u.dtype) + phi[..., 1]
    cz = z.to(u.dtype) + phi[..., 2]
    fx, fy, fz = torch.floor(cx), torch.floor(cy), torch.floor(cz)
    wx, wy, wz = _cr_weights(cx - fx), _cr_weights(cy - fy), _cr_weights(cz - fz)
    fx, fy, fz = fx.long(), fy.long(), fz.long()
    flat = u.reshape(B, C, -1)
    out = torch.zeros_like(u)
    for a in range(4):
        iz = _mirror(fz + a - 1, NZ)
        for b in range(4):
            iy = _mirror(fy + b - 1, NY)
            for c in range(4):
                ix = _mirror(fx + c - 1, NX)
                idx = ((iz * NY + iy) * NX + ix).reshape(B, 1, -1).expand(B, C, -1)
                vals = torch.gather(flat, 2, idx).reshape(u.shape)
                out = out + (wz[a] * wy[b] * wx[c]).unsqueeze(1) * vals
    return out
