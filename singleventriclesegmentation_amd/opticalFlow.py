"""Drop-in for the reference's ``opticalFlow_cuda_ext.opticalFlow`` module (3D part).

Same class, method and enum names, argument meaning, shapes and fresh-output
semantics as the pybind11 module (svs/src/bindings.cpp:36-167), backed by the
HIP kernels of libsvsflow.so.  Sizes come from the input tensors and the mesh
spacing from the MeshInfo3D, as in the reference launchers (e.g.
differentialOps.cu:1452-1457).  Differences, all additive:

* every tensor argument is validated (CUDA device, contiguous, float32, rank,
  matching spatial shape) and a violation raises RuntimeError; the reference
  checks only some arguments (coreDefines.h:10-12) and never launch errors;
* operators copy the mesh values instead of holding a reference to a
  MeshInfo3D that may be a temporary (differentialOps.h:77);
* work is launched on torch's current stream instead of the legacy default
  stream.
"""
import enum

import numpy as np
import torch

from . import _lib
from ._lib import check, on_device, ptr, stream_ptr


class BoundaryType(enum.IntEnum):
    """boundary.cuh:16-24 / bindings.cpp:51-55."""

    BOUNDARY_NEAREST = 0
    BOUNDARY_MIRROR = 1
    BOUNDARY_REFLECT = 2


class InterpolationType(enum.IntEnum):
    """coreDefines.h:94-100 / bindings.cpp:99-104."""

    INTERPOLATE_NEAREST = 0
    INTERPOLATE_LINEAR = 1
    INTERPOLATE_CUBIC_HERMITESPLINE = 2


# pybind11's export_values(): the enumerators are also module attributes.
BOUNDARY_NEAREST = BoundaryType.BOUNDARY_NEAREST
BOUNDARY_MIRROR = BoundaryType.BOUNDARY_MIRROR
BOUNDARY_REFLECT = BoundaryType.BOUNDARY_REFLECT
INTERPOLATE_NEAREST = InterpolationType.INTERPOLATE_NEAREST
INTERPOLATE_LINEAR = InterpolationType.INTERPOLATE_LINEAR
INTERPOLATE_CUBIC_HERMITESPLINE = InterpolationType.INTERPOLATE_CUBIC_HERMITESPLINE


class MeshInfo3D:
    """coreDefines.h:62-90: sizes N, lengths L (FP32) and spacing h = L/(N-1) in FP32."""

    def __init__(self, NZ: int, NY: int, NX: int, LZ: float, LY: float, LX: float):
        self._N = (int(NZ), int(NY), int(NX))
        self._L = tuple(np.float32(v) for v in (LZ, LY, LX))
        self._h = tuple(np.float32(L / np.float32(N - 1)) for L, N in zip(self._L, self._N))

    def getNZ(self): return self._N[0]
    def getNY(self): return self._N[1]
    def getNX(self): return self._N[2]
    def getLZ(self): return float(self._L[0])
    def getLY(self): return float(self._L[1])
    def getLX(self): return float(self._L[2])
    def gethZ(self): return float(self._h[0])
    def gethY(self): return float(self._h[1])
    def gethX(self): return float(self._h[2])

    def _hs(self):
        return float(self._h[0]), float(self._h[1]), float(self._h[2])

    def _Ls(self):
        return float(self._L[0]), float(self._L[1]), float(self._L[2])

    def __repr__(self):
        return f"MeshInfo3D(N={self._N}, L={tuple(float(v) for v in self._L)})"


def _check(t, name, ndim=None):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if t.dtype != torch.float32:
        raise RuntimeError(f"{name} must be a float32 tensor (got {t.dtype})")
    if ndim is not None and t.dim() != ndim:
        raise RuntimeError(f"Expected {ndim}d tensor for {name} (got {t.dim()}d)")


def _same_device(*ts):
    dev = ts[0].device
    for t in ts[1:]:
        if t.device != dev:
            raise RuntimeError("all tensors must be on the same device")
    return dev


def _shape_eq(t, shape, name):
    if tuple(t.shape) != tuple(shape):
        raise RuntimeError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")


def _enum(v, cls, what):
    try:
        return int(cls(int(v)))
    except ValueError:
        raise RuntimeError(f"invalid {what}: {v}") from None


class _MeshOp:
    def __init__(self, meshInfo: MeshInfo3D):
        if not isinstance(meshInfo, MeshInfo3D):
            raise TypeError("meshInfo must be a MeshInfo3D")
        self._N = meshInfo._N
        self._h = meshInfo._hs()
        self._L = meshInfo._Ls()


class Nabla3D_CD(_MeshOp):
    """Central-difference gradient and its adjoint (differentialOps.h:70-99)."""

    def __init__(self, meshInfo: MeshInfo3D, boundary):
        super().__init__(meshInfo)
        self._bc = _enum(boundary, BoundaryType, "boundary type")

    def _fwd(self, b, C):
        with on_device(b.device):
            out = torch.empty(tuple(b.shape) + (3,), dtype=torch.float32, device=b.device)
            check(_lib.lib().svs_nabla3d_forward(ptr(b), C, *b.shape[:3], *self._h, self._bc, ptr(out),
                                                 stream_ptr()))
        return out

    def _bwd(self, p, C):
        with on_device(p.device):
            out = torch.empty(tuple(p.shape[:-1]), dtype=torch.float32, device=p.device)
            check(_lib.lib().svs_nabla3d_backward(ptr(p), C, *p.shape[:3], *self._h, self._bc, ptr(out),
                                                  stream_ptr()))
        return out

    def forward(self, b):
        """[Z,Y,X] -> [Z,Y,X,3]."""
        _check(b, "b", 3)
        return self._fwd(b, 1)

    def backward(self, b):
        """[Z,Y,X,3] -> [Z,Y,X]."""
        _check(b, "b", 4)
        if b.shape[3] != 3:
            raise RuntimeError("Expected last dimension 3")
        return self._bwd(b, 1)

    def forwardVectorField(self, b):
        """[Z,Y,X,3] -> [Z,Y,X,3,3]."""
        _check(b, "b", 4)
        if b.shape[3] != 3:
            raise RuntimeError("Expected last dimension 3")
        return self._fwd(b, 3)

    def backwardVectorField(self, b):
        """[Z,Y,X,3,3] -> [Z,Y,X,3]."""
        _check(b, "b", 5)
        if tuple(b.shape[3:]) != (3, 3):
            raise RuntimeError("Expected trailing dimensions (3, 3)")
        return self._bwd(b, 3)


class AnisotropicNabla3D(_MeshOp):
    """Anisotropic diffusion tensor D = s n n^T + t1 t1^T (anisotropicDifferentialOps.h:62-84)."""

    def __init__(self, meshInfo: MeshInfo3D, alpha: float, beta: float):
        super().__init__(meshInfo)
        self._alpha = float(np.float32(alpha))
        self._beta = float(np.float32(beta))

    def computeTangentVecs(self, NablaInputImage):
        """[Z,Y,X,3] -> [scalars [Z,Y,X], normals, tangents1, tangents2 [Z,Y,X,3]]."""
        g = NablaInputImage
        _check(g, "NablaInputImage", 4)
        if g.shape[3] != 3:
            raise RuntimeError("Expected last dimension 3")
        with on_device(g.device):
            s = torch.empty(tuple(g.shape[:3]), dtype=torch.float32, device=g.device)
            n, t1, t2 = (torch.empty_like(g) for _ in range(3))
            check(_lib.lib().svs_aniso_tangents(ptr(g), s.numel(), self._alpha, self._beta, ptr(s), ptr(n), ptr(t1),
                                                ptr(t2), stream_ptr()))
        return [s, n, t1, t2]

    def _apply(self, b, scalars, normals, tangents1, tangents2):
        _check(b, "b", 5)
        for t, nm in ((scalars, "scalars"), (normals, "normals"), (tangents1, "tangents1"),
                      (tangents2, "tangents2")):
            _check(t, nm)
        _same_device(b, scalars, normals, tangents1, tangents2)
        sp = tuple(b.shape[:3])
        if tuple(b.shape[3:]) != (3, 3):
            raise RuntimeError("Expected trailing dimensions (3, 3)")
        _shape_eq(scalars, sp, "scalars")
        for t, nm in ((normals, "normals"), (tangents1, "tangents1"), (tangents2, "tangents2")):
            _shape_eq(t, sp + (3,), nm)
        with on_device(b.device):
            out = torch.empty_like(b)
            check(_lib.lib().svs_aniso_apply(ptr(b), ptr(scalars), ptr(normals), ptr(tangents1), ptr(tangents2),
                                             scalars.numel(), ptr(out), stream_ptr()))
        return out

    def forwardVectorField(self, b, scalars, normals, tangents1, tangents2):
        return self._apply(b, scalars, normals, tangents1, tangents2)

    def backwardVectorField(self, b, scalars, normals, tangents1, tangents2):
        # The reference calls the same kernel: D is symmetric (anisotropicDifferentialOps.h:77-83).
        return self._apply(b, scalars, normals, tangents1, tangents2)


class Warping3D(_MeshOp):
    """Backward warping out(i) = u(i*h + phi(i)) (warpingOps.h:84-111)."""

    def __init__(self, meshInfo: MeshInfo3D, interpolation, boundary):
        super().__init__(meshInfo)
        self._ip = _enum(interpolation, InterpolationType, "interpolation type")
        self._bc = _enum(boundary, BoundaryType, "boundary type")

    def _check_phi(self, u, phi):
        _check(phi, "phi", 4)
        _same_device(u, phi)
        _shape_eq(phi, tuple(u.shape[:3]) + (3,), "phi")

    def forward(self, u, phi):
        """u [Z,Y,X], phi [Z,Y,X,3] -> [Z,Y,X]."""
        _check(u, "u", 3)
        self._check_phi(u, phi)
        with on_device(u.device):
            out = torch.empty_like(u)
            check(_lib.lib().svs_warp3d(ptr(u), 1, ptr(phi), *u.shape, *self._h, self._ip, self._bc, ptr(out),
                                        stream_ptr()))
        return out

    def forwardVectorField(self, u, phi):
        """u [Z,Y,X,C], phi [Z,Y,X,3] -> [Z,Y,X,C]."""
        _check(u, "u", 4)
        self._check_phi(u, phi)
        with on_device(u.device):
            out = torch.empty_like(u)
            check(_lib.lib().svs_warp3d(ptr(u), u.shape[3], ptr(phi), *u.shape[:3], *self._h, self._ip, self._bc,
                                        ptr(out), stream_ptr()))
        return out

    def backward(self, u, phi, forward_out):
        """Adjoint: (u [Z,Y,X], phi [Z,Y,X,3], grad_out [Z,Y,X]) -> [grad_u, grad_phi]."""
        _check(u, "u", 3)
        self._check_phi(u, phi)
        _check(forward_out, "forward_out", 3)
        _shape_eq(forward_out, u.shape, "forward_out")
        with on_device(u.device):
            gu = torch.empty_like(u)
            gphi = torch.empty_like(phi)
            check(_lib.lib().svs_warp3d_backward(ptr(u), ptr(phi), ptr(forward_out), 1, 1, *u.shape, *self._h,
                                                 self._ip, self._bc, ptr(gu), ptr(gphi), stream_ptr()))
        return [gu, gphi]


class WarpingCNN3D(_MeshOp):
    """Batched warping u [B,C,Z,Y,X] by phi [B,Z,Y,X,3] (warpingOpsCNN.h:15-34).

    ``backward`` is new: the reference leaves the CNN adjoint commented out
    (warpingOpsCNN.h:36-41).  It mirrors Warping3D.backward per (batch,
    channel) and sums the flow gradient over channels."""

    def __init__(self, meshInfo: MeshInfo3D, interpolation, boundary):
        super().__init__(meshInfo)
        self._ip = _enum(interpolation, InterpolationType, "interpolation type")
        self._bc = _enum(boundary, BoundaryType, "boundary type")

    def _check_args(self, u, phi):
        _check(u, "u", 5)
        _check(phi, "phi", 5)
        _same_device(u, phi)
        B = u.shape[0]
        _shape_eq(phi, (B,) + tuple(u.shape[2:]) + (3,), "phi")

    def forward(self, u, phi):
        self._check_args(u, phi)
        with on_device(u.device):
            out = torch.empty_like(u)
            check(_lib.lib().svs_warp3d_cnn(ptr(u), ptr(phi), u.shape[0], u.shape[1], *u.shape[2:], *self._h,
                                            self._ip, self._bc, ptr(out), stream_ptr()))
        return out

    def backward(self, u, phi, grad_out):
        self._check_args(u, phi)
        _check(grad_out, "grad_out", 5)
        _shape_eq(grad_out, u.shape, "grad_out")
        with on_device(u.device):
            gu = torch.empty_like(u)
            gphi = torch.empty_like(phi)
            check(_lib.lib().svs_warp3d_backward(ptr(u), ptr(phi), ptr(grad_out), u.shape[0], u.shape[1],
                                                 *u.shape[2:], *self._h, self._ip, self._bc, ptr(gu), ptr(gphi),
                                                 stream_ptr()))
        return [gu, gphi]


class Prolongation3D:
    """Resampling from one mesh to another (prolongationOps.h:76-104)."""

    def __init__(self, meshInfo: MeshInfo3D, meshInfoProlongated: MeshInfo3D, interpolation, boundary):
        if not isinstance(meshInfo, MeshInfo3D) or not isinstance(meshInfoProlongated, MeshInfo3D):
            raise TypeError("meshInfo and meshInfoProlongated must be MeshInfo3D")
        self._src_N = meshInfo._N
        self._src_L = meshInfo._Ls()
        self._dst_N = meshInfoProlongated._N
        self._dst_L = meshInfoProlongated._Ls()
        self._ip = _enum(interpolation, InterpolationType, "interpolation type")
        self._bc = _enum(boundary, BoundaryType, "boundary type")

    def _run(self, u, C, trailing):
        # h of the source comes from the mesh in the reference launcher
        # (prolongationOps.cu:1000-1010); a tensor of another size would be
        # sampled with a foreign spacing, so it is rejected here.
        _shape_eq(u, tuple(self._src_N) + trailing, "u")
        with on_device(u.device):
            out = torch.empty(tuple(self._dst_N) + trailing, dtype=torch.float32, device=u.device)
            check(_lib.lib().svs_prolongate3d(ptr(u), C, *u.shape[:3], *self._src_L, *self._dst_N, *self._dst_L,
                                              self._ip, self._bc, ptr(out), stream_ptr()))
        return out

    def forward(self, u):
        """[Z,Y,X] -> [Z',Y',X']."""
        _check(u, "u", 3)
        return self._run(u, 1, ())

    def forwardVectorField(self, u):
        """[Z,Y,X,C] -> [Z',Y',X',C]."""
        _check(u, "u", 4)
        return self._run(u, u.shape[3], (u.shape[3],))

    def forwardMatrixField(self, u):
        """[Z,Y,X,I,J] -> [Z',Y',X',I,J]."""
        _check(u, "u", 5)
        return self._run(u, u.shape[3] * u.shape[4], (u.shape[3], u.shape[4]))


def TVL1OF3D_proxPrimal(primalVariable, primalStepSize_tau, primalFctWeight_Matching, rho, I1_warped_grad,
                        meshInfo):
    """L1 data-term prox (opticalFlowOps.h:103-113): [Z,Y,X,3] -> [Z,Y,X,3]."""
    _check(primalVariable, "primalVariable", 4)
    _check(rho, "rho", 3)
    _check(I1_warped_grad, "I1_warped_grad", 4)
    _same_device(primalVariable, rho, I1_warped_grad)
    sp = (meshInfo.getNZ(), meshInfo.getNY(), meshInfo.getNX())
    _shape_eq(primalVariable, sp + (3,), "primalVariable")
    _shape_eq(rho, sp, "rho")
    _shape_eq(I1_warped_grad, sp + (3,), "I1_warped_grad")
    with on_device(primalVariable.device):
        out = torch.empty_like(primalVariable)
        check(_lib.lib().svs_prox_primal(ptr(primalVariable), float(np.float32(primalStepSize_tau)),
                                         float(np.float32(primalFctWeight_Matching)), ptr(rho), ptr(I1_warped_grad),
                                         rho.numel(), ptr(out), stream_ptr()))
    return out


def TVL1OF3D_proxDual(dualVariable, dualStepSize_sigma, dualFctWeight_TV, meshInfo):
    """Row-wise projection onto the l2 ball of radius weight_tv (opticalFlowOps.h:116-123)."""
    _check(dualVariable, "dualVariable", 5)
    sp = (meshInfo.getNZ(), meshInfo.getNY(), meshInfo.getNX())
    _shape_eq(dualVariable, sp + (3, 3), "dualVariable")
    with on_device(dualVariable.device):
        out = torch.empty_like(dualVariable)
        check(_lib.lib().svs_prox_dual(ptr(dualVariable), float(np.float32(dualStepSize_sigma)),
                                       float(np.float32(dualFctWeight_TV)), dualVariable.numel() // 9, ptr(out),
                                       stream_ptr()))
    return out
