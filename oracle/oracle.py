"""numpy front end of the CPU oracle (libsvs_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg.  Nothing in the product package
imports this module.

Every function takes and returns C-contiguous numpy arrays in the reference's
layouts ([Z,Y,X], [Z,Y,X,3], [Z,Y,X,3,3], [B,C,Z,Y,X]).  ``dtype`` selects the
literal FP32 restatement (np.float32) or the FP64 twin (np.float64).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SVS_ORACLE_LIB selects another build of the same source (the sanitizer
# build of `make sanitize`, for tests/test_oracle_sanitized.py).
_LIB_PATH = os.environ.get("SVS_ORACLE_LIB") or os.path.join(_HERE, "libsvs_oracle.so")
_lib = None

BOUNDARY = {"NEAREST": 0, "MIRROR": 1, "REFLECT": 2}
INTERP = {"NEAREST": 0, "LINEAR": 1, "CUBIC_HERMITESPLINE": 2}


def build():
    """Compile the oracle with its Makefile (g++, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def use_library(path=None):
    """Routes every following call to the oracle build at ``path`` (one of the
    rounding-sensitivity variants that ``make variants`` writes to
    oracle/_variants/), or back to the default build when ``path`` is None."""
    global _lib
    _lib = None if path is None else ctypes.CDLL(path)


def _sfx(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def _ct(dtype):
    return ctypes.c_float if np.dtype(dtype) == np.float32 else ctypes.c_double


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _arr(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _fn(name, dtype):
    return getattr(lib(), f"orc_{name}_{_sfx(dtype)}")


def mesh_h(N, L, dtype=np.float32):
    """h = L / (N - 1) as MeshInfo3D computes it (coreDefines.h:72-75)."""
    if np.dtype(dtype) == np.float32:
        return float(np.float32(L) / np.float32(N - 1))
    return float(L) / float(N - 1)


def bdry_index(i, N, bc):
    return lib().orc_bdry_index_f32(int(i), int(N), int(bc))


def cr_local(buf, lc, dtype=np.float32):
    f = _fn("cr_local", dtype)
    f.restype = _ct(dtype)
    b = _arr(buf, dtype)
    return f(_p(b), _ct(dtype)(lc))


def _hs(h, dtype):
    ct = _ct(dtype)
    return [ct(h[0]), ct(h[1]), ct(h[2])]


def warp(u, phi, h=(1.0, 1.0, 1.0), ip=2, bc=1, dtype=np.float32):
    """Warping3D.forward (u [Z,Y,X]) or .forwardVectorField (u [Z,Y,X,C])."""
    u = _arr(u, dtype)
    phi = _arr(phi, dtype)
    NZ, NY, NX = u.shape[:3]
    C = 1 if u.ndim == 3 else u.shape[3]
    out = np.empty_like(u)
    _fn("warp", dtype)(_p(u), C, _p(phi), NZ, NY, NX, *_hs(h, dtype), ip, bc, _p(out))
    return out


def warp_cnn(u, phi, h=(1.0, 1.0, 1.0), ip=2, bc=1, dtype=np.float32):
    u = _arr(u, dtype)
    phi = _arr(phi, dtype)
    B, C, NZ, NY, NX = u.shape
    out = np.empty_like(u)
    _fn("warp_cnn", dtype)(_p(u), _p(phi), B, C, NZ, NY, NX, *_hs(h, dtype), ip, bc, _p(out))
    return out


def warp_backward(u, phi, g, h=(1.0, 1.0, 1.0), ip=2, bc=1, dtype=np.float32):
    u, phi, g = _arr(u, dtype), _arr(phi, dtype), _arr(g, dtype)
    NZ, NY, NX = u.shape
    gu = np.empty_like(u)
    gphi = np.empty_like(phi)
    _fn("warp_backward", dtype)(_p(u), _p(phi), _p(g), NZ, NY, NX, *_hs(h, dtype), ip, bc, _p(gu), _p(gphi))
    return gu, gphi


def warp_cnn_backward(u, phi, g, h=(1.0, 1.0, 1.0), ip=2, bc=1, dtype=np.float32):
    u, phi, g = _arr(u, dtype), _arr(phi, dtype), _arr(g, dtype)
    B, C, NZ, NY, NX = u.shape
    gu = np.empty_like(u)
    gphi = np.empty_like(phi)
    _fn("warp_cnn_backward", dtype)(_p(u), _p(phi), _p(g), B, C, NZ, NY, NX, *_hs(h, dtype), ip, bc,
                                    _p(gu), _p(gphi))
    return gu, gphi


def prolongate(u, src_L, dst_shape, dst_L, ip=2, bc=1, dtype=np.float32):
    """Prolongation3D.forward / forwardVectorField / forwardMatrixField.

    ``u`` is [Z,Y,X], [Z,Y,X,3] or [Z,Y,X,3,3]; lengths are (LZ, LY, LX)."""
    u = _arr(u, dtype)
    NZ, NY, NX = u.shape[:3]
    trailing = u.shape[3:]
    C = int(np.prod(trailing)) if trailing else 1
    out = np.empty(tuple(dst_shape) + trailing, dtype=dtype)
    ct = _ct(dtype)
    _fn("prolongate", dtype)(_p(u), C, NZ, NY, NX, ct(src_L[0]), ct(src_L[1]), ct(src_L[2]),
                             int(dst_shape[0]), int(dst_shape[1]), int(dst_shape[2]),
                             ct(dst_L[0]), ct(dst_L[1]), ct(dst_L[2]), ip, bc, _p(out))
    return out


def nabla_fwd(b, h=(1.0, 1.0, 1.0), bc=1, dtype=np.float32):
    """Nabla3D_CD.forward ([Z,Y,X] -> [..,3]) or .forwardVectorField ([..,3] -> [..,3,3])."""
    b = _arr(b, dtype)
    NZ, NY, NX = b.shape[:3]
    C = 1 if b.ndim == 3 else b.shape[3]
    out = np.empty(b.shape + (3,), dtype=dtype)
    _fn("nabla_fwd", dtype)(_p(b), C, NZ, NY, NX, *_hs(h, dtype), bc, _p(out))
    return out


def nabla_bwd(p, h=(1.0, 1.0, 1.0), bc=1, dtype=np.float32):
    """Nabla3D_CD.backward ([..,3] -> [Z,Y,X]) or .backwardVectorField ([..,3,3] -> [..,3])."""
    p = _arr(p, dtype)
    NZ, NY, NX = p.shape[:3]
    C = 1 if p.ndim == 4 else p.shape[3]
    out = np.empty(p.shape[:-1], dtype=dtype)
    _fn("nabla_bwd", dtype)(_p(p), C, NZ, NY, NX, *_hs(h, dtype), bc, _p(out))
    return out


def aniso_tangents(G, alpha, beta, dtype=np.float32):
    G = _arr(G, dtype)
    N = G.size // 3
    sh = G.shape[:-1]
    s = np.empty(sh, dtype=dtype)
    n, t1, t2 = (np.empty_like(G) for _ in range(3))
    ct = _ct(dtype)
    _fn("aniso_tangents", dtype)(_p(G), ctypes.c_longlong(N), ct(alpha), ct(beta), _p(s), _p(n), _p(t1), _p(t2))
    return s, n, t1, t2


def aniso_apply(b, s, n, t1, t2, dtype=np.float32):
    b, s, n, t1, t2 = (_arr(a, dtype) for a in (b, s, n, t1, t2))
    out = np.empty_like(b)
    _fn("aniso_apply", dtype)(_p(b), _p(s), _p(n), _p(t1), _p(t2), ctypes.c_longlong(s.size), _p(out))
    return out


def prox_primal(primal, tau, lam, rho, g, dtype=np.float32):
    primal, rho, g = _arr(primal, dtype), _arr(rho, dtype), _arr(g, dtype)
    out = np.empty_like(primal)
    ct = _ct(dtype)
    _fn("prox_primal", dtype)(_p(primal), ct(tau), ct(lam), _p(rho), _p(g), ctypes.c_longlong(rho.size), _p(out))
    return out


def prox_dual(dual, wtv, dtype=np.float32):
    dual = _arr(dual, dtype)
    out = np.empty_like(dual)
    _fn("prox_dual", dtype)(_p(dual), _ct(dtype)(wtv), ctypes.c_longlong(dual.size // 9), _p(out))
    return out


def blur(vol, weights, dtype=np.float32):
    vol = _arr(vol, dtype)
    w = _arr(weights, dtype)
    K = w.shape[0]
    out = np.empty_like(vol)
    _fn("blur", dtype)(_p(vol), *vol.shape, _p(w), K, _p(out))
    return out


def median3d(u, ks, dtype=np.float32):
    """scipy.ndimage.median_filter(u, size=(ks,ks,ks,1)) (flow_utils.py:115-132) on
    [Z,Y,X,C] or a batch [B,Z,Y,X,C]."""
    u = _arr(u, dtype)
    batched = u.ndim == 5
    ub = u if batched else u[None]
    out = np.empty_like(ub)
    B, NZ, NY, NX, C = ub.shape
    _fn("median3d", dtype)(_p(ub), B, NZ, NY, NX, C, int(ks), _p(out))
    return out if batched else out[0]


def mask_counts(a, b):
    """(|A|, |B|, |A and B|) of two binary masks (value > 0.5 is set)."""
    a = _arr(a, np.float32)
    b = _arr(b, np.float32)
    out = np.zeros(3, np.int64)
    lib().orc_mask_counts(_p(a), _p(b), ctypes.c_longlong(a.size), _p(out))
    return tuple(int(v) for v in out)


def dice(a, b):
    """monai compute_dice on one binary channel: 2|A and B| / (|A| + |B|) in float32,
    NaN when the ground truth ``b`` is empty (ignore_empty=True)."""
    na, nb, nab = mask_counts(a, b)
    if nb == 0:
        return float("nan")
    return float(np.float32(2.0 * nab) / np.float32(na + nb))


def hausdorff2(a, b):
    """Directed squared Hausdorff distances (A->B, B->A) between the 6-connected
    edge voxels of two [Z,Y,X] masks; -1: the source has no edge, -2: only the
    target has none."""
    a = _arr(a, np.float32)
    b = _arr(b, np.float32)
    out = np.zeros(2, np.int64)
    lib().orc_mask_hausdorff2(_p(a), _p(b), *a.shape, _p(out))
    return int(out[0]), int(out[1])


def hausdorff(a, b):
    """Symmetric Hausdorff distance as monai returns it (float32 of the float64
    root); NaN when either mask has no edge voxel."""
    d = hausdorff2(a, b)
    if min(d) < 0:
        return float("nan")
    return float(np.float32(np.sqrt(np.float64(max(d)))))


def edge_count(m):
    m = _arr(m, np.float32)
    fn = lib().orc_mask_edge_count
    fn.restype = ctypes.c_longlong
    return int(fn(_p(m), *m.shape))


PADDING = {"zeros": 0, "border": 1, "reflection": 2}
SAMPLE_MODE = {"bilinear": 0, "nearest": 1}


def rotation_grid(axes, R, rounding=0):
    """identity_grid @ R (transforms/spatial.py:56, :93-105) as [Z,Y,X,3] (x, y, z),
    each component a 3-term dot product rounded as the host BLAS does
    (``rounding``: 0 MKL's FMA chain, 1 plain, 2 plain from +0).

    ``axes`` = (lz, ly, lx), the identity grid's axes: torch.linspace(-1, 1, n)
    per axis as the reference computes them on the host CPU (spatial.py:100).
    They are an input here, not restated: torch's vectorised CPU linspace
    restarts its count at every SIMD chunk, so the values depend on the host's
    vector width."""
    lz, ly, lx = (_arr(a, np.float32) for a in axes)
    OZ, OY, OX = lz.size, ly.size, lx.size
    R = _arr(R, np.float32)
    grid = np.empty((OZ, OY, OX, 3), np.float32)
    lib().orc_rotation_grid(_p(lz), _p(ly), _p(lx), OZ, OY, OX, _p(R), int(rounding), _p(grid))
    return grid


def grid_sample3d(x, grid, mode="bilinear", padding="zeros", align_corners=False):
    """F.grid_sample(x, grid repeated over the batch) on x [B,C,Z,Y,X] and one
    grid [OZ,OY,OX,3]."""
    x = _arr(x, np.float32)
    grid = _arr(grid, np.float32)
    B, C, IZ, IY, IX = x.shape
    OZ, OY, OX = grid.shape[:3]
    out = np.empty((B, C, OZ, OY, OX), np.float32)
    lib().orc_grid_sample3d(_p(x), B, C, IZ, IY, IX, _p(grid), OZ, OY, OX, SAMPLE_MODE[mode], PADDING[padding],
                            int(bool(align_corners)), _p(out))
    return out


def rotate_vectors(v, R, rounding=0):
    """einsum('ij,tjxyz->tixyz', R, v) (transforms/spatial.py:127) on v [B,3,Z,Y,X],
    rounded as in rotation_grid."""
    v = np.array(v, dtype=np.float32, order="C", copy=True)
    R = _arr(R, np.float32)
    lib().orc_rotate_vectors(_p(v), v.shape[0], ctypes.c_longlong(int(np.prod(v.shape[2:]))), _p(R),
                             int(rounding))
    return v


FLIP_COMPONENT = {2: 2, 3: 1, 4: 0}   # transforms/spatial.py:166


def spatial_augment(x, R=None, flip_axes=(), axes=None, mode="bilinear", padding="zeros",
                    align_corners=False, vector=False, grid=None, grid_rounding=0, vector_rounding=0):
    """RandomRotate with rotation R (None: not drawn) followed by RandomFlip on
    each axis of ``flip_axes`` in order (transforms/spatial.py:53-90, :142-168),
    on one key x [B,C,Z,Y,X]; ``vector`` marks a flow key (C == 3); ``axes``
    as in rotation_grid, or ``grid`` = identity_grid @ R as given; the two
    roundings as in rotation_grid."""
    x = _arr(x, np.float32)
    if R is not None:
        g = rotation_grid(axes, R, grid_rounding) if grid is None else grid
        x = grid_sample3d(x, g, mode, padding, align_corners)
        if vector:
            x = rotate_vectors(x, R, vector_rounding)
    for axis in flip_axes:
        x = np.ascontiguousarray(np.flip(x, axis=axis))
        if vector:
            x[:, FLIP_COMPONENT[axis]] *= np.float32(-1.0)
    return x


def solve_params(num_scales=4, max_warps=10, max_outer_iterations=50, weight_matching=1.0, weight_tv=0.025,
                 pd_type=2, pd_params=(0.5, 0.5, 0.25, 1.0), use_aniso=True, aniso_params=(5.0, 0.5),
                 interpolation="CUBIC_HERMITESPLINE", boundary="MIRROR", length_type="numDofs",
                 length=(1, 1, 1), huber_eps=0.0, relax_theta=0.0):
    """Packs the TVL13DOpticalFlow constructor arguments (tvl13d.py:14-67).

    ``length`` is (lx, ly, lz) as in conf/flow.yaml; the packed order is LZ, LY, LX.
    ``huber_eps`` (not in the reference, whose TV is eps = 0) selects the
    Huber-TV dual step (svs_oracle.cpp, prox_dual); ``relax_theta`` (not in the
    reference, whose z = u) the over-relaxation of the dual step's point
    (svs_oracle.cpp, solve_level)."""
    lx, ly, lz = length
    return np.array([num_scales, max_warps, max_outer_iterations, weight_matching, weight_tv, pd_type,
                     *pd_params, 1.0 if use_aniso else 0.0, *aniso_params, INTERP[interpolation],
                     BOUNDARY[boundary], 1.0 if length_type == "fixed" else 0.0, lz, ly, lx, huber_eps,
                     relax_theta],
                    dtype=np.float64)


def step_schedule(params):
    n = int(params[2])
    sig = np.empty(n, np.float64)
    tau = np.empty(n, np.float64)
    lib().orc_step_schedule(_p(params), _p(sig), _p(tau))
    return sig, tau


def solve(I0, I1, params, u=None, p=None, dtype=np.float32, stats=False):
    """TVL13DOpticalFlow.compute on already-blurred I0, I1 (tvl13d.py:122-162).

    With ``stats`` it also returns float64 [num_scales, max_warps,
    max_outer_iterations, 3] per-iteration statistics (level 0 the finest):
    ||u^{n+1} - u^n||, ||p^{n+1} - p^n|| (the reference's residual norms,
    tvl13d.py:222, 237) and the energy E(u^n) of the iterate entering the
    iteration (svs_oracle.cpp, IterStats)."""
    I0, I1 = _arr(I0, dtype), _arr(I1, dtype)
    NZ, NY, NX = I0.shape
    u = np.zeros((NZ, NY, NX, 3), dtype) if u is None else _arr(u, dtype)
    p = np.zeros((NZ, NY, NX, 3, 3), dtype) if p is None else _arr(p, dtype)
    uo = np.empty_like(u)
    po = np.empty_like(p)
    params = np.ascontiguousarray(params, dtype=np.float64)
    if params.size < 21:  # packed before huber_eps / relax_theta existed (golden fixtures): 0
        params = np.append(params, np.zeros(21 - params.size))
    if not stats:
        _fn("solve", dtype)(_p(params), _p(I0), _p(I1), NZ, NY, NX, _p(u), _p(p), _p(uo), _p(po))
        return uo, po
    st = np.zeros((int(params[0]), int(params[1]), int(params[2]), 3), np.float64)
    _fn("solve_stats", dtype)(_p(params), _p(I0), _p(I1), NZ, NY, NX, _p(u), _p(p), _p(uo), _p(po), _p(st))
    return uo, po, st
