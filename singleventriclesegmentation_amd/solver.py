"""Python handle on a native solve plan (svs_tvl1_plan_* in include/svsflow.h).

A plan fixes the solver parameters and the list of instance sizes; it owns a
device workspace (a torch uint8 tensor) that holds every pyramid level's
state and the descriptor tables the kernels read.  Instances are bound to
input frames and output tensors before each run; re-binding keeps a captured
hipGraph valid because the kernels read addresses from the tables.
"""
import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import TVL1Params, check, ptr, stream_ptr
from .flow_utils import gaussian_weights3d


@dataclass(frozen=True)
class SolverConfig:
    """TVL13DOpticalFlow constructor arguments (tvl13d.py:14-67) in normalised form.

    Defaults are conf/flow.yaml:17-56 except the schedule, which defaults to
    the 4 x 10 x 50 benchmark schedule; ``length`` is (lx, ly, lz)."""

    num_scales: int = 4
    max_warps: int = 10
    max_outer_iterations: int = 50
    weight_matching: float = 1.0
    weight_tv: float = 0.025
    pd_type: int = 2
    pd_params: Tuple[float, float, float, float] = (0.5, 0.5, 0.25, 1.0)  # sigma, tau, theta, gamma
    use_aniso: bool = True
    aniso_params: Tuple[float, float] = (5.0, 0.5)
    use_blur: bool = True
    blur_sigma: float = 1.0
    interpolation: int = 2
    boundary: int = 1
    length_fixed: bool = False
    length: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    sweep_shape: Tuple[int, int, int] = (0, 0, 0)
    # Huber-TV smoothing (not in the reference, whose TV is eps = 0): each dual
    # row becomes v_c / max(1 + sigma_n eps / weight_tv, ||v_c|| / weight_tv).
    huber_eps: float = 0.0
    # Over-relaxation (not in the reference, whose z = u): the dual step
    # differentiates z = u^n + theta (u^n - u^{n-1}), reset to u every warp.
    relax_theta: float = 0.0

    def native(self, use_initial_state=False):
        p = TVL1Params()
        p.num_scales = self.num_scales
        p.max_warps = self.max_warps
        p.max_outer_iterations = self.max_outer_iterations
        p.weight_matching = float(np.float32(self.weight_matching))
        p.weight_tv = float(np.float32(self.weight_tv))
        p.pd_type = int(self.pd_type)
        p.sigma, p.tau, p.theta, p.gamma = (float(v) for v in self.pd_params)
        p.use_aniso = 1 if self.use_aniso else 0
        p.alpha, p.beta = (float(np.float32(v)) for v in self.aniso_params)
        weights = None
        if self.use_blur:
            weights = gaussian_weights3d(self.blur_sigma)
            p.use_blur = 1
            p.blur_K = weights.shape[0]
            p.blur_weights = weights.ctypes.data_as(ctypes.c_void_p)
        p.interpolation = int(self.interpolation)
        p.boundary = int(self.boundary)
        p.length_fixed = 1 if self.length_fixed else 0
        lx, ly, lz = self.length
        p.LZ, p.LY, p.LX = float(np.float32(lz)), float(np.float32(ly)), float(np.float32(lx))
        p.use_initial_state = 1 if use_initial_state else 0
        p.sweep_bx, p.sweep_by, p.sweep_cz = (int(v) for v in self.sweep_shape)
        p.huber_eps = float(self.huber_eps)
        p.relax_theta = float(self.relax_theta)
        return p, weights  # keep `weights` alive while `p` is used

    def level_shapes(self, shape):
        nz, ny, nx = shape
        out = [(nz, ny, nx)]
        for _ in range(1, self.num_scales):
            nz, ny, nx = (nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2
            out.append((nz, ny, nx))
        return out

    def voxel_iterations(self, shape):
        """Unit of work: one PD iteration at one voxel of one level."""
        return sum(a * b * c for a, b, c in self.level_shapes(shape)) * self.max_warps * self.max_outer_iterations

    def voxel_warps(self, shape):
        return sum(a * b * c for a, b, c in self.level_shapes(shape)) * self.max_warps


# Algorithmic HBM bytes (reads + writes) of the fused path.
BYTES_PER_VOXEL_ITERATION = 124  # u 12 + p 36 + rho_c 4 + g 12 + grad I1 12 read; u 12 + p 36 written
BYTES_PER_VOXEL_WARP = 48        # u 12, I0 4, I1 4, grad I1 12 read; rho_c 4, g 12 written


def workspace_bytes(cfg: SolverConfig, shapes) -> int:
    """Device workspace a plan for these instance sizes needs (0 if invalid)."""
    params, _w = cfg.native()
    dims = (ctypes.c_int * (3 * len(shapes)))(*[int(v) for s in shapes for v in s])
    return int(_lib.lib().svs_tvl1_workspace_bytes(ctypes.byref(params), len(shapes), dims))


def algorithmic_bytes(cfg: SolverConfig, shape):
    return BYTES_PER_VOXEL_ITERATION * cfg.voxel_iterations(shape) + BYTES_PER_VOXEL_WARP * cfg.voxel_warps(shape)


class Plan:
    """Native plan for ``len(shapes)`` instances on one device."""

    def __init__(self, cfg: SolverConfig, shapes: Sequence[Tuple[int, int, int]], device=None,
                 use_initial_state=False, workspace: Optional[torch.Tensor] = None):
        self.cfg = cfg
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        dev = torch.device(device if device is not None else "cuda")
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.n = len(self.shapes)
        self._params, self._weights = cfg.native(use_initial_state)
        dims = (ctypes.c_int * (3 * self.n))(*[v for s in self.shapes for v in s])
        self._dims = dims
        L = _lib.lib()
        nbytes = L.svs_tvl1_workspace_bytes(ctypes.byref(self._params), self.n, dims)
        if nbytes == 0:
            raise RuntimeError(L.svs_last_error().decode() or "invalid solver configuration")
        self.workspace_bytes = int(nbytes)
        with torch.cuda.device(self.device):
            if workspace is None:
                workspace = torch.empty(self.workspace_bytes, dtype=torch.uint8, device=self.device)
            elif workspace.numel() < self.workspace_bytes:
                raise RuntimeError("workspace too small")
            self.workspace = workspace
            h = ctypes.c_void_p()
            check(L.svs_tvl1_plan_create(ctypes.byref(self._params), self.n, dims, ptr(workspace),
                                         ctypes.c_size_t(workspace.numel()), stream_ptr(), ctypes.byref(h)))
        self._h = h
        self._bound = [None] * self.n
        self.stats = None

    def info(self, level):
        out = (ctypes.c_int * 8)()
        check(_lib.lib().svs_tvl1_plan_info(self._h, level, out))
        return {"bx": out[0], "by": out[1], "cz": out[2], "tiles": out[3], "blocks": out[4], "resident": out[7]}

    def bind(self, k, I0, I1, u_out=None, p_out=None, u_in=None, p_in=None):
        """Binds instance k; the tensors must stay alive until the run completes."""
        shp = self.shapes[k]
        for t, nm in ((I0, "I0"), (I1, "I1")):
            if t.device != self.device or t.dtype != torch.float32 or not t.is_contiguous() or tuple(t.shape) != shp:
                raise RuntimeError(f"{nm} must be a contiguous float32 {shp} tensor on {self.device}")
        for t, nm, sh in ((u_out, "u_out", shp + (3,)), (p_out, "p_out", shp + (3, 3)),
                          (u_in, "u_in", shp + (3,)), (p_in, "p_in", shp + (3, 3))):
            if t is not None and (t.device != self.device or t.dtype != torch.float32 or not t.is_contiguous()
                                  or tuple(t.shape) != sh):
                raise RuntimeError(f"{nm} must be a contiguous float32 {sh} tensor on {self.device}")
        check(_lib.lib().svs_tvl1_plan_bind(self._h, k, ptr(I0), ptr(I1), ptr(u_in), ptr(p_in), ptr(u_out),
                                            ptr(p_out)))
        self._bound[k] = (I0, I1, u_out, p_out, u_in, p_in)

    def run(self, use_graph=True):
        """Enqueues the solve on the current stream and returns without waiting;
        it starts after the plan's previous work, whatever stream ran it."""
        with torch.cuda.device(self.device):
            check(_lib.lib().svs_tvl1_plan_run(self._h, 1 if use_graph else 0, stream_ptr()))

    def autotune(self, reps=3):
        """Picks the fastest fused-sweep shape per level (needs one prior run).

        Returns [(bx, by, cz)] per level; later plans of the same geometry in
        this process start from these shapes."""
        rep = (ctypes.c_int * (3 * self.cfg.num_scales))()
        with torch.cuda.device(self.device):
            check(_lib.lib().svs_tvl1_plan_autotune(self._h, reps, stream_ptr(), rep))
        return [tuple(rep[3 * i:3 * i + 3]) for i in range(self.cfg.num_scales)]

    def set_shapes(self, shapes):
        """Sets the fused-sweep shape policy (bx, by, cz) of every level, e.g. to
        reproduce a configuration the autotuner picked earlier: tiles at most
        bx - 2 voxels wide within bx * by threads, marching cz planes (an
        instance of width bx - 2 gets bx x by work-groups)."""
        if len(shapes) != self.cfg.num_scales:
            raise RuntimeError(f"need one (bx, by, cz) per level ({self.cfg.num_scales})")
        with torch.cuda.device(self.device):
            for lv, (bx, by, cz) in enumerate(shapes):
                check(_lib.lib().svs_tvl1_plan_set_shape(self._h, lv, int(bx), int(by), int(cz), stream_ptr()))

    def set_resident(self, level, mode):
        """How a level runs its iterations: 1 all of a warp's iterations in one
        resident launch (levels of at most 2048 voxels per instance), 0 the
        marching sweep once per iteration, -1 the automatic choice."""
        with torch.cuda.device(self.device):
            check(_lib.lib().svs_tvl1_plan_set_resident(self._h, int(level), int(mode), stream_ptr()))

    def set_stats(self, enabled=True):
        """Per-iteration statistics for every later run (or none with False).

        ``self.stats`` becomes a float32 tensor [levels, warps, iterations,
        instances, 3] on the plan's device, level 0 the finest, which each run
        fills in stream order: [..., 0] = ||u^{n+1} - u^n||, [..., 1] =
        ||p^{n+1} - p^n|| (the residual norms tvl13d.py:222, 237, 253 computes
        and discards), [..., 2] = the energy lambda sum|rho(u^n)| + lambda_tv
        sum_c ||D grad u^n_c|| of the iterate entering iteration n (NaN with
        over-relaxation).  The values are the same on every run; u and p do not
        change.  Replacing or dropping the tensor waits for the plan's enqueued
        runs."""
        c = self.cfg
        out = None
        if enabled:
            out = torch.zeros((c.num_scales, c.max_warps, c.max_outer_iterations, self.n, 3), dtype=torch.float32,
                              device=self.device)
        with torch.cuda.device(self.device):
            check(_lib.lib().svs_tvl1_plan_set_stats(self._h, ptr(out), stream_ptr()))
        self.stats = out

    def sweeps(self, level, count):
        """Enqueues ``count`` fused sweeps of one level (timing aid; needs a prior run)."""
        with torch.cuda.device(self.device):
            check(_lib.lib().svs_tvl1_plan_sweeps(self._h, level, count, stream_ptr()))

    def close(self):
        if getattr(self, "_h", None):
            # Waits for the plan's last enqueued work (on whichever stream it
            # ran), so the workspace can go back to the allocator.
            _lib.lib().svs_tvl1_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
