"""Drop-in for svs/modules/flow/tvl13d.py: TVL13DOpticalFlow with the same
constructor and ``compute(I0, I1, u, p) -> (u, p)``.

``compute`` runs the whole coarse-to-fine solve on the GPU through a cached
native plan (one fused primal-dual sweep per iteration, no host syncs inside
the solve).  ``compute_unfused`` walks the reference's step-by-step driver
(tvl13d.py:69-255) through the drop-in operators of ``opticalFlow``, with
torch's own elementwise and sum kernels between them, as the reference does.
The two give the same bits: the fused kernels evaluate torch's 3-term
sum(dim=3) in the order torch's GPU reduction rounds it, (a + c) + b (checked
at 80^3 with the 4 x 10 x 50 schedule).
"""
import math
from typing import Literal, Tuple

import numpy as np
import torch

from . import opticalFlow
from .flow_utils import apply_gaussian_blur3d, get_interpolation_type, get_mesh_length
from .solver import Plan, SolverConfig


def _stats_dict(t):
    """[levels, warps, iterations, 3] -> the named statistics (views of t)."""
    return {"primal": t[..., 0], "dual": t[..., 1], "update": t[..., 0], "energy": t[..., 2]}


class TVL13DOpticalFlow:
    def __init__(
        self,
        num_scales: int,
        max_warps: int,
        max_outer_iterations: int,
        weight_matching: float,
        weight_tv: float,
        primaldual_algorithm_type: Literal[1, 2],
        primaldual_algorithm_params: Tuple[float, float, float, float],
        use_anisotropic_diff: bool,
        anisotropic_diff_params: Tuple[float, float],
        use_gaussian_blur: bool,
        gaussian_blur_sigma: float,
        interpolation_type: Literal['NEAREST', 'LINEAR', 'CUBIC_HERMITESPLINE'],
        boundary_type: Literal['NEAREST', 'MIRROR', 'REFLECT'],
        length_type: Literal['numDofs', 'fixed'],
        length: Tuple[int, int, int],
        device: torch.device,
        use_graph: bool = True,
        autotune: bool = False,
        record_stats: bool = False,
        huber_eps: float = 0.0,
        relax_theta: float = 0.0,
    ):
        """Arguments as in tvl13d.py:14-52; ``length`` is (lx, ly, lz).

        Additive options: ``use_graph`` replays the fixed level x warp x
        iteration schedule as a hipGraph; ``autotune`` times the fused-sweep
        launch shapes after the first solve of each volume size and keeps the
        fastest (results do not depend on the shape); ``record_stats`` makes
        every compute (and compute_unfused) leave per-iteration statistics in
        ``self.stats``: a dict of float32 device tensors [num_scales, max_warps,
        max_outer_iterations] (level 0 the finest) holding the residual norms
        the reference computes and discards each iteration, "primal" =
        ||u^{n+1} - u^n|| (tvl13d.py:237; also its ||z - z_old||, :253, since
        z = u) and "dual" = ||p^{n+1} - p^n|| (:222), and "energy" = lambda
        sum|rho(u^n)| + lambda_tv sum_c ||D grad u^n_c|| of the iterate
        entering each iteration; ``huber_eps`` > 0 turns the TV term into a
        Huber-TV one (the reference's TV is eps = 0): each dual row becomes
        v_c / max(1 + sigma_n eps / weight_tv, ||v_c|| / weight_tv), and the
        energy uses the Huber function of the row norms.  compute_unfused
        follows the reference's operators and supports eps = 0 only.
        ``relax_theta`` != 0 over-relaxes the point the dual step
        differentiates, z = u^n + theta (u^n - u^{n-1}) (the reference keeps
        z = u, tvl13d.py:251-252), restarting from z = u at every warp like the
        step sizes; the energy statistic is then NaN."""
        self.num_scales = num_scales
        self.max_warps = max_warps
        self.max_outer_iterations = max_outer_iterations
        self.weight_matching = weight_matching
        self.weight_tv = weight_tv
        self.primaldual_algorithm_type = primaldual_algorithm_type
        self.sigma, self.tau, self.theta, self.gamma = primaldual_algorithm_params
        self.use_anisotropic_diff = use_anisotropic_diff
        self.anisotropic_diff_alpha, self.anistropic_diff_beta = anisotropic_diff_params
        self.use_gaussian_blur = use_gaussian_blur
        self.gaussian_blur_sigma = gaussian_blur_sigma
        self.length_type = length_type
        self.length_x, self.length_y, self.length_z = length
        if length_type not in ("numDofs", "fixed"):
            raise Exception("Wrong length type")
        self.interpolation_type, self.boundary_type = get_interpolation_type(interpolation_type, boundary_type)
        self.device = device
        self.use_graph = use_graph
        self.autotune = autotune
        self.record_stats = record_stats
        self.stats = None
        if not huber_eps >= 0.0:
            raise ValueError("huber_eps must be >= 0")
        self.huber_eps = float(huber_eps)
        if not math.isfinite(relax_theta):
            raise ValueError("relax_theta must be finite")
        self.relax_theta = float(relax_theta)
        self.config = SolverConfig(
            num_scales=num_scales, max_warps=max_warps, max_outer_iterations=max_outer_iterations,
            weight_matching=weight_matching, weight_tv=weight_tv, pd_type=primaldual_algorithm_type,
            pd_params=tuple(float(v) for v in primaldual_algorithm_params), use_aniso=bool(use_anisotropic_diff),
            aniso_params=tuple(float(v) for v in anisotropic_diff_params), use_blur=bool(use_gaussian_blur),
            blur_sigma=float(gaussian_blur_sigma), interpolation=int(self.interpolation_type),
            boundary=int(self.boundary_type), length_fixed=(length_type == "fixed"),
            length=(float(self.length_x), float(self.length_y), float(self.length_z)), huber_eps=self.huber_eps,
            relax_theta=self.relax_theta)
        self._plans = {}

    # ------------------------------------------------------------ fused path
    def _plan(self, shape, device, initial):
        key = (tuple(shape), str(device), initial)
        plan = self._plans.get(key)
        if plan is None:
            plan = Plan(self.config, [tuple(shape)], device=device, use_initial_state=initial)
            self._plans[key] = plan
        return plan

    def compute(self, I0: torch.Tensor, I1: torch.Tensor, u: torch.Tensor, p: torch.Tensor):
        """TV-L1 flow u [Z,Y,X,3] with I1(x + u(x)) ~ I0(x), and the dual p [Z,Y,X,3,3].

        As in the reference, ``u`` and ``p`` only matter when num_scales == 1
        (tvl13d.py:135-139)."""
        for t, nm in ((I0, "I0"), (I1, "I1")):
            if t.device.type != "cuda" or t.dtype != torch.float32 or t.dim() != 3:
                raise RuntimeError(f"{nm} must be a float32 [Z,Y,X] CUDA tensor")
        if I0.shape != I1.shape or I0.device != I1.device:
            raise RuntimeError("I0 and I1 must have the same shape and device")
        I0 = I0.contiguous()
        I1 = I1.contiguous()
        shape = tuple(I0.shape)
        initial = self.num_scales == 1
        u_in = p_in = None
        if initial:
            if u is None or p is None:
                raise RuntimeError("num_scales == 1 starts from the given u and p")
            u_in = u.to(device=I0.device, dtype=torch.float32).contiguous()
            p_in = p.to(device=I0.device, dtype=torch.float32).contiguous()
        plan = self._plan(shape, I0.device, initial)
        if self.record_stats != (plan.stats is not None):
            plan.set_stats(self.record_stats)
        u_out = torch.empty(shape + (3,), dtype=torch.float32, device=I0.device)
        p_out = torch.empty(shape + (3, 3), dtype=torch.float32, device=I0.device)
        plan.bind(0, I0, I1, u_out=u_out, p_out=p_out, u_in=u_in, p_in=p_in)
        plan.run(use_graph=self.use_graph)
        if self.record_stats:
            self.stats = _stats_dict(plan.stats[:, :, :, 0].clone())
        if self.autotune and not getattr(plan, "tuned", False):
            plan.autotune()
            plan.tuned = True
        return u_out, p_out

    # ------------------------------------------------------- operator path
    def generate_pyramid(self, I0, I1, u, p):
        """tvl13d.py:69-120 through the drop-in operators."""
        if self.use_gaussian_blur:
            I0 = apply_gaussian_blur3d(I0, self.gaussian_blur_sigma)
            I1 = apply_gaussian_blur3d(I1, self.gaussian_blur_sigma)
        NZ, NY, NX = I0.shape
        LZ, LY, LX = get_mesh_length(self.length_type, NZ, NY, NX, self.length_z, self.length_y, self.length_x)
        meshes = [opticalFlow.MeshInfo3D(NZ, NY, NX, LZ, LY, LX)]
        nz, ny, nx = NZ, NY, NX
        for _ in range(1, self.num_scales):
            nz, ny, nx = math.ceil(0.5 * nz), math.ceil(0.5 * ny), math.ceil(0.5 * nx)
            lz, ly, lx = get_mesh_length(self.length_type, nz, ny, nx, self.length_z, self.length_y, self.length_x)
            meshes.append(opticalFlow.MeshInfo3D(nz, ny, nx, lz, ly, lx))
        I0s, I1s, us, ps = [I0], [I1], [u], [p]
        for s in range(1, self.num_scales):
            op = opticalFlow.Prolongation3D(meshes[s - 1], meshes[s], self.interpolation_type, self.boundary_type)
            I0s.append(op.forward(I0s[s - 1]))
            I1s.append(op.forward(I1s[s - 1]))
            m = meshes[s]
            us.append(torch.zeros([m.getNZ(), m.getNY(), m.getNX(), 3], dtype=torch.float32, device=I0.device))
            ps.append(torch.zeros([m.getNZ(), m.getNY(), m.getNX(), 3, 3], dtype=torch.float32, device=I0.device))
        return I0s, I1s, us, ps, meshes

    def compute_unfused(self, I0, I1, u, p):
        """tvl13d.py:122-162 step by step (one operator call per reference call)."""
        if self.huber_eps != 0.0:
            raise NotImplementedError("compute_unfused runs the reference's operators, whose TV has no Huber eps")
        I0s, I1s, us, ps, meshes = self.generate_pyramid(I0, I1, u, p)
        rec = None
        if self.record_stats:
            rec = torch.zeros((self.num_scales, self.max_warps, self.max_outer_iterations, 3), dtype=torch.float32,
                              device=I0.device)
        self._rec = rec
        for s in range(self.num_scales - 1, -1, -1):
            us[s], ps[s] = self.step(s, I0s[s], I1s[s], us[s], ps[s], meshes[s])
            if s == 0:
                break
            op = opticalFlow.Prolongation3D(meshes[s], meshes[s - 1], self.interpolation_type, self.boundary_type)
            us[s - 1] = op.forwardVectorField(us[s])
            us[s - 1][..., 0] *= meshes[s - 1].getLX() / meshes[s].getLX()
            us[s - 1][..., 1] *= meshes[s - 1].getLY() / meshes[s].getLY()
            us[s - 1][..., 2] *= meshes[s - 1].getLZ() / meshes[s].getLZ()
            ps[s - 1] = op.forwardMatrixField(ps[s])
        self._rec = None
        if rec is not None:
            self.stats = _stats_dict(rec)
        return us[0], ps[0]

    def step(self, s, I0, I1, u, p, meshInfo):
        """One pyramid level (tvl13d.py:164-255) without the residual .item() syncs,
        whose values the reference never reads; with ``record_stats`` the same
        norms (torch.norm of the differences, as tvl13d.py:222, 237 compute
        them) and the energy stay on the device in ``self.stats``."""
        rec = getattr(self, "_rec", None)
        nabla = opticalFlow.Nabla3D_CD(meshInfo, self.boundary_type)
        warping = opticalFlow.Warping3D(meshInfo, self.interpolation_type, self.boundary_type)
        I1_grad = nabla.forward(I1)
        aniso = None
        if self.use_anisotropic_diff:
            aniso = opticalFlow.AnisotropicNabla3D(meshInfo, self.anisotropic_diff_alpha, self.anistropic_diff_beta)
            tangents = aniso.computeTangentVecs(I1_grad)
        theta = float(np.float32(self.relax_theta))
        for w in range(self.max_warps):
            z = u  # tvl13d.py:194; with over-relaxation, every warp restarts from z = u
            I1_warped = warping.forward(I1, u)
            I1_warped_grad = warping.forwardVectorField(I1_grad, u)
            rho_c = I1_warped - torch.sum(u * I1_warped_grad, dim=3) - I0
            sigma, tau, gamma = self.sigma, self.tau, self.gamma
            for n in range(self.max_outer_iterations):
                pold, uold = p, u
                z_grad = nabla.forwardVectorField(z)
                Dz_grad = aniso.forwardVectorField(z_grad, *tangents) if aniso else z_grad
                p = opticalFlow.TVL1OF3D_proxDual(p + sigma * Dz_grad, sigma, self.weight_tv, meshInfo)
                if self.weight_tv == 0.:
                    p = torch.zeros_like(p)
                rho = rho_c + torch.sum(u * I1_warped_grad, dim=3)
                Dp = aniso.backwardVectorField(p, *tangents) if aniso else p
                Dp_div = nabla.backwardVectorField(Dp)
                u = opticalFlow.TVL1OF3D_proxPrimal(u - tau * Dp_div, tau, self.weight_matching, rho,
                                                    I1_warped_grad, meshInfo)
                if rec is not None:
                    rec[s, w, n, 0] = torch.norm(u - uold)
                    rec[s, w, n, 1] = torch.norm(p - pold)
                    rec[s, w, n, 2] = (float("nan") if theta != 0.0 else
                                       float(np.float32(self.weight_matching)) * rho.abs().sum()
                                       + float(np.float32(self.weight_tv))
                                       * torch.linalg.vector_norm(Dz_grad, dim=-1).sum())
                if self.primaldual_algorithm_type == 2:
                    th = 1. / math.sqrt(1. + 2. * gamma * tau)
                    tau *= th
                    sigma /= th
                z = u if theta == 0.0 else u + theta * (u - uold)
        return u, p
