"""Per-iteration statistics of the fused solve (the reference's residual norms
tvl13d.py:222, 237, 253 and the energy of each iterate): against the oracle's,
against the reference-shaped driver's torch.norm values, across batch
layouts and launch modes, and without effect on u and p."""
import numpy as np
import pytest
import torch

from singleventriclesegmentation_amd.flow_utils import gaussian_weights3d
from singleventriclesegmentation_amd.solver import Plan, SolverConfig

from test_gpu_solver import make_solver, pair

pytestmark = pytest.mark.gpu

# The fused path adds each work-group's terms in float and the work-groups in
# double; the oracle adds everything in double: measured relative differences
# ~6e-8 at these sizes.
RTOL = 1e-6


def oracle_stats(O, I0, I1, S, W, K, boundary="MIRROR", aniso=True):
    w = gaussian_weights3d(1.0)
    prm = O.solve_params(num_scales=S, max_warps=W, max_outer_iterations=K, boundary=boundary, use_aniso=aniso)
    return O.solve(O.blur(I0, w), O.blur(I1, w), prm, stats=True)


def plan_stats(dev, frames, S, W, K, boundary="MIRROR", aniso=True, resident=None, graph=True, stats=True,
               huber_eps=0.0):
    cfg = SolverConfig(num_scales=S, max_warps=W, max_outer_iterations=K, boundary={"NEAREST": 0, "MIRROR": 1,
                                                                                    "REFLECT": 2}[boundary],
                       use_aniso=aniso, huber_eps=huber_eps)
    shapes = [f[0].shape for f in frames]
    plan = Plan(cfg, shapes, device=dev)
    outs = []
    for k, (a, b) in enumerate(frames):
        I0, I1 = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
        u = torch.empty(a.shape + (3,), device=dev)
        p = torch.empty(a.shape + (3, 3), device=dev)
        plan.bind(k, I0, I1, u_out=u, p_out=p)
        outs.append((I0, I1, u, p))
    if resident is not None:
        for lv, mode in resident.items():
            plan.set_resident(lv, mode)
    if stats:
        plan.set_stats(True)
    plan.run(use_graph=graph)
    torch.cuda.synchronize()
    st = plan.stats.cpu().numpy().astype(np.float64) if stats else None
    res = [(u.cpu().numpy(), p.cpu().numpy()) for (_, _, u, p) in outs]
    plan.close()
    return st, res


@pytest.mark.parametrize("boundary,aniso", [("MIRROR", True), ("NEAREST", False), ("REFLECT", True)])
def test_stats_match_oracle(O, dev, boundary, aniso):
    """20x24x28 at two levels: level 0 runs the marching sweep, level 1
    (10x12x14) the resident sweep with two voxels per thread."""
    I0, I1, _ = pair((20, 24, 28))
    S, W, K = 2, 3, 8
    uo, po, so = oracle_stats(O, I0, I1, S, W, K, boundary, aniso)
    st, [(u, p)] = plan_stats(dev, [(I0, I1)], S, W, K, boundary, aniso)
    assert st.shape == (S, W, K, 1, 3)
    assert np.array_equal(u, uo) and np.array_equal(p, po)
    np.testing.assert_allclose(st[:, :, :, 0], so, rtol=RTOL)
    print(f"{boundary}: max rel diff {np.max(np.abs(st[:, :, :, 0] - so) / np.maximum(so, 1e-30)):.2e}")


def test_stats_do_not_change_the_solve_and_repeat_exactly(dev):
    I0, I1, _ = pair((24, 20, 28), seed=3)
    S, W, K = 3, 2, 6
    st1, [(u1, p1)] = plan_stats(dev, [(I0, I1)], S, W, K)
    st2, [(u2, p2)] = plan_stats(dev, [(I0, I1)], S, W, K, graph=False)
    _, [(u0, p0)] = plan_stats(dev, [(I0, I1)], S, W, K, stats=False)
    assert np.array_equal(u1, u0) and np.array_equal(p1, p0)
    assert np.array_equal(u2, u0) and np.array_equal(p2, p0)
    assert np.array_equal(st1, st2)  # fixed summation order: the same bits, graph or not


def test_resident_and_marching_levels_give_the_same_stats(O, dev):
    """A 10^3 level runs resident (one voxel per thread) or marching; the
    statistics agree with each other and with the oracle."""
    I0, I1, _ = pair((20, 20, 20), seed=5)
    S, W, K = 2, 2, 7
    _, _, so = oracle_stats(O, I0, I1, S, W, K)
    sr, _ = plan_stats(dev, [(I0, I1)], S, W, K, resident={1: 1})
    sm, _ = plan_stats(dev, [(I0, I1)], S, W, K, resident={1: 0})
    np.testing.assert_allclose(sr[:, :, :, 0], so, rtol=RTOL)
    np.testing.assert_allclose(sm[:, :, :, 0], so, rtol=RTOL)
    np.testing.assert_allclose(sr, sm, rtol=RTOL)


def test_uneven_batch_stats_per_instance(O, dev):
    """Three sizes in one plan: each instance's statistics match its own
    oracle solve (tiles of several instances interleave in the launches)."""
    sizes = [(16, 18, 20), (24, 16, 12), (12, 12, 30)]
    frames = []
    for i, sh in enumerate(sizes):
        a, b, _ = pair(sh, seed=10 + i)
        frames.append((a, b))
    S, W, K = 2, 2, 5
    st, res = plan_stats(dev, frames, S, W, K)
    for k, (a, b) in enumerate(frames):
        uo, po, so = oracle_stats(O, a, b, S, W, K)
        assert np.array_equal(res[k][0], uo) and np.array_equal(res[k][1], po)
        np.testing.assert_allclose(st[:, :, :, k], so, rtol=RTOL)


def test_stats_match_the_reference_driver(dev):
    """compute (fused) against compute_unfused, which records torch.norm of the
    differences as the reference's tvl13d.py:222, 237 compute them."""
    I0, I1, _ = pair((32, 32, 32), seed=7)
    a, b = torch.from_numpy(I0).to(dev), torch.from_numpy(I1).to(dev)
    solver = make_solver(2, 3, 10, record_stats=True)
    u_f, p_f = solver.compute(a, b, None, None)
    fused = {k: v.cpu().numpy().astype(np.float64) for k, v in solver.stats.items()}
    u_r, p_r = solver.compute_unfused(a, b, None, None)
    ref = {k: v.cpu().numpy().astype(np.float64) for k, v in solver.stats.items()}
    assert torch.equal(u_f, u_r) and torch.equal(p_f, p_r)
    assert fused["primal"].shape == (2, 3, 10)
    for k in ("primal", "dual", "update", "energy"):
        np.testing.assert_allclose(fused[k], ref[k], rtol=1e-4)
    assert np.array_equal(fused["update"], fused["primal"])


@pytest.mark.parametrize("boundary,aniso,eps", [("MIRROR", True, 0.05), ("NEAREST", False, 0.2),
                                                ("REFLECT", True, 0.01)])
def test_huber_tv_matches_oracle(O, dev, boundary, aniso, eps):
    """Huber-TV (eps > 0, not in the reference): the fused solve (a marching and
    a resident level) equals the oracle's bits, its statistics the oracle's
    Huber energies and residuals, and TVL13DOpticalFlow(huber_eps=...) the plan."""
    I0, I1, _ = pair((20, 24, 28), seed=9)
    S, W, K = 2, 3, 8
    w = gaussian_weights3d(1.0)
    prm = O.solve_params(num_scales=S, max_warps=W, max_outer_iterations=K, boundary=boundary, use_aniso=aniso,
                         huber_eps=eps)
    uo, po, so = O.solve(O.blur(I0, w), O.blur(I1, w), prm, stats=True)
    st, [(u, p)] = plan_stats(dev, [(I0, I1)], S, W, K, boundary, aniso, huber_eps=eps)
    assert np.array_equal(u, uo) and np.array_equal(p, po)
    np.testing.assert_allclose(st[:, :, :, 0], so, rtol=RTOL)
    solver = make_solver(S, W, K, boundary_type=boundary, use_anisotropic_diff=aniso, huber_eps=eps)
    uc, pc = solver.compute(torch.from_numpy(I0).to(dev), torch.from_numpy(I1).to(dev), None, None)
    assert np.array_equal(uc.cpu().numpy(), uo) and np.array_equal(pc.cpu().numpy(), po)


def test_stats_follow_shape_changes_and_toggling(O, dev):
    """Statistics stay right when the launch shapes change after they were
    turned on (more tiles: the plan grows its partial-sum buffer and captures
    its graph again) and when they are turned off and on again."""
    I0, I1, _ = pair((24, 24, 24), seed=13)
    S, W, K = 2, 2, 6
    _, _, so = oracle_stats(O, I0, I1, S, W, K)
    cfg = SolverConfig(num_scales=S, max_warps=W, max_outer_iterations=K)
    plan = Plan(cfg, [I0.shape], device=dev)
    a, b = torch.from_numpy(I0).to(dev), torch.from_numpy(I1).to(dev)
    u = torch.empty(I0.shape + (3,), device=dev)
    plan.bind(0, a, b, u_out=u)
    plan.set_stats(True)
    plan.run()
    first = plan.stats.clone()
    tiles0 = plan.info(0)["tiles"]
    plan.set_shapes([(8, 4, 1), (8, 4, 1)])  # 6 x 2 interiors, one plane: more tiles
    assert plan.info(0)["tiles"] > 2 * tiles0
    plan.run()
    torch.cuda.synchronize()
    np.testing.assert_allclose(plan.stats.cpu().numpy()[:, :, :, 0], so, rtol=RTOL)
    np.testing.assert_allclose(plan.stats.cpu().numpy(), first.cpu().numpy(), rtol=RTOL)
    plan.set_stats(False)
    plan.run()
    plan.set_stats(True)
    plan.run(use_graph=False)
    torch.cuda.synchronize()
    np.testing.assert_allclose(plan.stats.cpu().numpy()[:, :, :, 0], so, rtol=RTOL)
    plan.close()


@pytest.mark.parametrize("boundary,aniso,theta,eps", [("MIRROR", True, 0.5, 0.0), ("NEAREST", False, 1.0, 0.0),
                                                      ("REFLECT", True, 0.8, 0.05)])
def test_over_relaxation_matches_oracle(O, dev, boundary, aniso, theta, eps):
    """Over-relaxation (relax_theta != 0, not in the reference), alone and with
    Huber-TV: the fused solve through a marching and a resident level equals
    the oracle's bits, the residual statistics agree and the energy is NaN,
    and TVL13DOpticalFlow(relax_theta=...) returns the plan's bits."""
    I0, I1, _ = pair((20, 24, 28), seed=17)
    S, W, K = 2, 3, 8
    w = gaussian_weights3d(1.0)
    prm = O.solve_params(num_scales=S, max_warps=W, max_outer_iterations=K, boundary=boundary, use_aniso=aniso,
                         huber_eps=eps, relax_theta=theta)
    uo, po, so = O.solve(O.blur(I0, w), O.blur(I1, w), prm, stats=True)
    cfg = SolverConfig(num_scales=S, max_warps=W, max_outer_iterations=K,
                       boundary={"NEAREST": 0, "MIRROR": 1, "REFLECT": 2}[boundary], use_aniso=aniso,
                       huber_eps=eps, relax_theta=theta)
    plan = Plan(cfg, [I0.shape], device=dev)
    a, b = torch.from_numpy(I0).to(dev), torch.from_numpy(I1).to(dev)
    u = torch.empty(I0.shape + (3,), device=dev)
    p = torch.empty(I0.shape + (3, 3), device=dev)
    plan.bind(0, a, b, u_out=u, p_out=p)
    plan.set_stats(True)
    plan.run()
    torch.cuda.synchronize()
    assert plan.info(1)["resident"] > 0
    assert np.array_equal(u.cpu().numpy(), uo) and np.array_equal(p.cpu().numpy(), po)
    st = plan.stats.cpu().numpy().astype(np.float64)[:, :, :, 0]
    np.testing.assert_allclose(st[..., :2], so[..., :2], rtol=RTOL)
    assert np.all(np.isnan(st[..., 2])) and np.all(np.isnan(so[..., 2]))
    plan.close()
    solver = make_solver(S, W, K, boundary_type=boundary, use_anisotropic_diff=aniso, huber_eps=eps,
                         relax_theta=theta)
    uc, pc = solver.compute(a, b, None, None)
    assert np.array_equal(uc.cpu().numpy(), uo) and np.array_equal(pc.cpu().numpy(), po)
    # the operator-by-operator driver with the same relaxation (Huber has no
    # reference operator, so only eps = 0 runs there)
    if eps == 0.0:
        ur, pr = solver.compute_unfused(a, b, None, None)
        assert torch.equal(ur, uc) and torch.equal(pr, pc)


def test_over_relaxation_marching_equals_resident(dev):
    """The 10^3 level with theta != 0 gives the same bits resident or marching."""
    I0, I1, _ = pair((20, 20, 20), seed=19)
    outs = []
    for mode in (1, 0):
        cfg = SolverConfig(num_scales=2, max_warps=2, max_outer_iterations=6, relax_theta=0.7)
        plan = Plan(cfg, [I0.shape], device=dev)
        a, b = torch.from_numpy(I0).to(dev), torch.from_numpy(I1).to(dev)
        u = torch.empty(I0.shape + (3,), device=dev)
        plan.bind(0, a, b, u_out=u)
        plan.set_resident(1, mode)
        plan.run()
        torch.cuda.synchronize()
        outs.append(u.cpu().numpy())
        plan.close()
    assert np.array_equal(outs[0], outs[1])


def test_extended_model_and_stats_on_a_fixed_mesh(O, dev):
    """Non-unit spacing (fixed lengths: h != 1, the kernels' general-h variant)
    with Huber-TV, over-relaxation and statistics together: bits and residuals
    as the oracle's."""
    I0, I1, _ = pair((18, 20, 22), seed=23)
    S, W, K = 2, 2, 6
    length = (1.3, 0.9, 1.1)   # (lx, ly, lz)
    w = gaussian_weights3d(1.0)
    prm = O.solve_params(num_scales=S, max_warps=W, max_outer_iterations=K, length_type="fixed", length=length,
                         huber_eps=0.03, relax_theta=0.6)
    uo, po, so = O.solve(O.blur(I0, w), O.blur(I1, w), prm, stats=True)
    cfg = SolverConfig(num_scales=S, max_warps=W, max_outer_iterations=K, length_fixed=True, length=length,
                       huber_eps=0.03, relax_theta=0.6)
    plan = Plan(cfg, [I0.shape], device=dev)
    a, b = torch.from_numpy(I0).to(dev), torch.from_numpy(I1).to(dev)
    u = torch.empty(I0.shape + (3,), device=dev)
    p = torch.empty(I0.shape + (3, 3), device=dev)
    plan.bind(0, a, b, u_out=u, p_out=p)
    plan.set_stats(True)
    plan.run()
    torch.cuda.synchronize()
    assert np.array_equal(u.cpu().numpy(), uo) and np.array_equal(p.cpu().numpy(), po)
    st = plan.stats.cpu().numpy().astype(np.float64)[:, :, :, 0]
    np.testing.assert_allclose(st[..., :2], so[..., :2], rtol=RTOL)
    plan.close()
