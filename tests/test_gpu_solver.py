"""The fused GPU solve against the FP32 oracle (the reference's arithmetic)
and against the operator-by-operator path, plus batch/graph equivalence."""
import numpy as np
import pytest
import torch

from singleventriclesegmentation_amd import TVL13DOpticalFlow, phantom
from singleventriclesegmentation_amd.flow_utils import gaussian_weights3d
from singleventriclesegmentation_amd.solver import Plan, SolverConfig

pytestmark = pytest.mark.gpu

# Parity bar (flow in voxels): max |u - u_ref| <= 1e-3 and mean endpoint
# difference <= 1e-5 against the FP32 oracle of the same iteration.
MAX_TOL = 1e-3
MEAN_TOL = 1e-5


def make_solver(S=4, W=10, K=50, **kw):
    args = dict(num_scales=S, max_warps=W, max_outer_iterations=K, weight_matching=1.0, weight_tv=0.025,
                primaldual_algorithm_type=2, primaldual_algorithm_params=(0.5, 0.5, 0.25, 1.0),
                use_anisotropic_diff=True, anisotropic_diff_params=(5.0, 0.5), use_gaussian_blur=True,
                gaussian_blur_sigma=1.0, interpolation_type="CUBIC_HERMITESPLINE", boundary_type="MIRROR",
                length_type="numDofs", length=(1, 1, 1), device=torch.device("cuda:0"))
    args.update(kw)
    return TVL13DOpticalFlow(**args)


def oracle_solve(O, f_I0, f_I1, S, W, K, interpolation="CUBIC_HERMITESPLINE", boundary="MIRROR", aniso=True,
                 dtype=np.float32, length_type="numDofs", length=(1, 1, 1)):
    w = gaussian_weights3d(1.0)
    I0 = O.blur(f_I0, w, dtype=dtype)
    I1 = O.blur(f_I1, w, dtype=dtype)
    params = O.solve_params(num_scales=S, max_warps=W, max_outer_iterations=K, interpolation=interpolation,
                            boundary=boundary, use_aniso=aniso, length_type=length_type, length=length)
    return O.solve(I0, I1, params, dtype=dtype)


def epe(a, b):
    return np.sqrt(((a.astype(np.float64) - b.astype(np.float64)) ** 2).sum(-1))


def pair(shape, seed=0):
    f0, f1, d = phantom.frame_pair(shape, seed0=seed, seed1=seed + 1, disp_seed=seed + 2, peak=2.0)
    return f1, f0, d  # I0 = later frame, I1 = earlier frame


@pytest.mark.parametrize("boundary", ["MIRROR", "NEAREST", "REFLECT"])
@pytest.mark.parametrize("aniso", [True, False])
def test_fused_matches_oracle_small(O, dev, boundary, aniso):
    shape = (20, 24, 28)
    I0, I1, _ = pair(shape)
    S, W, K = 2, 3, 10
    solver = make_solver(S, W, K, boundary_type=boundary, use_anisotropic_diff=aniso)
    u, p = solver.compute(torch.from_numpy(I0).to(dev), torch.from_numpy(I1).to(dev), None, None)
    u, p = u.cpu().numpy(), p.cpu().numpy()
    uo, po = oracle_solve(O, I0, I1, S, W, K, boundary=boundary, aniso=aniso)
    e = epe(u, uo)
    print(f"{boundary} aniso={aniso}: max|du|={np.abs(u - uo).max():.3e} mean EPE diff={e.mean():.3e}")
    assert np.abs(u - uo).max() <= MAX_TOL and e.mean() <= MEAN_TOL
    # Same operation order, no fused multiply-adds, shared expf/powf: identical bits.
    assert np.array_equal(u, uo) and np.array_equal(p, po)


@pytest.mark.parametrize("boundary", ["MIRROR", "NEAREST", "REFLECT"])
def test_four_levels_down_to_four_voxels_match_oracle(O, dev, boundary):
    """Odd level sizes all the way to the 4-voxel minimum (25x30x27 -> 13x15x14 ->
    7x8x7 -> 4x4x4): restriction, the separable prolongation between odd sizes
    and every boundary rule, bit for bit."""
    shape = (25, 30, 27)
    I0, I1, _ = pair(shape, seed=21)
    S, W, K = 4, 2, 6
    solver = make_solver(S, W, K, boundary_type=boundary)
    u, p = solver.compute(torch.from_numpy(I0).to(dev), torch.from_numpy(I1).to(dev), None, None)
    uo, po = oracle_solve(O, I0, I1, S, W, K, boundary=boundary)
    assert np.array_equal(u.cpu().numpy(), uo) and np.array_equal(p.cpu().numpy(), po)


@pytest.mark.parametrize("interpolation,length_type,length", [
    ("NEAREST", "numDofs", (1, 1, 1)), ("LINEAR", "numDofs", (1, 1, 1)),
    ("CUBIC_HERMITESPLINE", "fixed", (30.0, 22.5, 17.0)), ("LINEAR", "fixed", (9.0, 31.0, 20.0))])
def test_fused_matches_oracle_interpolation_and_mesh(O, dev, interpolation, length_type, length):
    """The other interpolation types, and fixed mesh lengths (h != 1 on every axis
    and level, the general sweep kernel), equal the oracle bit for bit."""
    shape = (20, 24, 28)
    I0, I1, _ = pair(shape, seed=3)
    S, W, K = 2, 3, 10
    solver = make_solver(S, W, K, interpolation_type=interpolation, length_type=length_type, length=length)
    u, p = solver.compute(torch.from_numpy(I0).to(dev), torch.from_numpy(I1).to(dev), None, None)
    uo, po = oracle_solve(O, I0, I1, S, W, K, interpolation=interpolation, length_type=length_type, length=length)
    assert np.array_equal(u.cpu().numpy(), uo) and np.array_equal(p.cpu().numpy(), po)


def test_fused_equals_operator_path(dev):
    shape = (16, 18, 20)
    I0, I1, _ = pair(shape, seed=5)
    solver = make_solver(2, 2, 6)
    a, b = torch.from_numpy(I0).to(dev), torch.from_numpy(I1).to(dev)
    u1, p1 = solver.compute(a, b, None, None)
    z = torch.zeros(shape + (3,), device=dev)
    zp = torch.zeros(shape + (3, 3), device=dev)
    u2, p2 = solver.compute_unfused(a, b, z, zp)
    du = (u1 - u2).abs().max().item()
    print("fused vs operator path max |du| =", du)
    assert torch.equal(u1, u2) and torch.equal(p1, p2)


def test_graph_replay_matches_direct(dev):
    shape = (20, 20, 20)
    I0, I1, _ = pair(shape, seed=7)
    a, b = torch.from_numpy(I0).to(dev), torch.from_numpy(I1).to(dev)
    s_graph = make_solver(2, 2, 5, use_graph=True)
    s_direct = make_solver(2, 2, 5, use_graph=False)
    ud, pd = s_direct.compute(a, b, None, None)
    for _ in range(3):  # several replays, with eager work in between
        ug, pg = s_graph.compute(a, b, None, None)
        torch.cuda.synchronize()
        assert torch.equal(ug, ud) and torch.equal(pg, pd)
        _ = (a * 2).sum().item()


def test_batch_equals_alone(dev):
    cfg = SolverConfig(num_scales=2, max_warps=2, max_outer_iterations=5)
    shapes = [(20, 24, 28), (16, 16, 16), (24, 20, 18)]
    pairs = [pair(s, seed=10 + i) for i, s in enumerate(shapes)]
    inputs = [(torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)) for a, b, _ in pairs]
    batch = Plan(cfg, shapes)
    outs = []
    for k, (a, b) in enumerate(inputs):
        u = torch.empty(shapes[k] + (3,), device=dev)
        batch.bind(k, a, b, u_out=u)
        outs.append(u)
    batch.run(use_graph=True)
    torch.cuda.synchronize()
    for k, (a, b) in enumerate(inputs):
        alone = Plan(cfg, [shapes[k]])
        u = torch.empty(shapes[k] + (3,), device=dev)
        alone.bind(0, a, b, u_out=u)
        alone.run(use_graph=False)
        torch.cuda.synchronize()
        assert torch.equal(u, outs[k]), f"instance {k} differs"


def test_single_scale_uses_initial_state(O, dev):
    shape = (12, 14, 16)
    I0, I1, _ = pair(shape, seed=3)
    r = np.random.default_rng(0)
    u0 = (r.standard_normal(shape + (3,)) * 0.3).astype(np.float32)
    p0 = (r.standard_normal(shape + (3, 3)) * 0.01).astype(np.float32)
    solver = make_solver(1, 2, 5)
    u, p = solver.compute(torch.from_numpy(I0).to(dev), torch.from_numpy(I1).to(dev),
                          torch.from_numpy(u0).to(dev), torch.from_numpy(p0).to(dev))
    w = gaussian_weights3d(1.0)
    params = O.solve_params(num_scales=1, max_warps=2, max_outer_iterations=5)
    uo, po = O.solve(O.blur(I0, w), O.blur(I1, w), params, u=u0, p=p0)
    assert np.abs(u.cpu().numpy() - uo).max() <= 1e-5


@pytest.mark.slow
def test_pair_80cubed_matches_oracle(O, dev):
    """The benchmark case: 80^3 phantom pair, 4 levels x 10 warps x 50 iterations."""
    f0, f1, d = phantom.frame_pair((80, 80, 80))
    solver = make_solver(4, 10, 50)
    u, _ = solver.compute(torch.from_numpy(f1).to(dev), torch.from_numpy(f0).to(dev), None, None)
    u = u.cpu().numpy()
    uo, _ = oracle_solve(O, f1, f0, 4, 10, 50)
    e = epe(u, uo)
    print(f"80^3: max|du|={np.abs(u - uo).max():.3e} mean EPE diff={e.mean():.3e} "
          f"EPE vs truth: gpu {epe(u, d).mean():.4f} oracle {epe(uo, d).mean():.4f}")
    assert np.abs(u - uo).max() <= MAX_TOL
    assert e.mean() <= MEAN_TOL
    assert np.array_equal(u, uo)


@pytest.mark.parametrize("sweep_shape", [(0, 0, 0), (18, 18, 4), (30, 17, 3), (10, 6, 5), (7, 5, 1), (34, 4, 16)])
def test_sweep_tilings_bit_exact(O, dev, sweep_shape):
    """Any work-group tiling of the fused sweep gives the oracle's bits (the
    anisotropic weight is off so that no expf/powf is involved)."""
    shape = (16, 17, 18)
    I0, I1, _ = pair(shape, seed=21)
    cfg = SolverConfig(num_scales=2, max_warps=2, max_outer_iterations=6, use_aniso=False, sweep_shape=sweep_shape)
    plan = Plan(cfg, [shape])
    u = torch.empty(shape + (3,), device=dev)
    p = torch.empty(shape + (3, 3), device=dev)
    plan.bind(0, torch.from_numpy(I0).to(dev), torch.from_numpy(I1).to(dev), u_out=u, p_out=p)
    plan.run(use_graph=False)
    uo, po = oracle_solve(O, I0, I1, 2, 2, 6, aniso=False)
    assert np.array_equal(u.cpu().numpy(), uo)
    assert np.array_equal(p.cpu().numpy(), po)


def test_gpu_reproduces_golden_solve(dev):
    """HIP solve of the committed fixture pair equals the oracle's FP32 bits."""
    import os
    g = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "solve_10x11x12.npz"), allow_pickle=False))
    params = g["params"]
    solver = make_solver(int(params[0]), int(params[1]), int(params[2]))
    u, p = solver.compute(torch.from_numpy(g["I0"]).to(dev), torch.from_numpy(g["I1"]).to(dev), None, None)
    assert np.array_equal(u.cpu().numpy(), g["u_f32"]) and np.array_equal(p.cpu().numpy(), g["p_f32"])
    err64 = np.abs(u.cpu().numpy().astype(np.float64) - g["u_f64"]).max()
    print(f"golden: GPU == FP32 oracle; max |u - u_fp64| = {err64:.2e}")


def test_autotune_keeps_results(dev):
    cfg = SolverConfig(num_scales=2, max_warps=2, max_outer_iterations=4)
    shapes = [(16, 18, 20), (12, 12, 24)]
    pairs = [pair(s, seed=30 + i) for i, s in enumerate(shapes)]
    plan = Plan(cfg, shapes)
    outs = []
    for k, (a, b, _) in enumerate(pairs):
        u = torch.empty(shapes[k] + (3,), device=dev)
        plan.bind(k, torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev), u_out=u)
        outs.append(u)
    plan.run(use_graph=True)
    before = [u.clone() for u in outs]
    tuned = plan.autotune(reps=2)
    assert len(tuned) == 2 and all(bx >= 3 and by >= 3 and cz >= 1 for bx, by, cz in tuned)
    plan.run(use_graph=True)
    torch.cuda.synchronize()
    assert all(torch.equal(a, b) for a, b in zip(before, outs))


def test_set_shapes_keeps_results_and_rejects_bad_shapes(dev):
    """Fixed per-level sweep shapes (bench --sweep-shapes) re-tile the launch
    and drop the graph; the results stay bit-identical."""
    cfg = SolverConfig(num_scales=2, max_warps=2, max_outer_iterations=4)
    shape = (20, 22, 24)
    a, b, _ = pair(shape, seed=41)
    plan = Plan(cfg, [shape])
    u = torch.empty(shape + (3,), device=dev)
    plan.bind(0, torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev), u_out=u)
    plan.run(use_graph=True)
    before = u.clone()
    for shapes in ([(26, 6, 8), (14, 6, 3)], [(10, 10, 20), (8, 8, 2)]):
        plan.set_shapes(shapes)
        assert plan.info(0)["bx"] == shapes[0][0] and plan.info(1)["cz"] == shapes[1][2]
        plan.run(use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(u, before)
    with pytest.raises(RuntimeError):
        plan.set_shapes([(2, 6, 8), (14, 6, 3)])        # bx < 3
    with pytest.raises(RuntimeError):
        plan.set_shapes([(40, 40, 8), (14, 6, 3)])      # more than 1024 threads
    with pytest.raises(RuntimeError):
        plan.set_shapes([(26, 6, 8)])                   # one shape per level


@pytest.mark.parametrize("boundary", ["MIRROR", "NEAREST", "REFLECT"])
@pytest.mark.parametrize("aniso", [True, False])
def test_resident_levels_equal_marching_sweep(O, dev, boundary, aniso):
    """The resident sweep (all iterations of a warp in one launch, one
    work-group per instance) gives the marching sweep's bits: one voxel per
    thread (8^3, 10^3 levels) and two (10x12x14 = 1680 voxels), in one batch;
    and the oracle's."""
    shapes = [(20, 24, 28), (16, 16, 16), (20, 20, 20)]
    cfg = SolverConfig(num_scales=2, max_warps=2, max_outer_iterations=7, use_aniso=aniso,
                       boundary={"NEAREST": 0, "MIRROR": 1, "REFLECT": 2}[boundary])
    pairs = [pair(s, seed=50 + i) for i, s in enumerate(shapes)]
    inputs = [(torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)) for a, b, _ in pairs]
    outs = {}
    for mode in (-1, 0):
        plan = Plan(cfg, shapes)
        if mode == 0:
            plan.set_resident(1, 0)
        assert plan.info(1)["resident"] == (2 if mode == -1 else 0)
        assert plan.info(0)["resident"] == 0
        us, ps = [], []
        for k, (a, b) in enumerate(inputs):
            u = torch.empty(shapes[k] + (3,), device=dev)
            p = torch.empty(shapes[k] + (3, 3), device=dev)
            plan.bind(k, a, b, u_out=u, p_out=p)
            us.append(u)
            ps.append(p)
        plan.run(use_graph=True)
        torch.cuda.synchronize()
        outs[mode] = ([u.cpu().numpy() for u in us], [p.cpu().numpy() for p in ps])
        plan.close()
    for k in range(len(shapes)):
        assert np.array_equal(outs[-1][0][k], outs[0][0][k]), f"u of instance {k}"
        assert np.array_equal(outs[-1][1][k], outs[0][1][k]), f"p of instance {k}"
    uo, po = oracle_solve(O, pairs[1][0], pairs[1][1], 2, 2, 7, boundary=boundary, aniso=aniso)
    assert np.array_equal(outs[-1][0][1], uo) and np.array_equal(outs[-1][1][1], po)


def test_resident_single_level_and_odd_iteration_count(O, dev):
    """A one-level plan whose only level is resident, with an odd number of
    iterations per warp (the state ends in the other buffer), equals the
    oracle; forcing the resident mode on a level that is too large fails."""
    shape = (11, 12, 13)
    I0, I1, _ = pair(shape, seed=61)
    solver = make_solver(1, 3, 5)
    z = torch.zeros(shape + (3,), device=dev)
    zp = torch.zeros(shape + (3, 3), device=dev)
    u, p = solver.compute(torch.from_numpy(I0).to(dev), torch.from_numpy(I1).to(dev), z, zp)
    plan = next(iter(solver._plans.values()))
    assert plan.info(0)["resident"] == 2
    w = gaussian_weights3d(1.0)
    uo, po = O.solve(O.blur(I0, w), O.blur(I1, w), O.solve_params(num_scales=1, max_warps=3, max_outer_iterations=5))
    assert np.array_equal(u.cpu().numpy(), uo) and np.array_equal(p.cpu().numpy(), po)
    big = Plan(SolverConfig(num_scales=1, max_warps=1, max_outer_iterations=2), [(16, 16, 16)])
    with pytest.raises(RuntimeError):
        big.set_resident(0, 1)
    big.close()
