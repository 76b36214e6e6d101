"""The oracle's whole solve: fixture regression, FP32 vs FP64, invariants. CPU only."""
import os

import numpy as np

from conftest import REPO

GOLD = os.path.join(REPO, "tests", "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


def test_oracle_reproduces_solve_fixture(O):
    g = load("solve_10x11x12.npz")
    w = g["blur_w"]
    for name, dt in (("f32", np.float32), ("f64", np.float64)):
        u, p = O.solve(O.blur(g["I0"], w, dtype=dt), O.blur(g["I1"], w, dtype=dt), g["params"], dtype=dt)
        assert np.array_equal(u, g[f"u_{name}"]) and np.array_equal(p, g[f"p_{name}"])


def test_oracle_reproduces_ops_fixture(O):
    g = load("ops_10x11x12.npz")
    assert np.array_equal(O.warp(g["u"], g["phi"], ip=2, bc=1), g["warp_cubic_mirror"])
    assert np.array_equal(O.warp(g["v"], g["phi"], ip=1, bc=2), g["warp_vf_linear_reflect"])
    assert np.array_equal(O.nabla_fwd(g["v"], bc=1), g["nabla_fwd_vf_mirror"])
    assert np.array_equal(O.nabla_bwd(g["q"], bc=1), g["nabla_bwd_vf_mirror"])
    s, n, t1, t2 = O.aniso_tangents(g["G"], 5.0, 0.5)
    assert all(np.array_equal(a, g[k]) for a, k in ((s, "aniso_s"), (n, "aniso_n"), (t1, "aniso_t1"),
                                                     (t2, "aniso_t2")))
    assert np.array_equal(O.aniso_apply(g["q"], s, n, t1, t2), g["aniso_apply"])


def test_fp32_restatement_close_to_fp64(O):
    """The FP32 restatement's error against the FP64 oracle of the same iteration."""
    g = load("solve_10x11x12.npz")
    d = np.abs(g["u_f32"].astype(np.float64) - g["u_f64"])
    assert d.max() < 1e-4, d.max()


def test_identical_frames_give_zero_flow(O):
    """I0 == I1: rho_c = 0 at u = 0, grad u = 0, so every step keeps u = 0 and p = 0 exactly."""
    from singleventriclesegmentation_amd import phantom
    f = phantom.add_noise(phantom.clean_frame((12, 12, 12)), 3)
    u, p = O.solve(f, f, O.solve_params(num_scales=2, max_warps=2, max_outer_iterations=5))
    assert not u.any() and not p.any()


def test_flow_follows_the_known_displacement(O):
    """On a phantom pair the FP32 flow correlates with the true displacement."""
    from singleventriclesegmentation_amd import phantom
    from singleventriclesegmentation_amd.flow_utils import gaussian_weights3d
    f0, f1, d = phantom.frame_pair((24, 24, 24), peak=1.5)
    w = gaussian_weights3d(1.0)
    u, _ = O.solve(O.blur(f1, w), O.blur(f0, w), O.solve_params(num_scales=2, max_warps=5, max_outer_iterations=20))
    m = np.linalg.norm(d, axis=-1) > 0.5
    cos = np.sum(u[m] * d[m]) / (np.linalg.norm(u[m]) * np.linalg.norm(d[m]))
    assert cos > 0.5, cos
