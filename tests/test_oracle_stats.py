"""The oracle's per-iteration statistics: the reference's residual norms
||u^{n+1} - u^n||, ||p^{n+1} - p^n|| (tvl13d.py:222, 237, 253) and the energy
E(u^n) = lambda sum|rho(u^n)| + lambda_tv sum_c ||D grad u^n_c||.  Pinned on
the CPU by the oracle's own final states (solves of n and n + 1 iterations
differ by exactly one iteration), by the level and warp structure (a level of
a pyramid solve equals a one-level solve on that level's images), and by
closed forms of the energy and the first dual step."""
import numpy as np
import pytest

from singleventriclesegmentation_amd import phantom


def pair(shape=(9, 10, 11), seed=5):
    f = phantom.add_noise(phantom.clean_frame(shape), seed)
    g = phantom.add_noise(phantom.clean_frame(shape), seed + 1)
    return f, g


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("aniso,boundary", [(True, "MIRROR"), (False, "NEAREST"), (True, "REFLECT")])
def test_residuals_are_distances_between_consecutive_iterates(O, dtype, aniso, boundary):
    """One warp, one level, from a nonzero initial state: the statistics of a
    K-iteration solve at iteration n are the distances between the final
    states of solves with n and n + 1 iterations."""
    I0, I1 = pair()
    r = np.random.default_rng(2)
    u0 = (0.3 * r.standard_normal(I0.shape + (3,))).astype(dtype)
    p0 = (0.01 * r.standard_normal(I0.shape + (3, 3))).astype(dtype)
    K = 6
    mk = lambda k: O.solve_params(num_scales=1, max_warps=1, max_outer_iterations=k, use_aniso=aniso,
                                  boundary=boundary)
    _, _, st = O.solve(I0, I1, mk(K), u=u0, p=p0, dtype=dtype, stats=True)
    assert st.shape == (1, 1, K, 3)
    states = [O.solve(I0, I1, mk(k), u=u0, p=p0, dtype=dtype) for k in range(K + 1)]
    for n in range(K):
        du = states[n + 1][0].astype(np.float64) - states[n][0].astype(np.float64)
        dp = states[n + 1][1].astype(np.float64) - states[n][1].astype(np.float64)
        np.testing.assert_allclose(st[0, 0, n, 0], np.sqrt(np.sum(du * du)), rtol=1e-12)
        np.testing.assert_allclose(st[0, 0, n, 1], np.sqrt(np.sum(dp * dp)), rtol=1e-12)
    assert np.all(st[..., 2] > 0)


def test_levels_and_warps_are_indexed_as_documented(O):
    """stats[s, w, n]: level 0 is the finest; the coarse level of a two-level
    solve is a one-level solve from zero on the restricted images, and there
    the first warp of a two-warp solve is the one-warp solve."""
    I0, I1 = pair((12, 13, 14))
    W, K = 2, 4
    _, _, st = O.solve(I0, I1, O.solve_params(num_scales=2, max_warps=W, max_outer_iterations=K), stats=True)
    assert st.shape == (2, W, K, 3)
    coarse = tuple(int(np.ceil(n / 2)) for n in I0.shape)
    L = lambda shape: tuple(float(n - 1) for n in shape)  # numDofs lengths, (LZ, LY, LX)
    c0 = O.prolongate(I0, L(I0.shape), coarse, L(coarse))
    c1 = O.prolongate(I1, L(I1.shape), coarse, L(coarse))
    _, _, sc = O.solve(c0, c1, O.solve_params(num_scales=1, max_warps=W, max_outer_iterations=K), stats=True)
    assert np.array_equal(st[1], sc[0])
    _, _, s1 = O.solve(I0, I1, O.solve_params(num_scales=2, max_warps=1, max_outer_iterations=K), stats=True)
    assert np.array_equal(st[1, 0], s1[1, 0])
    assert not np.array_equal(st[1, 1], st[1, 0])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_energy_at_zero_flow_is_the_data_term(O, dtype):
    """u^0 = 0: rho = I1 - I0 (the warp at zero displacement samples the grid,
    grad 0 = 0), so E(u^0) = lambda sum |I1 - I0| in the solve's precision."""
    I0, I1 = pair()
    lam = 0.75
    prm = O.solve_params(num_scales=1, max_warps=1, max_outer_iterations=2, weight_matching=lam)
    z = np.zeros(I0.shape + (3,), dtype)
    _, _, st = O.solve(I0, I1, prm, u=z, p=np.zeros(I0.shape + (3, 3), dtype), dtype=dtype, stats=True)
    want = lam * np.sum(np.abs((I1.astype(dtype) - I0.astype(dtype)).astype(np.float64)))
    np.testing.assert_allclose(st[0, 0, 0, 2], want, rtol=1e-12)


def test_energy_and_first_dual_step_of_a_ramp(O):
    """I0 = I1 = constant, u^0 = alpha_c x (powers of two), no anisotropy: rho = 0
    and grad u_c = (alpha_c, 0, 0) away from the x faces (0 on them, MIRROR), so
    E(u^0) = lambda_tv |interior| sum_c |alpha_c|; from p^0 = 0 the first dual
    step projects (sigma_0 alpha_c, 0, 0) onto the ball of radius lambda_tv, so
    ||p^1 - p^0||^2 = |interior| sum_c min(sigma_0 |alpha_c|, lambda_tv)^2."""
    shape = (6, 7, 9)
    I = np.full(shape, 0.5, np.float32)
    alpha = np.array([0.25, -0.5, 0.125])
    x = np.arange(shape[2], dtype=np.float32)
    u0 = np.zeros(shape + (3,), np.float32)
    for c in range(3):
        u0[..., c] = alpha[c] * x
    for wtv, sigma in ((0.025, 0.5), (1.0, 0.5), (0.1, 0.25)):
        prm = O.solve_params(num_scales=1, max_warps=1, max_outer_iterations=1, weight_tv=wtv, use_aniso=False,
                             pd_params=(sigma, 0.5, 0.25, 1.0))
        _, _, st = O.solve(I, I, prm, u=u0, p=np.zeros(shape + (3, 3), np.float32), stats=True)
        interior = shape[0] * shape[1] * (shape[2] - 2)
        np.testing.assert_allclose(st[0, 0, 0, 2], float(np.float32(wtv)) * interior * np.abs(alpha).sum(),
                                   rtol=1e-12)
        want_dp = np.sqrt(interior * np.sum(np.minimum(sigma * np.abs(alpha), wtv) ** 2))
        np.testing.assert_allclose(st[0, 0, 0, 1], want_dp, rtol=1e-6)


def test_identical_frames_have_zero_statistics(O):
    """I0 == I1 from zero: every iterate stays 0, so residuals and energy are 0."""
    f = phantom.add_noise(phantom.clean_frame((10, 10, 10)), 3)
    _, _, st = O.solve(f, f, O.solve_params(num_scales=2, max_warps=2, max_outer_iterations=3), stats=True)
    assert not st.any()


def test_huber_dual_step_and_energy_of_a_ramp(O):
    """Huber-TV (eps > 0, not in the reference): from p^0 = 0 the first dual row
    is v_c / max(hub, |v_c| / lambda_tv) with v_c = (sigma_0 alpha_c, 0, 0) and
    hub = 1 + sigma_0 eps / lambda_tv -- scaled by 1/hub inside the ball,
    projected outside -- and the energy of the ramp is lambda_tv |interior|
    sum_c H_eps(|alpha_c|)."""
    shape = (5, 6, 8)
    I = np.full(shape, 0.25, np.float32)
    alpha = np.array([0.25, -0.5, 0.125])
    u0 = np.zeros(shape + (3,), np.float32)
    u0[...] = (alpha[None, None, None, :] * np.arange(shape[2], dtype=np.float32)[None, None, :, None])
    interior = shape[0] * shape[1] * (shape[2] - 2)
    H = lambda t, e: t * t / (2 * e) if t <= e else t - e / 2
    for wtv, sigma, eps in ((0.1, 0.5, 0.05), (1.0, 0.5, 0.5), (0.025, 0.5, 0.3)):
        prm = O.solve_params(num_scales=1, max_warps=1, max_outer_iterations=1, weight_tv=wtv, use_aniso=False,
                             pd_params=(sigma, 0.5, 0.25, 1.0), huber_eps=eps)
        _, p1, st = O.solve(I, I, prm, u=u0, p=np.zeros(shape + (3, 3), np.float32), stats=True)
        w = float(np.float32(wtv))
        hub = 1.0 + sigma * eps / w
        rows = sigma * np.abs(alpha) / np.maximum(hub, sigma * np.abs(alpha) / w)
        np.testing.assert_allclose(st[0, 0, 0, 1], np.sqrt(interior * np.sum(rows ** 2)), rtol=1e-6)
        # an interior voxel's rows: (sigma alpha_c / den_c, 0, 0)
        got = p1[2, 3, 4].astype(np.float64)
        np.testing.assert_allclose(got[:, 0], sigma * alpha / np.maximum(hub, sigma * np.abs(alpha) / w), rtol=1e-6)
        assert not got[:, 1:].any()
        want_e = w * interior * sum(H(abs(a), eps) for a in alpha)
        np.testing.assert_allclose(st[0, 0, 0, 2], want_e, rtol=1e-12)


def test_huber_eps_zero_is_the_reference_step(O):
    """eps = 0 and a 19-entry parameter pack (before eps existed) give the same bits."""
    I0, I1 = pair()
    prm = O.solve_params(num_scales=2, max_warps=2, max_outer_iterations=4)
    a = O.solve(I0, I1, prm)
    b = O.solve(I0, I1, prm[:19])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    c = O.solve(I0, I1, O.solve_params(num_scales=2, max_warps=2, max_outer_iterations=4, huber_eps=0.05))
    assert not np.array_equal(a[0], c[0])


def test_over_relaxation_second_dual_step(O):
    """Over-relaxation (relax_theta, not in the reference): with constant images
    (rho = 0, g = 0), no anisotropy and a ball too large to project onto, the
    second dual step differs from the unrelaxed one by exactly sigma_1 grad of
    theta (u^1 - u^0); the first iteration does not depend on theta."""
    shape = (7, 8, 9)
    I = np.full(shape, 0.5, np.float32)
    r = np.random.default_rng(3)
    zz, yy, xx = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    u0 = np.stack([np.sin(0.4 * xx + 0.1 * c) * np.cos(0.3 * yy) + 0.2 * zz / shape[0] for c in range(3)],
                  -1).astype(np.float32)
    p0 = np.zeros(shape + (3, 3), np.float32)
    theta = 0.75
    mk = lambda k, th: O.solve_params(num_scales=1, max_warps=1, max_outer_iterations=k, use_aniso=False,
                                      weight_tv=1000.0, relax_theta=th)
    u1, _ = O.solve(I, I, mk(1, theta), u=u0, p=p0)
    u1_plain, _ = O.solve(I, I, mk(1, 0.0), u=u0, p=p0)
    assert np.array_equal(u1, u1_plain)                      # iteration 0 uses z = u^0
    _, p2r = O.solve(I, I, mk(2, theta), u=u0, p=p0)
    _, p2 = O.solve(I, I, mk(2, 0.0), u=u0, p=p0)
    sig, _ = O.step_schedule(mk(2, 0.0))
    d = (u1.astype(np.float64) - u0.astype(np.float64))
    want = sig[1] * theta * O.nabla_fwd(d, bc=1, dtype=np.float64)
    got = p2r.astype(np.float64) - p2.astype(np.float64)
    assert np.abs(want).max() > 1e-3
    np.testing.assert_allclose(got, want, atol=2e-6 * np.abs(p2).max())


def test_over_relaxation_restarts_each_warp(O):
    """z is reset to u at every warp's first iteration: with one iteration per
    warp theta changes nothing, over several warps and levels."""
    I0, I1 = pair()
    a = O.solve(I0, I1, O.solve_params(num_scales=2, max_warps=3, max_outer_iterations=1, relax_theta=0.9))
    b = O.solve(I0, I1, O.solve_params(num_scales=2, max_warps=3, max_outer_iterations=1))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    _, _, st = O.solve(I0, I1, O.solve_params(num_scales=1, max_warps=1, max_outer_iterations=3, relax_theta=0.5),
                       stats=True)
    assert np.all(np.isfinite(st[..., :2])) and np.all(np.isnan(st[..., 2]))
