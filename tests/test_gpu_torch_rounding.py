"""Pins the oracle's restatement of the torch steps in the reference's driver
(tvl13d.py:201, 215-234, 151-153) to what torch computes on the GPU: the
3-term sum(u * g, dim=3) rounds as (a + c) + b, scalar products round the
Python float step size to FP32 first, and the in-place rescaling of a flow
component multiplies by the FP32 rounding of the Python ratio."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_sum_over_three_components_rounds_a_plus_c_then_b(dev):
    r = np.random.default_rng(0)
    for shape in ((80, 80, 80, 3), (10, 11, 12, 3), (7, 5, 3, 3)):
        u = r.standard_normal(shape).astype(np.float32)
        g = r.standard_normal(shape).astype(np.float32)
        got = torch.sum(torch.from_numpy(u).to(dev) * torch.from_numpy(g).to(dev), dim=3).cpu().numpy()
        a, b, c = u[..., 0] * g[..., 0], u[..., 1] * g[..., 1], u[..., 2] * g[..., 2]
        assert np.array_equal(got, (a + c) + b), shape
        # the other orders differ on a good fraction of these inputs, so the check has teeth
        assert np.count_nonzero(got != (a + b) + c) > got.size // 10


def test_step_size_products_and_rescaling(dev):
    r = np.random.default_rng(1)
    p = r.standard_normal((20, 20, 20, 3, 3)).astype(np.float32)
    q = r.standard_normal((20, 20, 20, 3, 3)).astype(np.float32)
    sigma = 0.5 / 0.9354143466934853                        # a float64 step size
    got = (torch.from_numpy(p).to(dev) + sigma * torch.from_numpy(q).to(dev)).cpu().numpy()
    assert np.array_equal(got, p + np.float32(sigma) * q)
    got = (torch.from_numpy(p).to(dev) - sigma * torch.from_numpy(q).to(dev)).cpu().numpy()
    assert np.array_equal(got, p - np.float32(sigma) * q)
    u = r.standard_normal((40, 40, 40, 3)).astype(np.float32)
    t = torch.from_numpy(u).to(dev)
    ratio = 79.0 / 39.0                                     # L_fine / L_coarse as Python floats
    t[..., 0] *= ratio
    assert np.array_equal(t[..., 0].cpu().numpy(), u[..., 0] * np.float32(ratio))
