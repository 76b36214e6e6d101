"""The median post-filter (flow_utils.py:115-132; applied per pair at
processor.py:116-117) on the GPU against the oracle, which is pinned to
scipy.ndimage.median_filter.  The result is an element of the window, so the
comparison is exact."""
import numpy as np
import pytest
import torch

from singleventriclesegmentation_amd.flow_utils import apply_median_filter, median_filter3d

pytestmark = pytest.mark.gpu


def g(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


@pytest.mark.parametrize("shape,ks", [((9, 10, 11, 3), 3), ((2, 3, 4, 3), 3), ((1, 1, 1, 3), 3), ((5, 7, 33, 1), 3),
                                      ((6, 5, 4, 3), 1), ((6, 5, 4, 3), 2), ((6, 5, 4, 2), 4), ((7, 6, 5, 3), 5),
                                      ((3, 3, 3, 1), 7)])
def test_median_matches_oracle(O, dev, shape, ks):
    r = np.random.default_rng(sum(shape) + ks)
    u = r.standard_normal(shape).astype(np.float32)
    u[..., 0] = np.round(u[..., 0])  # ties
    out = median_filter3d(g(u, dev), ks)
    assert out.shape == u.shape and out.dtype == torch.float32
    assert np.array_equal(out.cpu().numpy(), O.median3d(u, ks))


def test_median_batch_equals_single(O, dev):
    r = np.random.default_rng(5)
    u = r.standard_normal((4, 12, 13, 14, 3)).astype(np.float32)
    out = median_filter3d(g(u, dev), 3).cpu().numpy()
    assert np.array_equal(out, O.median3d(u, 3))
    for b in range(4):
        assert np.array_equal(median_filter3d(g(u[b], dev), 3).cpu().numpy(), out[b])


def test_apply_median_filter_dropin(O, dev):
    """Same call as processor.py:117 on a full-size 80^3 flow."""
    r = np.random.default_rng(6)
    u = (3.0 * r.standard_normal((80, 80, 80, 3))).astype(np.float32)
    got = apply_median_filter(g(u, dev), 3, dev)
    assert got.device.type == "cuda"
    assert np.array_equal(got.cpu().numpy(), O.median3d(u, 3))


def test_median_rejects_bad_arguments(dev):
    u = torch.zeros(4, 4, 4, 3, device=dev)
    for ks in (0, 8):
        with pytest.raises(RuntimeError):
            median_filter3d(u, ks)
    with pytest.raises(RuntimeError):
        median_filter3d(u.double(), 3)
    with pytest.raises(RuntimeError):
        median_filter3d(torch.zeros(4, 4, 4, 3), 3)  # host tensor
    with pytest.raises(RuntimeError):
        apply_median_filter(u[..., 0], 3, dev)  # size (ks,ks,ks,1) needs 4 dims
