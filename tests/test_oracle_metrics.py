"""Oracle of the warp-harness metrics (warping.py:245-270: monai compute_dice and
compute_hausdorff_distance) pinned to what monai computes them from --
scipy.ndimage binary_erosion (default cross) and distance_transform_edt -- and to
closed forms; and the flow-file layout of processor.py:108-120 / datasets.py:279-291."""
import numpy as np
import pytest
import torch
from scipy import ndimage

R = np.random.default_rng(123)


def scipy_metrics(a, b):
    a = a > 0.5
    b = b > 0.5
    ea = ndimage.binary_erosion(a) ^ a
    eb = ndimage.binary_erosion(b) ^ b
    hd = max(ndimage.distance_transform_edt(~eb)[ea].max(), ndimage.distance_transform_edt(~ea)[eb].max())
    dice = 2.0 * (a & b).sum() / (a.sum() + b.sum())
    return float(np.float32(hd)), float(np.float32(dice)), int(ea.sum()), int(eb.sum())


@pytest.mark.parametrize("it", range(6))
def test_metrics_match_scipy_construction(O, it):
    shape = tuple(int(v) for v in R.integers(4, 16, 3))
    a = ndimage.binary_dilation(R.random(shape) > 0.92, iterations=1).astype(np.float32)
    b = ndimage.binary_dilation(R.random(shape) > 0.92, iterations=1).astype(np.float32)
    if not a.any() or not b.any():
        pytest.skip("empty draw")
    hd, dice, ea, eb = scipy_metrics(a, b)
    assert O.hausdorff(a, b) == hd
    assert O.dice(a, b) == pytest.approx(dice, rel=1e-7)
    assert O.edge_count(a) == ea and O.edge_count(b) == eb


def cube(shape, lo, hi):
    m = np.zeros(shape, np.float32)
    m[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = 1
    return m


def test_metrics_closed_forms(O):
    s = (20, 20, 20)
    a = cube(s, (5, 5, 5), (11, 11, 11))                  # 6^3 cube
    assert O.dice(a, a) == 1.0 and O.hausdorff(a, a) == 0.0
    b = cube(s, (5, 5, 8), (11, 11, 14))                  # shifted by 3 along x
    assert O.hausdorff(a, b) == 3.0
    assert O.dice(a, b) == float(np.float32(2 * 6 * 6 * 3) / np.float32(2 * 216))
    c = cube(s, (7, 7, 7), (9, 9, 9))                     # 2^3 cube centred inside (voxels 7..8)
    # every voxel of the 2^3 cube is an edge; the corners of the 6^3 cube's edge
    # shell (5 and 10 per axis) are sqrt(2^2 * 3) from the nearest of them
    assert O.edge_count(c) == 8 and O.edge_count(a) == 216 - 64
    assert O.hausdorff(a, c) == float(np.float32(np.sqrt(12.0)))
    # a mask touching the volume face: its face voxels are edges (outside counts as unset)
    f = cube(s, (0, 0, 0), (20, 20, 20))
    assert O.edge_count(f) == 20 ** 3 - 18 ** 3
    assert np.isnan(O.hausdorff(np.zeros(s, np.float32), a)) and np.isnan(O.dice(a, np.zeros(s, np.float32)))


def test_flow_file_layout_is_the_processors():
    """processor.py:117-120 writes us[j] = u.permute(3, 2, 1, 0); the reader reorders
    with t3xyz_to_tzyx3 (datasets.py:34)."""
    import tempfile
    from singleventriclesegmentation_amd import flowio
    flows = torch.from_numpy(R.standard_normal((4, 5, 6, 7, 3)).astype(np.float32))
    us = torch.zeros((4, 3, 7, 6, 5))
    for j in range(4):
        us[j] = flows[j].permute(3, 2, 1, 0)
    arr = flowio.to_file_layout(flows)
    assert arr.shape == (4, 3, 7, 6, 5) and arr.dtype == np.float32
    assert np.array_equal(arr, us.numpy())
    with tempfile.TemporaryDirectory() as d:
        path = flowio.save_flows(d, "Adult_1", "forward", flows)
        assert path.endswith("Adult_1_forward_flow.npy")
        back = flowio.load_flows(d, "Adult_1", "forward")
    assert torch.equal(back, flows)
    m = (R.random((7, 6, 5)) > 0.5)
    t = flowio.mask_to_zyx(m)
    assert tuple(t.shape) == (1, 1, 5, 6, 7) and np.array_equal(t[0, 0].numpy(), m.transpose(2, 1, 0))
