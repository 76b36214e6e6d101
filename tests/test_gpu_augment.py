"""Flow-consistent spatial augmentation on the GPU (svs_spatial_augment behind
transforms.RandomRotate / RandomFlip / Compose) against the oracle, which
tests/test_oracle_augment.py pins to torch's CPU grid_sample / matmul / einsum
and to closed forms.  Every comparison is bit for bit (signed zeros included),
with the oracle and with the reference's own CPU calls.
"""
import ctypes

import numpy as np
import pytest
import torch

from singleventriclesegmentation_amd import _lib
from singleventriclesegmentation_amd import transforms as T

from test_oracle_augment import axes_of, bits, reference_sequence, rotation

pytestmark = pytest.mark.gpu

MODES = {"image": "bilinear", "med": "nearest", "mes": "nearest", "fwd_flow": "bilinear", "bwd_flow": "bilinear"}
KEYS = list(MODES)


def sample(shape, T_frames, seed):
    r = np.random.default_rng(seed)
    return {"image": r.standard_normal((T_frames, 1) + shape).astype(np.float32),
            "med": (r.random((1,) + shape) > 0.6).astype(np.float32),
            "mes": (r.random((1,) + shape) > 0.3).astype(np.float32),
            "fwd_flow": r.standard_normal((T_frames - 1, 3) + shape).astype(np.float32),
            "bwd_flow": r.standard_normal((T_frames - 1, 3) + shape).astype(np.float32)}


def to_dev(data, dev):
    return {k: torch.from_numpy(v).to(dev) for k, v in data.items()}


def oracle_key(O, key, x, rot, flip_axes, grid=None):
    xb = x[None] if x.ndim == 4 else x
    if rot is None:
        out = O.spatial_augment(xb, None, flip_axes, vector=key in T.FLOWS_KEYS)
    else:
        out = O.spatial_augment(xb, rot.R.numpy(), flip_axes, axes_of(rot.deform_shape), rot.modes[key],
                                rot.boundary, rot.align_corners, vector=key in T.FLOWS_KEYS, grid=grid,
                                grid_rounding=rot.grid_rounding, vector_rounding=rot.vector_rounding)
    return out[0] if x.ndim == 4 else out


FLIPS = {0: (), 1: (4,), 2: (3,), 4: (2,), 5: (2, 4), 6: (2, 3), 7: (2, 3, 4)}


@pytest.mark.parametrize("padding", ["zeros", "border", "reflection"])
@pytest.mark.parametrize("align", [False, True])
@pytest.mark.parametrize("grid", ["host", "device"])
@pytest.mark.parametrize("flipmask", [0, 5, 7])
def test_rotate_and_flip_match_oracle(O, dev, padding, align, grid, flipmask):
    shape, out_shape = (9, 11, 13), (7, 10, 12)
    data = sample(shape, 3, seed=flipmask + 10 * align)
    modes = dict(MODES, fwd_flow="nearest") if align else MODES
    rot = T.RandomRotate(KEYS, 1.0, out_shape, [(0, 360)] * 3, padding, modes, grid=grid, align_corners=align)
    rot.R = rotation(*np.random.default_rng(flipmask).uniform(0, 360, 3))
    got = T.spatial_pass(to_dev(data, dev), KEYS, rotate=rot, flips=flipmask)
    host_grid = rot.rotated_grid.numpy() if grid == "host" else None
    for key in KEYS:
        want = oracle_key(O, key, data[key], rot, FLIPS[flipmask], grid=host_grid)
        assert got[key].shape == want.shape, key
        assert np.array_equal(bits(got[key].cpu().numpy()), bits(want)), (key, padding, align, grid, flipmask)


@pytest.mark.parametrize("grid", ["host", "device"])
def test_resampling_to_a_coarser_grid_matches_oracle(O, dev, grid):
    """40^3 sampled onto 12^3: the source box of an output tile is too large
    to stage, so the tiles gather straight from global memory."""
    data = sample((40, 40, 40), 3, seed=17)
    rot = T.RandomRotate(KEYS, 1.0, (12, 12, 12), [(0, 360)] * 3, "zeros", MODES, grid=grid, align_corners=False)
    rot.R = rotation(*np.random.default_rng(17).uniform(0, 360, 3))
    got = T.spatial_pass(to_dev(data, dev), KEYS, rotate=rot, flips=6)
    host_grid = rot.rotated_grid.numpy() if grid == "host" else None
    for key in KEYS:
        want = oracle_key(O, key, data[key], rot, FLIPS[6], grid=host_grid)
        assert np.array_equal(bits(got[key].cpu().numpy()), bits(want)), (key, grid)


@pytest.mark.parametrize("flipmask", [1, 2, 4, 7])
def test_flips_alone_match_oracle(O, dev, flipmask):
    data = sample((6, 7, 8), 4, seed=flipmask)
    got = T.spatial_pass(to_dev(data, dev), KEYS, flips=flipmask)
    for key in KEYS:
        assert np.array_equal(bits(got[key].cpu().numpy()), bits(oracle_key(O, key, data[key], None,
                                                                            FLIPS[flipmask]))), key


@pytest.mark.parametrize("grid", ["host", "device"])
@pytest.mark.parametrize("shape", [(7, 9, 11), (48, 48, 48)])
def test_matches_reference_sequence(dev, grid, shape):
    """The reference's own CPU calls (matmul, grid_sample, einsum, flip, sign
    change) on the same inputs, bit for bit, with the rounding of this host's
    BLAS (the small grid is one torch.matmul rounds without FMA on every
    host)."""
    data = sample(shape, 3, seed=3)
    R = rotation(31.0, -12.0, 250.0)
    flips = (3, 4)
    ref = reference_sequence(data, R, flips, MODES, "zeros", False, T.MASKS_KEYS, T.FLOWS_KEYS)
    rot = T.RandomRotate(KEYS, 1.0, shape, [(0, 360)] * 3, "zeros", MODES, grid=grid, align_corners=False)
    rot.R = R
    got = T.spatial_pass(to_dev(data, dev), KEYS, rotate=rot, flips=T.FLIP_BIT[3] | T.FLIP_BIT[4])
    for key in KEYS:
        assert np.array_equal(bits(got[key].cpu().numpy()), bits(ref[key])), (key, grid, shape)


def transforms_warper_pipeline(grid, shape=(80, 80, 80)):
    """The spatial part of TransformsWarper (svs/modules/flow/warping.py:341-364)."""
    rot = T.RandomRotate(keys=KEYS, p=0.5, spatial_shape=shape, rot_ranges=[(0, 360)] * 3, boundary="zeros",
                         modes=MODES, grid=grid, align_corners=False)
    flips = [T.RandomFlip(keys=KEYS, p=0.5, axis=a) for a in (2, 3, 4)]
    return rot, flips


@pytest.mark.parametrize("grid", ["host", "device"])
def test_training_sample_80cubed_matches_oracle(O, dev, grid):
    """A full TransformsWarper sample: 25 frames, two masks, two 24-pair
    flows at 80^3, rotation and all three flips drawn (one launch)."""
    data = sample((80, 80, 80), 25, seed=21)
    rot, flips = transforms_warper_pipeline(grid)
    for t in [rot] + flips:
        t.p = 1.0
    np.random.seed(5)
    got = T.Compose([rot] + flips)(to_dev(data, dev))
    torch.cuda.synchronize()
    for key in KEYS:
        want = oracle_key(O, key, data[key], rot, (2, 3, 4),
                          grid=rot.rotated_grid.numpy() if grid == "host" else None)
        assert np.array_equal(bits(got[key].cpu().numpy()), bits(want)), key


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 4, 5])
def test_compose_fused_equals_sequential(dev, seed):
    """Compose's single launch equals the reference's one-transform-at-a-time
    application (each transform's own __call__), random draws included."""
    data = to_dev(sample((20, 24, 28), 3, seed=seed), dev)
    rot, flips = transforms_warper_pipeline("device", (20, 24, 28))
    np.random.seed(seed)
    fused = T.Compose([rot] + flips)(dict(data))
    R_fused = rot.R
    np.random.seed(seed)
    seq = dict(data)
    rot.R = None
    for t in [rot] + flips:
        seq = t(seq)
    assert (R_fused is None) == (rot.R is None)
    for key in KEYS:
        assert torch.equal(fused[key].view(torch.int32), seq[key].view(torch.int32)), (seed, key)


def test_identity_rotation_and_double_flip_at_80cubed(dev):
    data = to_dev(sample((80, 80, 80), 4, seed=8), dev)
    rot = T.RandomRotate(KEYS, 1.0, (80, 80, 80), [(0, 0)] * 3, "zeros", {k: "nearest" for k in KEYS},
                         grid="device", align_corners=True)
    rot.R = torch.eye(3)
    out = T.spatial_pass(dict(data), KEYS, rotate=rot)
    for key in KEYS:
        assert torch.equal(out[key], data[key]), key
    twice = T.spatial_pass(T.spatial_pass(dict(data), KEYS, flips=7), KEYS, flips=7)
    for key in KEYS:
        assert torch.equal(twice[key].view(torch.int32), data[key].view(torch.int32)), key


def test_flow_rotation_preserves_vector_length(dev):
    """Nearest sampling moves vectors without mixing them: each output vector
    is R times a source vector, so lengths are kept to rounding."""
    v = torch.randn(2, 3, 16, 16, 16, device=dev)
    rot = T.RandomRotate(["fwd_flow"], 1.0, (16, 16, 16), [(0, 360)] * 3, "zeros", {"fwd_flow": "nearest"},
                         grid="device", align_corners=False)
    rot.R = rotation(40.0, 75.0, 130.0)
    out = T.spatial_pass({"fwd_flow": v}, ["fwd_flow"], rotate=rot)["fwd_flow"]
    n_out = out.norm(dim=1)
    inside = n_out > 0
    assert inside.float().mean() > 0.5
    # Every non-zero output length is one of the source lengths.
    src_len = np.sort(v.norm(dim=1).flatten().cpu().numpy())
    got_len = n_out[inside].cpu().numpy()
    idx = np.clip(np.searchsorted(src_len, got_len), 1, len(src_len) - 1)
    nearest = np.minimum(np.abs(src_len[idx] - got_len), np.abs(src_len[idx - 1] - got_len))
    assert nearest.max() < 1e-5


def _vol(src, dst, batch=1, channels=1, mode=0, vector=0):
    return _lib.AugVolume(src, dst, batch, channels, mode, vector)


def test_entry_point_rejects_bad_arguments(dev):
    lib = _lib.lib()
    a = torch.zeros(2, 3, 4, 5, 6, device=dev)
    b = torch.zeros_like(a)
    R = np.eye(3, dtype=np.float32)
    Rp = R.ctypes.data_as(ctypes.c_void_p)
    axes = torch.zeros(6 + 5 + 4, device=dev)
    s = _lib.stream_ptr(dev)

    def call(vols, R=Rp, dims=(4, 5, 6, 4, 5, 6), grid=None, ax=axes, padding=0, flips=0, rounding=(0, 0)):
        arr = (_lib.AugVolume * max(1, len(vols)))(*vols)
        return lib.svs_spatial_augment(arr, len(vols), *dims, R, _lib.ptr(grid), _lib.ptr(ax), padding, 0, flips,
                                       rounding[0], rounding[1], s)

    good = _vol(a.data_ptr(), b.data_ptr(), 2, 3, 0, 1)
    assert call([good]) == 0
    torch.cuda.synchronize()
    assert call([]) == 1
    assert call([good] * 9) == 1
    assert call([_vol(a.data_ptr(), a.data_ptr(), 2, 3)]) == 1                      # in place
    assert "overlaps" in lib.svs_last_error().decode()
    assert call([_vol(a.data_ptr(), b.data_ptr(), 2, 2, 0, 1)]) == 1                # vector, 2 channels
    assert call([good], padding=3) == 1
    assert call([good], flips=8) == 1
    assert call([good], ax=None) == 1                                               # no grid source
    assert call([good], R=ctypes.c_void_p(0), dims=(4, 5, 6, 4, 5, 7)) == 1        # size change, no rotation
    assert call([_vol(a.data_ptr(), b.data_ptr(), 1, 1, 2)]) == 1                   # mode 2
    assert call([good], rounding=(3, 0)) == 1
    assert call([good], rounding=(0, -1)) == 1


def test_python_front_end_rejects_host_tensors_and_bad_layouts(dev):
    rot = T.RandomRotate(["image"], 1.0, (4, 4, 4), [(0, 360)] * 3, "zeros", {"image": "bilinear"})
    rot.R = torch.eye(3)
    with pytest.raises(RuntimeError, match="CUDA tensor"):
        T.spatial_pass({"image": torch.zeros(1, 1, 4, 4, 4)}, ["image"], rotate=rot)
    with pytest.raises(ValueError):
        T.spatial_pass({"image": torch.zeros(1, 4, 4, 4, device=dev)}, ["image"], rotate=rot)
    with pytest.raises(ValueError):
        T.spatial_pass({"fwd_flow": torch.zeros(1, 2, 4, 4, 4, device=dev)}, ["fwd_flow"], flips=1)
    with pytest.raises(ValueError):
        T.RandomRotate(["image"], 1.0, (4, 4, 4), [(0, 1)] * 3, "zeros", {"image": "bicubic"}).mode_of("image")
