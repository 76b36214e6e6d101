"""Synthetic cardiac-like test volumes (FP32, [Z,Y,X], x fastest).

A blob phantom: background 0.2, an ellipsoidal "ventricle" at 0.8 surrounded
by a shell at 0.5, Gaussian-smoothed, plus Gaussian noise, clipped to [0, 1].
A known smooth displacement (a sum of Gaussian radial basis functions) warps
the clean frame to make the second frame of a pair, so the solver's flow can
be compared with the ground truth.  Semi-axes scale with the volume
(18/22/20 voxels at 80^3 for x/y/z).
"""
import numpy as np
from scipy import ndimage

BASE_AXES = (18.0, 22.0, 20.0)  # x, y, z at 80^3
SHELL = 4.0


def _grid(shape):
    NZ, NY, NX = shape
    z, y, x = np.meshgrid(np.arange(NZ, dtype=np.float64), np.arange(NY, dtype=np.float64),
                          np.arange(NX, dtype=np.float64), indexing="ij")
    return z, y, x


def _axes(shape, scale=1.0):
    NZ, NY, NX = shape
    ax, ay, az = BASE_AXES
    return ax * NX / 80.0 * scale, ay * NY / 80.0 * scale, az * NZ / 80.0 * scale


def clean_frame(shape, scale=1.0, smooth=1.5):
    """Noise-free phantom; ``scale`` multiplies the ventricle semi-axes."""
    z, y, x = _grid(shape)
    cz, cy, cx = ((n - 1) / 2.0 for n in shape)
    ax, ay, az = _axes(shape, scale)
    r_in = ((x - cx) / ax) ** 2 + ((y - cy) / ay) ** 2 + ((z - cz) / az) ** 2
    r_out = ((x - cx) / (ax + SHELL)) ** 2 + ((y - cy) / (ay + SHELL)) ** 2 + ((z - cz) / (az + SHELL)) ** 2
    vol = np.full(shape, 0.2, dtype=np.float64)
    vol[r_out <= 1.0] = 0.5
    vol[r_in <= 1.0] = 0.8
    if smooth > 0:
        vol = ndimage.gaussian_filter(vol, sigma=smooth, mode="nearest")
    return vol.astype(np.float32)


def ventricle_mask(shape, scale=1.0):
    """uint8 interior of the ventricle ellipsoid."""
    z, y, x = _grid(shape)
    cz, cy, cx = ((n - 1) / 2.0 for n in shape)
    ax, ay, az = _axes(shape, scale)
    r_in = ((x - cx) / ax) ** 2 + ((y - cy) / ay) ** 2 + ((z - cz) / az) ** 2
    return (r_in <= 1.0).astype(np.uint8)


def add_noise(vol, seed, sigma=0.02):
    rng = np.random.default_rng(seed)
    return np.clip(vol + rng.normal(0.0, sigma, size=vol.shape), 0.0, 1.0).astype(np.float32)


def displacement(shape, seed=2, peak=3.0, rbf_sigma=10.0, n_rbf=3, spread=20.0):
    """Smooth field d [Z,Y,X,3] (components x, y, z) with max |d| = ``peak``."""
    rng = np.random.default_rng(seed)
    z, y, x = _grid(shape)
    c = np.array([(n - 1) / 2.0 for n in shape])  # z, y, x
    sp = spread * np.array(shape) / 80.0
    d = np.zeros(shape + (3,), dtype=np.float64)
    for _ in range(n_rbf):
        ck = c + rng.uniform(-1.0, 1.0, size=3) * sp
        direction = rng.normal(size=3)
        direction /= np.linalg.norm(direction)
        w = np.exp(-((z - ck[0]) ** 2 + (y - ck[1]) ** 2 + (x - ck[2]) ** 2) / (2.0 * rbf_sigma ** 2))
        d += w[..., None] * direction[None, None, None, :]
    mag = np.sqrt((d ** 2).sum(-1)).max()
    return (d * (peak / mag)).astype(np.float32)


def warp_scipy(vol, d):
    """vol(x + d(x)) with cubic splines and half-sample mirror boundaries."""
    z, y, x = _grid(vol.shape)
    coords = np.stack([z + d[..., 2], y + d[..., 1], x + d[..., 0]])
    return ndimage.map_coordinates(vol.astype(np.float64), coords, order=3, mode="reflect").astype(np.float32)


def frame_pair(shape=(80, 80, 80), seed0=0, seed1=1, disp_seed=2, peak=3.0, warp=None):
    """(frame0, frame1, d): frame1(x) = frame0_clean(x + d(x)) + noise.

    Solve with compute(I0=frame1, I1=frame0) so that u ~ d."""
    clean = clean_frame(shape)
    d = displacement(shape, seed=disp_seed, peak=peak)
    moved = (warp or warp_scipy)(clean, d)
    return add_noise(clean, seed0), add_noise(moved, seed1), d


def sequence(shape=(80, 80, 80), frames=35, seed_base=100):
    """A periodic contraction: semi-axes x (1 - 0.25 sin^2(pi t / frames)), fresh noise per frame."""
    out = np.empty((frames,) + tuple(shape), dtype=np.float32)
    for t in range(frames):
        s = 1.0 - 0.25 * np.sin(np.pi * t / frames) ** 2
        out[t] = add_noise(clean_frame(shape, scale=s), seed_base + t)
    return out


def sequence_masks(shape=(80, 80, 80), frames=35, which=(0, 12)):
    return {t: ventricle_mask(shape, scale=1.0 - 0.25 * np.sin(np.pi * t / frames) ** 2) for t in which}


UNEVEN_SIZES = ((48, 48, 48), (40, 64, 64), (80, 80, 80), (64, 96, 96), (48, 128, 128))  # (Z, Y, X)


def uneven_shapes(count=96, seed=3):
    rng = np.random.default_rng(seed)
    return [UNEVEN_SIZES[i] for i in rng.integers(0, len(UNEVEN_SIZES), size=count)]
