"""Flow-consistent spatial augmentation of svs/modules/transforms on the GPU.

RandomRotate (transforms/spatial.py:13-128), RandomFlip (:131-168), the
rotation helpers rotx/roty/rotz (transforms/functional.py:129-153) and
Compose (transforms/base.py:66-80), with the reference's constructor
arguments, random draws (np.random, same order) and layouts:

    images (T, C, Z, Y, X)    masks (C, Z, Y, X)    flows (T, 3, Z, Y, X)

held as float32 CUDA tensors; flow keys hold the (x, y, z) components.  Every
key of a transform goes through one launch of the fused gather kernel
(svs_spatial_augment), and Compose merges a RandomRotate with the RandomFlips
that follow it (same keys) into one launch: out(o) = rotated(flip(o)) with
the flipped flow components negated.  Flips are index maps and sign changes,
so the merged launch equals the sequential application bit for bit.

Rounding.  The reference forms identity_grid @ R with torch.matmul and
rotates flow vectors with torch.einsum, on the CPU, so their 3-term dot
products round as the host's BLAS does: MKL accumulates
fma(a2,b2,fma(a1,b1,fma(a0,b0,+0))) on Intel hosts (einsum at every size, the
grid from ~110K points up) and plain (a0*b0 + a1*b1) + a2*b2 on AMD hosts.
``grid="host"`` (the default) makes the reference's own matmul call and
uploads the product; ``grid="device"`` forms it in the kernel.  Either way the
rounding of the products the kernel forms is the one this host's torch uses:
it is read off once per process (and grid size) by comparing torch's own
results on probe data with the candidate roundings (``host_rounding``).

ElasticDeformation (elasticdeform, a third-party C extension) and the
intensity transforms (images only) are not part of this module.
"""
import ctypes
import warnings
from typing import Any, Dict, Iterable, Optional, Tuple

import numpy as np
import torch

from . import _lib

# svs/utils/constants.py:1-21
IMAGE_KEY = "image"
MI_KEY, MF_KEY, MES_KEY, MED_KEY = "mi", "mf", "mes", "med"
FWD_FLOW_KEY, BWD_FLOW_KEY = "fwd_flow", "bwd_flow"
METADATA_SUBFIX = "_meta"
MASKS_KEYS = {MI_KEY, MF_KEY, MED_KEY, MES_KEY}
FLOWS_KEYS = {FWD_FLOW_KEY, BWD_FLOW_KEY}

PADDING = {"zeros": 0, "border": 1, "reflection": 2}
MODES = {"bilinear": 0, "nearest": 1}
FLIP_BIT = {4: 1, 3: 2, 2: 4}           # axis X, Y, Z -> kernel flip bit
MAX_KEYS_PER_LAUNCH = 8


def rotx(deg):
    rad = np.deg2rad(deg)
    c, s = np.cos(rad), np.sin(rad)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def roty(deg):
    rad = np.deg2rad(deg)
    c, s = np.cos(rad), np.sin(rad)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def rotz(deg):
    rad = np.deg2rad(deg)
    c, s = np.cos(rad), np.sin(rad)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


ROUNDINGS = {0: "fma chain from +0", 1: "plain", 2: "plain from +0"}
_PROBE_ANGLES = (23.0, -41.0, 200.0)
_rounding_cache: Dict[Any, Tuple[int, bool]] = {}


def _dot3_roundings(terms) -> Dict[int, np.ndarray]:
    """a0*b0 + a1*b1 + a2*b2 (float32 operands) in each candidate rounding.
    The FMA chain is evaluated in float64 (products exact), rounding to
    float32 after each step."""
    prods = [np.float32(1) * a * b for a, b in terms]
    plain = (prods[0] + prods[1]) + prods[2]
    plain0 = ((np.float32(0) + prods[0]) + prods[1]) + prods[2]
    acc = np.zeros(plain.shape, np.float64)
    for a, b in terms:
        acc = (np.asarray(a, np.float64) * np.asarray(b, np.float64) + acc).astype(np.float32).astype(np.float64)
    return {0: acc.astype(np.float32), 1: plain.astype(np.float32), 2: plain0.astype(np.float32)}


def _classify(ref: np.ndarray, cands: Dict[int, np.ndarray], what: str) -> Tuple[int, bool]:
    miss = {m: int(np.count_nonzero(ref.view(np.int32) != np.ascontiguousarray(c).view(np.int32)))
            for m, c in cands.items()}
    best = min(miss, key=miss.get)
    # The float64 evaluation of the FMA chain can round twice in rare cases.
    exact = miss[best] <= max(1, ref.size // 1_000_000)
    if not exact:
        warnings.warn(f"{what}: no candidate rounding reproduces this host's BLAS ({miss}); using "
                      f"'{ROUNDINGS[best]}', results may differ from the reference in the last bit")
    return best, exact


def _probe_rotation() -> torch.Tensor:
    ax, ay, az = _PROBE_ANGLES
    return torch.from_numpy(rotz(az) @ roty(ay) @ rotx(ax)).float()


def host_rounding(kind: str, spatial_shape=None) -> Tuple[int, bool]:
    """(rounding, exact) of this host's torch.einsum('ij,tjxyz->tixyz')
    (kind "einsum") or of torch.matmul(identity_grid, R) at ``spatial_shape``
    (kind "matmul"), from torch's own results on probe data."""
    key = (kind, tuple(spatial_shape) if spatial_shape is not None else None)
    if key in _rounding_cache:
        return _rounding_cache[key]
    R = _probe_rotation()
    Rn = R.numpy()
    if kind == "einsum":
        g = torch.Generator().manual_seed(0)
        f = torch.randn(2, 3, 12, 12, 12, generator=g)
        signs = (torch.randint(0, 2, f.shape, generator=g) * 2 - 1).float()
        f[:, :, :4] = 0.0 * signs[:, :, :4]                  # signed zeros
        ref = torch.einsum("ij,tjxyz->tixyz", R, f).numpy()
        fn = f.numpy()
        per = [_dot3_roundings([(Rn[i, 0], fn[:, 0]), (Rn[i, 1], fn[:, 1]), (Rn[i, 2], fn[:, 2])])
               for i in range(3)]
        cands = {m: np.stack([per[i][m] for i in range(3)], 1) for m in per[0]}
    elif kind == "matmul":
        ig = RandomRotate.identity_grid_of(spatial_shape)
        ref = torch.matmul(ig, R).float().numpy()
        gn = ig.numpy()
        per = [_dot3_roundings([(gn[..., 0], Rn[0, i]), (gn[..., 1], Rn[1, i]), (gn[..., 2], Rn[2, i])])
               for i in range(3)]
        cands = {m: np.stack([per[i][m] for i in range(3)], -1) for m in per[0]}
    else:
        raise ValueError(kind)
    _rounding_cache[key] = _classify(np.ascontiguousarray(ref), cands, f"torch.{kind}")
    return _rounding_cache[key]


def _as_batched(x: torch.Tensor, key: str) -> torch.Tensor:
    """The 5-D view grid_sample / flip see (spatial.py:72-75: masks gain a batch dim)."""
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise RuntimeError(f"spatial augmentation of '{key}' needs a CUDA tensor (there is no CPU path)")
    if x.dtype != torch.float32:
        raise TypeError(f"'{key}': float32 expected (EnsureFloat), got {x.dtype}")
    if x.ndim == 4 and key in MASKS_KEYS:
        x = x.unsqueeze(0)
    if x.ndim != 5:
        raise ValueError(f"'{key}': expected (N, C, Z, Y, X), got shape {tuple(x.shape)}")
    if key in FLOWS_KEYS and x.shape[1] != 3:
        raise ValueError(f"flow '{key}' must have 3 components, got {x.shape[1]}")
    return x.contiguous()


def spatial_pass(data: Dict[str, Any], keys: Iterable[str], rotate: Optional["RandomRotate"] = None,
                 flips: int = 0) -> Dict[str, Any]:
    """Applies rotate.R (when ``rotate`` is given) and then the flips of the
    bit mask ``flips`` (X = 1, Y = 2, Z = 4) to data[key] for every key."""
    keys = list(keys)
    if rotate is None and flips == 0:
        return data
    prepared = [(k, _as_batched(data[k], k)) for k in keys]
    groups: Dict[tuple, list] = {}
    for k, x in prepared:
        groups.setdefault((x.device, tuple(x.shape[2:])), []).append((k, x))
    for (device, in_shape), members in groups.items():
        for i in range(0, len(members), MAX_KEYS_PER_LAUNCH):
            _launch(data, members[i:i + MAX_KEYS_PER_LAUNCH], device, in_shape, rotate, flips)
    return data


def _launch(data, members, device, in_shape, rotate, flips):
    IZ, IY, IX = in_shape
    if rotate is not None:
        OZ, OY, OX = (int(s) for s in rotate.deform_shape)
        R = np.ascontiguousarray(rotate.R.detach().cpu().numpy(), dtype=np.float32)
        grid, axes = rotate.device_inputs(device)
        Rp = R.ctypes.data_as(ctypes.c_void_p)
        padding, align = PADDING[rotate.boundary], int(rotate.align_corners)
        grid_rounding, vector_rounding = rotate.grid_rounding, rotate.vector_rounding
    else:
        OZ, OY, OX = IZ, IY, IX
        R, grid, axes, Rp = None, None, None, ctypes.c_void_p(0)
        padding, align, grid_rounding, vector_rounding = 0, 0, 0, 0
    vols = (_lib.AugVolume * len(members))()
    outs = []
    for j, (key, x) in enumerate(members):
        out = torch.empty((x.shape[0], x.shape[1], OZ, OY, OX), dtype=torch.float32, device=device)
        mode = MODES[rotate.mode_of(key)] if rotate is not None else 0
        vols[j] = _lib.AugVolume(x.data_ptr(), out.data_ptr(), int(x.shape[0]), int(x.shape[1]), mode,
                                 int(key in FLOWS_KEYS))
        outs.append((key, out))
    with torch.cuda.device(device):
        _lib.check(_lib.lib().svs_spatial_augment(
            vols, len(members), IZ, IY, IX, OZ, OY, OX, Rp, _lib.ptr(grid), _lib.ptr(axes), padding, align, flips,
            grid_rounding, vector_rounding, _lib.stream_ptr(device)))
    for key, out in outs:
        data[key] = out.squeeze(0) if data[key].ndim == 4 else out


class BaseTransform:
    """transforms/base.py:15-64: a transform applied to data[key] for each key."""

    def __init__(self, keys: Iterable[str]):
        self.keys = keys

    def apply_transform(self, data: Dict[str, Any]) -> Dict[str, Any]:
        for key in self.keys:
            metadata = data.get(key + METADATA_SUBFIX)
            data[key] = self._transform_impl(data[key], key, metadata)
        return data

    def _transform_impl(self, x, key, metadata=None):
        raise NotImplementedError

    def class_name(self):
        return str(type(self).__name__)


class RandomRotate(BaseTransform):
    def __init__(self, keys: Iterable[str], p: float, spatial_shape: Iterable[int],
                 rot_ranges: Iterable[Iterable[float]], boundary: str, modes: Dict[str, str], grid: str = "host",
                 **kwargs):
        """spatial.py:14-51.  ``kwargs`` are grid_sample's: align_corners
        (default False, grid_sample's default).  ``grid`` selects where
        identity_grid @ R is formed ("host": torch.matmul on the CPU as the
        reference; "device": in the kernel)."""
        super().__init__(keys)
        assert len(spatial_shape) == 3, "Spatial shape must have exactly 3 elements."
        assert len(rot_ranges) == 3, "Rotation ranges must have exactly 3 elements."
        unknown = set(kwargs) - {"align_corners"}
        if unknown:
            raise TypeError(f"unsupported grid_sample arguments: {sorted(unknown)}")
        if boundary not in PADDING:
            raise ValueError(f"boundary must be one of {sorted(PADDING)}, got {boundary!r}")
        if grid not in ("host", "device"):
            raise ValueError("grid must be 'host' or 'device'")
        self.p = p
        self.deform_shape = tuple(int(s) for s in spatial_shape)
        self.range_x, self.range_y, self.range_z = rot_ranges
        self.boundary = boundary
        self.modes = modes
        self.kwargs = kwargs
        self.align_corners = bool(kwargs.get("align_corners") or False)
        self.grid = grid
        self.identity_grid = self.get_identity_grid()
        self._axes = {}
        self.R = None
        self.vector_rounding, _ = host_rounding("einsum")
        self.grid_rounding = host_rounding("matmul", self.deform_shape)[0] if grid == "device" else 0

    def __call__(self, data: Dict[str, Any]) -> Dict[str, Any]:
        R = self.draw()
        if R is not None:
            self.R = R
            data = spatial_pass(data, self.keys, rotate=self)
        return data

    def draw(self) -> Optional[torch.Tensor]:
        """The random decision of __call__ (spatial.py:53-56): R or None."""
        if np.random.uniform() < self.p:
            return self.sample_rotation_matrix()
        return None

    def _transform_impl(self, x, key, metadata=None):
        return spatial_pass({key: x}, [key], rotate=self)[key]

    def mode_of(self, key: str) -> str:
        mode = self.modes[key]
        if mode not in MODES:
            raise ValueError(f"grid_sample mode {mode!r} is not available for 5-D inputs")
        return mode

    def get_identity_grid(self) -> torch.Tensor:
        """spatial.py:93-105: (Z, Y, X, 3) with components (x, y, z), on the CPU."""
        return self.identity_grid_of(self.deform_shape)

    @staticmethod
    def identity_grid_of(spatial_shape) -> torch.Tensor:
        space = [torch.linspace(-1, 1, int(s)) for s in tuple(spatial_shape)[::-1]]
        grid = torch.stack(torch.meshgrid(space, indexing="ij"), -1)
        return grid.permute(2, 1, 0, 3)

    def sample_rotation_matrix(self) -> torch.Tensor:
        """spatial.py:107-112: R = Rz @ Ry @ Rx in float64, then float32."""
        Rx = rotx(np.random.uniform(self.range_x[0], self.range_x[1]))
        Ry = roty(np.random.uniform(self.range_y[0], self.range_y[1]))
        Rz = rotz(np.random.uniform(self.range_z[0], self.range_z[1]))
        return torch.from_numpy(Rz @ Ry @ Rx).float()

    @property
    def rotated_grid(self) -> torch.Tensor:
        """identity_grid @ R as the reference holds it (spatial.py:56), on the CPU."""
        return torch.matmul(self.identity_grid, self.R).float()

    def device_inputs(self, device):
        """(grid, axes) for the kernel: the uploaded host product, or the
        identity grid's axes (lx ++ ly ++ lz) for the in-kernel product."""
        if self.grid == "host":
            return self.rotated_grid.contiguous().to(device), None
        key = str(device)
        if key not in self._axes:
            lz, ly, lx = (torch.linspace(-1, 1, s) for s in self.deform_shape)
            self._axes[key] = torch.cat([lx, ly, lz]).to(device)
        return None, self._axes[key]


class RandomFlip(BaseTransform):
    def __init__(self, keys: Iterable[str], p: float, axis: int):
        """spatial.py:132-145.  axis 2, 3, 4 flips depth, vertical, horizontal."""
        super().__init__(keys)
        assert axis in {2, 3, 4}, "Axis must be 2, 3 or 4"
        self.p = p
        self.axis = axis

    def __call__(self, data: Dict[str, Any]) -> Dict[str, Any]:
        if self.draw():
            data = spatial_pass(data, self.keys, flips=FLIP_BIT[self.axis])
        return data

    def draw(self) -> bool:
        return bool(np.random.uniform() < self.p)

    def _transform_impl(self, x, key, metadata=None):
        return spatial_pass({key: x}, [key], flips=FLIP_BIT[self.axis])[key]


class Compose:
    """transforms/base.py:66-80, with each RandomRotate and the RandomFlips
    that follow it on the same keys (or a run of RandomFlips) applied as one
    launch.  Random draws happen in the reference's order; other transforms
    are called as they are."""

    def __init__(self, transforms: Iterable[Any]):
        self.transforms = list(transforms)

    def __call__(self, data: Dict[str, Any]) -> Dict[str, Any]:
        ts, i = self.transforms, 0
        while i < len(ts):
            t = ts[i]
            if not isinstance(t, (RandomRotate, RandomFlip)):
                data = t(data)
                i += 1
                continue
            j = i + 1
            while j < len(ts) and isinstance(ts[j], RandomFlip) and list(ts[j].keys) == list(t.keys):
                j += 1
            rotate, flips = None, 0
            for g in ts[i:j]:
                if isinstance(g, RandomRotate):
                    R = g.draw()
                    if R is not None:
                        g.R = R
                        rotate = g
                elif g.draw():
                    flips ^= FLIP_BIT[g.axis]
            data = spatial_pass(data, t.keys, rotate=rotate, flips=flips)
            i = j
        return data

    def __repr__(self) -> str:
        return "Compose(" + "".join(f"\n    {t!r}" for t in self.transforms) + "\n)"
