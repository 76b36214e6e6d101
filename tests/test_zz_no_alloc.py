"""svs_mask_hausdorff2_noalloc (fixed device scratch, no allocation) against
svs_mask_hausdorff2 (caller's workspace), through the C ABI: two calls in a
row, the second on a smaller volume, then a volume beyond its limit."""
import ctypes

import numpy as np
import pytest
import torch
from scipy import ndimage

from singleventriclesegmentation_amd import _lib

pytestmark = pytest.mark.gpu


def _masks(shape, seed, dev):
    r = np.random.default_rng(seed)
    a = ndimage.binary_dilation(r.random(shape) > 0.97, iterations=2).astype(np.float32)
    b = ndimage.binary_dilation(r.random(shape) > 0.97, iterations=2).astype(np.float32)
    return torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)


def _with_workspace(a, b):
    L = _lib.lib()
    nbytes = L.svs_mask_hausdorff_workspace_bytes(*a.shape)
    ws = torch.empty((nbytes + 7) // 8, dtype=torch.int64, device=a.device)
    out = torch.empty(2, dtype=torch.int64, device=a.device)
    _lib.check(L.svs_mask_hausdorff2(_lib.ptr(a), _lib.ptr(b), *a.shape, _lib.ptr(ws), ws.numel() * 8,
                                     _lib.ptr(out), _lib.stream_ptr()))
    return out.cpu().tolist(), ws[:2].cpu().tolist()


def _no_alloc(a, b, out, counts):
    L = _lib.lib()
    _lib.check(L.svs_mask_hausdorff2_noalloc(_lib.ptr(a), _lib.ptr(b), *a.shape, _lib.ptr(out),
                                             _lib.ptr(counts), _lib.stream_ptr()))
    return out.cpu().tolist(), counts.cpu().tolist()


def test_no_alloc_matches_workspace_entry_point(dev):
    out = torch.full((2,), -1, dtype=torch.int64, device=dev)
    counts = torch.full((2,), -1, dtype=torch.int64, device=dev)
    for shape, seed in (((48, 64, 56), 0), ((20, 24, 28), 1)):   # second call smaller
        a, b = _masks(shape, seed, dev)
        want_d, want_e = _with_workspace(a, b)
        got_d, got_e = _no_alloc(a, b, out, counts)
        print(f"{shape}: distances^2 {got_d} (workspace entry point {want_d}), edge voxels {got_e}")
        assert got_d == want_d and got_e == want_e
        assert min(want_e) > 0
    # edge counts are optional
    a, b = _masks((16, 18, 20), 2, dev)
    L = _lib.lib()
    _lib.check(L.svs_mask_hausdorff2_noalloc(_lib.ptr(a), _lib.ptr(b), *a.shape, _lib.ptr(out), None,
                                             _lib.stream_ptr()))
    assert out.cpu().tolist() == _with_workspace(a, b)[0]


def test_no_alloc_rejects_volumes_beyond_its_scratch(dev):
    a = torch.zeros((129, 128, 128), device=dev)   # 2^21 + 2^14 voxels
    out = torch.zeros(2, dtype=torch.int64, device=dev)
    L = _lib.lib()
    rc = L.svs_mask_hausdorff2_noalloc(_lib.ptr(a), _lib.ptr(a), *a.shape, _lib.ptr(out), None, _lib.stream_ptr())
    assert rc != 0
    assert "beyond" in L.svs_last_error().decode()
    torch.cuda.synchronize()
