"""The C-ABI library loads without a GPU and exports every entry point that
include/svsflow.h declares (no compute calls here)."""
import ctypes
import os
import re

from conftest import REPO

HEADER = os.path.join(REPO, "include", "svsflow.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(svs_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "svs_tvl1_solve" in syms and "svs_warp3d" in syms and len(syms) >= 20


def test_library_exports_every_declared_symbol():
    from singleventriclesegmentation_amd import _lib
    lib = _lib.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_header():
    from singleventriclesegmentation_amd import _lib
    assert set(declared_symbols()) <= set(_lib.exported_symbols())


def test_workspace_size_and_validation():
    from singleventriclesegmentation_amd import _lib
    from singleventriclesegmentation_amd.solver import SolverConfig
    p, _w = SolverConfig().native()
    dims = (ctypes.c_int * 3)(80, 80, 80)
    n = _lib.lib().svs_tvl1_workspace_bytes(ctypes.byref(p), 1, dims)
    assert 60e6 < n < 120e6
    dims = (ctypes.c_int * 3)(8, 8, 8)  # level 3 would be 1^3
    assert _lib.lib().svs_tvl1_workspace_bytes(ctypes.byref(p), 1, dims) == 0
    assert b"pyramid level" in _lib.lib().svs_last_error()


def test_plan_rejects_instances_beyond_32bit_offsets():
    """The fused sweep addresses an instance's 9 dual fields with 32-bit byte
    offsets: 36 N must stay below 2^32 (N = 119,304,647 voxels)."""
    from singleventriclesegmentation_amd import _lib
    from singleventriclesegmentation_amd.solver import SolverConfig
    p, _w = SolverConfig().native()
    ok = (ctypes.c_int * 3)(488, 488, 500)        # 119,072,000 voxels
    assert _lib.lib().svs_tvl1_workspace_bytes(ctypes.byref(p), 1, ok) > 0
    big = (ctypes.c_int * 3)(490, 490, 500)       # 120,050,000 voxels
    assert _lib.lib().svs_tvl1_workspace_bytes(ctypes.byref(p), 1, big) == 0
    assert b"exceeds" in _lib.lib().svs_last_error()
    odd = (ctypes.c_int * 6)(16, 16, 16, 0, 16, 16)  # a zero-size second instance
    assert _lib.lib().svs_tvl1_workspace_bytes(ctypes.byref(p), 2, odd) == 0


def test_one_dimensional_launch_limits_are_rejected_up_front():
    """Element counts beyond a 1-D launch's 2^32 - 256 work-items are refused
    by argument validation, before any HIP call (so this runs without a
    GPU; the pointers are never dereferenced)."""
    from singleventriclesegmentation_amd import _lib
    lib = _lib.lib()
    fake, fake_out = 256, 512
    # median filter: 2 x 1024^3 x 3 elements (the volume itself is legal)
    assert lib.svs_median3d(fake, 2, 1024, 1024, 1024 - 1, 3, 3, fake_out, None) != 0
    assert "work-items" in lib.svs_last_error().decode()
    # CNN warp: 3 x 1290^3 work-items
    assert lib.svs_warp3d_cnn(fake, fake, 3, 1, 1290, 1290, 1290, 1.0, 1.0, 1.0, 2, 1, fake_out, None) != 0
    assert "work-items" in lib.svs_last_error().decode()
    assert lib.svs_prox_dual(fake, 0.5, 0.025, (1 << 32), fake_out, None) != 0
    assert "work-items" in lib.svs_last_error().decode()
