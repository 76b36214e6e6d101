"""OpticalFlowProcessor (processor.py:13-133): the batched patient path equals the
reference's pair loop bit for bit, in both modes, with and without the median
post-filter, and writes the (T'-1, 3, X, Y, Z) layout."""
import copy

import numpy as np
import pytest
import torch

from singleventriclesegmentation_amd import phantom
from singleventriclesegmentation_amd.processor import OpticalFlowProcessor

pytestmark = pytest.mark.gpu

FLOW_CFG = {  # conf/flow.yaml's flow section with a short schedule
    "mode": "forward", "indices": "(-1, -1)", "num_scales": 2, "max_warps": 2, "max_outer_iterations": 5,
    "weight_matching": 1.0, "weight_tv": 0.025,
    "primal_dual": {"type": 2, "params": [0.5, 0.5, 0.25, 1.0]},
    "anisotropic_diff": {"enabled": True, "params": [5.0, 0.5]},
    "gaussian_blur": {"enabled": True, "sigma": 1.0},
    "median_filter": {"enabled": True, "kernel": 3},
    "interpolation": {"type": "CUBIC_HERMITESPLINE", "boundary": "MIRROR"},
    "mesh": {"type": "numDofs", "length": [1, 1, 1]},
}


def patient_xyzt(frames=7, shape=(16, 20, 24)):
    seq = phantom.sequence(shape, frames, seed_base=40)          # [T, Z, Y, X]
    return np.ascontiguousarray(seq.transpose(3, 2, 1, 0))      # (X, Y, Z, T) as patient.img_array()


@pytest.mark.parametrize("mode", ["forward", "backward"])
@pytest.mark.parametrize("median", [True, False])
def test_batched_patient_equals_pair_loop(dev, mode, median):
    cfg = copy.deepcopy(FLOW_CFG)
    cfg["mode"] = mode
    cfg["median_filter"]["enabled"] = median
    img = patient_xyzt()
    a = OpticalFlowProcessor(cfg, dev, batched=True)
    b = OpticalFlowProcessor(cfg, dev, batched=False)
    us_a = a.process_patient(img, tdia=1, tsys=5)
    us_b = b.process_patient(img, tdia=1, tsys=5)
    assert us_a.shape == (4, 3, 24, 20, 16) and us_a.dtype == np.float32
    assert np.array_equal(us_a, us_b)
    # the first pair of the forward chain is (I0 = frame 2, I1 = frame 1) in the reference's order
    frames = a.frames(img)
    t0, t1 = (1, 2) if mode == "forward" else (5, 4)
    u, _ = a.optflow.compute(frames[t1], frames[t0], None, None)
    if median:
        from singleventriclesegmentation_amd.flow_utils import apply_median_filter
        u = apply_median_filter(u, 3, dev)
    assert np.array_equal(us_a[0], u.permute(3, 2, 1, 0).cpu().numpy())
    a.close()


def test_single_level_keeps_the_sequential_loop(dev):
    cfg = copy.deepcopy(FLOW_CFG)
    cfg["num_scales"] = 1
    img = patient_xyzt(frames=5)
    a = OpticalFlowProcessor(cfg, dev, batched=True)
    b = OpticalFlowProcessor(cfg, dev, batched=False)
    assert np.array_equal(a.process_patient(img, 0, 3), b.process_patient(img, 0, 3))


def test_save_layout(dev, tmp_path):
    cfg = copy.deepcopy(FLOW_CFG)
    proc = OpticalFlowProcessor(cfg, dev)
    us = proc.process_patient(patient_xyzt(frames=4), 0, 3)
    path = proc.save(str(tmp_path), "Child_3", us)
    assert path.endswith("Child_3_forward_flow.npy")
    assert np.array_equal(np.load(path, allow_pickle=False), us)
    proc.close()


@pytest.mark.parametrize("batched", [True, False])
def test_equal_end_phases_give_no_pairs(dev, batched):
    """tdia == tsys: compute_timepoints yields one frame, so the reference's loop
    runs no pair and saves its empty (0, 3, X, Y, Z) array."""
    proc = OpticalFlowProcessor(copy.deepcopy(FLOW_CFG), dev, batched=batched)
    us = proc.process_patient(patient_xyzt(frames=4), 2, 2)
    assert us.shape == (0, 3, 24, 20, 16) and us.dtype == np.float32
    proc.close()
