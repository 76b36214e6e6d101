"""OpticalFlowProcessor of svs/modules/flow/processor.py:13-133 on the GPU path.

Per patient: the flows between consecutive frames from end-diastole to
end-systole (or back), each optionally median-filtered, stored as one
(T'-1, 3, X, Y, Z) float32 array (``us[j] = u.permute(3, 2, 1, 0)``,
processor.py:108-120).  The dataset, directory and tqdm plumbing of the
reference stay with the caller: ``process_patient`` takes the patient's 4D
image as the (X, Y, Z, T) array that ``patient.img_array()`` returns.

When the configuration has more than one pyramid level, TVL13DOpticalFlow
discards the carried (u, p) (tvl13d.py:135-139), so the pairs are independent
and ``process_patient`` solves all of them in one batched plan and filters
them in one launch; the result is bit-identical to the reference's pair loop
(``batched=False``).  With a single level each pair starts from the previous
pair's (filtered) flow and the loop is kept.
"""
from typing import Any, Optional

import numpy as np
import torch

from . import flowio
from .flow_utils import apply_median_filter, compute_timepoints, median_filter3d
from .solver import Plan
from .tvl13d import TVL13DOpticalFlow

XYZT_TO_TZYX = (3, 2, 1, 0)   # datasets.py:30


def _get(cfg: Any, path: str):
    """cfg.a.b for attribute-style configs (ml_collections) and nested dicts."""
    for key in path.split("."):
        cfg = cfg[key] if isinstance(cfg, dict) else getattr(cfg, key)
    return cfg


class OpticalFlowProcessor:
    def __init__(self, flow_cfg: Any, device: Optional[torch.device] = None, batched: bool = True):
        """``flow_cfg`` is the ``flow:`` section of conf/flow.yaml (a nested dict
        or an attribute-style config).  The solver is built as
        processor.py:45-69 builds it."""
        self.cfg = flow_cfg
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        if self.device.type != "cuda":
            raise RuntimeError("Optical flow computation requires a CUDA device.")
        self.batched = batched
        self.optflow = TVL13DOpticalFlow(
            _get(flow_cfg, "num_scales"), _get(flow_cfg, "max_warps"), _get(flow_cfg, "max_outer_iterations"),
            _get(flow_cfg, "weight_matching"), _get(flow_cfg, "weight_tv"), _get(flow_cfg, "primal_dual.type"),
            tuple(_get(flow_cfg, "primal_dual.params")), _get(flow_cfg, "anisotropic_diff.enabled"),
            tuple(_get(flow_cfg, "anisotropic_diff.params")), _get(flow_cfg, "gaussian_blur.enabled"),
            _get(flow_cfg, "gaussian_blur.sigma"), _get(flow_cfg, "interpolation.type"),
            _get(flow_cfg, "interpolation.boundary"), _get(flow_cfg, "mesh.type"),
            tuple(_get(flow_cfg, "mesh.length")), self.device)
        self.mode = str(_get(flow_cfg, "mode"))
        self.median = bool(_get(flow_cfg, "median_filter.enabled"))
        self.median_ks = int(_get(flow_cfg, "median_filter.kernel"))
        self._plans = {}

    def frames(self, img_xyzt: np.ndarray) -> torch.Tensor:
        """processor.py:103: the (X, Y, Z, T) image as float32 [T, Z, Y, X] on the device."""
        arr = np.ascontiguousarray(np.asarray(img_xyzt).transpose(XYZT_TO_TZYX), dtype=np.float32)
        return torch.from_numpy(arr).to(self.device)

    def _times(self, tdia: int, tsys: int):
        return compute_timepoints(tdia, tsys, None, None, self.mode)[2]

    def process_patient(self, img_xyzt: np.ndarray, tdia: int, tsys: int) -> np.ndarray:
        """processor.py:95-121 for one patient: the (T'-1, 3, X, Y, Z) flows."""
        img = self.frames(img_xyzt)
        times = self._times(tdia, tsys)
        if len(times) < 2:
            # tdia == tsys: no pair, an empty (0, 3, X, Y, Z) array as processor.py:106 allocates.
            _, nz, ny, nx = img.shape
            return np.zeros((0, 3, nx, ny, nz), dtype=np.float32)
        if self.batched and self.optflow.num_scales > 1:
            flows = self._pairs_batched(img, times)
        else:
            flows = self._pairs_loop(img, times)
        return flowio.to_file_layout(flows)

    def _pairs_loop(self, img: torch.Tensor, times) -> torch.Tensor:
        """The reference's loop (processor.py:108-119): u and p carried from pair to pair."""
        _, nz, ny, nx = img.shape
        u = torch.zeros((nz, ny, nx, 3), dtype=torch.float32, device=self.device)
        p = torch.zeros((nz, ny, nx, 3, 3), dtype=torch.float32, device=self.device)
        out = []
        for j in range(len(times) - 1):
            u, p = self.optflow.compute(img[times[j + 1]], img[times[j]], u, p)
            if self.median:
                u = apply_median_filter(u, self.median_ks, self.device)
            out.append(u)
        return torch.stack(out)

    def _pairs_batched(self, img: torch.Tensor, times) -> torch.Tensor:
        """All pairs of the patient as one plan (one fused sweep launch per
        iteration for all of them), then one median-filter launch."""
        shape = tuple(img.shape[1:])
        npairs = len(times) - 1
        key = (shape, npairs)
        plan = self._plans.get(key)
        if plan is None:
            plan = Plan(self.optflow.config, [shape] * npairs, device=self.device)
            self._plans[key] = plan
        flows = torch.empty((npairs,) + shape + (3,), dtype=torch.float32, device=self.device)
        for j in range(npairs):
            plan.bind(j, img[times[j + 1]], img[times[j]], u_out=flows[j])
        plan.run(use_graph=self.optflow.use_graph)
        if self.median:
            flows = median_filter3d(flows, self.median_ks)
        return flows

    def save(self, directory: str, patient: str, us: np.ndarray) -> str:
        """processor.py:120: ``{patient}_{mode}_flow.npy``."""
        import os
        path = os.path.join(directory, flowio.flow_filename(patient, self.mode))
        np.save(path, np.ascontiguousarray(us, dtype=np.float32))
        return path

    def close(self):
        for plan in self._plans.values():
            plan.close()
        self._plans.clear()
