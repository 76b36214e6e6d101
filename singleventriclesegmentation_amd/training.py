"""Mask-propagation training of svs/models/lit_unet.py on the GPU path.

    PropagationLoss     svs/modules/loss.py:8-64
    PropagationModule   svs/models/lit_unet.py:13-194 (forward, propagation_step,
                        forward_and_loss, optimiser and schedule, Dice and
                        Hausdorff per direction)
    collate_fn          svs/modules/datasets.py:352-409 (NNDataset.collate_fn)
    fit                 scripts/train.py:12-39 with the Trainer behaviour it
                        asks for: a sanity pass over the whole validation set,
                        per-epoch training and validation, epoch means
                        weighted by batch size, the cosine schedule stepped
                        per epoch, the best ``loss_val/t`` checkpoint kept
                        and early stopping on it (not before ``min_epochs``).

Masks are propagated with the drop-in ``Warp`` (WarpingCNN3D on the HIP
path): as in the reference, its output carries no gradient, so each U-Net
call learns from its own inputs.  Dice and Hausdorff come from the GPU mask
metrics (monai's compute_dice / compute_hausdorff_distance on one channel).
Lightning, torchmetrics and TensorBoard are not used; the history is a list
of dicts and the checkpoint a ``torch.save`` of the state dicts.

``SyntheticPropagationData`` builds samples in the layout the reference's
transforms produce (images (T,1,Z,Y,X), masks (1,Z,Y,X), flows
(T',Z,Y,X,3)) from phantom sequences, with flows computed by this package's
TV-L1 solver and training samples augmented by ``transforms`` on the GPU.
"""
import math
import random
from collections import defaultdict
from typing import Any, Callable, Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np
import torch
import torch.nn.functional as F
from torch import nn

from . import metrics as gpu_metrics
from .unet import UNet3D
from .warping import Warp

# svs/utils/constants.py:1-28
IMAGE_KEY, MI_KEY, MF_KEY, TI_KEY, TF_KEY = "image", "mi", "mf", "ti", "tf"
FWD_FLOW_KEY, BWD_FLOW_KEY = "fwd_flow", "bwd_flow"
FWD_TS_KEY, BWD_TS_KEY, OFFSET_KEY = "fwd_times", "bwd_times", "offset"
TOTAL_LOSS_KEY, SUPERVISED_LOSS_KEY, UNSUPERVISED_LOSS_KEY, PENALIZATION_LOSS_KEY = "t", "s", "u", "p"


class PropagationLoss(nn.Module):
    """loss.py:8-64: supervised MSE at the two labelled frames, the
    forward/backward agreement on the interior frames and a penalty on the
    raw U-Net outputs, the last two divided by each sample's chain length."""

    def __init__(self, lambda_u: float, gamma_p: float):
        super().__init__()
        self.lambda_u = lambda_u
        self.gamma_p = gamma_p
        self.mse_fn = nn.MSELoss(reduction="sum")

    def forward(self, y: Tuple[torch.Tensor, ...], offsets: torch.Tensor) -> Dict[str, torch.Tensor]:
        mts, mtts, mhs, mhhs = y
        bs = mts.shape[0]
        bi = torch.arange(bs)
        sup_1 = self.mse_fn(mtts[bi, offsets[bi]], mts[bi, 0])
        sup_2 = self.mse_fn(mts[bi, -offsets[bi] - 1], mtts[bi, -1])
        kl = mts.shape[1] - offsets
        n_tilde, n_hat = mts.shape[1], mhs.shape[1]
        unsup = torch.as_tensor(0.0, device=mts.device)
        pen = torch.as_tensor(0.0, device=mts.device)
        for b in range(bs):
            o = int(offsets[b])
            unsup = unsup + self.mse_fn(mts[b, 1:n_tilde - o - 1], mtts[b, 1 + o:-1]) / kl[b]
            mh, mhh = mhs[b, 0:n_hat - o], mhhs[b, 0:n_hat - o]
            pen = pen + (mh.pow(2).sum() + mhh.pow(2).sum()) / kl[b]
        sup = (sup_1 + sup_2) / bs
        unsup = unsup * (self.lambda_u / bs)
        pen = pen * (self.gamma_p / bs)
        return {TOTAL_LOSS_KEY: sup + unsup + pen, SUPERVISED_LOSS_KEY: sup, UNSUPERVISED_LOSS_KEY: unsup,
                PENALIZATION_LOSS_KEY: pen}


def collate_fn(batch: List[Dict[str, Any]]) -> Dict[str, Any]:
    """datasets.py:352-409: flows padded at the end of their time axis to the
    longest in the batch (``offset`` = the padding), tensors stacked, and the
    frame indices of the forward (ti -> tf) and backward (tf -> ti) chains,
    clamped at the sample's own end."""
    max_t = max(sample[FWD_FLOW_KEY].shape[0] for sample in batch)
    out: Dict[str, Any] = defaultdict(list)
    for sample in batch:
        pad_t = None
        for key, value in sample.items():
            if key in (FWD_FLOW_KEY, BWD_FLOW_KEY):
                pad_t = max_t - value.shape[0]
                out[key].append(F.pad(value, (0, 0, 0, 0, 0, 0, 0, 0, 0, pad_t)))
            else:
                out[key].append(value)
        if pad_t is not None:
            out[OFFSET_KEY].append(pad_t)
    for key in list(out.keys()):
        if isinstance(out[key][0], torch.Tensor):
            out[key] = torch.stack(out[key], dim=0)
    ti, tf = out[TI_KEY], out[TF_KEY]
    gap = int((tf - ti).max().item())
    fwd, bwd = [ti], [tf]
    for _ in range(gap):
        nxt = fwd[-1] + 1
        fwd.append(torch.where(nxt > tf, tf, nxt))
        prv = bwd[-1] - 1
        bwd.append(torch.where(prv < ti, ti, prv))
    out[FWD_TS_KEY] = torch.stack(fwd, dim=1)
    out[BWD_TS_KEY] = torch.stack(bwd, dim=1)
    out[OFFSET_KEY] = torch.tensor(out[OFFSET_KEY]).to(torch.long)
    return dict(out)


def to_device(batch: Dict[str, Any], device) -> Dict[str, Any]:
    """LitNNDataset.transfer_batch_to_device (datasets.py:443-447)."""
    for k in (IMAGE_KEY, MI_KEY, MF_KEY, FWD_FLOW_KEY, BWD_FLOW_KEY):
        batch[k] = batch[k].to(device)
    return batch


class _MaskMetric:
    """Dice / Hausdorff of svs/modules/metrics.py: predictions and targets
    collected over an epoch, then the per-sample value averaged."""

    def __init__(self):
        self.pairs: List[Tuple[torch.Tensor, torch.Tensor]] = []

    def update(self, pred: torch.Tensor, target: torch.Tensor):
        self.pairs.append((pred, target))

    def compute(self) -> Dict[str, float]:
        dice, hsdf = [], []
        for pred, target in self.pairs:
            for b in range(pred.shape[0]):
                dice.append(gpu_metrics.compute_dice(pred[b, 0], target[b, 0]))
                hsdf.append(gpu_metrics.compute_hausdorff_distance(pred[b, 0], target[b, 0]))
        mean = lambda v: float(np.mean(np.asarray(v, np.float64))) if v else float("nan")  # noqa: E731
        return {"dice": mean(dice), "hsdf": mean(hsdf)}

    def reset(self):
        self.pairs = []


class PropagationModule(nn.Module):
    """lit_unet.py:13-194 without Lightning: the U-Net, the warp, the loss,
    the optimiser, and the metric state of each stage and direction."""

    def __init__(self, epochs: int, lambda_u: float, gamma_p: float, lr: float = 1e-4,
                 betas: Tuple[float, float] = (0.9, 0.999), weight_decay: float = 1e-4,
                 model_kwargs: Optional[Dict[str, Any]] = None, warp_kwargs: Optional[Dict[str, Any]] = None,
                 conv_search: bool = True):
        """lit_unet.py:14-50.  ``conv_search`` lets MIOpen search its
        convolution algorithms once per shape (torch's cudnn.benchmark): its
        default heuristics pick naive direct kernels for several FP32 3D
        convolutions of this U-Net."""
        super().__init__()
        if conv_search:
            torch.backends.cudnn.benchmark = True
        model_kwargs = dict(model_kwargs or {})
        warp_kwargs = dict(warp_kwargs or {})
        self.hparams = dict(epochs=epochs, lambda_u=lambda_u, gamma_p=gamma_p, lr=lr, betas=tuple(betas),
                            weight_decay=weight_decay, model_kwargs=model_kwargs, warp_kwargs=warp_kwargs)
        self.model = UNet3D(**model_kwargs)
        self.warp = Warp(**warp_kwargs)
        self.loss_fn = PropagationLoss(lambda_u, gamma_p)
        self.epochs, self.lr, self.betas, self.weight_decay = epochs, lr, tuple(betas), weight_decay
        self.metrics = {stage: [_MaskMetric(), _MaskMetric()] for stage in ("trn", "val")}   # [b, f]

    def configure_optimizers(self, lr: Optional[float] = None):
        """lit_unet.py:60-64: Adam and a cosine schedule down to lr / 100."""
        lr = lr or self.lr
        opt = torch.optim.Adam(self.parameters(), lr=lr, betas=self.betas, weight_decay=self.weight_decay)
        sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=self.epochs, eta_min=lr / 100)
        return opt, sched

    def forward(self, batch: Dict[str, Any]) -> Tuple[torch.Tensor, ...]:
        """lit_unet.py:66-92: both propagation chains over the padded length."""
        img, mi, mf = batch[IMAGE_KEY], batch[MI_KEY], batch[MF_KEY]
        fflow, bflow = batch[FWD_FLOW_KEY], batch[BWD_FLOW_KEY]
        ftimes, btimes = batch[FWD_TS_KEY], batch[BWD_TS_KEY]
        bs, ts, nz, ny, nx, _ = fflow.shape
        bi = torch.arange(bs)
        mts = torch.empty((bs, ts + 1, 1, nz, ny, nx), dtype=mi.dtype, device=mi.device)
        mts[:, 0] = mi
        mtts = torch.empty_like(mts)
        mtts[:, -1] = mf
        mhs = torch.empty((bs, ts, 1, nz, ny, nx), dtype=mi.dtype, device=mi.device)
        mhhs = torch.empty_like(mhs)
        for i in range(ts):
            mts[:, i + 1], mhs[:, i] = self.propagation_step(mts[:, i], img, fflow[:, i], ftimes, i, bi)
            mtts[:, ts - i - 1], mhhs[:, ts - i - 1] = self.propagation_step(mtts[:, ts - i], img, bflow[:, i],
                                                                              btimes, i, bi)
        return mts, mtts, mhs, mhhs

    def propagation_step(self, mask, img, flow, times, i, bi):
        """lit_unet.py:104-120: warp the mask, stack it with the target frame
        and let the U-Net refine it."""
        warped = self.warp(mask, flow)
        x = torch.cat((img[bi, times[bi, i + 1]], warped), dim=1)
        return self.model(x)

    def forward_and_loss(self, batch):
        y = self(batch)
        return self.loss_fn(y, batch[OFFSET_KEY]), y

    @torch.no_grad()
    def update_metrics(self, stage: str, y, offsets: torch.Tensor):
        """lit_unet.py:178-194: the labelled end of each chain against the
        other chain's prediction there, thresholded at 0.5."""
        mts, mtts = y[0], y[1]
        bi = torch.arange(mts.shape[0])
        mi, mitt = mts[bi, 0], mtts[bi, offsets[bi]]
        mf, mft = mtts[bi, -1], mts[bi, -offsets[bi] - 1]
        for k, (target, pred) in enumerate(((mi, mitt), (mf, mft))):
            self.metrics[stage][k].update(torch.where(pred > 0.5, 1.0, 0.0), target)

    def compute_metrics(self, stage: str) -> Dict[str, float]:
        """compute_and_log_mean_metrics (lit_unet.py:152-170)."""
        out = {}
        per = []
        for k, d in enumerate(("b", "f")):
            m = self.metrics[stage][k].compute()
            per.append(m)
            for name, v in m.items():
                out[f"{name}_{stage}/{d}"] = v
            self.metrics[stage][k].reset()
        for name in ("dice", "hsdf"):
            out[f"{name}_{stage}/mean"] = 0.5 * (per[0][name] + per[1][name])
        return out


def seed_everything(seed: int = 42):
    random.seed(seed)
    np.random.seed(seed)
    torch.manual_seed(seed)
    torch.cuda.manual_seed_all(seed)


def _run_epoch(module: PropagationModule, batches: Iterable[Dict[str, Any]], stage: str, device,
               opt: Optional[torch.optim.Optimizer]) -> Dict[str, float]:
    sums: Dict[str, float] = defaultdict(float)
    count = 0
    for batch in batches:
        batch = to_device(batch, device)
        if opt is not None:
            opt.zero_grad(set_to_none=True)
            loss, y = module.forward_and_loss(batch)
            loss[TOTAL_LOSS_KEY].backward()
            opt.step()
        else:
            with torch.no_grad():
                loss, y = module.forward_and_loss(batch)
        n = len(batch[OFFSET_KEY])
        for k, v in loss.items():
            sums[k] += float(v.detach()) * n
        count += n
        module.update_metrics(stage, y, batch[OFFSET_KEY])
    out = {f"loss_{stage}/{k}": v / max(count, 1) for k, v in sums.items()}
    out.update(module.compute_metrics(stage))
    return out


def fit(module: PropagationModule, train_batches: Callable[[int], Iterable[Dict[str, Any]]],
        val_batches: Callable[[], Iterable[Dict[str, Any]]], device, max_epochs: int = 1000,
        min_epochs: int = 1000, patience: int = 50, checkpoint_path: Optional[str] = None,
        sanity_check: bool = True, log: Optional[Callable[[Dict[str, Any]], None]] = None) -> List[Dict[str, Any]]:
    """scripts/train.py:21-39.  ``train_batches(epoch)`` and ``val_batches()``
    yield collated batches.  Returns the per-epoch history."""
    module.to(device)
    opt, sched = module.configure_optimizers()
    if sanity_check:                      # num_sanity_val_steps=-1: every validation batch, nothing kept
        module.eval()
        _run_epoch(module, val_batches(), "val", device, None)
    best, wait, history = math.inf, 0, []
    for epoch in range(max_epochs):
        module.train()
        row: Dict[str, Any] = {"epoch": epoch, "lr": opt.param_groups[0]["lr"]}
        row.update(_run_epoch(module, train_batches(epoch), "trn", device, opt))
        module.eval()
        row.update(_run_epoch(module, val_batches(), "val", device, None))
        sched.step()
        current = row["loss_val/t"]
        if current < best:                 # ModelCheckpoint(save_top_k=1) and EarlyStopping (min_delta=0)
            best, wait = current, 0
            row["checkpoint"] = True
            if checkpoint_path:
                torch.save({"epoch": epoch, "state_dict": module.state_dict(), "optimizer": opt.state_dict(),
                            "hparams": module.hparams, "loss_val/t": current}, checkpoint_path)
        else:
            wait += 1
            row["checkpoint"] = False
        history.append(row)
        if log is not None:
            log(row)
        if wait >= patience and epoch + 1 >= min_epochs:
            break
    return history


class SyntheticPropagationData:
    """Patients in the layout of LitNNDataset after its transforms (without
    the final flow reorder): images (T,1,Z,Y,X), masks (1,Z,Y,X) at the
    end-diastolic and end-systolic frames, forward flows frame t -> t+1 and
    backward flows frame t+1 -> t over [ti, tf] as (T',3,Z,Y,X).  Images and
    masks are phantom sequences; the flows come from this package's TV-L1
    solver (batched); everything lives on ``device``."""

    def __init__(self, shape: Sequence[int], frames: int, n: int, device, gaps: Sequence[int],
                 solver_cfg=None, seed: int = 0):
        from . import phantom
        from .solver import Plan, SolverConfig
        shape = tuple(int(s) for s in shape)
        cfg = solver_cfg or SolverConfig(num_scales=2, max_warps=3, max_outer_iterations=20)
        self.samples: List[Dict[str, Any]] = []
        pairs, slots = [], []
        for k in range(n):
            seq = torch.from_numpy(phantom.sequence(shape, frames, seed_base=100 + 37 * (seed + k))).to(device)
            ti = 0
            tf = int(gaps[k % len(gaps)])
            masks = phantom.sequence_masks(shape, frames, which=(ti, tf))
            s = {IMAGE_KEY: seq[:, None].contiguous(),
                 MI_KEY: torch.from_numpy(masks[ti].astype(np.float32))[None].to(device),
                 MF_KEY: torch.from_numpy(masks[tf].astype(np.float32))[None].to(device),
                 TI_KEY: torch.tensor(ti, dtype=torch.int), TF_KEY: torch.tensor(tf, dtype=torch.int)}
            for d, key in enumerate((FWD_FLOW_KEY, BWD_FLOW_KEY)):
                s[key] = torch.empty((tf - ti,) + shape + (3,), device=device)
                for j in range(tf - ti):
                    if d == 0:       # I0 = frame t+1, I1 = frame t (processor.py:108-111)
                        pairs.append((seq[ti + j + 1], seq[ti + j]))
                    else:            # backward chain: I0 = frame t-1, I1 = frame t
                        pairs.append((seq[tf - j - 1], seq[tf - j]))
                    slots.append((len(self.samples), key, j))
            self.samples.append(s)
        plan = Plan(cfg, [shape] * len(pairs), device=device)
        for idx, ((i0, i1), (si, key, j)) in enumerate(zip(pairs, slots)):
            plan.bind(idx, i0, i1, u_out=self.samples[si][key][j])
        plan.run(use_graph=False)
        plan.close()
        for s in self.samples:        # (T', Z, Y, X, 3) -> (T', 3, Z, Y, X), the layout before the final reorder
            for key in (FWD_FLOW_KEY, BWD_FLOW_KEY):
                s[key] = s[key].permute(0, 4, 1, 2, 3).contiguous()

    def __len__(self):
        return len(self.samples)

    def item(self, idx: int, transform=None) -> Dict[str, Any]:
        """One sample, optionally through ``transform`` (a transforms.Compose),
        then flows moved to (T', Z, Y, X, 3) (datasets.py:523)."""
        s = {k: (v.clone() if isinstance(v, torch.Tensor) and v.is_cuda else v) for k, v in self.samples[idx].items()}
        if transform is not None:
            s = transform(s)
        for key in (FWD_FLOW_KEY, BWD_FLOW_KEY):
            s[key] = s[key].permute(0, 2, 3, 4, 1).contiguous()
        return s

    def batches(self, batch_size: int, shuffle: bool = False, transform=None, generator=None):
        order = list(range(len(self)))
        if shuffle:
            rng = generator or random.Random(0)
            rng.shuffle(order)
        for i in range(0, len(order), batch_size):
            yield collate_fn([self.item(k, transform) for k in order[i:i + batch_size]])
