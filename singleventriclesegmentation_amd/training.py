"""Mask-propagation training of svs/models/lit_unet.py on the GPU path.

    PropagationLoss     svs/modules/loss.py:8-64 (the same objective, written
                        over the two chains aligned position by position)
    PropagationModule   svs/models/lit_unet.py:13-194 (both propagation
                        chains, advanced together with one batched warp and
                        one U-Net call per step; forward_and_loss, optimiser
                        and schedule, Dice and Hausdorff per direction)
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


def _align_backward(bwd: torch.Tensor, offsets: torch.Tensor) -> torch.Tensor:
    """The backward chain re-indexed onto the forward chain's positions.

    A sample whose flows were padded by o steps holds its L = n - o real
    frames at positions 0..L-1 of the forward chain (the padding steps come
    after them) and at o..n-1 of the backward chain (the padding steps come
    first, since the backward chain is stored end-first).  Position j of the
    result is position min(j + o, n - 1) of ``bwd``."""
    bs, n = bwd.shape[:2]
    pos = torch.arange(n, device=bwd.device)
    src = torch.clamp(pos[None, :] + offsets[:, None], max=n - 1)
    return torch.gather(bwd, 1, src.view(bs, n, *([1] * (bwd.dim() - 2))).expand_as(bwd))


def _volume_sums(x: torch.Tensor) -> torch.Tensor:
    """Sum over every axis after the first two: [B, T, ...] -> [B, T]."""
    return x.flatten(2).sum(2)


class PropagationLoss(nn.Module):
    """The propagation objective of svs/modules/loss.py:8-64.

    With the backward chain aligned onto the forward one (_align_backward),
    every term is a per-position volume sum over a sample's L real frames:

    * supervised: the squared error at the first position (the backward
      chain's estimate of the given initial mask) plus the one at position
      L - 1 (the forward chain's estimate of the given final mask), summed
      over the batch and divided by its size;
    * unsupervised: the squared disagreement of the two chains at the
      interior positions 1..L-2, divided by L per sample, times
      lambda_u / batch size;
    * penalty: the squared raw U-Net outputs of the L - 1 real steps of both
      chains, divided by L per sample, times gamma_p / batch size.
    """

    def __init__(self, lambda_u: float, gamma_p: float):
        super().__init__()
        self.lambda_u = lambda_u
        self.gamma_p = gamma_p

    def forward(self, y: Tuple[torch.Tensor, ...], offsets: torch.Tensor) -> Dict[str, torch.Tensor]:
        fwd, bwd, raw_fwd, raw_bwd = y
        bs, n = fwd.shape[:2]
        dev = fwd.device
        off = offsets.to(device=dev, dtype=torch.long)
        length = n - off
        err = _volume_sums((fwd - _align_backward(bwd, off)).pow(2))          # [bs, n]
        last = (length - 1)[:, None]
        # Both ends are counted, also when they coincide (a single real frame).
        supervised = (err[:, 0] + err.gather(1, last)[:, 0]).sum() / bs
        pos = torch.arange(n, device=dev)[None, :]
        interior = (pos > 0) & (pos < last)
        agreement = torch.where(interior, err, torch.zeros_like(err)).sum(1) / length
        unsupervised = agreement.sum() * (self.lambda_u / bs)
        steps = raw_fwd.shape[1]
        real = torch.arange(steps, device=dev)[None, :] < (steps - off)[:, None]
        raw = _volume_sums(raw_fwd.pow(2)) + _volume_sums(raw_bwd.pow(2))       # [bs, steps]
        penalty = (torch.where(real, raw, torch.zeros_like(raw)).sum(1) / length).sum() * (self.gamma_p / bs)
        return {TOTAL_LOSS_KEY: supervised + unsupervised + penalty, SUPERVISED_LOSS_KEY: supervised,
                UNSUPERVISED_LOSS_KEY: unsupervised, PENALIZATION_LOSS_KEY: penalty}


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
        """Adam with the module's betas and weight decay, and a cosine
        schedule from ``lr`` down to lr / 100 over the epochs (stepped once
        per epoch), as lit_unet.py:60-64 configures them."""
        base = lr or self.lr
        optimiser = torch.optim.Adam(self.parameters(), lr=base, betas=self.betas, weight_decay=self.weight_decay)
        schedule = torch.optim.lr_scheduler.CosineAnnealingLR(optimiser, T_max=self.epochs, eta_min=base / 100)
        return optimiser, schedule

    def forward(self, batch: Dict[str, Any]) -> Tuple[torch.Tensor, ...]:
        """Both propagation chains over the padded length (lit_unet.py:66-120).

        The chains advance together: step i warps the forward chain's
        latest mask by forward flow i and the backward chain's latest mask
        by backward flow i as one batch of 2 x batch volumes (one warp
        launch), stacks each with the frame its chain reaches next, and
        refines all of them in one U-Net call.  The U-Net normalises per
        sample (instance norm), so batching the directions changes no
        sample's arithmetic.  Returns (forward chain, backward chain, raw
        forward outputs, raw backward outputs): chains [B, T+1, 1, Z, Y, X]
        indexed by frame from the initial to the final mask, raw outputs
        [B, T, 1, Z, Y, X] indexed the same way (the backward chain's step
        that produced frame t is stored at t)."""
        img = batch[IMAGE_KEY]
        flows = torch.cat((batch[FWD_FLOW_KEY], batch[BWD_FLOW_KEY]))        # [2B, T, Z, Y, X, 3]
        times = torch.cat((batch[FWD_TS_KEY], batch[BWD_TS_KEY]))            # [2B, T+1]
        bs, steps = batch[FWD_FLOW_KEY].shape[:2]
        sample = torch.arange(2 * bs, device=times.device) % bs
        current = torch.cat((batch[MI_KEY], batch[MF_KEY]))                  # latest mask of each chain
        refined, raw = [current], []
        for i in range(steps):
            target = img[sample, times[:, i + 1]]
            nxt, out = self.model(torch.cat((target, self.warp(current, flows[:, i])), dim=1))
            refined.append(nxt)
            raw.append(out)
            current = nxt
        fwd = torch.stack([m[:bs] for m in refined], dim=1)
        bwd = torch.stack([m[bs:] for m in reversed(refined)], dim=1)
        raw_fwd = torch.stack([m[:bs] for m in raw], dim=1)
        raw_bwd = torch.stack([m[bs:] for m in reversed(raw)], dim=1)
        return fwd, bwd, raw_fwd, raw_bwd

    def forward_and_loss(self, batch):
        y = self(batch)
        return self.loss_fn(y, batch[OFFSET_KEY]), y

    @torch.no_grad()
    def update_metrics(self, stage: str, y, offsets: torch.Tensor):
        """Scores each chain's estimate of the other chain's given mask
        (lit_unet.py:178-194): the backward chain at the initial frame
        against the initial mask ("b"), the forward chain at the final frame
        against the final mask ("f"), thresholded at 0.5."""
        fwd, bwd = y[0], y[1]
        n = fwd.shape[1]
        off = offsets.to(device=fwd.device, dtype=torch.long)
        rows = torch.arange(fwd.shape[0], device=fwd.device)
        estimates = (bwd[rows, off], fwd[rows, n - 1 - off])
        given = (fwd[:, 0], bwd[:, n - 1])
        for k in range(2):
            self.metrics[stage][k].update(torch.where(estimates[k] > 0.5, 1.0, 0.0), given[k])

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
