"""Host-side parts of the propagation training (svs/modules/loss.py,
svs/modules/datasets.py collate_fn): the collate rules on hand-made samples
and the loss against a float64 loop restatement of loss.py:21-64 plus
closed forms."""
import numpy as np
import pytest
import torch

from singleventriclesegmentation_amd import training as TR


def sample(tp, shape=(3, 4, 5), frames=7, ti=0, seed=0):
    g = torch.Generator().manual_seed(seed)
    return {"image": torch.randn(frames, 1, *shape, generator=g), "mi": torch.rand(1, *shape, generator=g),
            "mf": torch.rand(1, *shape, generator=g), "ti": torch.tensor(ti, dtype=torch.int),
            "tf": torch.tensor(ti + tp, dtype=torch.int), "fwd_flow": torch.randn(tp, *shape, 3, generator=g),
            "bwd_flow": torch.randn(tp, *shape, 3, generator=g), "patient": f"p{tp}"}


def test_collate_pads_flows_and_builds_clamped_time_chains():
    s = [sample(3, seed=1), sample(5, seed=2), sample(4, ti=1, seed=3)]
    b = TR.collate_fn(s)
    assert b["fwd_flow"].shape == (3, 5, 3, 4, 5, 3) and b["bwd_flow"].shape == (3, 5, 3, 4, 5, 3)
    assert b["offset"].tolist() == [2, 0, 1] and b["offset"].dtype == torch.long
    for k, x in enumerate(s):
        n = x["fwd_flow"].shape[0]
        assert torch.equal(b["fwd_flow"][k, :n], x["fwd_flow"])
        assert torch.count_nonzero(b["fwd_flow"][k, n:]) == 0
        assert torch.equal(b["image"][k], x["image"])
    assert b["patient"] == ["p3", "p5", "p4"]
    assert b["fwd_times"].tolist() == [[0, 1, 2, 3, 3, 3], [0, 1, 2, 3, 4, 5], [1, 2, 3, 4, 5, 5]]
    assert b["bwd_times"].tolist() == [[3, 2, 1, 0, 0, 0], [5, 4, 3, 2, 1, 0], [5, 4, 3, 2, 1, 1]]


def loss_restated(mts, mtts, mhs, mhhs, offsets, lam, gam):
    """loss.py:21-64 as float64 loops over the batch."""
    mts, mtts, mhs, mhhs = (np.asarray(a, np.float64) for a in (mts, mtts, mhs, mhhs))
    bs, n = mts.shape[0], mts.shape[1]
    nh = mhs.shape[1]
    sup = 0.0
    for b in range(bs):
        o = int(offsets[b])
        sup += ((mtts[b, o] - mts[b, 0]) ** 2).sum()
        sup += ((mts[b, n - o - 1] - mtts[b, n - 1]) ** 2).sum()
    unsup = pen = 0.0
    for b in range(bs):
        o = int(offsets[b])
        kl = n - o
        unsup += ((mts[b, 1:n - o - 1] - mtts[b, 1 + o:n - 1]) ** 2).sum() / kl
        pen += ((mhs[b, :nh - o] ** 2).sum() + (mhhs[b, :nh - o] ** 2).sum()) / kl
    sup /= bs
    unsup *= lam / bs
    pen *= gam / bs
    return {"t": sup + unsup + pen, "s": sup, "u": unsup, "p": pen}


@pytest.mark.parametrize("offsets", [[2, 0, 1], [5, 0, 3], [0, 0, 0], [4, 1, 5]])
def test_loss_matches_loop_restatement(offsets):
    """Including samples with a single real frame (offset = T), whose two
    supervised ends coincide and are both counted."""
    g = torch.Generator().manual_seed(4)
    bs, ts, shape = 3, 5, (4, 5, 6)
    mts = torch.rand(bs, ts + 1, 1, *shape, generator=g)
    mtts = torch.rand(bs, ts + 1, 1, *shape, generator=g)
    mhs = torch.randn(bs, ts, 1, *shape, generator=g)
    mhhs = torch.randn(bs, ts, 1, *shape, generator=g)
    offsets = torch.tensor(offsets)
    got = TR.PropagationLoss(1.0, 0.1)((mts, mtts, mhs, mhhs), offsets)
    want = loss_restated(mts, mtts, mhs, mhhs, offsets, 1.0, 0.1)
    for k in "tsup":
        np.testing.assert_allclose(float(got[k]), want[k], rtol=2e-6)


def test_loss_closed_forms():
    """Chains that agree everywhere and raw outputs of zero cost nothing; a
    constant offset c on the interior of one chain costs lambda_u c^2 per
    interior voxel over the chain length."""
    bs, ts, shape = 2, 4, (3, 3, 3)
    chain = torch.rand(bs, ts + 1, 1, *shape)
    zeros = torch.zeros(bs, ts, 1, *shape)
    offsets = torch.tensor([0, 0])
    out = TR.PropagationLoss(1.0, 0.1)((chain, chain.clone(), zeros, zeros), offsets)
    assert all(float(out[k]) == 0.0 for k in "tsup")
    c = 0.25
    shifted = chain.clone()
    shifted[:, 1:-1] += c
    out = TR.PropagationLoss(2.0, 0.1)((chain, shifted, zeros, zeros), offsets)
    n_interior = (ts - 1) * 27
    np.testing.assert_allclose(float(out["u"]), 2.0 * c * c * n_interior / (ts + 1), rtol=1e-5)
    assert float(out["s"]) == 0.0
