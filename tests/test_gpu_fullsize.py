"""The benchmark workloads at their full sizes.

* 80^3 pair, 4 x 10 x 50: the fused solve against the reference's own driver
  (tvl13d.py:122-255) run operator by operator through the drop-in;
* 35-frame 80^3 sequence: 68 solves in one batched plan against 68 plain
  solves, the backward flows in the order the reference's processor writes
  them (processor.py:95-121), and mask propagation 0 -> 12 and 12 -> 0
  against the oracle's warp chain after the 0.5 threshold
  (warping.py:205-243);
* the five-size uneven batch: every instance against its solve alone, and
  each size against the oracle at a short schedule;
* the serving loop with an 80^3 sequence and 128x128x48 pairs in the mix,
  every result against the plain drop-in alone.
"""
import numpy as np
import pytest
import torch

from singleventriclesegmentation_amd import TVL13DOpticalFlow, phantom
from singleventriclesegmentation_amd.flow_utils import compute_timepoints, gaussian_weights3d, median_filter3d
from singleventriclesegmentation_amd.serve import FlowServer, Request, run_service
from singleventriclesegmentation_amd.solver import Plan, SolverConfig
from singleventriclesegmentation_amd.warping import Warp, propagate

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

FULL = dict(num_scales=4, max_warps=10, max_outer_iterations=50)


def solver(dev, S=4, W=10, K=50, use_graph=True):
    return TVL13DOpticalFlow(S, W, K, 1.0, 0.025, 2, (0.5, 0.5, 0.25, 1.0), True, (5.0, 0.5), True, 1.0,
                             "CUBIC_HERMITESPLINE", "MIRROR", "numDofs", (1, 1, 1), dev, use_graph=use_graph)


def diff_stats(a, b):
    a = a.double()
    b = b.double()
    d = (a - b).abs()
    e = torch.sqrt(((a - b) ** 2).sum(-1))
    return {"max_abs": d.max().item(), "mean_epe": e.mean().item(),
            "voxels_above_1e-3": int((d.amax(-1) > 1e-3).sum().item())}


def test_reference_driver_equals_fused_80cubed(dev):
    """The reference's step-by-step driver over the drop-in operators (one
    operator call per reference call, torch's own elementwise and sum
    kernels between them) gives the fused solve's bits at the benchmark size."""
    f0, f1, d = phantom.frame_pair((80, 80, 80))
    I0, I1 = torch.from_numpy(f1).to(dev), torch.from_numpy(f0).to(dev)
    s = solver(dev)
    u_fused, p_fused = s.compute(I0, I1, None, None)
    u_op, p_op = s.compute_unfused(I0, I1, None, None)
    st = diff_stats(u_op, u_fused)
    print(f"80^3 4x10x50 operator path vs fused: {st}")
    assert st["max_abs"] <= 1e-3 and st["mean_epe"] <= 1e-5
    assert torch.equal(u_op, u_fused) and torch.equal(p_op, p_fused)


def _sequence_flows(dev, T, shape):
    seq = torch.from_numpy(phantom.sequence(shape, T)).to(dev)
    n = 2 * (T - 1)
    plan = Plan(SolverConfig(**FULL), [shape] * n, device=dev)
    flows = torch.empty((n,) + shape + (3,), device=dev)
    for t in range(T - 1):
        plan.bind(t, seq[t + 1], seq[t], u_out=flows[t])                  # forward: I0 = f[t+1], I1 = f[t]
        plan.bind(T - 1 + t, seq[t], seq[t + 1], u_out=flows[T - 1 + t])  # backward: I0 = f[t], I1 = f[t+1]
    plan.run(use_graph=True)
    torch.cuda.synchronize()
    plan.close()
    return seq, flows


def _oracle_chain(O, mask, flows):
    """warping.py:222-243 with the oracle's tricubic warp: the continuous mask
    carried through every step; returns the chain [T+1, Z, Y, X]."""
    out = [mask.astype(np.float32)]
    for f in flows:
        out.append(O.warp_cnn(out[-1][None, None], f[None])[0, 0])
    return np.stack(out)


def test_sequence_35x80cubed_batch_and_propagation(O, dev):
    T, shape, last = 35, (80, 80, 80), 12
    seq, flows = _sequence_flows(dev, T, shape)

    # 68 batched solves == 68 plain drop-in solves, forward pairs and the
    # backward pairs of the reference processor's backward mode.
    s = solver(dev)
    for t in range(T - 1):
        assert torch.equal(s.compute(seq[t + 1], seq[t], None, None)[0], flows[t]), f"forward pair {t}"
    _, _, times = compute_timepoints(0, last, None, None, "backward")          # [12, 11, ..., 0]
    bwd_ref = [s.compute(seq[times[j + 1]], seq[times[j]], None, None)[0] for j in range(len(times) - 1)]
    for j in range(last):
        assert torch.equal(bwd_ref[j], flows[T - 1 + last - 1 - j]), f"backward pair {j} of the processor's loop"
    for t in range(last, T - 1):
        assert torch.equal(s.compute(seq[t], seq[t + 1], None, None)[0], flows[T - 1 + t]), f"backward pair {t}"

    # The flow files: median filter (conf/flow.yaml), forward flows 0->12 and
    # the backward flows in the processor's order, which the batch holds
    # reversed at [T-1, T-1+12).
    filt = median_filter3d(flows, 3)
    fwd = filt[0:last]
    bwd = filt[T - 1:T - 1 + last].flip(0)
    for j in range(last):
        assert torch.equal(bwd[j], median_filter3d(bwd_ref[j], 3)), f"backward flow {j} out of the processor's order"

    masks = phantom.sequence_masks(shape, T, which=(0, last))
    m0 = torch.from_numpy(masks[0].astype(np.float32)).to(dev)[None, None]
    m12 = torch.from_numpy(masks[last].astype(np.float32)).to(dev)[None, None]
    warper = Warp("CUBIC_HERMITESPLINE", "MIRROR", "numDofs", (1, 1, 1))
    fm, bm = propagate(warper, m0, m12, fwd, bwd)
    assert fm.shape == (last + 1, 1) + shape and bm.shape == (last + 1, 1) + shape

    f_np, b_np = fwd.cpu().numpy(), bwd.cpu().numpy()
    chain_f = _oracle_chain(O, masks[0], f_np)
    chain_b = _oracle_chain(O, masks[last], b_np)[::-1]      # index 0 estimates the frame-0 mask
    for name, got, chain in (("forward", fm, chain_f), ("backward", bm, chain_b)):
        want = np.where(chain > 0.5, 1.0, 0.0).astype(np.float32)
        got = got[:, 0].cpu().numpy()
        flips = np.argwhere(got != want)
        margin = np.abs(chain - 0.5)
        print(f"{name} chain: {flips.shape[0]} flipped voxels; closest value to the threshold "
              f"|m - 0.5| = {margin.min():.3e}; voxels within 1e-4 of it: {int((margin < 1e-4).sum())}")
        for idx in flips[:10]:
            print(f"  flip at step/voxel {tuple(idx)}: oracle {chain[tuple(idx)]:.9g}")
        assert flips.shape[0] == 0
    # the ends of the chains are the given masks
    assert torch.equal(fm[0, 0], m0[0, 0]) and torch.equal(bm[last, 0], m12[0, 0])


def test_uneven_five_size_batch(O, dev):
    sizes = list(phantom.UNEVEN_SIZES)
    pairs = []
    for i, sh in enumerate(sizes):
        f0, f1, _ = phantom.frame_pair(sh, seed0=10 * i, seed1=10 * i + 1, disp_seed=10 * i + 2)
        pairs.append((torch.from_numpy(f1).to(dev), torch.from_numpy(f0).to(dev), f1, f0))
    # one persistent plan over the five sizes (two instances of 80^3)
    order = sizes + [sizes[2]]
    inst = pairs + [pairs[2]]
    plan = Plan(SolverConfig(**FULL), order, device=dev)
    outs = []
    for k, (a, b, _, _) in enumerate(inst):
        u = torch.empty(order[k] + (3,), device=dev)
        plan.bind(k, a, b, u_out=u)
        outs.append(u)
    plan.run(use_graph=True)
    plan.autotune(reps=2)
    plan.run(use_graph=True)
    torch.cuda.synchronize()
    s = solver(dev)
    for k, (a, b, _, _) in enumerate(inst):
        alone, _ = s.compute(a, b, None, None)
        assert torch.equal(alone, outs[k]), f"instance {k} {order[k]} differs from its solve alone"
    plan.close()

    # every size against the oracle at a short schedule, in one batch
    short = dict(num_scales=4, max_warps=2, max_outer_iterations=5)
    plan = Plan(SolverConfig(**short), sizes, device=dev)
    outs = []
    for k, (a, b, _, _) in enumerate(pairs):
        u = torch.empty(sizes[k] + (3,), device=dev)
        plan.bind(k, a, b, u_out=u)
        outs.append(u)
    plan.run(use_graph=False)
    torch.cuda.synchronize()
    w = gaussian_weights3d(1.0)
    params = O.solve_params(**short)
    for k, (_, _, f1, f0) in enumerate(pairs):
        uo, _ = O.solve(O.blur(f1, w), O.blur(f0, w), params)
        assert np.array_equal(outs[k].cpu().numpy(), uo), f"{sizes[k]} differs from the oracle"
    plan.close()


def test_serving_mix_with_80cubed_sequence(dev):
    """Continuous batching with an 80^3 sequence (68 solves + propagation)
    and 128x128x48 / 48^3 pairs arriving while it runs: every result equals
    the plain drop-in's alone."""
    cfg = SolverConfig(**FULL)
    T = 35
    big, small = (48, 128, 128), (48, 48, 48)
    seq = torch.from_numpy(phantom.sequence((80, 80, 80), T)).to(dev)
    masks = phantom.sequence_masks((80, 80, 80), T, which=(0, 12))
    m0 = torch.from_numpy(masks[0].astype(np.float32)).to(dev)[None, None]
    m12 = torch.from_numpy(masks[12].astype(np.float32)).to(dev)[None, None]
    frames = {}
    for i, sh in enumerate((big, small)):
        f0, f1, _ = phantom.frame_pair(sh, seed0=50 + 3 * i, seed1=51 + 3 * i, disp_seed=52 + 3 * i)
        frames[sh] = (torch.from_numpy(f1).to(dev), torch.from_numpy(f0).to(dev))
    reqs = [Request(0, "sequence", (80, 80, 80), 0.0, [seq[t] for t in range(T)], masks=(m0, m12, 12))]
    arrivals = [0.05, 0.1, 0.15, 0.2, 0.25]
    for j, t in enumerate(arrivals):
        sh = big if j % 2 == 0 else small
        reqs.append(Request(j + 1, "pair", sh, t, frames[sh]))
    server = FlowServer(cfg, device=dev, slots=4)
    lat, makespan = run_service(server, reqs)
    print(f"served {len(reqs)} requests in {makespan:.2f} s; latencies (s): {np.round(lat, 3).tolist()}")
    assert len(server.completed) == len(reqs)
    s = solver(dev)
    for r in reqs[1:]:
        assert torch.equal(s.compute(r.frames[0], r.frames[1], None, None)[0], r.flows), f"request {r.rid}"
    r = reqs[0]
    for t in range(T - 1):
        assert torch.equal(s.compute(seq[t + 1], seq[t], None, None)[0], r.flows[t])
        assert torch.equal(s.compute(seq[t], seq[t + 1], None, None)[0], r.flows[T - 1 + t])
    fm, bm = propagate(Warp("CUBIC_HERMITESPLINE", "MIRROR", "numDofs", (1, 1, 1)), m0, m12,
                       r.flows[0:12], r.flows[T - 1:T - 1 + 12].flip(0))
    assert torch.equal(fm, r.propagated[0]) and torch.equal(bm, r.propagated[1])
    server.close()


def test_instance_at_the_size_limit(dev):
    """An instance at the plan's size limit (488 x 488 x 500 = 119,072,000
    voxels: its nine dual fields span 36 N bytes, just below the 2^32 the
    sweep's 32-bit offsets address) solved in one plan with a small instance
    placed after it in the workspace: identical frames give exactly zero flow
    and dual (as the oracle gives at any size), and the small instance equals
    its solve alone."""
    big, small = (500, 488, 488), (20, 24, 28)
    assert 36 * big[0] * big[1] * big[2] < 2 ** 32
    cfg = SolverConfig(num_scales=2, max_warps=1, max_outer_iterations=2)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    frame = torch.rand(big, device=dev, generator=g)
    f0, f1, _ = phantom.frame_pair(small, seed0=70, seed1=71, disp_seed=72, peak=2.0)
    a, b = torch.from_numpy(f1).to(dev), torch.from_numpy(f0).to(dev)
    u_big = torch.full(big + (3,), 7.0, device=dev)
    p_big = torch.full(big + (3, 3), 7.0, device=dev)
    u_s = torch.empty(small + (3,), device=dev)
    p_s = torch.empty(small + (3, 3), device=dev)
    plan = Plan(cfg, [big, small], device=dev)
    plan.bind(0, frame, frame, u_out=u_big, p_out=p_big)
    plan.bind(1, a, b, u_out=u_s, p_out=p_s)
    plan.run(use_graph=False)
    torch.cuda.synchronize()
    assert int(torch.count_nonzero(u_big)) == 0 and int(torch.count_nonzero(p_big)) == 0
    plan.close()
    del u_big, p_big, frame
    alone = Plan(cfg, [small], device=dev)
    u_a = torch.empty(small + (3,), device=dev)
    p_a = torch.empty(small + (3, 3), device=dev)
    alone.bind(0, a, b, u_out=u_a, p_out=p_a)
    alone.run(use_graph=False)
    torch.cuda.synchronize()
    assert torch.equal(u_a, u_s) and torch.equal(p_a, p_s)
    assert float(u_s.abs().max()) > 0.01   # the small pair does move
    alone.close()
