"""Continuous-batching service: every request's result is bit-identical to
the plain drop-in solving it alone, whatever it was batched with, and the
HBM budget is respected."""
import numpy as np
import pytest
import torch

from singleventriclesegmentation_amd import TVL13DOpticalFlow
from singleventriclesegmentation_amd.serve import FlowServer, Request, make_pool, poisson_requests, run_service
from singleventriclesegmentation_amd.solver import SolverConfig, workspace_bytes

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 16), (12, 20, 16), (20, 16, 24)]
FRAMES = 6


def plain_solver(dev):
    return TVL13DOpticalFlow(2, 2, 4, 1.0, 0.025, 2, (0.5, 0.5, 0.25, 1.0), True, (5.0, 0.5), True, 1.0,
                             "CUBIC_HERMITESPLINE", "MIRROR", "numDofs", (1, 1, 1), dev, use_graph=False)


def check_against_alone(reqs, dev):
    solver = plain_solver(dev)
    for r in reqs:
        if r.kind == "pair":
            u, _ = solver.compute(r.frames[0], r.frames[1], None, None)
            assert torch.equal(u, r.flows), f"request {r.rid} differs"
        else:
            T = len(r.frames)
            for t in range(T - 1):
                assert torch.equal(solver.compute(r.frames[t + 1], r.frames[t], None, None)[0], r.flows[t])
                assert torch.equal(solver.compute(r.frames[t], r.frames[t + 1], None, None)[0], r.flows[T - 1 + t])
            assert r.propagated[0].shape[0] == r.masks[2] + 1


@pytest.mark.parametrize("budget_scale", [None, 1.2])
def test_mixed_requests_equal_alone(dev, budget_scale):
    cfg = SolverConfig(num_scales=2, max_warps=2, max_outer_iterations=4)
    pool = make_pool(SHAPES, FRAMES, dev)
    reqs = poisson_requests(0, 12, SHAPES, pool, FRAMES, seed=5, p_pair=0.75, mask_last=3)  # all arrive at t=0
    budget = 64 << 30
    if budget_scale:
        # room for roughly one sequence plan at a time: forces waiting and eviction
        budget = int(budget_scale * max(workspace_bytes(cfg, [s] * 2 * (FRAMES - 1)) for s in SHAPES))
    server = FlowServer(cfg, device=dev, hbm_budget_bytes=budget, slots=3, pair_buckets=(1, 2, 4))
    lat, makespan = run_service(server, reqs)
    assert len(server.completed) == len(reqs) and np.all(lat >= 0)
    assert server.used <= server.budget
    check_against_alone(reqs, dev)
    server.close()


def test_poisson_arrivals_complete(dev):
    cfg = SolverConfig(num_scales=2, max_warps=1, max_outer_iterations=3)
    pool = make_pool(SHAPES[:2], FRAMES, dev)
    reqs = poisson_requests(200.0, 10, SHAPES[:2], pool, FRAMES, seed=6, mask_last=3)
    server = FlowServer(cfg, device=dev, slots=2)
    lat, makespan = run_service(server, reqs)
    print(f"10 requests at 200/s: p50 {np.percentile(lat, 50) * 1e3:.1f} ms, p99 {np.percentile(lat, 99) * 1e3:.1f} ms")
    assert all(r.done is not None and r.done >= r.arrival for r in reqs)
    server.close()


def test_warmed_and_tuned_plans_keep_results(dev):
    """warm() creates and autotunes the plans of every size and bucket up front;
    requests served from those plans still equal the plain drop-in alone, and
    warm() stays inside the HBM budget."""
    cfg = SolverConfig(num_scales=2, max_warps=2, max_outer_iterations=4)
    pool = make_pool(SHAPES, FRAMES, dev)
    server = FlowServer(cfg, device=dev, slots=2, pair_buckets=(1, 2, 4))
    server.warm({sh: (pool[sh][0][1], pool[sh][0][0]) for sh in SHAPES}, frames=FRAMES, reps=2)
    assert len(server.idle_plans) == len(SHAPES) * 4 and server.used <= server.budget
    reqs = poisson_requests(0, 8, SHAPES, pool, FRAMES, seed=8, p_pair=0.75, mask_last=3)
    run_service(server, reqs)
    check_against_alone(reqs, dev)
    server.close()
    small = FlowServer(cfg, device=dev, hbm_budget_bytes=int(1.5 * workspace_bytes(cfg, [SHAPES[2]] * 4)),
                       pair_buckets=(1, 2, 4))
    small.warm({sh: (pool[sh][0][1], pool[sh][0][0]) for sh in SHAPES}, frames=FRAMES, autotune=False)
    assert small.used <= small.budget
    small.close()


def test_unit_over_budget_raises(dev):
    """A unit that cannot fit even with every idle plan evicted and nothing
    running is an error, not an endless wait."""
    cfg = SolverConfig(num_scales=2, max_warps=1, max_outer_iterations=2)
    pool = make_pool(SHAPES[:1], FRAMES, dev)
    reqs = poisson_requests(0, 1, SHAPES[:1], pool, FRAMES, seed=9, p_pair=1.0)
    server = FlowServer(cfg, device=dev, hbm_budget_bytes=1 << 16)
    with pytest.raises(RuntimeError, match="HBM budget"):
        run_service(server, reqs)
    with pytest.raises(ValueError, match="kind"):
        server.submit(Request(99, "triple", SHAPES[0], 0.0, reqs[0].frames))
    server.close()


def test_sequence_without_masks_from_fresh_frames(dev):
    """Frames written on the caller's stream just before submit are what the
    server solves; a sequence request without masks returns flows only."""
    cfg = SolverConfig(num_scales=2, max_warps=2, max_outer_iterations=3)
    pool = make_pool(SHAPES[:1], FRAMES, dev)
    seq = pool[SHAPES[0]][0]
    frames = [(f * 0.5 + 0.25).contiguous() for f in seq]   # produced on the current stream
    req = Request(0, "sequence", SHAPES[0], 0.0, frames)
    server = FlowServer(cfg, device=dev, slots=2)
    run_service(server, [req])
    assert req.propagated is None and req.flows.shape == (2 * (FRAMES - 1),) + SHAPES[0] + (3,)
    solver = TVL13DOpticalFlow(2, 2, 3, 1.0, 0.025, 2, (0.5, 0.5, 0.25, 1.0), True, (5.0, 0.5), True, 1.0,
                               "CUBIC_HERMITESPLINE", "MIRROR", "numDofs", (1, 1, 1), dev, use_graph=False)
    for t in range(FRAMES - 1):
        assert torch.equal(solver.compute(frames[t + 1], frames[t], None, None)[0], req.flows[t])
        assert torch.equal(solver.compute(frames[t], frames[t + 1], None, None)[0], req.flows[FRAMES - 1 + t])
    server.close()


def test_cancelled_requests_leave_the_others_unchanged(dev):
    """Requests withdrawn mid-flight (one pair of a batched unit, one sequence
    unit) or while pending never complete and return their memory; every other
    request still equals the plain drop-in solving it alone."""
    cfg = SolverConfig(num_scales=2, max_warps=2, max_outer_iterations=4)
    pool = make_pool(SHAPES, FRAMES, dev)
    reqs = poisson_requests(0, 10, SHAPES, pool, FRAMES, seed=11, p_pair=0.7, mask_last=3)
    server = FlowServer(cfg, device=dev, slots=2, pair_buckets=(1, 2, 4))
    for r in reqs:
        server.submit(r)
    server.step(0.0)                       # fills both slots; the rest stay pending
    assert len(server.running) == 2 and server.pending
    running = [r for u in server.running for r in u.requests]
    mid = [running[0]] + ([running[-1]] if running[-1] is not running[0] else [])
    pend = list(server.pending)[:2]
    for r in mid + pend:
        assert server.cancel(r, now=0.0)
    while not server.idle():
        server.step(0.0)
    torch.cuda.synchronize(dev)
    withdrawn = {id(r) for r in mid + pend}
    assert {id(r) for r in server.cancelled} == withdrawn
    assert all(r.flows is None for r in mid + pend)
    done = [r for r in reqs if id(r) not in withdrawn]
    assert {id(r) for r in server.completed} == {id(r) for r in done}
    assert server.used == sum(p.workspace_bytes for _, p in server.idle_plans)
    check_against_alone(done, dev)
    server.close()
