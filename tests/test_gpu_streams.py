"""The library launches on torch's current stream: the raw-stream query behind
_lib.stream_ptr agrees with torch.cuda.current_stream on the default stream and
inside a side-stream context, and an operator run on a side stream is ordered
after work queued there."""
import pytest
import torch

from singleventriclesegmentation_amd import _lib
from singleventriclesegmentation_amd import opticalFlow as OF

pytestmark = pytest.mark.gpu


def test_stream_ptr_is_torch_current_stream(dev):
    assert _lib.stream_ptr() == torch.cuda.current_stream().cuda_stream
    assert _lib.stream_ptr(dev) == torch.cuda.current_stream(dev).cuda_stream
    side = torch.cuda.Stream(device=dev)
    with torch.cuda.stream(side):
        assert _lib.stream_ptr() == side.cuda_stream
        assert _lib.stream_ptr(dev) == side.cuda_stream
    assert _lib.stream_ptr() == torch.cuda.current_stream().cuda_stream


def test_operator_on_side_stream_follows_its_queue(dev):
    m = OF.MeshInfo3D(32, 32, 32, 31.0, 31.0, 31.0)
    op = OF.Nabla3D_CD(m, OF.BOUNDARY_MIRROR)
    side = torch.cuda.Stream(device=dev)
    x = torch.zeros(32, 32, 32, device=dev)
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        ramp = torch.arange(32, device=dev, dtype=torch.float32).view(1, 1, 32).expand(32, 32, 32)
        x.copy_(ramp)                       # queued on the side stream
        g = op.forward(x)                   # must see the copy
    side.synchronize()
    assert torch.equal(g[:, :, 1:31, 0], torch.ones(32, 32, 30, device=dev))


def _solver(dev, S=4, W=10, K=50):
    from singleventriclesegmentation_amd import TVL13DOpticalFlow
    return TVL13DOpticalFlow(S, W, K, 1.0, 0.025, 2, (0.5, 0.5, 0.25, 1.0), True, (5.0, 0.5), True, 1.0,
                             "CUBIC_HERMITESPLINE", "MIRROR", "numDofs", (1, 1, 1), dev)


def _pair(dev, shape, seed):
    from singleventriclesegmentation_amd import phantom
    f0, f1, _ = phantom.frame_pair(shape, seed0=seed, seed1=seed + 1, disp_seed=seed + 2)
    return torch.from_numpy(f1).to(dev), torch.from_numpy(f0).to(dev)


def test_compute_returns_before_the_solve_completes(dev):
    """A cached plan re-bound to new frames uploads its tables in stream order:
    compute() enqueues the ~25 ms 80^3 solve and returns while it runs."""
    s = _solver(dev)
    a, b = _pair(dev, (80, 80, 80), 0)
    c, d = _pair(dev, (80, 80, 80), 10)
    ref_ab = s.compute(a, b, None, None)[0].clone()     # creates and captures the plan
    ref_cd = s.compute(c, d, None, None)[0].clone()
    torch.cuda.synchronize()
    pending = []
    for x, y in ((a, b), (c, d), (a, b), (c, d), (a, b), (c, d)):
        u, _ = s.compute(x, y, None, None)               # re-binds every call
        pending.append((u, torch.cuda.current_stream().query()))
    busy = [not done for _, done in pending]
    torch.cuda.synchronize()
    print(f"stream still busy right after each compute(): {busy}")
    assert busy[0], "compute() waited for its solve"
    for i, (u, _) in enumerate(pending):
        assert torch.equal(u, ref_ab if i % 2 == 0 else ref_cd), f"call {i}"


def test_one_plan_on_two_streams(dev):
    """The same solver (one cached plan, one workspace) called on two streams
    in turn: each run waits for the plan's previous one, so every result
    equals the solve alone."""
    shape = (32, 36, 40)
    s = _solver(dev, 3, 3, 20)
    pairs = [_pair(dev, shape, 20 + 3 * i) for i in range(4)]
    alone = [s.compute(x, y, None, None)[0].clone() for x, y in pairs]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)]
    outs = []
    for rep in range(3):
        for i, (x, y) in enumerate(pairs):
            st = streams[i % 2]
            st.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(st):
                u, _ = s.compute(x, y, None, None)
            outs.append((i, u, st))
    for st in streams:
        st.synchronize()
    for i, u, st in outs:
        assert torch.equal(u, alone[i]), f"pair {i} on stream {st.cuda_stream:#x}"
