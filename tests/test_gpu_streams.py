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
