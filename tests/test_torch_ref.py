"""The test-only PyTorch Catmull-Rom warp agrees with the FP64 oracle (CPU)."""
import numpy as np
import torch

from torch_warp_ref import warp_ref


def test_torch_reference_warp_equals_oracle(O):
    r = np.random.default_rng(3)
    u = r.standard_normal((2, 2, 9, 10, 11))
    phi = r.uniform(-3, 3, (2, 9, 10, 11, 3))
    ref = warp_ref(torch.from_numpy(u), torch.from_numpy(phi)).numpy()
    np.testing.assert_allclose(ref, O.warp_cnn(u, phi, dtype=np.float64), atol=1e-12)
