"""MI355X-native 3D TV-L1 optical flow and tricubic Catmull-Rom warping.

Drop-in for the flow/warp path of singleVentricleSegmentation:
``opticalFlow`` mirrors the ``opticalFlow_cuda_ext.opticalFlow`` operator
module, ``TVL13DOpticalFlow`` the solver driver, ``Warp``/``propagate`` the
mask warping harness and ``warp`` the differentiable warp.  All compute runs
in the HIP kernels of libsvsflow.so (built by ``build.build()``).
"""
from . import opticalFlow, transforms  # noqa: F401
from .flow_utils import (apply_gaussian_blur3d, apply_median_filter, compute_timepoints,  # noqa: F401
                         get_interpolation_type, get_mesh_length)
from .processor import OpticalFlowProcessor  # noqa: F401
from .solver import Plan, SolverConfig  # noqa: F401
from .tvl13d import TVL13DOpticalFlow  # noqa: F401
from .warping import Warp, propagate  # noqa: F401
from .warpingOps import WarpingOpFunction, warp  # noqa: F401
