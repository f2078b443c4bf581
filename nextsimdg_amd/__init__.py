"""MI355X (gfx950) implementation of the neXtSIM_DG dynamics hot path.

The mEVP momentum/rheology subcycle (and the brittle BBM substep) and the DG advection
of ice thickness and concentration run as hand-written HIP kernels in libnsdg_hip.so, behind the C ABI
of include/nsdg_dyn.h. This package is the Python mirror of the reference's
DynamicsKernel / DGTransport / ParametricMesh interfaces over that ABI, plus the
synthetic benchmark inputs. There is no CPU fallback.
"""
from .dynamics import (  # noqa: F401
    PREC_FP32, PREC_FP64, BBMDynamicsKernel, DGTransport, LimitMax, LimitMin, MEVPDynamicsKernel, ParametricMesh,
    default_meb_params, default_params,
)
from ._lib import NsdgError  # noqa: F401
from . import inputs  # noqa: F401
