"""MI355X-native implicit yarn-level time step.

The compute path is libyarnstep.so (HIP kernels for gfx950 behind the C ABI of
include/yarnstep.h).  This module is the host-side mirror of the reference's step
API: ``Cloth`` + ``PhysicsCloth`` (src/Cloth.h, src/PhysicsCloth.h) and ``SimStep``
(src/TrainHelper.cpp:15-97), with state held as numpy arrays at the boundary.
There is no CPU fallback: importing works without a GPU, but every compute call
raises if the HIP library is missing or no device is present.
"""
from . import cloth_spec  # noqa: F401
from ._abi import (YcError, NotConverged, NonFinite, YC_OK, YC_ERR_NONFINITE,  # noqa: F401
                   YC_ERR_NOT_CONVERGED, YC_ERR_HIP, YC_ERR_ARG, lib_path, load_library)
from .cloth import Cloth, EigenSolver, PhysicsCloth, SimStep  # noqa: F401
from . import grad  # noqa: F401
from .batch import ClothBatch, autotune_batch, autotune_cloth  # noqa: F401
from . import learn  # noqa: F401
from . import h5io  # noqa: F401
from . import render  # noqa: F401
from . import control  # noqa: F401
