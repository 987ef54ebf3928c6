"""sph -- Python host side of the SPH path of libmpm355 (MI355X SPH step).

``SPHSimulator`` mirrors the reference class of sph/sph.ts; the compute runs
in the gfx950 HIP kernels of libmpm355.so.  Importing this package does not
touch the GPU; constructing a simulator does.
"""
from ._capi import SPHError, Params, load, LIB_PATH, STAGES  # noqa: F401
from .simulator import SPHSimulator, default_params, PARTICLE_STRUCT_SIZE  # noqa: F401
from . import scenes  # noqa: F401
