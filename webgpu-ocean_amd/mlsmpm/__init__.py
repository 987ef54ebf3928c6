"""mlsmpm -- Python host side of libmpm355 (MI355X MLS-MPM substep).

``MLSMPMSimulator`` mirrors the reference class of mls-mpm/mls-mpm.ts; the
compute runs in the gfx950 HIP kernels of libmpm355.so.  Importing this
package does not touch the GPU; constructing a simulator does.
"""
from ._capi import MPMError, Params, load, LIB_PATH, STAGES  # noqa: F401
from .simulator import MLSMPMSimulator, default_params, PARTICLE_STRUCT_SIZE  # noqa: F401
from .batch import MLSMPMBatch  # noqa: F401
from . import scenes  # noqa: F401
