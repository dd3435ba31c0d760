"""MI355X-native MGBF localization operator B = W A W^T (MGBF of the
reference Standalone_MGBF_loc), HIP kernels behind a C ABI (include/mgbf.h).

Python surface:
  Config / derive / MGBF ........ ctypes binding of the C ABI (mgbf.py)
  read_namelist ................. &PARAMETERS_MGBETA reader (namelist.py)
"""
from .mgbf import (Config, Dims, MGBF, MGBFError, derive, load_library, version,  # noqa: F401
                   OP_WAWT, OP_WT, OP_W, OP_A, OP_ASQRT, OP_BSQRT, OP_BSQRTT,
                   STAGE_VF, STAGE_VFT, STAGE_F, STAGE_FT, STAGE_R, STAGE_P, STAGE_A)
