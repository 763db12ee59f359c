"""galaxy_amd — MI355X-native LBVH gravity step (HIP/CDNA4) behind a C ABI.

The compute path is libgalaxy_lbvh.so (galaxy_amd/csrc, built by galaxy_amd.build);
this package is its ctypes front end, mirroring the reference's solver interfaces.
"""
from .lbvh import (  # noqa: F401
    GlxError,
    GpuResidentLeapfrog,
    Lbvh,
    State,
    aggregate,
    flatten,
    device_info,
    direct_sum_f64,
    karras,
    lib,
    morton,
    sort,
    PREC_ABS32,
    PREC_F32,
    PREC_KAHAN,
    NO_PARENT,
    WALK_GROUP,
    WALK_REFERENCE,
)
