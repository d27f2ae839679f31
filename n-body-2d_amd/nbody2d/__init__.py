"""nbody2d — host-side Python access to the gfx950 2-D N-body library.

The compute path is the HIP library ``libnbody2d.so`` behind the C ABI of
include/nbody.h; this package only loads it (``abi``), builds it
(``_build``) and generates seeded workloads (``ics``).
"""
from . import ics  # noqa: F401
from .abi import Context, NBodyError, direct_sum_f64, lib, params  # noqa: F401
