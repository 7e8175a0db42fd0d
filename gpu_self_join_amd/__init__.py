"""MI355X-native exact Euclidean epsilon-grid self-join.

Entry points:
  gdsjoin(DATASET, EPSILON, NDIM, DTYPE, verbose=False)  -- reference drop-in
  SelfJoin(device).join(points, eps, ...)                -- device-resident API
The compute path is libgdsjoin.so (gfx950 HIP kernels behind a C ABI).
"""
from ._lib import GdsjError, default_params, load  # noqa: F401
from .gdsjoingpu import gdsjoin, gdsjoinmain  # noqa: F401
from .selfjoin import BatchJoin, BatchResult, JoinResult, SelfJoin  # noqa: F401
