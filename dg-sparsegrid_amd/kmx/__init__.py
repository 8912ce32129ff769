"""kmx -- MI355X (gfx950) operator application for ASGarD-style sparse-grid DG.

Host-side mirror of ASGarD-MATLAB's operator-application interface
(project-asgard/DG-SparseGrid apply_A.m / kron_multd.m / Kronmult/) over the
C ABI of libkmx.so (include/kmx.h).
"""
from ._lib import KmxError, PREC_BF16, PREC_FP32, PREC_FP64  # noqa: F401
from .plan import Batch, Plan, RealSpace, gemm_vbatched, kronmult, kronmult_batch_plan, version  # noqa: F401
