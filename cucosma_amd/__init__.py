"""cucosma_amd -- MI355X (gfx950) COSMA-tiled GEMM: FP32 / FP64 / BF16x3 on MFMA.

The compute lives in libcosma_amd.so (hand-written HIP kernels behind the C
ABI of include/cosma/cosma.h).  Importing this package loads that library and
raises ImportError when it has not been built; nothing falls back to another
implementation.
"""
from ._lib import (  # noqa: F401
    MODE_BF16X3,
    MODE_BF16X4,
    MODE_FP32,
    MODE_FP64,
    EPILOGUES,
    CosmaError,
    check,
    device_topology,
    heuristic,
    peak_tflops,
    lib,
    tile_configs,
)
from .gemm import (  # noqa: F401
    dgemm,
    error_stats,
    gemm,
    normwise_error,
    fp64_product,
    release_workspace,
    sgemm,
    sgemm_bf16x3,
    sgemm_bf16x4,
    split_operand,
    sumsq,
    to_f64,
    workspace_bytes,
)

from .grouped import GroupedGemm, group_families, grouped_autotune, grouped_gemm  # noqa: F401,E402
from .autograd import CosmaMatmul, matmul  # noqa: F401,E402

__version__ = lib.cosma_version().decode()
