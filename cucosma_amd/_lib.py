"""ctypes binding of libcosma_amd.so (the C ABI in include/cosma/cosma.h).

The library is built in-tree by ``python cucosma_amd/build.py``.  Importing
this module fails loudly when it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcosma_amd.so")

STATUS_SUCCESS = 0
STATUS_INVALID_VALUE = -1
STATUS_MISALIGNED = -2
STATUS_NOT_SUPPORTED = -3
STATUS_WORKSPACE = -4
STATUS_SCRATCH = -5

MODE_FP32 = 0
MODE_FP64 = 1
MODE_BF16X3 = 2
MODE_BF16X4 = 3
MODES = {"fp32": MODE_FP32, "fp64": MODE_FP64, "bf16x3": MODE_BF16X3, "bf16x4": MODE_BF16X4}

EPILOGUE_NONE = 0
EPILOGUE_RELU = 1
EPILOGUE_SIGMOID = 2
EPILOGUES = {None: EPILOGUE_NONE, "none": EPILOGUE_NONE, "relu": EPILOGUE_RELU,
             "sigmoid": EPILOGUE_SIGMOID}


class CosmaError(RuntimeError):
    """A non-zero status returned by the C ABI."""

    def __init__(self, status: int, what: str = ""):
        self.status = status
        text = _lib.cosma_status_string(status).decode() if _lib is not None else str(status)
        super().__init__(f"{what}: status {status} ({text})" if what else f"status {status} ({text})")


class GemmDesc(C.Structure):
    """Mirror of struct cosma_gemm_desc."""

    _fields_ = [
        ("mode", C.c_int),
        ("m", C.c_int),
        ("n", C.c_int),
        ("k", C.c_int),
        ("trans_a", C.c_int),
        ("trans_b", C.c_int),
        ("a", C.c_void_p),
        ("lda", C.c_int64),
        ("b", C.c_void_p),
        ("ldb", C.c_int64),
        ("c", C.c_void_p),
        ("ldc", C.c_int64),
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("split_k", C.c_int),
        ("tile_cfg", C.c_int),
        ("group_m", C.c_int),
        ("workspace", C.c_void_p),
        ("workspace_bytes", C.c_size_t),
        ("stream", C.c_void_p),
        ("epilogue", C.c_int),
        ("a_planes", C.c_void_p),
        ("b_planes", C.c_void_p),
    ]


class GroupProblem(C.Structure):
    """Mirror of struct cosma_group_problem."""

    _fields_ = [
        ("m", C.c_int),
        ("n", C.c_int),
        ("k", C.c_int),
        ("a", C.c_void_p),
        ("lda", C.c_int64),
        ("b", C.c_void_p),
        ("ldb", C.c_int64),
        ("c", C.c_void_p),
        ("ldc", C.c_int64),
        ("alpha", C.c_double),
        ("beta", C.c_double),
    ]


class GroupTuneResult(C.Structure):
    """Mirror of struct cosma_group_tune_result."""

    _fields_ = [
        ("bucket", C.c_int),
        ("tile_cfg", C.c_int),
        ("split_k", C.c_int),
        ("status", C.c_int),
        ("ms", C.c_double),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first (python cucosma_amd/build.py)"
        )
    # torch bundles its own libamdhip64 (file libamdhip64.so, soname
    # libamdhip64.so.7).  Loading it first lets this library's
    # libamdhip64.so.7 dependency resolve to that same runtime; loading
    # /opt/rocm's first would put two HIP runtimes in one process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    f32p, f64p, vp = C.POINTER(C.c_float), C.POINTER(C.c_double), C.c_void_p
    i, i64, d, f = C.c_int, C.c_int64, C.c_double, C.c_float
    sig = {
        "cosma_gemm_desc_init": (None, [C.POINTER(GemmDesc)]),
        "cosma_gemm_ex": (i, [C.POINTER(GemmDesc)]),
        "cosma_gemm_ex_static": (i, [C.POINTER(GemmDesc)]),
        "cosma_static_scratch_bytes": (C.c_size_t, []),
        "cosma_workspace_query": (i, [C.POINTER(GemmDesc), C.POINTER(C.c_size_t)]),
        "cosma_workspace_set": (i, [vp, C.c_size_t]),
        "cosma_workspace_release": (i, [C.POINTER(C.c_size_t)]),
        "cosma_sgemm": (i, [f, vp, i, vp, i, f, vp, i]),
        "cosma_dgemm": (i, [d, vp, i, vp, i, d, vp, i]),
        "cosma_fixed_shape": (None, [C.POINTER(i), C.POINTER(i), C.POINTER(i)]),
        "cosma_sgemm_1024x1024x1024": (i, [f, vp, i, vp, i, f, vp, i]),
        "cosma_sgemm_8192x8192x8192": (i, [f, vp, i, vp, i, f, vp, i]),
        "cosma_sgemm_256x256x1048576": (i, [f, vp, i, vp, i, f, vp, i]),
        "cosma_sgemm_bf16x3_8192x8192x8192": (i, [f, vp, i, vp, i, f, vp, i]),
        "cosma_sgemm_bf16x3_256x256x1048576": (i, [f, vp, i, vp, i, f, vp, i]),
        "cosma_dgemm_32768x1024x1024": (i, [d, vp, i, vp, i, d, vp, i]),
        "cosma_dgemm_1024x65536x1024": (i, [d, vp, i, vp, i, d, vp, i]),
        "cosma_sgemm_mnk": (i, [i, i, i, f, vp, i, vp, i, f, vp, i, vp]),
        "cosma_dgemm_mnk": (i, [i, i, i, d, vp, i, vp, i, d, vp, i, vp]),
        "cosma_sgemm_bf16x3_mnk": (i, [i, i, i, f, vp, i, vp, i, f, vp, i, vp]),
        "cosma_num_tile_configs": (i, [i]),
        "cosma_tile_config": (i, [i, i, C.POINTER(i)]),
        "cosma_heuristic": (i, [i, i, i, i, C.POINTER(i), C.POINTER(i)]),
        "cosma_heuristic_topo": (i, [i, i, i, i, i, i, C.POINTER(i), C.POINTER(i)]),
        "cosma_device_topology": (i, [C.POINTER(i), C.POINTER(i), C.POINTER(i)]),
        "cosma_error_stats_f32": (i, [i, i, vp, i64, vp, i64, f64p, vp]),
        "cosma_error_stats_f64": (i, [i, i, vp, i64, vp, i64, f64p, vp]),
        "cosma_error_stats_ex_f32": (i, [i, i, vp, i64, vp, i64, d, f64p, vp]),
        "cosma_error_stats_ex_f64": (i, [i, i, vp, i64, vp, i64, d, f64p, vp]),
        "cosma_sumsq_f32": (i, [i, i, vp, i64, f64p, vp]),
        "cosma_sumsq_f64": (i, [i, i, vp, i64, f64p, vp]),
        "cosma_convert_f32_f64": (i, [i, i, vp, i64, vp, i64, vp]),
        "cosma_grouped_plan_create": (i, [i, i, C.POINTER(GroupProblem), C.POINTER(i), i,
                                          C.POINTER(vp)]),
        "cosma_grouped_plan_create_ex": (i, [i, i, C.POINTER(GroupProblem), C.POINTER(i), C.POINTER(i), i,
                                             C.POINTER(vp)]),
        "cosma_grouped_plan_choice": (i, [vp, i, C.POINTER(i), C.POINTER(i)]),
        "cosma_group_families": (i, [i, i, C.POINTER(i)]),
        "cosma_grouped_autotune": (i, [i, i, C.POINTER(GroupProblem), i, i, C.POINTER(i), C.POINTER(i),
                                       C.POINTER(GroupTuneResult), i, C.POINTER(i)]),
        "cosma_grouped_plan_run": (i, [vp, vp]),
        "cosma_grouped_plan_run_graph": (i, [vp, vp]),
        "cosma_grouped_plan_info": (i, [vp, C.POINTER(i), C.POINTER(i), C.POINTER(d)]),
        "cosma_grouped_plan_bucket": (i, [vp, i, C.POINTER(i), C.POINTER(i), C.POINTER(i),
                                          C.POINTER(i)]),
        "cosma_grouped_plan_destroy": (i, [vp]),
        "cosma_benchmark": (i, [C.POINTER(GemmDesc), i, i, C.c_size_t, C.POINTER(f)]),
        "cosma_tune_set": (i, [i] * 9),
        "cosma_tune_lookup": (i, [i] * 6 + [C.POINTER(i)] * 3),
        "cosma_tune_clear": (i, []),
        "cosma_planes_bytes": (C.c_size_t, [i, i]),
        "cosma_split_operand": (i, [i, i, i, i, vp, i64, vp, vp]),
        "cosma_status_string": (C.c_char_p, [i]),
        "cosma_version": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    del f32p
    return lib


_lib = None
_lib = _load()
lib = _lib


def check(status: int, what: str = "") -> None:
    if status != STATUS_SUCCESS:
        raise CosmaError(status, what)


def tile_configs(mode: int) -> list[dict]:
    """Tile configurations built into the library for a mode."""
    out = []
    arr = (C.c_int * 8)()
    for cfg in range(lib.cosma_num_tile_configs(mode)):
        check(lib.cosma_tile_config(mode, cfg, arr), "cosma_tile_config")
        bm, bn, bk, wm, wn, thr, lds, pipe = list(arr)
        out.append(dict(cfg=cfg, BM=bm, BN=bn, BK=bk, WM=wm, WN=wn, threads=thr, lds_bytes=lds,
                        pipe=pipe))
    return out


def heuristic(mode: int, m: int, n: int, k: int, *, cus: int | None = None,
              trans_a: bool = False) -> tuple[int, int]:
    """(tile_cfg, split_k) the library picks for a shape: on the current
    device, or on a device of `cus` compute units."""
    cfg, split = C.c_int(), C.c_int()
    if cus is None and not trans_a:
        check(lib.cosma_heuristic(mode, m, n, k, C.byref(cfg), C.byref(split)), "cosma_heuristic")
    else:
        if cus is None:
            cus = device_topology()["cus"]
        check(lib.cosma_heuristic_topo(mode, m, n, k, int(trans_a), int(cus), C.byref(cfg), C.byref(split)),
              "cosma_heuristic_topo")
    return cfg.value, split.value


def device_topology() -> dict:
    """Compute units, XCDs and peak engine clock (kHz) of the current device;
    queried=False means no device answered and the nominal MI355X figures
    stand in."""
    cus, xcds, clk = C.c_int(), C.c_int(), C.c_int()
    st = lib.cosma_device_topology(C.byref(cus), C.byref(xcds), C.byref(clk))
    return dict(cus=cus.value, xcds=xcds.value, clock_khz=clk.value, queried=st == 0)


# Dense MFMA flops per clock per CU on gfx950 (FP32 v_mfma_f32_16x16x4_f32,
# FP64 v_mfma_f64_16x16x4_f64, BF16 v_mfma_f32_16x16x32_bf16).
MFMA_FLOPS_PER_CLK_CU = {"fp32": 256, "fp64": 128, "bf16": 4096}


def peak_tflops(kind: str, topo: dict | None = None) -> float:
    """Dense MFMA peak of `kind` (fp32, fp64, bf16, bf16x3 = bf16/3 useful,
    bf16x4 = bf16/4) from the device's CU count and peak clock: the published
    MI355X figures (157.3 / 78.6 / 2516.6 TFLOP/s) for 256 CUs at 2.4 GHz."""
    topo = topo or device_topology()
    div = {"bf16x3": 3, "bf16x4": 4}.get(kind, 1)
    base = MFMA_FLOPS_PER_CLK_CU["bf16" if kind.startswith("bf16") else kind]
    return base * topo["cus"] * topo["clock_khz"] * 1e3 / div / 1e12
