"""Torch-facing wrappers over the C ABI.

torch supplies device memory and the current stream; every FLOP runs in the
HIP kernels of libcosma_amd.so.  The operand convention is the reference's:
row-major matrices, C = alpha * A @ B + beta * C
(Kernel/cuda/src/cuCOSMAV100.cuh:2751-2776).
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import EPILOGUES, MODES, GemmDesc, check, lib

_DTYPE = {"fp32": torch.float32, "bf16x3": torch.float32, "bf16x4": torch.float32,
          "fp64": torch.float64}


def _stream_ptr(stream) -> int:
    if stream is None:
        stream = torch.cuda.current_stream()
    return int(stream.cuda_stream)


def _check_operand(t: torch.Tensor, name: str, dt, mode: str, device) -> None:
    """dtype of the mode, on the GPU (on `device` when given): the kernels
    read and write raw device pointers."""
    if t.dtype != dt:
        raise TypeError(f"{name} must be {dt} for mode {mode}, got {t.dtype}")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if device is not None and t.device != device:
        raise ValueError(f"{name} is on {t.device}, A on {device}")


def _ld(t: torch.Tensor, name: str) -> int:
    if t.dim() != 2:
        raise ValueError(f"{name} must be 2-D, got {tuple(t.shape)}")
    if t.stride(1) != 1 and t.shape[1] > 1:
        raise ValueError(f"{name} must have unit stride along its rows")
    return max(int(t.stride(0)), int(t.shape[1]), 1)


def gemm(
    A: torch.Tensor,
    B: torch.Tensor,
    C_out: torch.Tensor | None = None,
    *,
    alpha: float = 1.0,
    beta: float = 0.0,
    mode: str = "fp32",
    trans_a: bool = False,
    trans_b: bool = False,
    split_k: int = 0,
    tile_cfg: int = -1,
    group_m: int = 0,
    workspace: torch.Tensor | None = None,
    stream=None,
    epilogue: str | None = None,
    a_planes: torch.Tensor | None = None,
    b_planes: torch.Tensor | None = None,
) -> torch.Tensor:
    """C = act(alpha * op(A) @ op(B) + beta * C) on the GPU.

    op(A) is A (shape M x K) or, with trans_a, A.T of a stored K x M tensor;
    op(B) is B (K x N) or, with trans_b, B.T of a stored N x K tensor.
    mode: "fp32" (FP32 MFMA), "fp64" (FP64 MFMA), "bf16x3" (FP32 operands,
    bf16 hi/lo split, three bf16 MFMA products, FP32 accumulation) or
    "bf16x4" (the same split with the lo*lo product added).
    split_k <= 0 and tile_cfg < 0 select the library heuristic.
    epilogue: None, "relu" (max(x, 0)) or "sigmoid" (1 / (1 + exp(-x))), fused
    into the kernel's store (after the split-K reduction when K is split).
    a_planes / b_planes: split-mode operands already split by split_operand
    (op(A) and op(B) respectively); A and B still describe the problem.
    """
    if mode not in MODES:
        raise ValueError(f"unknown mode {mode!r}")
    if epilogue not in EPILOGUES:
        raise ValueError(f"unknown epilogue {epilogue!r}")
    dt = _DTYPE[mode]
    _check_operand(A, "A", dt, mode, None)
    _check_operand(B, "B", dt, mode, A.device)
    m, k = (A.shape[1], A.shape[0]) if trans_a else (A.shape[0], A.shape[1])
    kb, n = (B.shape[1], B.shape[0]) if trans_b else (B.shape[0], B.shape[1])
    if k != kb:
        raise ValueError(f"inner dimensions differ: {k} vs {kb}")
    if C_out is None:
        if beta != 0.0:
            raise ValueError("beta != 0 needs an input C")
        C_out = torch.empty((m, n), dtype=dt, device=A.device)
    if C_out.dtype != dt or tuple(C_out.shape) != (m, n):
        raise ValueError(f"C must be {dt} of shape {(m, n)}")
    _check_operand(C_out, "C", dt, mode, A.device)
    d = GemmDesc()
    lib.cosma_gemm_desc_init(C.byref(d))
    d.mode = MODES[mode]
    d.m, d.n, d.k = m, n, k
    d.trans_a, d.trans_b = int(trans_a), int(trans_b)
    d.a, d.lda = A.data_ptr(), _ld(A, "A")
    d.b, d.ldb = B.data_ptr(), _ld(B, "B")
    d.c, d.ldc = C_out.data_ptr(), _ld(C_out, "C")
    d.alpha, d.beta = float(alpha), float(beta)
    d.split_k, d.tile_cfg, d.group_m = int(split_k), int(tile_cfg), int(group_m)
    d.epilogue = EPILOGUES[epilogue]
    for t, name, rows in ((a_planes, "a_planes", m), (b_planes, "b_planes", n)):
        if t is None:
            continue
        if not (t.is_cuda and t.is_contiguous()) or t.device != A.device:
            raise ValueError(f"{name} must be a contiguous tensor on {A.device} from split_operand")
        need = int(lib.cosma_planes_bytes(rows, k))
        if t.numel() * t.element_size() < need:
            raise ValueError(f"{name} holds {t.numel() * t.element_size()} bytes, the planes of a "
                             f"{rows} x {k} operand need {need}")
    d.a_planes = a_planes.data_ptr() if a_planes is not None else None
    d.b_planes = b_planes.data_ptr() if b_planes is not None else None
    if workspace is not None:
        d.workspace = workspace.data_ptr()
        d.workspace_bytes = workspace.numel() * workspace.element_size()
    d.stream = _stream_ptr(stream)
    check(lib.cosma_gemm_ex(C.byref(d)), "cosma_gemm_ex")
    return C_out


def workspace_bytes(m: int, n: int, k: int, *, mode: str = "fp32", split_k: int = 0,
                    tile_cfg: int = -1) -> int:
    """Split-K workspace a call of this shape needs (no device work)."""
    d = GemmDesc()
    lib.cosma_gemm_desc_init(C.byref(d))
    d.mode, d.m, d.n, d.k = MODES[mode], m, n, k
    # Placeholder operands that satisfy the contract: the query never touches them.
    d.a = d.b = d.c = 16
    d.lda = ((max(k, 1) + 3) // 4) * 4
    d.ldb = ((max(n, 1) + 3) // 4) * 4
    d.ldc = max(n, 1)
    d.split_k, d.tile_cfg = split_k, tile_cfg
    out = C.c_size_t()
    check(lib.cosma_workspace_query(C.byref(d), C.byref(out)), "cosma_workspace_query")
    return int(out.value)


def release_workspace() -> int:
    """Free the current device's library-owned workspace buffers (one per
    stream that ran a split-K or pre-split call); returns the bytes freed.
    Synchronises the device."""
    out = C.c_size_t()
    check(lib.cosma_workspace_release(C.byref(out)), "cosma_workspace_release")
    return int(out.value)


def sgemm(A, B, C_out=None, alpha=1.0, beta=0.0, **kw):
    return gemm(A, B, C_out, alpha=alpha, beta=beta, mode="fp32", **kw)


def sgemm_bf16x4(A, B, C_out=None, alpha=1.0, beta=0.0, **kw):
    return gemm(A, B, C_out, alpha=alpha, beta=beta, mode="bf16x4", **kw)


def sgemm_bf16x3(A, B, C_out=None, alpha=1.0, beta=0.0, **kw):
    return gemm(A, B, C_out, alpha=alpha, beta=beta, mode="bf16x3", **kw)


def dgemm(A, B, C_out=None, alpha=1.0, beta=0.0, **kw):
    return gemm(A, B, C_out, alpha=alpha, beta=beta, mode="fp64", **kw)


def error_stats(Cm: torch.Tensor, R: torch.Tensor, stream=None, rel_floor: float = 1e-3) -> dict:
    """FP64 residual statistics of C against a float64 reference R, on the
    GPU.  max_rel is max |C-R|/|R| over the entries with |R| >= rel_floor *
    max|R| (rel_floor 0: every R != 0), so near-zero reference entries do
    not dominate it; max_rel_all is the unfloored statistic."""
    if R.dtype != torch.float64:
        raise TypeError("R must be float64")
    out = (C.c_double * 5)()
    rows, cols = Cm.shape
    fn = lib.cosma_error_stats_ex_f32 if Cm.dtype == torch.float32 else lib.cosma_error_stats_ex_f64
    args = (rows, cols, Cm.data_ptr(), _ld(Cm, "C"), R.data_ptr(), _ld(R, "R"))
    check(fn(*args, 0.0, out, _stream_ptr(stream)), "cosma_error_stats")
    s2, r2, mx, mr_all, rmax = list(out)
    mr = mr_all
    if rel_floor > 0:
        check(fn(*args, rel_floor * rmax, out, _stream_ptr(stream)), "cosma_error_stats")
        mr = out[3]
    return dict(sum_sq_err=s2, sum_sq_ref=r2, max_abs=mx, max_rel=mr, max_rel_all=mr_all, max_ref=rmax,
                rel_floor=rel_floor)


def sumsq(X: torch.Tensor, stream=None) -> float:
    """Sum of squares of X in FP64, on the GPU."""
    out = (C.c_double * 5)()
    rows, cols = X.shape
    fn = lib.cosma_sumsq_f32 if X.dtype == torch.float32 else lib.cosma_sumsq_f64
    check(fn(rows, cols, X.data_ptr(), _ld(X, "X"), out, _stream_ptr(stream)), "cosma_sumsq")
    return float(out[0])


def to_f64(X: torch.Tensor, stream=None) -> torch.Tensor:
    """float32 -> float64 copy on the GPU."""
    rows, cols = X.shape
    Y = torch.empty((rows, cols), dtype=torch.float64, device=X.device)
    check(lib.cosma_convert_f32_f64(rows, cols, X.data_ptr(), _ld(X, "X"), Y.data_ptr(), cols,
                                    _stream_ptr(stream)), "cosma_convert_f32_f64")
    return Y


def normwise_error(A: torch.Tensor, B: torch.Tensor, Cm: torch.Tensor, R: torch.Tensor,
                   stream=None) -> dict:
    """||C - R||_F / (||A||_F ||B||_F) and friends, all in FP64 on the GPU."""
    st = error_stats(Cm, R, stream)
    na = sumsq(A, stream) ** 0.5
    nb = sumsq(B, stream) ** 0.5
    st["normwise"] = (st["sum_sq_err"] ** 0.5) / (na * nb) if na * nb > 0 else 0.0
    st["rel_fro"] = (st["sum_sq_err"] / st["sum_sq_ref"]) ** 0.5 if st["sum_sq_ref"] > 0 else 0.0
    return st


def fp64_product(A: torch.Tensor, B: torch.Tensor, C0: torch.Tensor | None = None,
               alpha: float = 1.0, beta: float = 0.0, stream=None) -> torch.Tensor:
    """FP64 product of FP32 operands on the GPU (FP64 MFMA on exact float->double
    copies): the device-side FP64 reference used for error reports at sizes the host
    oracle cannot finish; checked against the host oracle on row samples."""
    A64, B64 = to_f64(A, stream), to_f64(B, stream)
    R = to_f64(C0, stream) if (C0 is not None and beta != 0.0) else None
    return gemm(A64, B64, R, alpha=alpha, beta=beta if R is not None else 0.0, mode="fp64",
                stream=stream)


def split_operand(X: torch.Tensor, *, trans: bool = False, as_b: bool = False, mode: str = "bf16x3",
                  stream=None) -> torch.Tensor:
    """Split an FP32 operand once into the bf16 hi/lo planes of the pre-split
    path; pass the result as gemm(..., a_planes=) or b_planes=.

    As A (as_b False): X is A (M x K), or with trans its stored K x M form.
    As B (as_b True): X is B (K x N), or with trans its stored N x K form.
    Returns a bfloat16 tensor of rows x 2*Kp (rows = M or N, Kp = K rounded
    up to 32).
    """
    if mode not in ("bf16x3", "bf16x4"):
        raise ValueError("split_operand is for the split modes")
    if X.dtype != torch.float32 or not X.is_cuda:
        raise TypeError("X must be a float32 device tensor")
    if as_b:
        rows, k = (X.shape[0], X.shape[1]) if trans else (X.shape[1], X.shape[0])
        kmajor = trans
    else:
        rows, k = (X.shape[1], X.shape[0]) if trans else (X.shape[0], X.shape[1])
        kmajor = not trans
    kp = (k + 31) // 32 * 32
    out = torch.empty((rows, 2 * kp), dtype=torch.bfloat16, device=X.device)
    check(lib.cosma_split_operand(MODES[mode], int(kmajor), rows, k, X.data_ptr(), _ld(X, "X"),
                                  out.data_ptr(), _stream_ptr(stream)), "cosma_split_operand")
    return out
