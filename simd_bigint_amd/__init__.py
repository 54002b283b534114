"""simd_bigint_amd -- MI355X (gfx950) large-operand integer multiply.

Python mirror of the reference's mpn-style C entries over the C ABI of
``lib/libsimd_bigint_amd.so`` (declared in ``include/simd_bigint_amd.h``).
Operands are numpy ``uint64`` limb arrays, little-endian, as in the
reference (``include/ntmul.h:47-79``, ``include/u64cvt.h:94-132``,
``include/divrem.h:218-221``). Device-resident entries take torch CUDA
(HIP) tensors of dtype ``torch.uint64`` (or int64 viewed as limbs); torch is
plumbing for device memory and streams only.

The library has no CPU fallback for the GPU path: if the shared library is
missing, importing the loader raises; if no HIP device is present, the GPU
entries abort with a message. A HIP failure inside a host entry is reported
and passed up: nt_* return 0, simd_mpn_mul / divrem finish that product on
the library's CPU tier and count it (``gpu_fallbacks``).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsimd_bigint_amd.so")

SBA_OK = 1
SBA_BAND = 0
SBA_ERR_HIP = -1
SBA_ERR_WORKSPACE = -2
SBA_ERR_NODEV = -3
SBA_ERR_TOO_LARGE = -4
NOALLOC_MAX_LG = 20

_lib: Optional[ctypes.CDLL] = None

_u64p = ctypes.POINTER(ctypes.c_uint64)


def _arr(a: np.ndarray):
    return a.ctypes.data_as(_u64p)


def lib() -> ctypes.CDLL:
    """Load the shared library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"simd_bigint_amd: {LIB_PATH} is missing; build it first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    # One HIP runtime per process: when torch (ROCm build) is importable, load
    # it first so this library binds to the same libamdhip64.so.7 instance as
    # torch's device tensors and streams (both carry that SONAME; loading ours
    # first would leave torch with a second runtime that sees no GPU).
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    L = ctypes.CDLL(LIB_PATH)
    sp = ctypes.c_ssize_t
    u64 = ctypes.c_uint64
    vp = ctypes.c_void_p
    for name in ("nt_mul", "p30_mul", "p50_mul"):
        f = getattr(L, name)
        f.argtypes = [_u64p, _u64p, sp, _u64p, sp]
        f.restype = ctypes.c_int
    for name in ("nt_sqr", "p30_sqr", "p50_sqr"):
        f = getattr(L, name)
        f.argtypes = [_u64p, _u64p, sp]
        f.restype = ctypes.c_int
    for name in ("nt_mul_r", "p30_mul_r", "p50_mul_r"):
        f = getattr(L, name)
        f.argtypes = [_u64p, _u64p, sp, _u64p, sp, vp]
        f.restype = ctypes.c_int
    for name in ("nt_sqr_r", "p30_sqr_r", "p50_sqr_r"):
        f = getattr(L, name)
        f.argtypes = [_u64p, _u64p, sp, vp]
        f.restype = ctypes.c_int
    L.simd_mpn_mul.argtypes = [_u64p, _u64p, u64, _u64p, u64]
    L.simd_mpn_mul.restype = None
    L.sba_cpu_mul.argtypes = [_u64p, _u64p, u64, _u64p, u64]
    L.sba_cpu_mul.restype = None
    for name in ("divrem_u64", "divrem_u64_dc"):
        f = getattr(L, name)
        f.argtypes = [_u64p, _u64p, _u64p, u64, _u64p, u64, _u64p, _u64p]
        f.restype = None
    L.sba_gpu_mul_dev.argtypes = [vp, vp, ctypes.c_int64, vp, ctypes.c_int64, vp]
    L.sba_gpu_mul_dev.restype = ctypes.c_int
    L.sba_gpu_sqr_dev.argtypes = [vp, vp, ctypes.c_int64, vp]
    L.sba_gpu_sqr_dev.restype = ctypes.c_int
    L.sba_gpu_mul_dev_ws.argtypes = [vp, vp, ctypes.c_int64, vp, ctypes.c_int64, vp,
                                     ctypes.c_size_t, vp]
    L.sba_gpu_mul_dev_ws.restype = ctypes.c_int
    L.sba_gpu_sqr_dev_ws.argtypes = [vp, vp, ctypes.c_int64, vp, ctypes.c_size_t, vp]
    L.sba_gpu_sqr_dev_ws.restype = ctypes.c_int
    L.sba_gpu_mul_dev_noalloc.argtypes = [vp, vp, ctypes.c_int64, vp, ctypes.c_int64, vp]
    L.sba_gpu_mul_dev_noalloc.restype = ctypes.c_int
    L.sba_gpu_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.c_int64]
    L.sba_gpu_workspace_bytes.restype = ctypes.c_size_t
    L.sba_gpu_min_operand_cap.argtypes = []
    L.sba_gpu_min_operand_cap.restype = ctypes.c_uint64
    L.sba_gpu_max_lg.argtypes = []
    L.sba_gpu_max_lg.restype = ctypes.c_int
    L.sba_gpu_device_count.argtypes = []
    L.sba_gpu_device_count.restype = ctypes.c_int
    L.sba_batch_create.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.sba_batch_create.restype = vp
    L.sba_batch_run.argtypes = [vp, vp]
    L.sba_batch_run.restype = ctypes.c_int
    L.sba_batch_classes.argtypes = [vp]
    L.sba_batch_classes.restype = ctypes.c_int
    L.sba_batch_destroy.argtypes = [vp]
    L.sba_batch_destroy.restype = None
    L.sba_batch_ws_bytes.argtypes = [vp, ctypes.c_int]
    L.sba_batch_ws_bytes.restype = ctypes.c_size_t
    L.sba_batch_launch_ws.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t, vp]
    L.sba_batch_launch_ws.restype = ctypes.c_int
    L.sba_gpu_mul_loop.argtypes = [vp, ctypes.c_int, vp]
    L.sba_gpu_mul_loop.restype = ctypes.c_int
    ip = ctypes.POINTER(ctypes.c_int)
    fp = ctypes.POINTER(ctypes.c_float)
    L.sba_plan_bucket.argtypes = [ctypes.c_int]
    L.sba_plan_bucket.restype = ctypes.c_int
    L.sba_plan_set.argtypes = [ctypes.c_int] * 6
    L.sba_plan_set.restype = ctypes.c_int
    L.sba_plan_get.argtypes = [ctypes.c_int] * 3 + [ip] * 3
    L.sba_plan_get.restype = ctypes.c_int
    L.sba_plan_heuristic.argtypes = [ctypes.c_int] + [ip] * 3
    L.sba_plan_heuristic.restype = ctypes.c_int
    L.sba_plan_reset.argtypes = []
    L.sba_plan_reset.restype = None
    L.sba_plan_candidates.argtypes = [ctypes.c_int, ctypes.c_int, ip, ctypes.c_int]
    L.sba_plan_candidates.restype = ctypes.c_int
    L.sba_class_time.argtypes = [vp, ctypes.c_int] + [ctypes.c_int] * 4 + [vp, ctypes.c_size_t,
                                                                            vp, fp]
    L.sba_class_time.restype = ctypes.c_int
    L.sba_job_class.argtypes = [vp, ip, ip]
    L.sba_job_class.restype = ctypes.c_int
    L.sba_autotune.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp,
                               ctypes.c_size_t, vp, ip]
    L.sba_autotune.restype = ctypes.c_int
    L.sba_cm_nparams.argtypes = []
    L.sba_cm_nparams.restype = ctypes.c_int
    L.sba_cm_ws_floats.argtypes = [ctypes.c_int]
    L.sba_cm_ws_floats.restype = ctypes.c_size_t
    L.sba_cm_forward.argtypes = [vp, vp, ctypes.c_int, vp, vp]
    L.sba_cm_forward.restype = ctypes.c_int
    L.sba_cm_backward.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, vp]
    L.sba_cm_backward.restype = ctypes.c_int
    L.sba_cm_adam.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_float, vp]
    L.sba_cm_adam.restype = ctypes.c_int
    L.sba_cm_features.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, fp]
    L.sba_cm_features.restype = ctypes.c_int
    L.sba_cm_load.argtypes = [fp, fp, fp, ctypes.c_float, ctypes.c_float]
    L.sba_cm_load.restype = ctypes.c_int
    L.sba_cm_loaded.argtypes = []
    L.sba_cm_loaded.restype = ctypes.c_int
    L.sba_cm_predict.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, vp,
                                 ctypes.POINTER(ctypes.c_double)]
    L.sba_cm_predict.restype = ctypes.c_int
    L.sba_gpu_in_band.argtypes = [ctypes.c_int64, ctypes.c_int64]
    L.sba_gpu_in_band.restype = ctypes.c_int
    L.sba_ref_nt_band.argtypes = [ctypes.c_int64, ctypes.c_int64]
    L.sba_ref_nt_band.restype = ctypes.c_int
    L.sba_gpu_chunking.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ip,
                                   ctypes.POINTER(u64), ctypes.POINTER(u64)]
    L.sba_gpu_chunking.restype = ctypes.c_int
    L.sba_gpu_radix3.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
    L.sba_gpu_radix3.restype = ctypes.c_int
    L.sba_gpu_carry_blocks.argtypes = [vp, vp, u64, vp]
    L.sba_gpu_carry_blocks.restype = ctypes.c_int
    L.sba_gpu_carry_block_limbs.argtypes = []
    L.sba_gpu_carry_block_limbs.restype = ctypes.c_uint32
    L.sba_gpu_probe_butterflies.argtypes = [ctypes.POINTER(ctypes.c_double)]
    L.sba_gpu_probe_butterflies.restype = ctypes.c_int
    L.sba_set_gpu_threshold.argtypes = [u64]
    L.sba_set_gpu_threshold.restype = None
    L.sba_get_gpu_threshold.argtypes = []
    L.sba_get_gpu_threshold.restype = u64
    L.sba_set_school_max.argtypes = [u64]
    L.sba_set_school_max.restype = None
    L.sba_get_school_max.argtypes = []
    L.sba_get_school_max.restype = u64
    L.sba_tune_sizes.argtypes = [ctypes.POINTER(u64), ctypes.c_int]
    L.sba_tune_sizes.restype = ctypes.c_int
    L.sba_tune_gpu_threshold.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double)]
    L.sba_tune_gpu_threshold.restype = u64
    L.sba_gpu_school.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
    L.sba_gpu_school.restype = ctypes.c_int
    L.sba_set_school_long_ratio.argtypes = [ctypes.c_uint32]
    L.sba_set_school_long_ratio.restype = None
    L.sba_get_school_long_ratio.argtypes = []
    L.sba_get_school_long_ratio.restype = ctypes.c_uint32
    L.sba_set_r3_min_lg.argtypes = [ctypes.c_int]
    L.sba_set_r3_min_lg.restype = None
    L.sba_get_r3_min_lg.argtypes = []
    L.sba_get_r3_min_lg.restype = ctypes.c_int
    L.sba_gpu_truncated.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ip, ip, ip]
    L.sba_gpu_truncated.restype = ctypes.c_int
    L.sba_set_trunc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.sba_set_trunc.restype = None
    L.sba_get_trunc.argtypes = [ip, ip, ip]
    L.sba_get_trunc.restype = None
    L.sba_gpu_fallbacks.argtypes = []
    L.sba_gpu_fallbacks.restype = u64
    L.sba_test_inject_failure.argtypes = [ctypes.c_int, ctypes.c_int]
    L.sba_test_inject_failure.restype = None
    _lib = L
    return L


FAIL_ENTRY = 0
FAIL_ALLOC = 1


def truncated(an: int, bn: int, sqr: bool = False):
    """(L, last, rows) when a single product runs truncated, else None."""
    Lg, last, rows = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    if lib().sba_gpu_truncated(an, bn, 1 if sqr else 0, ctypes.byref(Lg), ctypes.byref(last),
                               ctypes.byref(rows)):
        return Lg.value, last.value, rows.value
    return None


def set_trunc(max_fill_pct: int, r3_min_pct: int, min_lg: int = 21) -> None:
    """Truncation thresholds (percent of last-pass tiles, percent of the
    power-of-two length, smallest log2 length); max_fill_pct 0 turns
    truncation off."""
    lib().sba_set_trunc(int(max_fill_pct), int(r3_min_pct), int(min_lg))


def get_trunc():
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().sba_get_trunc(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


def school_route(an: int, bn: int, sqr: bool = False) -> int:
    """1: one-workgroup schoolbook, 2: short x long schoolbook, 0: NTT."""
    return int(lib().sba_gpu_school(an, bn, 1 if sqr else 0))


def set_school_long_ratio(sixteenths: int) -> None:
    """Short x long schoolbook when an * bn <= sixteenths / 16 * N log2 N;
    0 turns it off."""
    lib().sba_set_school_long_ratio(int(sixteenths))


def get_school_long_ratio() -> int:
    return int(lib().sba_get_school_long_ratio())


def set_r3_min_lg(L: int) -> None:
    """Smallest radix-3 row length 2^L for single products."""
    lib().sba_set_r3_min_lg(int(L))


def get_r3_min_lg() -> int:
    return int(lib().sba_get_r3_min_lg())


def gpu_fallbacks() -> int:
    """Products simd_mpn_mul or the division finished on the CPU tier after
    a HIP failure, since the process started."""
    return int(lib().sba_gpu_fallbacks())


def inject_failure(where: int, count: int) -> None:
    """Test hook: the next `count` GPU multiplies of the host entries fail
    before touching the device (FAIL_ENTRY), or the next `count` staging
    allocations fail (FAIL_ALLOC)."""
    lib().sba_test_inject_failure(int(where), int(count))


def _as_limbs(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a


# ---------------------------------------------------------------------------
# host-memory entries (reference semantics)
# ---------------------------------------------------------------------------
def nt_mul(a, b) -> Tuple[int, np.ndarray]:
    """include/ntmul.h:71-74: returns (1, product) or (0, zeros) beyond band."""
    a, b = _as_limbs(a), _as_limbs(b)
    r = np.zeros(len(a) + len(b), dtype=np.uint64)
    ok = lib().nt_mul(_arr(r), _arr(a), len(a), _arr(b), len(b))
    return ok, r


def nt_sqr(a) -> Tuple[int, np.ndarray]:
    """include/ntmul.h:76-78."""
    a = _as_limbs(a)
    r = np.zeros(2 * len(a), dtype=np.uint64)
    ok = lib().nt_sqr(_arr(r), _arr(a), len(a))
    return ok, r


def p30_mul(a, b) -> Tuple[int, np.ndarray]:
    """include/p30.h:1940 (p30's band, include/p30.h:1851-1858)."""
    a, b = _as_limbs(a), _as_limbs(b)
    r = np.zeros(len(a) + len(b), dtype=np.uint64)
    return lib().p30_mul(_arr(r), _arr(a), len(a), _arr(b), len(b)), r


def p50_mul(a, b) -> Tuple[int, np.ndarray]:
    """include/p50.h:1096."""
    a, b = _as_limbs(a), _as_limbs(b)
    r = np.zeros(len(a) + len(b), dtype=np.uint64)
    return lib().p50_mul(_arr(r), _arr(a), len(a), _arr(b), len(b)), r


def p30_sqr(a) -> Tuple[int, np.ndarray]:
    """include/p30.h:2018."""
    a = _as_limbs(a)
    r = np.zeros(2 * len(a), dtype=np.uint64)
    return lib().p30_sqr(_arr(r), _arr(a), len(a)), r


def p50_sqr(a) -> Tuple[int, np.ndarray]:
    """include/p50.h:1176."""
    a = _as_limbs(a)
    r = np.zeros(2 * len(a), dtype=np.uint64)
    return lib().p50_sqr(_arr(r), _arr(a), len(a)), r


def mul_r(name: str, a, b=None) -> Tuple[int, np.ndarray]:
    """The reference's `_r` entries (nt_mul_r, nt_sqr_r, p30_mul_r,
    p30_sqr_r, p50_mul_r, p50_sqr_r: include/ntmul.h:47-69,
    include/p30.h:1847,1950, include/p50.h:1020,1106) with a NULL arena;
    b None for the squares."""
    a = _as_limbs(a)
    f = getattr(lib(), name)
    if b is None:
        r = np.zeros(2 * len(a), dtype=np.uint64)
        return f(_arr(r), _arr(a), len(a), None), r
    b = _as_limbs(b)
    r = np.zeros(len(a) + len(b), dtype=np.uint64)
    return f(_arr(r), _arr(a), len(a), _arr(b), len(b), None), r


def simd_mpn_mul(a, b) -> np.ndarray:
    """include/u64cvt.h:94-132: void, writes exactly an+bn limbs."""
    a, b = _as_limbs(a), _as_limbs(b)
    r = np.zeros(len(a) + len(b), dtype=np.uint64)
    lib().simd_mpn_mul(_arr(r), _arr(a), len(a), _arr(b), len(b))
    return r


def cpu_mul(a, b) -> np.ndarray:
    a, b = _as_limbs(a), _as_limbs(b)
    r = np.zeros(len(a) + len(b), dtype=np.uint64)
    lib().sba_cpu_mul(_arr(r), _arr(a), len(a), _arr(b), len(b))
    return r


def divrem_u64(n, d, dc: bool = False) -> Tuple[np.ndarray, np.ndarray]:
    """include/divrem.h:218-221 (dc=True: include/dc_div.h:154-157).
    Returns (Q trimmed to *qn64_out limbs, R trimmed to *rn64_out limbs)."""
    n, d = _as_limbs(n), _as_limbs(d)
    if len(d) == 0 or d[-1] == 0:
        raise ValueError("divisor's top limb must be nonzero")
    qm = max(len(n) - len(d) + 1, 1)
    q = np.zeros(qm, dtype=np.uint64)
    r = np.zeros(len(d), dtype=np.uint64)
    qn = ctypes.c_uint64()
    rn = ctypes.c_uint64()
    f = lib().divrem_u64_dc if dc else lib().divrem_u64
    f(_arr(q), _arr(r), _arr(n), len(n), _arr(d), len(d), ctypes.byref(qn), ctypes.byref(rn))
    return q[: qn.value], r[: rn.value]


def set_gpu_threshold(min_total_limbs: int) -> None:
    lib().sba_set_gpu_threshold(int(min_total_limbs))


def get_gpu_threshold() -> int:
    return int(lib().sba_get_gpu_threshold())


def set_school_max(max_product: int) -> None:
    """Products with an + bn <= 2048 take the schoolbook kernel when
    an * bn <= max_product * N / 1024 in a batch, half that in a single call
    (N = their NTT length, at least 1024); 0 turns that tier off."""
    lib().sba_set_school_max(int(max_product))


def get_school_max() -> int:
    return int(lib().sba_get_school_max())


def tune_gpu_threshold(apply: bool = False) -> dict:
    """Measure the CPU/GPU crossover of simd_mpn_mul (balanced products on
    host buffers, copies included) and optionally install it."""
    n = lib().sba_tune_sizes(None, 0)
    sizes = (ctypes.c_uint64 * n)()
    lib().sba_tune_sizes(sizes, n)
    cpu = (ctypes.c_double * n)()
    gpu = (ctypes.c_double * n)()
    t = int(lib().sba_tune_gpu_threshold(1 if apply else 0, cpu, gpu))
    return {"threshold": t, "totals": list(sizes), "cpu_ms": list(cpu), "gpu_ms": list(gpu)}


# ---------------------------------------------------------------------------
# device-resident entries (torch tensors as device buffers)
# ---------------------------------------------------------------------------
def _ptr(t) -> int:
    return int(t.data_ptr())


def _check_dev(r, need):
    if not r.is_contiguous() or r.numel() < need:
        raise ValueError("result tensor must be contiguous with >= %d limbs" % need)


def mul_dev(r, a, b, stream=None) -> int:
    """Enqueue r = a * b on device tensors (uint64/int64 limbs)."""
    _check_dev(r, a.numel() + b.numel())
    s = None if stream is None else ctypes.c_void_p(int(stream.cuda_stream))
    return lib().sba_gpu_mul_dev(_ptr(r), _ptr(a), a.numel(), _ptr(b), b.numel(), s)


def mul_dev_noalloc(r, a, b, stream=None) -> int:
    """Enqueue r = a * b through the allocation-free entry (products of up
    to 2^NOALLOC_MAX_LG coefficients; the device's static scratch is shared
    by every such call: overlapping calls race)."""
    _check_dev(r, a.numel() + b.numel())
    s = None if stream is None else ctypes.c_void_p(int(stream.cuda_stream))
    return lib().sba_gpu_mul_dev_noalloc(_ptr(r), _ptr(a), a.numel(), _ptr(b), b.numel(), s)


def probe_butterflies() -> float:
    """Radix-2 butterflies per second the current GPU issues from registers
    alone (the VALU ceiling of the transform kernels)."""
    r = ctypes.c_double(0.0)
    rc = lib().sba_gpu_probe_butterflies(ctypes.byref(r))
    if rc != 1:
        raise RuntimeError("simd_bigint_amd: butterfly probe failed (%d)" % rc)
    return r.value


def carry_block_limbs() -> int:
    """Limbs per block of the carry stage (the Garner stage's block)."""
    return int(lib().sba_gpu_carry_block_limbs())


def carry_blocks(r, bmaps, stream=None) -> int:
    """The carry stage between blocks of carry_block_limbs() limbs on device
    tensors: r holds limbs finished with block carry-in 0, bmaps the blocks'
    2-state carry maps (one 4-byte word per block)."""
    lpb = carry_block_limbs()
    nb = (r.numel() + lpb - 1) // lpb
    if not (r.is_cuda and bmaps.is_cuda):
        raise ValueError("carry_blocks needs device tensors")
    if bmaps.numel() < nb or bmaps.element_size() != 4 or r.element_size() != 8:
        raise ValueError("carry_blocks: r of 8-byte limbs, bmaps of nb 4-byte words")
    s = None if stream is None else ctypes.c_void_p(int(stream.cuda_stream))
    return lib().sba_gpu_carry_blocks(_ptr(r), _ptr(bmaps), r.numel(), s)


def sqr_dev(r, a, stream=None) -> int:
    _check_dev(r, 2 * a.numel())
    s = None if stream is None else ctypes.c_void_p(int(stream.cuda_stream))
    return lib().sba_gpu_sqr_dev(_ptr(r), _ptr(a), a.numel(), s)


def chunking(an: int, bn: int, sqr: bool = False):
    """(Lc, chunk limbs, chunks) of a single product; chunks == 0: one
    transform. None outside the device band."""
    Lc, C, n = ctypes.c_int(), ctypes.c_uint64(), ctypes.c_uint64()
    if lib().sba_gpu_chunking(an, bn, 1 if sqr else 0, ctypes.byref(Lc), ctypes.byref(C),
                              ctypes.byref(n)) != SBA_OK:
        return None
    return Lc.value, C.value, n.value


def min_operand_cap() -> int:
    return int(lib().sba_gpu_min_operand_cap())


def max_lg() -> int:
    return int(lib().sba_gpu_max_lg())


# ---------------------------------------------------------------------------
# batches
# ---------------------------------------------------------------------------
class SbaJob(ctypes.Structure):
    _fields_ = [("rp", ctypes.c_void_p), ("ap", ctypes.c_void_p), ("an", ctypes.c_int64),
                ("bp", ctypes.c_void_p), ("bn", ctypes.c_int64)]


def job_array(jobs):
    """(r, a, b) torch tensors -> ctypes array of sba_job (b None: square)."""
    arr = (SbaJob * len(jobs))()
    for i, (r, a, b) in enumerate(jobs):
        need = a.numel() + (b.numel() if b is not None else a.numel())
        _check_dev(r, need)
        arr[i] = SbaJob(_ptr(r), _ptr(a), a.numel(), None if b is None else _ptr(b),
                        0 if b is None else b.numel())
    return arr


def mul_loop(jobs, stream=None) -> int:
    """The plain per-call path over the jobs (one GPU call each)."""
    arr = job_array(jobs)
    s = None if stream is None else ctypes.c_void_p(int(stream.cuda_stream))
    return lib().sba_gpu_mul_loop(ctypes.cast(arr, ctypes.c_void_p), len(jobs), s)


class Batch:
    """Many independent device-resident multiplies (or squares) of uneven
    sizes, launched together (sba_batch_create / sba_batch_run). jobs is a
    list of (r, a, b) torch tensors; b = None means a square of a."""

    def __init__(self, jobs, use_graph: bool = True):
        self._keep = jobs
        arr = job_array(jobs)
        st = ctypes.c_int()
        self._h = lib().sba_batch_create(ctypes.cast(arr, ctypes.c_void_p), len(jobs),
                                         1 if use_graph else 0, ctypes.byref(st))
        if not self._h:
            raise RuntimeError("sba_batch_create failed with status %d" % st.value)
        self.classes = lib().sba_batch_classes(self._h)

    def run(self, stream=None) -> None:
        s = None if stream is None else ctypes.c_void_p(int(stream.cuda_stream))
        rc = lib().sba_batch_run(self._h, s)
        if rc != SBA_OK:
            raise RuntimeError("sba_batch_run returned %d" % rc)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().sba_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
