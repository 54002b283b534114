"""ctypes view of the CPU oracle (oracle/liboracle.so) -- the checker."""
import ctypes
import os

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
_o = None


def lib():
    global _o
    if _o is None:
        o = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
        o.orc_nt_mul.argtypes = [u64p, u64p, ctypes.c_ssize_t, u64p, ctypes.c_ssize_t]
        o.orc_nt_mul.restype = ctypes.c_int
        o.orc_nt_sqr.argtypes = [u64p, u64p, ctypes.c_ssize_t]
        o.orc_nt_sqr.restype = ctypes.c_int
        o.orc_nt_nprimes.argtypes = [ctypes.c_ssize_t, ctypes.c_ssize_t]
        o.orc_nt_nprimes.restype = ctypes.c_int
        o.orc_mul_school.argtypes = [u64p, u64p, ctypes.c_size_t, u64p, ctypes.c_size_t]
        o.orc_mul_school.restype = None
        o.orc_residues.argtypes = [u32p, u64p, ctypes.c_size_t, ctypes.c_uint32]
        o.orc_residues.restype = None
        o.orc_fingerprint.argtypes = [u64p, ctypes.c_size_t, ctypes.c_uint64]
        o.orc_fingerprint.restype = ctypes.c_uint64
        o.orc_divrem.argtypes = [u64p, u64p, u64p, ctypes.c_size_t, u64p, ctypes.c_size_t,
                                 ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        o.orc_divrem.restype = None
        o.orc_fill_xorshift.argtypes = [u64p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint64)]
        o.orc_fill_xorshift.restype = None
        _o = o
    return _o


def nt_mul(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    r = np.zeros(len(a) + len(b), np.uint64)
    ok = lib().orc_nt_mul(r, a, len(a), b, len(b))
    assert ok == 1, "oracle out of band"
    return r


def nt_sqr(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    r = np.zeros(2 * len(a), np.uint64)
    assert lib().orc_nt_sqr(r, a, len(a)) == 1
    return r


def school(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    r = np.zeros(len(a) + len(b), np.uint64)
    lib().orc_mul_school(r, a, len(a), b, len(b))
    return r


def fingerprint(a, q):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return int(lib().orc_fingerprint(a, len(a), q))


def divrem(n, d):
    n = np.ascontiguousarray(n, dtype=np.uint64)
    d = np.ascontiguousarray(d, dtype=np.uint64)
    q = np.zeros(max(len(n) - len(d) + 1, 1), np.uint64)
    r = np.zeros(len(d), np.uint64)
    qn = ctypes.c_size_t()
    rn = ctypes.c_size_t()
    lib().orc_divrem(q, r, n, len(n), d, len(d), ctypes.byref(qn), ctypes.byref(rn))
    return q[: qn.value], r[: rn.value]




