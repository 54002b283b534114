"""The C-ABI library loads and exports every symbol include/*.h declares;
band and size queries answer without a GPU (CPU only, no compute calls)."""
import glob
import os
import re
import subprocess

import pytest

import simd_bigint_amd as sba

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
PRIMES = [1107296257, 1711276033, 1811939329, 2013265921, 2113929217]


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \t\*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M):
            name = m.group(1)
            if name not in ("defined",):
                names.add(name)
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = sba.lib()
    names = declared_functions()
    assert "nt_mul" in names and "divrem_u64" in names and "simd_mpn_mul" in names
    out = subprocess.run(["nm", "-D", "--defined-only", sba.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(L, n)


def test_band_queries():
    P = 1
    for p in PRIMES:
        P *= p
    F = (2**64 - 1) ** 2
    assert sba.min_operand_cap() == P // F
    assert sba.max_lg() == 25
    for p in PRIMES:
        assert (p - 1) % (1 << 25) == 0
    assert PRIMES == sorted(PRIMES)
    L = sba.lib()
    assert L.sba_gpu_workspace_bytes(0, 5) == 0
    assert L.sba_gpu_workspace_bytes(1 << 22, 1 << 22) > 0
    assert L.sba_gpu_workspace_bytes(1 << 25, 1 << 25) == 0      # beyond 2^25 coefficients


def test_out_of_band_returns_zero_without_device():
    import numpy as np
    a = np.ones(4, np.uint64)
    e = np.zeros(0, np.uint64)
    L = sba.lib()
    r = np.zeros(8, np.uint64)
    p = sba._arr
    assert L.nt_mul(p(r), p(a), 0, p(a), 4) == 0
    assert L.nt_mul(p(r), p(a), 4, p(a), -1) == 0
    assert L.nt_sqr(p(r), p(a), 0) == 0
    # p30 band (include/p30.h:1851-1858): min operand cap 2,344,425
    assert L.p30_mul(p(r), p(a), 2344426, p(a), 2344426) == 0
    del e


def test_tuning_knobs_round_trip_without_device():
    """The CPU/GPU threshold, the schoolbook limit and the carry block size
    answer without a GPU; setters round-trip and are restored."""
    import ctypes
    L = sba.lib()
    n = L.sba_tune_sizes(None, 0)
    sizes = (ctypes.c_uint64 * n)()
    assert L.sba_tune_sizes(sizes, n) == n
    ladder = list(sizes)
    assert ladder == sorted(ladder) and ladder[0] >= 1 and len(set(ladder)) == n
    old = sba.get_gpu_threshold()
    assert old == 1024
    try:
        sba.set_gpu_threshold(12345)
        assert sba.get_gpu_threshold() == 12345
    finally:
        sba.set_gpu_threshold(old)
    olds = sba.get_school_max()
    assert olds == 32768
    try:
        sba.set_school_max(0)
        assert sba.get_school_max() == 0
    finally:
        sba.set_school_max(olds)
    lpb = sba.carry_block_limbs()
    assert lpb in (512, 1024, 2048) and lpb % 64 == 0
