"""The drop-in boundary on the GPU: every reference-named entry called in
band through the C ABI and checked against the oracle or the reference's
golden SHA-256 (include/ntmul.h:47-79, include/p30.h:1847-2020,
include/p50.h:1020-1178); p30's transform cap at its edge (p30.h:1851-1858);
the caller-workspace device entries with workspaces of exactly
sba_gpu_workspace_bytes and one byte short, on the direct, radix-3 and
chunked routes; and a compiled C caller linked against the library."""
import hashlib
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import simd_bigint_amd as sba
from simd_bigint_amd.workloads import pattern_fill, xorshift_fill
from tests import _oracle as orc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, ".."))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


MUL_R = ("nt_mul_r", "p30_mul_r", "p50_mul_r")
SQR_R = ("nt_sqr_r", "p30_sqr_r", "p50_sqr_r")


@pytest.mark.parametrize("an,bn", [(20000, 13000), (1 << 17, 5000), (70001, 70001)])
def test_reference_entries_in_band_vs_oracle(an, bn):
    st = [an + 3 * bn]
    a, b = xorshift_fill(an, st), xorshift_fill(bn, st)
    want = orc.nt_mul(a, b)
    want_sq = orc.nt_sqr(a)
    for name in MUL_R:
        ok, r = sba.mul_r(name, a, b)
        assert ok == 1 and (r == want).all(), name
    for name in SQR_R:
        ok, r = sba.mul_r(name, a)
        assert ok == 1 and (r == want_sq).all(), name
    for f in (sba.p30_mul, sba.p50_mul):
        ok, r = f(a, b)
        assert ok == 1 and (r == want).all(), f.__name__
    for f in (sba.p30_sqr, sba.p50_sqr):
        ok, r = f(a)
        assert ok == 1 and (r == want_sq).all(), f.__name__


def _gold(name):
    return next(e for e in GOLD["large_mul"] if e["name"] == name)


@pytest.mark.parametrize("gname", ["metric_4Mi", "sqr_2^22", "mul_3x2^20"])
def test_p50_entries_reproduce_reference_sha(gname):
    """Beyond p30's caps the reference's nt_mul runs p50 (include/ntmul.h:
    52-57): p50_mul_r / p50_sqr_r / p50_mul / p50_sqr give the reference's
    own products there."""
    names = [e["name"] for e in GOLD["large_mul"]]
    if gname not in names:
        pytest.skip("no golden entry %s" % gname)
    e = _gold(gname)
    say("case", gname)
    st = [e["seed"]]
    a = pattern_fill(e["an"], e["pat"], st)
    if e["sqr"]:
        for name in ("p50_sqr_r",):
            ok, r = sba.mul_r(name, a)
            assert ok == 1 and sha(r) == e["sha256"], name
        ok, r = sba.p50_sqr(a)
        assert ok == 1 and sha(r) == e["sha256"]
        ok, _ = sba.p30_sqr(a)                 # beyond p30's transform cap
        assert ok == 0
    else:
        b = pattern_fill(e["bn"], e["pat"], st)
        ok, r = sba.mul_r("p50_mul_r", a, b)
        assert ok == 1 and sha(r) == e["sha256"]
        ok, r = sba.p50_mul(a, b)
        assert ok == 1 and sha(r) == e["sha256"]
        ok, _ = sba.mul_r("p30_mul_r", a, b)
        assert ok == 0


def test_p30_transform_cap_edge():
    """tv = ceil((an+bn-1)/16) <= 2^18 (include/p30.h:1858): an + bn - 1 =
    2^22 is the last accepted product, 2^22 + 1 the first rejected; the
    square's edge is an = 2^21. Accepted products checked by fingerprints
    and against nt_mul."""
    st = [30]
    a = xorshift_fill(1 << 21, st)
    b = xorshift_fill((1 << 21) + 1, st)
    ok, r = sba.p30_mul(a, b)
    assert ok == 1
    q = (1 << 61) - 1
    assert orc.fingerprint(r, q) == orc.fingerprint(a, q) * orc.fingerprint(b, q) % q
    ok2, r2 = sba.nt_mul(a, b)
    assert ok2 == 1 and (r2 == r).all()
    b2 = np.append(b, np.uint64(5))
    ok, _ = sba.p30_mul(a, b2)
    assert ok == 0
    ok, _ = sba.mul_r("p30_mul_r", b2, a)
    assert ok == 0
    ok, r = sba.p30_sqr(a)
    assert ok == 1
    assert orc.fingerprint(r, q) == orc.fingerprint(a, q) ** 2 % q
    ok, _ = sba.p30_sqr(b)                     # 2^21 + 1 limbs: 2^22 + 1 coefficients
    assert ok == 0
    ok, _ = sba.mul_r("p30_sqr_r", b)
    assert ok == 0


# direct (one transform), radix-3 (3 x 2^19 length), chunked (3 chunks)
WS_SHAPES = [("direct", 1 << 19, 300000), ("radix3", 3 << 18, 3 << 18),
             ("chunked", 1 << 20, 1 << 17)]
GUARD = 4096


@pytest.mark.parametrize("route,an,bn", WS_SHAPES, ids=[s[0] for s in WS_SHAPES])
def test_ws_entries_exact_and_short_workspace(route, an, bn):
    torch = pytest.importorskip("torch")
    L = sba.lib()
    ch = sba.chunking(an, bn)
    assert (ch[2] > 0) == (route == "chunked")
    assert (L.sba_gpu_radix3(an, bn, 0) > 0) == (route == "radix3")
    st = [an ^ bn]
    a, b = xorshift_fill(an, st), xorshift_fill(bn, st)
    ta = torch.from_numpy(a.view(np.int64)).cuda()
    tb = torch.from_numpy(b.view(np.int64)).cuda()
    for sq in (False, True):
        nb = an if sq else bn
        need = L.sba_gpu_workspace_bytes(an, nb)
        assert need > 0
        ws = torch.full((need + GUARD,), 0x5A, dtype=torch.uint8, device="cuda")
        r = torch.zeros(an + nb, dtype=torch.int64, device="cuda")
        say("ws", route, an, nb, "sqr" if sq else "mul", need)
        if sq:
            rc = L.sba_gpu_sqr_dev_ws(r.data_ptr(), ta.data_ptr(), an, ws.data_ptr(), need, None)
        else:
            rc = L.sba_gpu_mul_dev_ws(r.data_ptr(), ta.data_ptr(), an, tb.data_ptr(), bn,
                                      ws.data_ptr(), need, None)
        assert rc == 1
        torch.cuda.synchronize()
        got = r.cpu().numpy().view(np.uint64)
        want = orc.nt_sqr(a) if sq else orc.nt_mul(a, b)
        assert (got == want).all(), (route, sq)
        assert bool((ws[need:] == 0x5A).all()), "workspace overrun"
        # one byte short, and no workspace at all
        r.fill_(3)
        if sq:
            rc = L.sba_gpu_sqr_dev_ws(r.data_ptr(), ta.data_ptr(), an, ws.data_ptr(), need - 1,
                                      None)
            rc0 = L.sba_gpu_sqr_dev_ws(r.data_ptr(), ta.data_ptr(), an, None, need, None)
        else:
            rc = L.sba_gpu_mul_dev_ws(r.data_ptr(), ta.data_ptr(), an, tb.data_ptr(), bn,
                                      ws.data_ptr(), need - 1, None)
            rc0 = L.sba_gpu_mul_dev_ws(r.data_ptr(), ta.data_ptr(), an, tb.data_ptr(), bn, None,
                                       need, None)
        assert rc == sba.SBA_ERR_WORKSPACE and rc0 == sba.SBA_ERR_WORKSPACE
        torch.cuda.synchronize()
        assert bool((r == 3).all())


def test_compiled_c_caller(tmp_path):
    """A C program that includes only include/simd_bigint_amd.h and links
    the shared library, as INTEGRATION.md shows, calls every reference-named
    entry in band and checks each result itself."""
    cc = shutil.which("cc") or shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = str(tmp_path / "drop_in_caller")
    libdir = os.path.join(ROOT, "simd_bigint_amd", "lib")
    subprocess.run([cc, "-O2", os.path.join(HERE, "c", "drop_in_caller.c"),
                    "-I" + os.path.join(ROOT, "include"), "-L" + libdir, "-lsimd_bigint_amd",
                    "-Wl,-rpath," + libdir, "-o", exe], check=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    say(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "all entries ok" in p.stdout


def test_error_contract_on_device():
    """A failed staging allocation inside a host entry (injected): nt_mul
    returns 0, simd_mpn_mul still delivers the product (CPU tier, counted),
    and the next call runs on the GPU again."""
    st = [99]
    a, b = xorshift_fill(30000, st), xorshift_fill(20000, st)
    want = orc.nt_mul(a, b)
    try:
        sba.inject_failure(sba.FAIL_ALLOC, 1)
        ok, _ = sba.nt_mul(a, b)
        assert ok == 0
        before = sba.gpu_fallbacks()
        sba.inject_failure(sba.FAIL_ALLOC, 1)
        r = sba.simd_mpn_mul(a, b)
        assert (r == want).all() and sba.gpu_fallbacks() == before + 1
    finally:
        sba.inject_failure(sba.FAIL_ALLOC, 0)
    ok, r = sba.nt_mul(a, b)
    assert ok == 1 and (r == want).all()
    r = sba.simd_mpn_mul(a, b)
    assert (r == want).all() and sba.gpu_fallbacks() == before + 1
