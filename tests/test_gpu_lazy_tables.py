"""GPU: the root tables are built when a call can first use them.

A fresh process builds the tables of the two small-prime sets with its first
GPU call and the wide set's only when a transform needs them (above 2^23
points, or with the small-prime sets switched off). This runs in a child
process, so that no earlier test has built any table: a first small product,
then a 2^24-point all-ones product (the wide set's first use, checked against
its closed form), then a product with the sets switched off, each bit-exact."""
import os
import subprocess
import sys
import textwrap

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = textwrap.dedent("""
    import sys
    import numpy as np
    sys.path.insert(0, %r)
    import simd_bigint_amd as sba
    from simd_bigint_amd.workloads import all_ones_product, xorshift_fill
    from tests import _oracle as orc
    st = [5]
    a, b = xorshift_fill(20000, st), xorshift_fill(13000, st)
    ok, r = sba.nt_mul(a, b)
    assert ok == 1 and (r == orc.nt_mul(a, b)).all(), "first product"
    print("first small product ok", flush=True)
    n = (1 << 23) + 5                     # 2^24-point transform: the wide set
    x = np.full(n, 2**64 - 1, np.uint64)
    ok, r = sba.nt_mul(x, x[: n - 9])
    assert ok == 1 and (r == all_ones_product(n, n - 9)).all(), "wide-set product"
    print("wide set on first use ok", flush=True)
    sba.lib().sba_set_prime_set_h(0)
    ok, r = sba.nt_mul(a, b)
    assert ok == 1 and (r == orc.nt_mul(a, b)).all(), "sets switched off"
    print("sets switched off ok", flush=True)
""") % ROOT


def test_tables_built_on_first_use_in_a_fresh_process():
    p = subprocess.run([sys.executable, "-c", CHILD], cwd=ROOT, capture_output=True, text=True,
                       timeout=240)
    print(p.stdout, p.stderr[-2000:], file=sys.stderr, flush=True)
    assert p.returncode == 0, p.stderr[-2000:]
    assert "sets switched off ok" in p.stdout
