"""The allocation-free inner product (HYPRE_MI355X_ParVectorInnerProdNoAlloc)
against HYPRE_ParVectorInnerProd, through the C-ABI.

Both run the same reduction (same grid, same per-block tree, same final
fold), so the bar is bit equality.  Two calls in a row, the second on shorter
vectors, show that the fixed array keeps nothing between calls that changes
a result; a third length reaches the 2048-block limit of the grid.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def inner(hx, name, x, y):
    d = C.c_double()
    hx.call(name, x.h, y.h, C.byref(d))
    return d.value


def test_no_alloc_inner_product_matches_the_allocating_one(hx, oracle):
    print("[gpu-no-alloc] inner products", flush=True)
    # 3,000,001 entries: 2048 blocks (the most the grid uses); then 77,777 and 5
    for n, seed in ((3_000_001, 1), (77_777, 2), (5, 3)):
        xh = 2 * oracle.rand_stream(seed, n) - 1
        yh = 2 * oracle.rand_stream(seed + 10, n) - 1
        x, y = hx.ParVector(n, xh), hx.ParVector(n, yh)
        got = inner(hx, "HYPRE_MI355X_ParVectorInnerProdNoAlloc", x, y)
        ref = inner(hx, "HYPRE_ParVectorInnerProd", x, y)
        assert np.float64(got).view(np.int64) == np.float64(ref).view(np.int64), (n, got, ref)
        np.testing.assert_allclose(got, float(xh @ yh), rtol=1e-12)
        x.destroy()
        y.destroy()


def test_no_alloc_inner_product_rejects_mismatched_lengths(hx):
    x, y = hx.ParVector(10), hx.ParVector(11)
    with pytest.raises(hx.HypreError) as e:
        inner(hx, "HYPRE_MI355X_ParVectorInnerProdNoAlloc", x, y)
    assert e.value.code & hx.ERROR_ARG
