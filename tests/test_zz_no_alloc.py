"""GPU test of the allocation-free entry point bvhcd_mlp_forward_noalloc through the C ABI.

Its scores must equal bvhcd_mlp_forward's (the same kernel on the same rows, bit for bit) on two
calls in a row, the second with fewer rows than the first (the static staging keeps the first
call's contents), and a call over the row limit must be refused.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu


def _forward_noalloc(pkg, ctx, x):
    out = np.full(x.shape[0], np.nan, np.float32)
    st = pkg.lib().bvhcd_mlp_forward_noalloc(ctx._h, x.ctypes.data_as(C.c_void_p), x.shape[0],
                                             out.ctypes.data_as(C.c_void_p))
    return st, out


def test_mlp_forward_noalloc_matches_forward():
    pkg = load_package()
    print("[gpu] mlp_forward_noalloc", flush=True)
    weights = os.path.join(os.path.dirname(pkg.__file__), "mlp_scheduler.npz")
    z = np.load(weights, allow_pickle=False)
    ctx = pkg.CollisionContext(64, 64, scheduler=pkg.SCHED_LEARNED)
    ctx.mlp_set_weights(z["params"])
    rng = np.random.default_rng(7)
    for rows in (9000, 777):  # second call smaller than the first
        x = rng.normal(size=(rows, 16)).astype(np.float32)
        want = ctx.mlp_forward(x)
        st, got = _forward_noalloc(pkg, ctx, x)
        assert st == pkg.OK
        np.testing.assert_array_equal(got, want, err_msg=f"rows {rows}")
    limit = 16384
    x = np.zeros((limit, 16), np.float32)
    st, got = _forward_noalloc(pkg, ctx, x)
    assert st == pkg.OK and np.array_equal(got, ctx.mlp_forward(x))
    st, _ = _forward_noalloc(pkg, ctx, np.zeros((limit + 1, 16), np.float32))
    assert st == pkg.E_INVALID
    ctx.close()
