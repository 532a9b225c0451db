"""The allocation-free dot products (wtg_fs_dots_noalloc) against
wtg_fs_dot on the same fields (needs a GPU).

Bar: bit-identical sums (the same block partials and final tree, the
partials in the library's file-scope scratch instead of the context's
buffers), on two calls in a row, the second on a smaller level with fewer
products; more than three products are refused.  The device arrays are a
context's state slot (wtg_state_device_ptrs): the products' operands in its
mass and temperature layers, the sums written over the first entries of its
velocity."""
import ctypes as C

import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W

pytestmark = pytest.mark.gpu


def _want(ctx, a, b):
    out = []
    for x, y in zip(a, b):
        ctx.fs_upload(W.FS_ETA, x)
        ctx.fs_upload(W.FS_RHS, y)
        out.append(ctx.fs_dot(W.FS_ETA, W.FS_RHS))
    return np.array(out)


def _noalloc(ctx, a, b):
    nc, nd = len(a[0]), len(a)
    ctx.upload_state(0, np.stack(a), np.stack(b), np.full((nd, 3, nc), np.nan))
    m, t, u = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert W.load_library().wtg_state_device_ptrs(ctx._h, 0, C.byref(m), C.byref(t), C.byref(u)) == 0
    ctx.fs_dots_noalloc([m.value + 8 * k * nc for k in range(nd)], [t.value + 8 * k * nc for k in range(nd)], u.value)
    ctx.synchronize()
    return ctx.download_state(0)[2].reshape(-1)[:nd].copy()


def test_noalloc_dots_equal_fs_dot():
    print("no-allocation dot products: J5 three products, then J4 one", flush=True)
    rng = np.random.default_rng(5)
    for level, nd in ((5, 3), (4, 1)):
        g = W.Grid(level=level, min_level=4, domain_level=1)
        ctx = W.Context(g, nlayers=nd)
        ctx.set_topography(np.full(g.layout.ncells, -1000.0))
        a = [rng.standard_normal(g.layout.ncells) for _ in range(nd)]
        b = [rng.standard_normal(g.layout.ncells) for _ in range(nd)]
        want = _want(ctx, a, b)
        got = _noalloc(ctx, a, b)
        print(f"  J{level}: {got} vs {want}", flush=True)
        assert np.array_equal(got, want)
        ctx.close()


def test_noalloc_refuses_four_products():
    g = W.Grid(level=4, min_level=4, domain_level=1)
    ctx = W.Context(g, nlayers=1)
    with pytest.raises(W.WtgError, match="at most 3"):
        ctx.fs_dots_noalloc([1, 1, 1, 1], [1, 1, 1, 1], 1)
    ctx.close()
