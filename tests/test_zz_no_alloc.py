"""dmtz_classify_noalloc_f32 (include/dmtz.h), the classifier that allocates nothing, against the
existing entry point dmtz_classify_f32 on the same fields: counts and per-vertex masks equal,
on two calls in a row with the second field smaller, through the C ABI."""
import ctypes

import numpy as np
import pytest

import dmtz_amd
from dmtz_amd import _native, devmem, fields

pytestmark = pytest.mark.gpu


def _noalloc(ctx, f):
    d_f = devmem.DeviceArray.from_numpy(f)
    d_lo = devmem.DeviceArray(f.shape, np.uint64)
    d_hi = devmem.DeviceArray(f.shape, np.uint32)
    counts = np.zeros(4, np.int64)
    nz, ny, nx = f.shape
    rc = _native.lib().dmtz_classify_noalloc_f32(ctx.handle, ctypes.c_void_p(d_f.ptr), nx, ny, nz,
                                                 counts.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(d_lo.ptr),
                                                 ctypes.c_void_p(d_hi.ptr))
    ctx.check(rc, "dmtz_classify_noalloc_f32")
    return tuple(int(x) for x in counts), d_lo.numpy(), d_hi.numpy()


def test_noalloc_classify_equals_classify():
    ctx = dmtz_amd.Context(0)
    cases = [fields.gaussian_mixture((40, 48, 64), 12, seed=5, noise=0.02),
             fields.fbm((25, 33, 40), seed=6)]
    for f in cases:
        print("noalloc classify", f.shape, flush=True)
        want, lo, hi = dmtz_amd.classify(f, ctx=ctx, masks=True)
        got, glo, ghi = _noalloc(ctx, f)
        assert got == want
        assert (glo == lo).all() and (ghi == hi).all()
    ctx.close()


def test_noalloc_rejects_too_large_grid():
    ctx = dmtz_amd.Context(0)
    nz, ny, nx = 129, 128, 128
    counts = np.zeros(4, np.int64)
    d_f = devmem.DeviceArray((4,), np.float32)
    rc = _native.lib().dmtz_classify_noalloc_f32(ctx.handle, ctypes.c_void_p(d_f.ptr), nx, ny, nz,
                                                 counts.ctypes.data_as(ctypes.c_void_p), None, None)
    assert rc == -3  # DMTZ_E_CAPACITY, refused before any device work
    ctx.close()
