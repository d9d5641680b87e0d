"""The measurement hook sfr_copy_time (no counterpart in the reference): the
streaming copy kernel behind the bench's HBM ceiling.  Bar: the copy is
exact, bit for bit, for an even and an odd number of doubles (the odd
trailing one is copied apart from the 16-byte stream) and a single double;
misaligned or overlapping buffers are refused with SFR_BAD_ARG and leave
the destination alone; the reported time is positive."""
import ctypes as C

import numpy as np
import pytest

import sfr3d_amd as S
from sfr3d_amd._lib import lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 1_000_001, 4_194_304])
def test_copy_is_exact(ctx, n):
    rng = np.random.default_rng(n)
    src = rng.standard_normal(n)
    print(f"copy_time n={n}", flush=True)
    ds, dd = ctx.array(src), ctx.array(np.full(n, -7.0))
    try:
        ms = C.c_double()
        rc = lib.sfr_copy_time(ctx.h, ds.ptr, dd.ptr, n, 3, C.byref(ms))
        assert rc == S._lib.OK
        assert ms.value > 0.0
        assert np.array_equal(dd.get().view(np.uint64), src.view(np.uint64))
    finally:
        ds.free(); dd.free()


def test_copy_refuses_bad_buffers(ctx):
    n = 4096
    ds, dd = ctx.array(np.ones(n)), ctx.array(np.zeros(n))
    try:
        ms = C.c_double()
        # 8-byte offset: not 16-byte aligned
        assert lib.sfr_copy_time(ctx.h, ds.ptr + 8, dd.ptr, n - 1, 1, C.byref(ms)) == S._lib.BAD_ARG
        # overlapping source and destination
        assert lib.sfr_copy_time(ctx.h, ds.ptr, ds.ptr + 16 * 8, 64, 1, C.byref(ms)) == S._lib.BAD_ARG
        assert lib.sfr_copy_time(ctx.h, ds.ptr, dd.ptr, 0, 1, C.byref(ms)) == S._lib.BAD_ARG
        assert lib.sfr_copy_time(ctx.h, ds.ptr, dd.ptr, n, 0, C.byref(ms)) == S._lib.BAD_ARG
        assert np.array_equal(dd.get(), np.zeros(n))
    finally:
        ds.free(); dd.free()
