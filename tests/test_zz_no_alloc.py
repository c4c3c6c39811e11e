"""az_gpu_sum_f32_sequential_noalloc, the entry point that allocates nothing
(its workspace is a fixed device array), against az_gpu_sum_f32_sequential
and the sequential float32 sum, through the C ABI: two calls in a row on the
same context, the second smaller, so the second reads a workspace the first
left behind; the size limit; the empty input."""
import ctypes as C

import numpy as np
import pytest

from azimuth_amd import native as N

pytestmark = pytest.mark.gpu

MAX_VALUES = 1 << 26


def sequential_sum(x):
    s = np.float32(0.0)
    for v in np.asarray(x, dtype=np.float32):   # only for short inputs
        s = np.float32(s + v)
    return s


def run(lib, ctx, fn, d_x, n, d_s):
    out = np.zeros(1, dtype=np.float32)
    assert fn(ctx, d_x, n, d_s) == 0, lib.az_gpu_last_error(ctx)
    assert lib.az_gpu_synchronize(ctx) == 0
    assert lib.az_gpu_download(ctx, out.ctypes.data, d_s, 4) == 0
    return out[0]


def test_noalloc_sum_matches_the_allocating_entry_point():
    lib = N.load()
    ctx = C.c_void_p()
    assert lib.az_gpu_create(C.byref(ctx), 0) == 0
    rng = np.random.default_rng(21)
    big = np.log(rng.exponential(0.3, 8294400).astype(np.float32) + np.float32(1e-5))   # a 4K image's log luminances
    small = (rng.integers(-40, 8, 100003) / 8.0).astype(np.float32)                     # many rounding ties
    d_x, d_a, d_b = C.c_void_p(), C.c_void_p(), C.c_void_p()
    try:
        assert lib.az_gpu_malloc(ctx, C.byref(d_x), big.nbytes) == 0
        assert lib.az_gpu_malloc(ctx, C.byref(d_a), 4) == 0
        assert lib.az_gpu_malloc(ctx, C.byref(d_b), 4) == 0
        for x in (big, small, small[:1025], small[:1]):
            assert lib.az_gpu_upload(ctx, d_x, x.ctypes.data, x.nbytes) == 0
            want = run(lib, ctx, lib.az_gpu_sum_f32_sequential, d_x, x.size, d_a)
            got = run(lib, ctx, lib.az_gpu_sum_f32_sequential_noalloc, d_x, x.size, d_b)
            print(f"n={x.size}: noalloc {got!r}, allocating {want!r}", flush=True)
            assert np.float32(got).tobytes() == np.float32(want).tobytes(), x.size
            if x.size <= 1025:
                assert np.float32(got).tobytes() == sequential_sum(x).tobytes()
        # The empty sum is +0; the limit is checked before any device work.
        assert lib.az_gpu_upload(ctx, d_b, np.float32([7.0]).ctypes.data, 4) == 0
        assert run(lib, ctx, lib.az_gpu_sum_f32_sequential_noalloc, d_x, 0, d_b) == 0.0
        assert lib.az_gpu_sum_f32_sequential_noalloc(ctx, d_x, MAX_VALUES + 1, d_b) == 1   # AZ_ERR_INVALID_ARGUMENT
        assert b"2^26" in lib.az_gpu_last_error(ctx)
        assert lib.az_gpu_sum_f32_sequential_noalloc(ctx, None, 5, d_b) == 1
    finally:
        for p in (d_x, d_a, d_b):
            if p.value:
                lib.az_gpu_free(ctx, p)
        lib.az_gpu_destroy(ctx)
