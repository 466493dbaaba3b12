"""GPU check of the allocation-free entry point stark_ntt_dev_noalloc against
stark_ntt_dev, called through the C ABI: two calls in a row (the second
smaller), forward and inverse, bit for bit; a size above the fixed scratch is
rejected with STARK_E_INVALID before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
P = 2**64 - 2**32 + 1
vp = ctypes.c_void_p


def test_noalloc_ntt_matches_ntt_dev():
    from stark_swarm_amd import api as S
    L = S.lib()
    L.stark_ntt_dev_noalloc.restype = ctypes.c_int
    L.stark_ntt_dev_noalloc.argtypes = [vp, vp, vp, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]
    ctx = S.Context()
    rng = np.random.default_rng(2024)
    for log_n, ncols in ((16, 32), (14, 4)):  # 2^21 then 2^16 elements; both two-pass (log_n > 12)
        for inverse in (0, 1):
            print("noalloc ntt 2^%d x %d inverse=%d" % (log_n, ncols, inverse), flush=True)
            x = rng.integers(0, P, size=(ncols, 1 << log_n), dtype=np.uint64)
            d_in, d_a, d_b = (S.DeviceBuffer(ctx, x.nbytes) for _ in range(3))
            d_in.upload(x)
            assert L.stark_ntt_dev(ctx.h, d_in.ptr, d_a.ptr, log_n, ncols, inverse) == 0
            assert L.stark_ntt_dev_noalloc(ctx.h, d_in.ptr, d_b.ptr, log_n, ncols, inverse) == 0
            a, b = d_a.download((ncols, 1 << log_n)), d_b.download((ncols, 1 << log_n))
            assert np.array_equal(a, b)
    # above the 2^22-element scratch: refused on the host
    d = S.DeviceBuffer(ctx, 8)
    assert L.stark_ntt_dev_noalloc(ctx.h, d.ptr, d.ptr, 20, 8, 0) == -1
