"""wl_mle_evaluate_noalloc (no device allocation, fixed per-GPU scratch) against
wl_mle_evaluate and the oracle, through the C ABI: two calls in a row, the second
on a smaller input, then a refused size.  The inputs and outputs live in device
memory this test allocates beforehand with the HIP runtime the library uses."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
P = 0x7F000001
H2D, D2H = 1, 2


def _hip():
    import whirlaway_amd
    whirlaway_amd.lib()  # loads the HIP runtime; the same instance is returned below
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def _dev_copy(h, arr):
    p = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(p), max(arr.nbytes, 4)) == 0
    assert h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
    return p


def _noalloc(ctx, h, ev, pt, is_ef):
    from whirlaway_amd import check, lib, u32ptr
    n = pt.size // 8
    d_ev = _dev_copy(h, ev)
    d_out = _dev_copy(h, np.zeros(8, dtype=np.uint32))
    try:
        check(lib().wl_mle_evaluate_noalloc(ctx.ptr, d_ev, int(is_ef), n, u32ptr(pt), d_out))
        out = np.zeros(8, dtype=np.uint32)
        assert h.hipMemcpy(out.ctypes.data, d_out, 32, D2H) == 0
        return out
    finally:
        h.hipFree(d_ev)
        h.hipFree(d_out)


@pytest.mark.parametrize("first,second", [((20, False), (9, False)), ((16, True), (5, True)),
                                          ((28, False), (12, True))])
def test_noalloc_matches_allocating_entry(gpu_ctx, oracle, first, second):
    from whirlaway_amd import utils
    h = _hip()
    rng = np.random.default_rng(first[0] * 31 + second[0])
    for n, is_ef in (first, second):
        print(f"noalloc evaluation n={n} ef={is_ef}", flush=True)
        w = 8 if is_ef else 1
        ev = rng.integers(0, P, (1 << n) * w, dtype=np.uint32)
        pt = rng.integers(0, P, n * 8, dtype=np.uint32)
        got = _noalloc(gpu_ctx, h, ev, pt, is_ef)
        want = utils.mle_evaluate(gpu_ctx, ev, pt, is_ef=is_ef)
        assert np.array_equal(got, want), (n, is_ef)
        if n <= 20:
            assert np.array_equal(got, oracle.mle_eval(ev, pt, is_ef=is_ef)), (n, is_ef)


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu_ctx):
    from whirlaway_amd import WhirlawayError, check, lib, u32ptr
    pt = np.zeros(29 * 8, dtype=np.uint32)
    with pytest.raises(WhirlawayError, match="at most 28"):
        # refused before any device access: the pointers are never dereferenced
        check(lib().wl_mle_evaluate_noalloc(gpu_ctx.ptr, ctypes.c_void_p(256), 0, 29, u32ptr(pt),
                                            ctypes.c_void_p(256)))
