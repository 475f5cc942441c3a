"""The no-allocation price epoch (apgx_price_update_no_alloc) against the
allocating one (apgx_price_update), through the C-ABI, on device planes made
with the HIP runtime the library itself uses.  Two calls in a row, the second
on a smaller input, each with a passing and a failing epoch; the size limit is
refused before any device access."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libapgx.so is linked to."""

    def __init__(self, apgx):
        apgx.lib()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.live = []

    def upload(self, a: np.ndarray) -> int:
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        self.live.append(p.value)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def download(self, p: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data, C.c_void_p(p), out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        for p in self.live:
            self.h.hipFree(C.c_void_p(p))
        self.live = []


def _planes(rng, cells):
    usage = rng.integers(0, 4, cells).astype(np.uint32)
    cap = rng.integers(0, 3, cells).astype(np.uint8)
    hist = np.where(rng.random(cells) < 0.3, rng.integers(1, 500, cells), 0).astype(np.uint32)
    return usage, cap, hist


def _one(apgx, hip, epoch, usage, cap, hist, **settings):
    want, w_present, w_hist, w_total = apgx.price_update(epoch, usage, hist, capacity=cap, **settings)
    d_usage, d_cap, d_hist = hip.upload(usage), hip.upload(cap), hip.upload(hist)
    zero = np.zeros_like(hist)
    d_present, d_total = hip.upload(zero), hip.upload(zero)
    got = apgx.price_update_no_alloc(epoch, usage.size, d_usage, d_hist, capacity=d_cap, present=d_present,
                                     total=d_total, **settings)
    assert got == want, (got, want)
    assert np.array_equal(hip.download(d_hist, hist), w_hist)
    assert np.array_equal(hip.download(d_present, hist), w_present)
    assert np.array_equal(hip.download(d_total, hist), w_total)
    return want


def test_no_alloc_matches_the_allocating_epoch(apgx, gpu):
    print("[case] no-allocation price epoch, two calls (3.1 M then 70 k cells)", flush=True)
    rng = np.random.default_rng(41)
    hip = Hip(apgx)
    try:
        for cells in (3_100_000, 70_001):
            usage, cap, hist = _planes(rng, cells)
            big = dict(maximum_price_entries=10_000_000, maximum_hot_resources=10_000_000)
            ok = _one(apgx, hip, 2, usage, cap, hist, **big)
            assert ok["outcome"] == 0 and ok["aggregate_price"] > 0
            # A failing epoch: the aggregate bound inside the planes; both leave them untouched.
            bad = _one(apgx, hip, 2, usage, cap, hist, **big, maximum_aggregate_price=ok["aggregate_price"] // 2)
            assert bad["outcome"] == 9 and bad["outcome_cell"] < cells
            bad = _one(apgx, hip, 3, usage, cap, hist, maximum_hot_resources=10_000_000,
                       maximum_price_entries=ok["price_entries"] // 3)
            assert bad["outcome"] == 3
    finally:
        hip.close()


def test_no_alloc_refuses_more_cells_than_its_scratch(apgx, gpu):
    hip = Hip(apgx)
    try:
        d = hip.upload(np.zeros(16, np.uint32))
        with pytest.raises(apgx.ApgxError) as err:
            apgx.price_update_no_alloc(0, (1 << 26) + 1, d, d)
        assert err.value.code == 2 and "67108864" in err.value.detail
    finally:
        hip.close()
