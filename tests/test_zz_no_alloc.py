"""hn_dot_dev_noalloc: the allocation-free variant of hn_dot_dev.

Called through the C ABI on device buffers; each result must equal
hn_dot_dev's bit for bit (same kernels, same grid), on two calls in a row
with the second one shorter, so a stale partial sum left in the shared
scratch array by the first call would show."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dot(H, fn, da, db, n):
    out = H.DeviceArray(1)
    H.check(fn(da.ptr, db.ptr, n, out.ptr, None))
    H.check(H.LIB.hn_stream_sync(None))
    return out.to_numpy()[0]


def test_noalloc_dot_matches_dot_dev_twice(hip):
    H = hip
    rng = np.random.default_rng(11)
    big, small = 3_000_001, 777
    a = rng.standard_normal(big)
    b = rng.standard_normal(big)
    da, db = H.DeviceArray.from_numpy(a), H.DeviceArray.from_numpy(b)
    # warm-up call: the runtime may set up per-module state on a first launch
    _dot(H, H.LIB.hn_dot_dev_noalloc, da, db, 5)
    free0, _ = H.mem_info()
    got = [_dot(H, H.LIB.hn_dot_dev_noalloc, da, db, n) for n in (big, small)]
    free1, _ = H.mem_info()
    want = [_dot(H, H.LIB.hn_dot_dev, da, db, n) for n in (big, small)]
    print(f"[noalloc] n={big}: {got[0]!r} vs {want[0]!r}; n={small}: {got[1]!r} vs {want[1]!r}", flush=True)
    for g, w, n in zip(got, want, (big, small)):
        assert g.tobytes() == w.tobytes(), (n, g, w)
        ref = float(np.dot(a[:n], b[:n]))
        assert abs(g - ref) <= 1e-12 * max(1.0, float(np.abs(a[:n] * b[:n]).sum())), (n, g, ref)
    assert free1 == free0, f"device free memory changed across the calls: {free0} -> {free1}"


def test_noalloc_dot_empty_and_short(hip):
    H = hip
    a = np.arange(1.0, 4.0)
    da, db = H.DeviceArray.from_numpy(a), H.DeviceArray.from_numpy(a)
    assert _dot(H, H.LIB.hn_dot_dev_noalloc, da, db, 3) == 14.0
    assert _dot(H, H.LIB.hn_dot_dev_noalloc, da, db, 0) == 0.0
