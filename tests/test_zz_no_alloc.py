"""spr_scan_i32_noalloc (its workspace is a static device array: no allocation) against
spr_scan_i32 (the handle's workspace), both through the C ABI: two calls in a row, the second on
a smaller input, each bit-identical to the existing entry point's output and to numpy; sizes
beyond the static workspace are rejected before any device work."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _exclusive(x: np.ndarray) -> np.ndarray:
    return np.concatenate([[0], np.cumsum(x.astype(np.int64))]).astype(np.uint64).astype(np.uint32).view(np.int32)


def test_noalloc_scan_matches_scan_twice(spr):
    from vkscanlinepr_amd import hip as H
    r = spr.Rasterizer(64, 64, device=0)
    st = H.Stream()
    for n, seed, stream in ((5_000_011, 1, st.handle), (300_007, 2, None)):  # the second is smaller
        print(f"[noalloc] n={n} stream={'caller' if stream else 'null'}", flush=True)
        x = np.random.default_rng(seed).integers(-1000, 1000, n, dtype=np.int32)
        xd, yd, zd = H.DeviceBuffer(4 * n), H.DeviceBuffer(4 * (n + 1)), H.DeviceBuffer(4 * (n + 1))
        xd.upload(x)
        yd.fill_bytes(0xA5)
        zd.fill_bytes(0x5A)
        spr.scan_i32_noalloc(xd.ptr, yd.ptr, n, stream)
        r.scan_i32(xd.ptr, zd.ptr, n, st.handle)
        H.hip().hipDeviceSynchronize()
        a, b = yd.download(np.int32, n + 1), zd.download(np.int32, n + 1)
        assert np.array_equal(a, b)
        assert np.array_equal(a, _exclusive(x))
    st.destroy()
    r.close()


def test_noalloc_scan_rejects_sizes_beyond_the_workspace(spr):
    with pytest.raises(spr.SprError, match="static workspace"):
        spr.scan_i32_noalloc(16, 16, (1 << 28) + 1)
    with pytest.raises(spr.SprError):
        spr.scan_i32_noalloc(0, 0, 10)
