"""The standalone exclusive scan (spr_scan_i32, the reference's naive_scan.comp:27-73) on device
memory against numpy: every tile-boundary size around both tile shapes (4096- and 8192-element
tiles), unaligned and in-place buffers, wrap-around of the int32 sum, a stream of the caller's,
and 10^8 elements.  Device memory and streams come from the HIP runtime libspr.so is linked
against (vkscanlinepr_amd.hip)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _exclusive(x: np.ndarray) -> np.ndarray:
    # int32 arithmetic wraps modulo 2^32, as the reference's does
    return np.concatenate([[0], np.cumsum(x.astype(np.int64))]).astype(np.uint64).astype(np.uint32).view(np.int32)


@pytest.fixture(scope="module")
def hip(spr):
    from vkscanlinepr_amd import hip as H
    assert H.device_count() > 0
    return H


@pytest.mark.parametrize("n", [0, 1, 2, 63, 4095, 4096, 4097, 8191, 8192, 8193, 3 * 4096 + 5,
                               (1 << 20) - 1, 1 << 20, (1 << 20) + 1, (1 << 20) + 8191, 5_000_003])
def test_scan_sizes(spr, hip, n):
    rng = np.random.default_rng(n)
    x = rng.integers(-1000, 1000, n, dtype=np.int32)
    xd = hip.DeviceBuffer(4 * n)
    yd = hip.DeviceBuffer(4 * (n + 1))
    if n:
        xd.upload(x)
    yd.fill_bytes(0xA5)
    st = hip.Stream()
    r = spr.Rasterizer(64, 64, device=0)
    print(f"[scan] n={n}", flush=True)
    r.scan_i32(xd.ptr, yd.ptr, n, st.handle)
    st.synchronize()
    assert np.array_equal(yd.download(np.int32, n + 1), _exclusive(x))
    r.close()
    st.destroy()


def test_scan_unaligned_in_place_and_wrapping(spr, hip):
    n = 3 * 8192 + 17
    rng = np.random.default_rng(9)
    x = rng.integers(1 << 28, 1 << 30, n, dtype=np.int32)  # the total wraps several times
    buf = hip.DeviceBuffer(4 * (n + 3))
    buf.fill_bytes(0)
    buf.upload(x, offset=4)
    r = spr.Rasterizer(64, 64, device=0)
    base = buf.ptr + 4  # 4-byte aligned, not 16: the scalar load path
    r.scan_i32(base, base, n)  # in place, on the handle's own stream
    r.synchronize()
    got = buf.download(np.int32, n + 3)
    assert np.array_equal(got[1:n + 2], _exclusive(x))
    assert got[0] == 0 and got[n + 2] == 0  # nothing written outside [d_out, d_out + n]
    r.close()


def test_scan_1e8(spr, hip):
    n = 100_000_000
    x = np.random.default_rng(3).integers(0, 1 << 10, n, dtype=np.int32)
    xd = hip.DeviceBuffer(4 * n)
    yd = hip.DeviceBuffer(4 * (n + 1))
    xd.upload(x)
    ref = _exclusive(x)
    st = hip.Stream()
    r = spr.Rasterizer(64, 64, device=0)
    for rep in range(2):  # the second call reuses the workspace
        print(f"[scan] n=1e8 rep={rep}", flush=True)
        yd.fill_bytes(0xFF)
        r.scan_i32(xd.ptr, yd.ptr, n, st.handle)
        st.synchronize()
        assert np.array_equal(yd.download(np.int32, n + 1), ref)
    r.close()
    st.destroy()


def test_scan_rejects_bad_arguments(spr, hip):
    r = spr.Rasterizer(64, 64, device=0)
    with pytest.raises(spr.SprError):
        r.scan_i32(0, 0, 10)
    with pytest.raises(spr.SprError):
        r.scan_i32(16, 16, 1 << 31)
    r.close()


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
def test_scan_tile_shapes(spr, hip, shape):
    """Every tile shape of the large scans (the autotuner's knob), around its tile boundaries."""
    r = spr.Rasterizer(64, 64, device=0)
    r.set_tuning(scan_shape=shape)
    st = hip.Stream()
    for n in ((1 << 20), (1 << 20) + 32767, 3 * 32768 + 1, 7_000_001):
        x = np.random.default_rng(n + shape).integers(-1000, 1000, n, dtype=np.int32)
        xd, yd = hip.DeviceBuffer(4 * n), hip.DeviceBuffer(4 * (n + 1))
        xd.upload(x)
        print(f"[scan] shape={shape} n={n}", flush=True)
        r.scan_i32(xd.ptr, yd.ptr, n, st.handle)
        st.synchronize()
        assert np.array_equal(yd.download(np.int32, n + 1), _exclusive(x))
    st.destroy()
    r.close()
