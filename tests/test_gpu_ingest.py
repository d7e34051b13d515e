"""TIFF straight into device memory (tops_tiff_read_device: file reads into
pinned staging buffers overlapped with their copies to HBM) against the
raster that was written, for whole and partial windows."""
import ctypes as C

import numpy as np
import pytest

import s1_files as S1

pytestmark = pytest.mark.gpu


def _cint16(lines, pixels, seed):
    rng = np.random.default_rng(seed)
    r = np.empty((lines, pixels), S1.TIFF_TYPES[8][2])
    r["re"] = rng.integers(-3000, 3000, (lines, pixels))
    r["im"] = rng.integers(-3000, 3000, (lines, pixels))
    return r


@pytest.mark.parametrize("window", [(0, 0, 1200, 2048), (100, 37, 900, 1500)])
def test_tiff_read_device(tops, dev, tmp_path, window):
    """1200 x 2048 CInt16 (9.8 MB); the staging buffers are 32 MiB, and a
    window whose rows are not adjacent in the file is read row by row."""
    N = tops.native
    L = N.load()
    torch = dev.torch
    r = _cint16(1200, 2048, 1)
    p = tmp_path / "slc.tif"
    S1.write_tiff(p, r, 8, rows_per_strip=16)
    x0, y0, lines, pixels = window
    t = L.tops_tiff_open(str(p).encode())
    assert t, N.last_error()
    try:
        d = dev.empty(lines * pixels * 4)
        d.fill_(0x5A)
        s = torch.cuda.Stream()
        N.check(L.tops_tiff_read_device(t, x0, y0, lines, pixels, 8, d.data_ptr(),
                                        C.c_void_p(s.cuda_stream)), "tops_tiff_read_device")
        got = dev.get(d, r.dtype, (lines, pixels))
        assert np.array_equal(got, r[y0:y0 + lines, x0:x0 + pixels])
        # the null stream as well, twice in a row (staging buffers reused)
        for _ in range(2):
            d.zero_()
            N.check(L.tops_tiff_read_device(t, x0, y0, lines, pixels, 8, d.data_ptr(), None))
            assert np.array_equal(dev.get(d, r.dtype, (lines, pixels)),
                                  r[y0:y0 + lines, x0:x0 + pixels])
    finally:
        L.tops_tiff_close(t)


def test_tiff_large_read_device(tops, dev, tmp_path):
    """A raster larger than both staging buffers together (3 x 32 MiB + a
    partial chunk): every chunk lands at its own offset."""
    N = tops.native
    L = N.load()
    lines, pixels = 3300, 8192          # 108 MB of CInt16
    r = _cint16(lines, pixels, 2)
    p = tmp_path / "big.tif"
    S1.write_tiff(p, r, 8)
    t = L.tops_tiff_open(str(p).encode())
    try:
        d = dev.empty(lines * pixels * 4)
        N.check(L.tops_tiff_read_device(t, 0, 0, lines, pixels, 8, d.data_ptr(), None))
        assert np.array_equal(dev.get(d, r.dtype, (lines, pixels)), r)
    finally:
        L.tops_tiff_close(t)
