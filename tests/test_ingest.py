"""Ingest without GDAL or tinyxml2 (the library's tops_annotation_read,
tops_eof_read and TIFF reader/writer; no GPU work): annotation and EOF
documents written from known values (tests/s1_files.py) are read back field
by field, with the rules of S1_Initialize / ComputeDerampPara /
S1_OrbitInitialize (Func.cpp:135-819, 1828-2031) worked by hand where they
do more than copy a value; TIFF files from an independent encoder are read,
and the library's writer is checked by an independent parser.
"""
import ctypes as C

import numpy as np
import pytest

import s1_files as S1


@pytest.fixture(scope="module")
def sw(tops):
    """3-burst IW-like swath metadata (400 lines, 13 invalid lines at each
    end of a burst)."""
    return tops.synth.subswath_iw1(n_bursts=3, lines=400, pixels=64).master


def _read_annotation(tops, path):
    N = tops.native
    a = N.Annotation()
    rc = N.load().tops_annotation_read(str(path).encode(), C.byref(a))
    return rc, a


def _arr(p, n):
    return np.ctypeslib.as_array(p, shape=(n,)).copy()


def test_utc2seconds(tops):
    L = tops.native.load()
    assert L.tops_utc2seconds(b"2026-01-03T12:34:56.789012") == 12 * 3600 + 34 * 60 + 56.789012
    # the date is not part of the result
    assert L.tops_utc2seconds(b"1999-12-31T00:00:01.5") == 1.5


def test_annotation_fields(tops, sw, tmp_path):
    p = tmp_path / "s1a-iw1-slc-vv.xml"
    p.write_text(S1.annotation_xml(sw, decorate=True))
    N = tops.native
    rc, a = _read_annotation(tops, p)
    assert rc == 0, N.last_error()
    nb, lpb, spb, dt = sw.n_bursts, sw.lines_per_burst, sw.samples_per_burst, sw.az_time_interval
    try:
        assert a.polarisation == b"VV"
        assert a.first_line_utc.decode() == S1.utc(sw.first_line_time)
        assert a.last_line_utc.decode() == S1.utc(sw.last_line_time)
        assert (a.n_bursts, a.lines_per_burst, a.samples_per_burst) == (nb, lpb, spb)
        assert (a.num_lines, a.num_samples) == (nb * lpb, spb)
        assert a.az_time_interval == dt and a.prf == 1.0 / dt
        assert a.range_pixel_spacing == sw.range_spacing
        assert a.first_line_time == S1.seconds(S1.utc(sw.first_line_time))
        assert a.ascending_node_time == S1.seconds(S1.utc(sw.first_line_time - 1234.5))
        # slant-range time halved (two way -> one way)
        assert a.slr_time_first_pixel == sw.slr_time_first_pixel
        assert a.slr_time_last_pixel == (sw.slr_time_first_pixel +
                                         (spb - 1) * sw.range_spacing / S1.SOL)
        assert abs(a.wavelength - sw.wavelength) <= 1e-15 * sw.wavelength
        assert a.azimuth_bandwidth == sw.azimuth_bandwidth
        assert a.range_bandwidth == 5.616665e+07 and a.heading == -12.345678
        # burst times: first from the list, last = first + (lpb - 1) dt
        first = np.array([S1.seconds(S1.utc(t)) for t in sw.burst_first_line_time])
        assert np.array_equal(_arr(a.burst_first_line_time, nb), first)
        assert np.array_equal(_arr(a.burst_last_line_time, nb), first + (lpb - 1) * dt)
        # valid lines: 13 .. 386 of each burst
        assert _arr(a.first_valid_line, nb).tolist() == [13] * nb
        assert _arr(a.last_valid_line, nb).tolist() == [386] * nb
        assert np.array_equal(_arr(a.burst_first_valid_line_time, nb), first + 13 * dt)
        assert np.array_equal(_arr(a.burst_last_valid_line_time, nb), first + 386 * dt)
        fvs = _arr(a.first_valid_sample, nb * lpb).reshape(nb, lpb)
        assert (fvs[:, :13] == -1).all() and fvs[0, 13] == 10 + 13 % 3
        # valid pixels: min of firstValidSample (10), max of lastValidSample
        assert (a.first_valid_pixel, a.last_valid_pixel) == (10, spb - 11)
        assert a.slr_time_first_valid_pixel == (sw.slr_time_first_pixel +
                                                10 * sw.range_spacing / S1.SOL)
        # geolocation grid: 4 lines of 6 points; corners of grid rows b, b+1
        assert (a.geo_lines, a.geo_points_per_line) == (nb + 1, 6)
        lat = _arr(a.geo_latitude, 24).reshape(4, 6)
        lon = _arr(a.geo_longitude, 24).reshape(4, 6)
        ll = _arr(a.burst_ll, nb * 4).reshape(nb, 4)
        for b in range(nb):
            cl = [lat[b, 0], lat[b, 5], lat[b + 1, 0], lat[b + 1, 5]]
            cn = [lon[b, 0], lon[b, 5], lon[b + 1, 0], lon[b + 1, 5]]
            assert ll[b].tolist() == [min(cl), max(cl), min(cn), max(cn)]
        assert lat[2, 3] == 30.0 + 0.1 * 2 - 0.01 * 3
        assert a.near_range_on_left == 1
        srt = _arr(a.geo_slant_range_time, 24)
        assert srt[0] == sw.slr_time_first_pixel
        # DC and FM per burst
        dc = _arr(a.dc, nb * 4).reshape(nb, 4)
        fm = _arr(a.fm, nb * 4).reshape(nb, 4)
        assert np.allclose(dc, sw.dc, rtol=1e-13, atol=0)
        assert np.array_equal(fm, sw.fm)
    finally:
        N.load().tops_annotation_free(C.byref(a))


def test_annotation_valid_line_counting(tops, sw, tmp_path):
    """A valid-line run with a gap: the reference counts the valid lines
    after the first (Func.cpp:373-392), so the 'last' valid line of a burst
    whose lines 20..29 and 40..49 are valid is 20 + 19 = 39."""
    N = tops.native
    text = S1.annotation_xml(sw)
    lpb = sw.lines_per_burst
    f = ["-1"] * lpb
    for y in list(range(20, 30)) + list(range(40, 50)):
        f[y] = "7"
    old = text.split("<firstValidSample", 1)[1].split("</firstValidSample>", 1)[0]
    text = text.replace("<firstValidSample" + old + "</firstValidSample>",
                        f'<firstValidSample count="{lpb}">{" ".join(f)}</firstValidSample>', 1)
    p = tmp_path / "a.xml"
    p.write_text(text)
    rc, a = _read_annotation(tops, p)
    assert rc == 0, N.last_error()
    try:
        assert a.first_valid_line[0] == 20 and a.last_valid_line[0] == 39
        assert a.first_valid_line[1] == 13 and a.first_valid_pixel == 7
    finally:
        N.load().tops_annotation_free(C.byref(a))


def test_annotation_dc_selection_and_interpolation(tops, sw, tmp_path):
    """computeDC (Func.cpp:1828-1866) at burst centre times: linear
    interpolation between the bracketing estimates, extrapolation from the
    outer pair; more estimates than bursts: the first n_bursts as they are;
    dcMethod 'Data Analysis' takes dataDcPolynomial, otherwise the geometry
    polynomial (written here with c0 + 1)."""
    N = tops.native
    nb = sw.n_bursts
    first = np.array([S1.seconds(S1.utc(t)) for t in sw.burst_first_line_time])
    centre = 0.5 * (first + first + (sw.lines_per_burst - 1) * sw.az_time_interval)
    # two estimates, 10 s apart, straddling burst 0's centre; c0 linear in time
    t_a, t_b = S1.seconds(S1.utc(centre[0] - 0.25)), S1.seconds(S1.utc(centre[0] + 9.75))
    est = [(t_a, 0.005, 100.0, 2.0, 0.5), (t_b, 0.006, 200.0, 4.0, 1.5)]
    for method, off in (("Data Analysis", 0.0), ("Geometry", 1.0)):
        p = tmp_path / f"dc_{method[0]}.xml"
        p.write_text(S1.annotation_xml(sw, dc=est, dc_method=method))
        rc, a = _read_annotation(tops, p)
        assert rc == 0, N.last_error()
        try:
            dc = _arr(a.dc, nb * 4).reshape(nb, 4)
        finally:
            N.load().tops_annotation_free(C.byref(a))
        for b in range(nb):
            mu = (centre[b] - t_a) / (t_b - t_a)
            assert dc[b, 0] == 0.005                 # t0 of the earlier estimate
            exp = [(1 - mu) * (100.0 + off) + mu * (200.0 + off), (1 - mu) * 2.0 + mu * 4.0,
                   (1 - mu) * 0.5 + mu * 1.5]
            assert np.allclose(dc[b, 1:], exp, rtol=1e-14), (method, b)
        # burst 0's centre is 0.25 s after t_a (to the microsecond rounding of
        # the written times): mu = 0.025
        assert abs(dc[0, 1] - (100.0 + off + 2.5)) < 1e-4
    # more estimates than bursts: taken in order, no interpolation
    many = [(first[0] + k, 0.001 * k, 10.0 * k, 1.0 * k, 0.1 * k) for k in range(nb + 2)]
    p = tmp_path / "dc_many.xml"
    p.write_text(S1.annotation_xml(sw, dc=many))
    rc, a = _read_annotation(tops, p)
    assert rc == 0, N.last_error()
    try:
        dc = _arr(a.dc, nb * 4).reshape(nb, 4)
        for b in range(nb):
            assert dc[b].tolist() == [0.001 * b, 10.0 * b, 1.0 * b, 0.1 * b]
    finally:
        N.load().tops_annotation_free(C.byref(a))


def test_annotation_fm_forms_and_errors(tops, sw, tmp_path):
    N = tops.native
    nb = sw.n_bursts
    fm = [(sw.burst_first_line_time[b], 0.0053 + b, -2300.0 - b, 400.0 + b, -7.0) for b in
          range(nb)]
    for form in ("polynomial", "elements"):
        p = tmp_path / f"fm_{form}.xml"
        p.write_text(S1.annotation_xml(sw, fm=fm, fm_form=form))
        rc, a = _read_annotation(tops, p)
        assert rc == 0, N.last_error()
        try:
            got = _arr(a.fm, nb * 4).reshape(nb, 4)
            assert np.array_equal(got, np.array([f[1:] for f in fm])), form
        finally:
            N.load().tops_annotation_free(C.byref(a))
    p = tmp_path / "fm_short.xml"
    p.write_text(S1.annotation_xml(sw, fm=fm[:1]))
    rc, a = _read_annotation(tops, p)
    assert rc == N.TOPS_ERR_ARG and "FM-rate" in N.last_error()[1]
    p = tmp_path / "broken.xml"
    p.write_text(S1.annotation_xml(sw).replace("</swathTiming>", ""))
    rc, a = _read_annotation(tops, p)
    assert rc == N.TOPS_ERR_ARG
    rc, a = _read_annotation(tops, tmp_path / "missing.xml")
    assert rc == N.TOPS_ERR_ARG and "cannot open" in N.last_error()[1]
    N.load().tops_clear_error()


# ---- EOF ----------------------------------------------------------------------
def _eof(tmp_path, t0=43200 - 120, t1=43200 + 600):
    pos = lambda t: (7.0e6 + t, -1.0e5 + 2 * t, 3.0e5 - t)
    vel = lambda t: (1.0 + t * 1e-3, 7500.0, -2.0)
    text, ts = S1.eof_xml(t0, t1, pos, vel)
    p = tmp_path / "S1A_OPER_AUX_POEORB.EOF"
    p.write_text(text)
    return p, ts, pos, vel


def _read_eof(tops, p, first, last):
    N = tops.native
    o = N.EofOrbit()
    rc = N.load().tops_eof_read(str(p).encode(), first.encode(), last.encode(), C.byref(o))
    return rc, o


def test_eof_window(tops, tmp_path):
    """The vector in the 10 s slot of the first time is found (12:00:03 ->
    the 12:00:00 vector), the window starts 4 vectors earlier and holds
    tt/10 + 10 vectors, tt = 12:00:30 - 12:00:03 = 27 s."""
    N = tops.native
    p, ts, pos, vel = _eof(tmp_path)
    rc, o = _read_eof(tops, p, "2026-01-03T12:00:03.512", "2026-01-03T12:00:30.2")
    assert rc == 0, N.last_error()
    try:
        assert o.n == 27 // 10 + 10
        t = _arr(o.t, o.n)
        assert t[0] == 43200 - 40 and np.array_equal(np.diff(t), np.full(o.n - 1, 10.0))
        assert o.time_interval == 10.0
        x = _arr(o.x, o.n)
        assert np.allclose(x, [pos(v)[0] for v in t], rtol=0, atol=1e-6)
        vy = _arr(o.vy, o.n)
        assert (vy == 7500.0).all()
    finally:
        N.load().tops_eof_free(C.byref(o))


def test_eof_errors(tops, tmp_path):
    N = tops.native
    p, *_ = _eof(tmp_path)
    # no vector in the slot
    rc, _ = _read_eof(tops, p, "2026-01-03T18:00:00", "2026-01-03T18:00:20")
    assert rc == N.TOPS_ERR_ARG and "slot" in N.last_error()[1]
    # the window would start before the list (first vector at 11:58:00)
    rc, _ = _read_eof(tops, p, "2026-01-03T11:58:05", "2026-01-03T11:58:20")
    assert rc == N.TOPS_ERR_ARG
    # the window would run past the end
    rc, _ = _read_eof(tops, p, "2026-01-03T12:08:00", "2026-01-03T12:09:50")
    assert rc == N.TOPS_ERR_ARG
    N.load().tops_clear_error()


# ---- TIFF ---------------------------------------------------------------------
def _raster(gdt, lines, pixels, seed=0):
    rng = np.random.default_rng(seed)
    if gdt == 8:
        r = np.empty((lines, pixels), S1.TIFF_TYPES[8][2])
        r["re"] = rng.integers(-3000, 3000, (lines, pixels))
        r["im"] = rng.integers(-3000, 3000, (lines, pixels))
        return r
    if gdt == 10:
        return (rng.standard_normal((lines, pixels)) +
                1j * rng.standard_normal((lines, pixels))).astype(np.complex64)
    return rng.standard_normal((lines, pixels)).astype(S1.TIFF_TYPES[gdt][2])


def _tiff_read(tops, path, x0, y0, lines, pixels, gdt, dtype):
    N = tops.native
    L = N.load()
    t = L.tops_tiff_open(str(path).encode())
    assert t, N.last_error()
    try:
        info = (C.c_int(), C.c_int(), C.c_int())
        N.check(L.tops_tiff_info(t, *[C.byref(v) for v in info]))
        out = np.zeros((lines, pixels), dtype)
        rc = L.tops_tiff_read(t, x0, y0, lines, pixels, gdt, out.ctypes.data)
        return rc, tuple(v.value for v in info), out
    finally:
        L.tops_tiff_close(t)


@pytest.mark.parametrize("gdt", [8, 6, 10, 7])
@pytest.mark.parametrize("layout", ["le", "be", "big", "strips"])
def test_tiff_reader_against_independent_encoder(tops, tmp_path, gdt, layout):
    lines, pixels = 37, 53
    r = _raster(gdt, lines, pixels, seed=gdt)
    p = tmp_path / f"r_{gdt}_{layout}.tif"
    S1.write_tiff(p, r, gdt, big_endian=layout == "be", bigtiff=layout == "big",
                  rows_per_strip=5 if layout == "strips" else 1,
                  strip_gap=3 if layout == "strips" else 0)
    rc, info, whole = _tiff_read(tops, p, 0, 0, lines, pixels, gdt, r.dtype)
    assert rc == 0 and info == (pixels, lines, gdt)
    assert np.array_equal(whole.view(np.uint8), r.view(np.uint8))
    rc, _, win = _tiff_read(tops, p, 7, 11, 13, 20, gdt, r.dtype)
    assert rc == 0 and np.array_equal(win.view(np.uint8), r[11:24, 7:27].view(np.uint8))


def test_tiff_writer_layout_and_round_trip(tops, tmp_path):
    N = tops.native
    L = N.load()
    r = _raster(10, 29, 41, seed=5)
    p = tmp_path / "w.tif"
    t = L.tops_tiff_create(str(p).encode(), 10, 41, 29)
    assert t, N.last_error()
    # two windows: the top rows whole, then the rest column block by block
    N.check(L.tops_tiff_write(t, 0, 0, 10, 41, 10, np.ascontiguousarray(r[:10]).ctypes.data))
    for x0, x1 in ((0, 17), (17, 41)):
        blk = np.ascontiguousarray(r[10:, x0:x1])
        N.check(L.tops_tiff_write(t, x0, 10, 19, x1 - x0, 10, blk.ctypes.data))
    assert L.tops_tiff_write(t, 0, 25, 5, 41, 10, r.ctypes.data) == N.TOPS_ERR_ARG
    N.check(L.tops_tiff_close(t))
    tags, raw = S1.parse_tiff(p)
    assert tags[256] == [41] and tags[257] == [29] and tags[258] == [64]
    assert tags[259] == [1] and tags[277] == [1] and tags[339] == [6]
    offs, cnts = tags[273], tags[279]
    assert len(offs) == 29 and cnts == [41 * 8] * 29
    back = np.stack([np.frombuffer(raw[o:o + c], np.complex64) for o, c in zip(offs, cnts)])
    assert np.array_equal(back, r)
    rc, _, rd = _tiff_read(tops, p, 0, 0, 29, 41, 10, np.complex64)
    assert rc == 0 and np.array_equal(rd, r)
    N.load().tops_clear_error()


def test_tiff_overlap_bands(tops, tmp_path):
    """ReadCpxShort's overlap form (DataTypes.cpp:456-481): band i forward =
    lines (i+1)*lpb - n_i .. , backward = (i+1)*lpb .. , concatenated."""
    N = tops.native
    L = N.load()
    lpb, nb, width = 20, 3, 16
    r = _raster(8, lpb * nb, width, seed=9)
    p = tmp_path / "ov.tif"
    S1.write_tiff(p, r, 8)
    sizes = np.array([4, 6], np.int32)
    t = L.tops_tiff_open(str(p).encode())
    fwd = np.zeros((10, width), r.dtype)
    bwd = np.zeros((10, width), r.dtype)
    try:
        N.check(L.tops_tiff_read_overlaps(t, fwd.ctypes.data, bwd.ctypes.data, 2,
                                          N.ptr(sizes, C.c_int), lpb, 0, 0, width, 8))
    finally:
        L.tops_tiff_close(t)
    ef = np.concatenate([r[(i + 1) * lpb - n:(i + 1) * lpb] for i, n in enumerate(sizes)])
    eb = np.concatenate([r[(i + 1) * lpb:(i + 1) * lpb + n] for i, n in enumerate(sizes)])
    assert np.array_equal(fwd, ef) and np.array_equal(bwd, eb)


def test_tiff_errors(tops, tmp_path):
    N = tops.native
    L = N.load()
    r = _raster(6, 8, 8)
    p = tmp_path / "e.tif"
    S1.write_tiff(p, r, 6)
    rc, _, out = _tiff_read(tops, p, 4, 0, 8, 8, 6, np.float32)       # outside
    assert rc == N.TOPS_ERR_ARG and not out.any()
    rc, _, _ = _tiff_read(tops, p, 0, 0, 8, 8, 10, np.complex64)      # wrong type
    assert rc == N.TOPS_ERR_ARG
    raw = bytearray(open(p, "rb").read())
    tags, _ = S1.parse_tiff(p)
    # compression 5 (LZW): refused
    bad = raw.replace(b"\x03\x01\x03\x00\x01\x00\x00\x00\x01\x00",
                      b"\x03\x01\x03\x00\x01\x00\x00\x00\x05\x00", 1)
    assert bad != raw
    q = tmp_path / "lzw.tif"
    q.write_bytes(bytes(bad))
    assert not L.tops_tiff_open(str(q).encode())
    assert "compressed" in N.last_error()[1]
    assert not L.tops_tiff_open(str(tmp_path / "none.tif").encode())
    (tmp_path / "txt.tif").write_text("not a tiff")
    assert not L.tops_tiff_open(str(tmp_path / "txt.tif").encode())
    assert not L.tops_tiff_create(str(tmp_path / "x.tif").encode(), 3, 4, 4)
    N.load().tops_clear_error()


# ---- corrupted files -------------------------------------------------------------
def _mutants(rng, data, n):
    """n seeded corruptions of data: truncations, byte changes, deleted and
    repeated ranges."""
    out = []
    for _ in range(n):
        b = bytearray(data)
        for _ in range(int(rng.integers(1, 4))):
            if not b:
                break
            i = int(rng.integers(0, len(b)))
            k = int(rng.integers(0, 4))
            if k == 0:
                del b[i:]
            elif k == 1:
                b[i] = int(rng.integers(0, 256))
            elif k == 2:
                del b[i:i + int(rng.integers(1, 64))]
            else:
                b[i:i] = b[i:i + int(rng.integers(1, 200))]
        out.append(bytes(b))
    return out


def test_readers_reject_corrupted_files(tops, sw, tmp_path):
    """Corrupted annotation, EOF and TIFF files (seeded truncations, changed
    bytes, deleted and repeated ranges) are either read or refused with
    TOPS_ERR_ARG and a message: no reader crashes, reads past its buffers or
    leaves a half-filled result (every successful read is freed again).  The
    same mutations at a larger count ran clean against a build whose host
    code had AddressSanitizer."""
    N = tops.native
    L = N.load()
    rng = np.random.default_rng(20260103)
    p = tmp_path / "f.bin"
    ann = S1.annotation_xml(sw, decorate=True).encode()
    for b in _mutants(rng, ann, 150):
        p.write_bytes(b)
        a = N.Annotation()
        rc = L.tops_annotation_read(str(p).encode(), C.byref(a))
        assert rc in (0, N.TOPS_ERR_ARG)
        if rc == 0:
            L.tops_annotation_free(C.byref(a))
        else:
            assert N.last_error()[1]
    eof, *_ = _eof(tmp_path)
    for b in _mutants(rng, eof.read_bytes(), 150):
        p.write_bytes(b)
        o = N.EofOrbit()
        rc = L.tops_eof_read(str(p).encode(), b"2026-01-03T12:00:03.512",
                             b"2026-01-03T12:00:30.2", C.byref(o))
        assert rc in (0, N.TOPS_ERR_ARG)
        if rc == 0:
            L.tops_eof_free(C.byref(o))
    for gdt, layout in [(8, "le"), (10, "be"), (6, "big"), (8, "strips")]:
        q = tmp_path / f"t{gdt}{layout}.tif"
        S1.write_tiff(q, _raster(gdt, 9, 7), gdt, big_endian=layout == "be",
                      bigtiff=layout == "big", rows_per_strip=2 if layout == "strips" else 1,
                      strip_gap=3 if layout == "strips" else 0)
        for b in _mutants(rng, q.read_bytes(), 60):
            p.write_bytes(b)
            t = L.tops_tiff_open(str(p).encode())
            if not t:
                assert N.last_error()[1]
                continue
            info = (C.c_int(), C.c_int(), C.c_int())
            if L.tops_tiff_info(t, *[C.byref(v) for v in info]) == 0:
                w, h, g = (v.value for v in info)
                if 0 < w <= 4096 and 0 < h <= 4096:
                    buf = np.zeros(w * h * 16, np.uint8)
                    assert L.tops_tiff_read(t, 0, 0, h, w, g, buf.ctypes.data) in (0, N.TOPS_ERR_ARG)
            L.tops_tiff_close(t)
    L.tops_clear_error()
