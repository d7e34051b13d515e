"""GPU parity of ESD (EstESD.cu:914-1217) and geometric coregistration
(Geocoreg_cuda.cu:211-855) against the FP64 oracle.

Tolerances: ESD shift |d| <= 1e-4 px; coregistration offsets (the fitted
slave line/pixel over the burst grid) |d| <= 1e-3 px.  Errors are printed.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_ESD = 1e-4
TOL_OFFSET = 1e-3


@pytest.fixture(scope="module")
def esd_case(tops, orc, dev):
    """3 bursts of 400 x 2048 with ~100-line overlaps and a 0.002-line
    misregistration; slave bursts resampled on the GPU, bands gathered as
    Func.cpp:2564-2577 / DataTypes.cpp:456-513 do."""
    N = tops.native
    L = N.load()
    pair = tops.synth.subswath_iw1(n_bursts=3, lines=400, pixels=2048, delta=0.002)
    m = pair.master
    lpb, spb = m.lines_per_burst, m.samples_per_burst
    kaz, krg = N.lut_init(12, m.prf, m.azimuth_bandwidth, 2 * m.range_sampling_rate)
    kt, tref, fdc, _ = pair.params()
    ctx = L.tops_ctx_create(N.ptr(kaz, C.c_float), N.ptr(krg, C.c_float), 12)
    masters, reslaves = [], []
    dk, dtr, dfd = dev.put(kt[0]), dev.put(tref[0]), dev.put(fdc[0])
    for b in range(m.n_bursts):
        M = pair.burst(b, "master")
        S = pair.burst(b, "slave")
        ca, cr = pair.cpm(b)
        box = [0, spb - 1, b * lpb, (b + 1) * lpb - 1]
        g = N.BurstGeom((C.c_int * 4)(*box), (C.c_int * 4)(*box), b * lpb, spb,
                        m.az_time_interval, 0.0, (C.c_double * 6)(*ca), (C.c_double * 6)(*cr))
        ds = dev.put(S)
        do = dev.empty(lpb * spb * 8)
        N.check(L.tops_resample_burst(ctx, C.byref(g), ds.data_ptr(), dk.data_ptr(),
                                      dtr.data_ptr(), dfd.data_ptr(), do.data_ptr(), 0, None))
        masters.append(M)
        reslaves.append(dev.get(do, np.complex64, (lpb, spb)))
    L.tops_ctx_destroy(ctx)
    sizes, total, mx = orc.overlap_sizes(m.burst_first_line_time, m.burst_last_line_time, 0,
                                         m.n_bursts - 1, m.az_time_interval)
    mf = np.concatenate([masters[i][lpb - n:] for i, n in enumerate(sizes)])
    mb = np.concatenate([masters[i + 1][:n] for i, n in enumerate(sizes)])
    sf = np.concatenate([reslaves[i][lpb - n:] for i, n in enumerate(sizes)])
    sb = np.concatenate([reslaves[i + 1][:n] for i, n in enumerate(sizes)])
    df = orc.spectral_sep(kt, lpb, m.az_time_interval)
    return dict(pair=pair, sizes=sizes, mf=mf, mb=mb, sf=sf, sb=sb, df=df, spb=spb,
                dt=m.az_time_interval)


def test_esd_device_and_reference_entry(tops, orc, dev, esd_case):
    d = esd_case
    N = tops.native
    L = N.load()
    sizes = np.ascontiguousarray(d["sizes"], dtype=np.int32)
    n = len(sizes)
    ref, rsums = orc.esd(d["mf"], d["mb"], d["sf"].astype(np.complex128),
                         d["sb"].astype(np.complex128), sizes, d["spb"], 6, 24, 0.3, d["df"],
                         d["dt"])
    shifts = np.zeros(n)
    sums = np.zeros(2 * n)
    dmf, dmb, dsf, dsb = (dev.put(a) for a in (d["mf"], d["mb"], d["sf"], d["sb"]))
    print(f"  esd: {n} overlaps of {list(sizes)} lines x {d['spb']}", flush=True)
    N.check(L.tops_esd_shifts(dmf.data_ptr(), dmb.data_ptr(), dsf.data_ptr(), dsb.data_ptr(),
                              sizes.ctypes.data_as(C.POINTER(C.c_int)), n, d["spb"], 6, 24, 0.3,
                              d["df"], d["dt"], N.ptr(shifts), N.ptr(sums), None))
    err = np.max(np.abs(shifts - ref))
    print(f"  esd shifts gpu={shifts} oracle={ref} max|d|={err:.2e} px (injected 0.002)")
    assert err <= TOL_ESD
    # host-buffer entry point of EstESD.cu:28-44
    host = np.zeros(n)
    mf, mb = np.ascontiguousarray(d["mf"]), np.ascontiguousarray(d["mb"])
    sf, sb = np.ascontiguousarray(d["sf"]), np.ascontiguousarray(d["sb"])
    L.EstESDShifts(mf.ctypes.data, mb.ctypes.data, sf.ctypes.data, sb.ctypes.data,
                   sizes.ctypes.data_as(C.POINTER(C.c_int)), N.ptr(host), n, int(sizes.sum()),
                   d["spb"], 6, 24, int(sizes.max()), 0.3, d["df"], d["dt"])
    assert L.tops_last_error() == 0, N.last_error()
    assert np.max(np.abs(host - ref)) <= TOL_ESD


@pytest.mark.parametrize("win", [(6, 24), (3, 7), (4, 9), (17, 31), (2, 5), (1, 3)])
def test_esd_windows(tops, orc, dev, esd_case, win):
    """Coherence windows other than the reference's 6 x 24: the shared window
    sums (winAz >= 3 and winRg >= 7, up to 17 rows) and the direct sums
    (smaller windows), against the oracle's shifts (1e-4 px) and band phasor
    sums (|d| <= 1e-4 of the band's samples: a few threshold flips)."""
    d = esd_case
    N = tops.native
    L = N.load()
    waz, wrg = win
    sizes = np.ascontiguousarray(d["sizes"], dtype=np.int32)
    n = len(sizes)
    ref, rsums = orc.esd(d["mf"], d["mb"], d["sf"].astype(np.complex128),
                         d["sb"].astype(np.complex128), sizes, d["spb"], waz, wrg, 0.3, d["df"],
                         d["dt"])
    shifts, sums = np.zeros(n), np.zeros(2 * n)
    dmf, dmb, dsf, dsb = (dev.put(a) for a in (d["mf"], d["mb"], d["sf"], d["sb"]))
    N.check(L.tops_esd_shifts(dmf.data_ptr(), dmb.data_ptr(), dsf.data_ptr(), dsb.data_ptr(),
                              sizes.ctypes.data_as(C.POINTER(C.c_int)), n, d["spb"], waz, wrg,
                              0.3, d["df"], d["dt"], N.ptr(shifts), N.ptr(sums), None))
    err = float(np.max(np.abs(shifts - ref)))
    serr = np.abs(sums.reshape(n, 2) - np.asarray(rsums)).max(axis=1)
    print(f"  esd window {waz}x{wrg}: shift |d| = {err:.2e} px, sum |d| = {serr}")
    assert err <= TOL_ESD
    # one sample masked or signed differently moves a sum by 1 or 2
    assert np.all(serr <= 0.25)


def test_esd_thin_and_narrow_bands(tops, orc, dev):
    """Bands of 1, 3 and 17 lines, 70 samples wide (a partial tile row and
    a partial tile column), built with a known double-difference phase:
    shifts and band sums against the oracle."""
    N = tops.native
    L = N.load()
    P, sizes = 70, np.array([1, 3, 17], np.int32)
    rng = np.random.default_rng(77)
    n = int(sizes.sum())
    sf = (rng.normal(size=(n, P)) + 1j * rng.normal(size=(n, P))) * 200
    sb = (rng.normal(size=(n, P)) + 1j * rng.normal(size=(n, P))) * 200
    noise = lambda: (rng.normal(size=(n, P)) + 1j * rng.normal(size=(n, P))) * 40
    q = lambda z: np.stack([np.round(z.real), np.round(z.imag)], -1).astype(np.int16)
    mf, mb = q(sf * np.exp(0.3j) + noise()), q(sb + noise())
    sf, sb = sf.astype(np.complex64), sb.astype(np.complex64)
    ref, rsums = orc.esd(mf, mb, sf.astype(np.complex128), sb.astype(np.complex128), sizes, P, 6,
                         24, 0.3, 5400.0, 2.055556e-3)
    d = [dev.put(np.ascontiguousarray(a)) for a in (mf, mb, sf, sb)]
    shifts, sums = np.zeros(3), np.zeros(6)
    N.check(L.tops_esd_shifts(d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr(),
                              sizes.ctypes.data_as(C.POINTER(C.c_int)), 3, P, 6, 24, 0.3, 5400.0,
                              2.055556e-3, N.ptr(shifts), N.ptr(sums), None))
    err = float(np.max(np.abs(shifts - ref)))
    serr = float(np.max(np.abs(sums.reshape(3, 2) - np.asarray(rsums))))
    print(f"  thin bands {sizes.tolist()} x {P}: shifts {shifts}, |d| = {err:.2e} px, sums |d| = {serr:.2e}")
    assert np.all(np.abs(np.asarray(rsums)).sum(axis=1) > 0)
    assert err <= TOL_ESD and serr <= 0.25


@pytest.mark.parametrize("sizes,P", [([16, 32], 300), ([100, 40], 300), ([20, 5, 21], 301)])
def test_esd_band_heights(tops, orc, dev, sizes, P):
    """The reference's 6 x 24 window over bands that end on a 16-row tile
    boundary ([16, 32]), inside a tile row ([100, 40]: 4 and 8 kept rows in
    the last one) and mixed ([20, 5, 21]), 300 or 301 samples wide (a partial
    tile column, an odd width), with a known double-difference phase: shifts
    and band sums against the oracle."""
    N = tops.native
    L = N.load()
    sizes = np.array(sizes, np.int32)
    rng = np.random.default_rng(int(sizes.sum()))
    n = int(sizes.sum())
    sf = (rng.normal(size=(n, P)) + 1j * rng.normal(size=(n, P))) * 200
    sb = (rng.normal(size=(n, P)) + 1j * rng.normal(size=(n, P))) * 200
    noise = lambda: (rng.normal(size=(n, P)) + 1j * rng.normal(size=(n, P))) * 40
    q = lambda z: np.stack([np.round(z.real), np.round(z.imag)], -1).astype(np.int16)
    mf, mb = q(sf * np.exp(0.25j) + noise()), q(sb + noise())
    sf, sb = sf.astype(np.complex64), sb.astype(np.complex64)
    k = len(sizes)
    ref, rsums = orc.esd(mf, mb, sf.astype(np.complex128), sb.astype(np.complex128), sizes, P, 6,
                         24, 0.3, 5400.0, 2.055556e-3)
    d = [dev.put(np.ascontiguousarray(a)) for a in (mf, mb, sf, sb)]
    shifts, sums = np.zeros(k), np.zeros(2 * k)
    N.check(L.tops_esd_shifts(d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr(),
                              sizes.ctypes.data_as(C.POINTER(C.c_int)), k, P, 6, 24, 0.3, 5400.0,
                              2.055556e-3, N.ptr(shifts), N.ptr(sums), None))
    err = float(np.max(np.abs(shifts - ref)))
    serr = float(np.max(np.abs(sums.reshape(k, 2) - np.asarray(rsums))))
    print(f"  bands {sizes.tolist()} x {P}: |d shift| = {err:.2e} px, sums |d| = {serr:.2e}")
    assert np.all(np.abs(np.asarray(rsums)).sum(axis=1) > 0)
    assert err <= TOL_ESD and serr <= 0.25


def test_esd_zero_products_take_the_sign_of_zero(tops, dev):
    """Invalid (all-zero) forward master lines, as at a real burst's edge:
    every sample passes the mask with a double difference of signed zeros,
    whose phasor is -1 where Re Sf > 0 and +1 elsewhere (the reference's
    atan2 of the products it forms); the band sum is that count exactly."""
    from conftest import zero_master_band
    N = tops.native
    L = N.load()
    n, P = 12, 2048
    mf, mb, sf, sb, expect = zero_master_band(n, P, 5)
    sizes = np.array([n], np.int32)
    d = [dev.put(np.ascontiguousarray(a)) for a in (mf, mb, sf.astype(np.complex64),
                                                     sb.astype(np.complex64))]
    shifts, sums = np.zeros(1), np.zeros(2)
    N.check(L.tops_esd_shifts(d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr(),
                              sizes.ctypes.data_as(C.POINTER(C.c_int)), 1, P, 6, 24, 0.3, 5400.0,
                              2.055556e-3, N.ptr(shifts), N.ptr(sums), None))
    print(f"  zero forward master band: sums {sums}, expected ({expect}, 0)")
    assert sums[0] == expect and sums[1] == 0.0


def _orbit_coefs(tops, orbit):
    t, x, y, z = tops.synth.orbit_polyfit_inputs(orbit)
    return t, tops.native.polyfit_orbit(t, x, y, z)


@pytest.fixture(scope="module")
def geo_case(tops, orc):
    pair = tops.synth.single_burst_small()
    m, s = pair.master, pair.slave
    tm, cm = _orbit_coefs(tops, pair.morbit)
    ts, cs = _orbit_coefs(tops, pair.sorbit)
    dem, lat0, lon0, dlat, dlon = pair.dem_window(*m.burst_ll[0], nodata_every=53)

    def og(sw, t, c, N):
        first, last = sw.burst_first_line_time[0], sw.burst_last_line_time[0]
        return dict(coefs=c, mid=t[len(t) // 2], dt=sw.az_time_interval, first=first,
                    slr=sw.slr_time_first_pixel, rs=sw.range_spacing, init=0.5 * (first + last))
    return dict(pair=pair, dem=dem, lat0=lat0, lon0=lon0, dlat=dlat, dlon=dlon,
                gm=og(m, tm, cm, 0), gs=og(s, ts, cs, 0), L=m.lines_per_burst,
                P=m.samples_per_burst)


def _surface(ca, cr):
    gl, gp = np.meshgrid(np.linspace(-2, 2, 17), np.linspace(-2, 2, 17))
    A = np.stack([np.ones_like(gl), gl, gp, gl * gl, gl * gp, gp * gp], axis=-1)
    return A @ ca, A @ cr


def test_geocoreg_parity(tops, orc, dev, geo_case):
    d = geo_case
    N = tops.native
    L = N.load()
    box = [0, d["P"] - 1, 0, d["L"] - 1]
    keep = []

    def ng(g):
        cx, cy, cz = (np.ascontiguousarray(c) for c in g["coefs"])
        keep.append((cx, cy, cz))
        return N.OrbitGeom(N.ptr(cx), N.ptr(cy), N.ptr(cz), len(cx), g["mid"], g["dt"], g["first"],
                           g["slr"], g["rs"], g["init"], 0)

    def og(g):
        return orc.orbit_geom(g["coefs"], g["mid"], g["dt"], g["first"], g["slr"], g["rs"],
                              g["init"], 0)

    dem = d["dem"]
    ddem = dev.put(dem)
    n = dem.size
    dlp = dev.empty(4 * n * 8)
    ca, cr = np.zeros(6), np.zeros(6)
    used = C.c_int64(0)
    gm, gs = ng(d["gm"]), ng(d["gs"])
    print(f"  geocoreg: DEM {dem.shape}", flush=True)
    N.check(L.tops_geocoreg(ddem.data_ptr(), dem.shape[0], dem.shape[1], d["lat0"], d["lon0"],
                            d["dlat"], d["dlon"], -32768, C.byref(gm), C.byref(gs), *box,
                            N.ptr(ca), N.ptr(cr), C.byref(used), dlp.data_ptr(), None))
    rc, oa, orr, oused = orc.geocoreg(dem, d["lat0"], d["lon0"], d["dlat"], d["dlon"], -32768,
                                      og(d["gm"]), og(d["gs"]), box)
    assert rc == 0 and used.value == oused
    ga, gr = _surface(ca, cr)
    ra, rr = _surface(oa, orr)
    e = max(np.max(np.abs(ga - ra)), np.max(np.abs(gr - rr)))
    print(f"  geocoreg: {oused} posts, max offset |d| = {e:.2e} px")
    assert e <= TOL_OFFSET
    lp = dev.get(dlp, np.float64, (4,) + dem.shape)
    ml, mp = orc.ell2lp(dem, d["lat0"], d["lon0"], d["dlat"], d["dlon"], -32768, og(d["gm"]))
    # the per-post planes: the FP32 DEM projection (the configs' precision
    # split) moves each post by under a millimetre, about 1e-4 px of range
    el, ep = float(np.max(np.abs(lp[0] - ml))), float(np.max(np.abs(lp[1] - mp)))
    print(f"  geocoreg: per-post master line |d| = {el:.2e}, pixel |d| = {ep:.2e} "
          f"(tol {TOL_OFFSET:g})")
    assert el <= TOL_OFFSET and ep <= TOL_OFFSET
    # host-buffer entry point (Geocoreg_cuda.cu:22-68)
    ha, hr = np.zeros(6), np.zeros(6)
    cmx, cmy, cmz = (np.ascontiguousarray(c) for c in d["gm"]["coefs"])
    csx, csy, csz = (np.ascontiguousarray(c) for c in d["gs"]["coefs"])
    tm = np.array([0.0] * 5 + [d["gm"]["mid"]] + [0.0] * 4)
    tsv = np.array([0.0] * 5 + [d["gs"]["mid"]] + [0.0] * 4)
    demc = np.ascontiguousarray(dem)
    L.GeoCoreg_warp_cuda(0, 0, *box, d["dlat"], d["dlon"], dem.shape[0], dem.shape[1],
                         N.ptr(ha), N.ptr(hr), d["lat0"], d["lon0"], -32768, d["gm"]["dt"],
                         d["gs"]["dt"], 0.0555, float(d["L"]), float(d["L"]), d["gm"]["first"],
                         d["gs"]["first"], d["P"], d["P"], N.ptr(tm), N.ptr(tsv), 10, len(cmx),
                         10, len(csx), d["gm"]["slr"], d["gs"]["slr"], d["gm"]["rs"],
                         d["gs"]["rs"], False, False,
                         demc.ctypes.data_as(C.POINTER(C.c_int)), N.ptr(cmx), N.ptr(cmy),
                         N.ptr(cmz), N.ptr(csx), N.ptr(csy), N.ptr(csz), d["gm"]["init"],
                         d["gs"]["init"])
    assert L.tops_last_error() == 0, N.last_error()
    assert np.allclose(ha, ca, rtol=0, atol=1e-9) and np.allclose(hr, cr, rtol=0, atol=1e-9)


def test_geocoreg_all_nodata_status(tops, dev, geo_case):
    d = geo_case
    N = tops.native
    L = N.load()
    dem = np.full((16, 16), -32768, np.int32)
    ddem = dev.put(dem)
    cx, cy, cz = (np.ascontiguousarray(c) for c in d["gm"]["coefs"])
    g = N.OrbitGeom(N.ptr(cx), N.ptr(cy), N.ptr(cz), len(cx), d["gm"]["mid"], d["gm"]["dt"],
                    d["gm"]["first"], d["gm"]["slr"], d["gm"]["rs"], d["gm"]["init"], 0)
    ca, cr = np.ones(6), np.ones(6)
    rc = L.tops_geocoreg(ddem.data_ptr(), 16, 16, d["lat0"], d["lon0"], d["dlat"], d["dlon"],
                         -32768, C.byref(g), C.byref(g), 0, 100, 0, 100, N.ptr(ca), N.ptr(cr),
                         None, None, None)
    assert rc == N.TOPS_ERR_NOT_PD
    assert np.all(ca == 0) and np.all(cr == 0)
    L.tops_clear_error()
