"""ctypes wrapper of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  Never imported by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


class OrbitGeom(C.Structure):
    _fields_ = [("cx", _dp), ("cy", _dp), ("cz", _dp), ("ncoef", C.c_int),
                ("mid_time", C.c_double), ("az_time_interval", C.c_double),
                ("first_line_time", C.c_double), ("slr_time_first_pixel", C.c_double),
                ("range_spacing", C.c_double), ("az_init", C.c_double),
                ("line_offset", C.c_int)]


class ResampleGeom(C.Structure):
    _fields_ = [("master_box", C.c_int * 4), ("slave_box", C.c_int * 4),
                ("first_line_in_burst", C.c_int), ("az_time_interval", C.c_double),
                ("az_shift", C.c_double), ("cpm_az", C.c_double * 6),
                ("cpm_rg", C.c_double * 6)]


class Poly(C.Structure):
    _fields_ = [("t0", C.c_double), ("c0", C.c_double), ("c1", C.c_double), ("c2", C.c_double)]


class MatchSwath(C.Structure):
    _fields_ = [("n_bursts", C.c_int), ("lpb", C.c_int), ("spb", C.c_int), ("dt", C.c_double),
                ("prf", C.c_double), ("rsr", C.c_double), ("slr_first", C.c_double),
                ("slr_last", C.c_double), ("first_line_time", C.c_double),
                ("last_line_time", C.c_double), ("burst_first_line_time", _dp),
                ("geo_lines", C.c_int), ("geo_ppl", C.c_int), ("geo_az", _dp), ("geo_rg", _dp),
                ("geo_lat", _dp), ("geo_lon", _dp)]


class MatchOrbit(C.Structure):
    _fields_ = [("num_points", C.c_int), ("t", _dp), ("ncoef", C.c_int), ("cx", _dp),
                ("cy", _dp), ("cz", _dp)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
    L = C.CDLL(LIB_PATH)
    L.orc_lut_init.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _fp, _fp]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_max_threads.restype = C.c_int
    L.orc_polyfit_orbit.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_polyfit_orbit.restype = C.c_int
    for f in ("orc_orbit_pos", "orc_orbit_vel", "orc_orbit_acc"):
        getattr(L, f).argtypes = [C.c_double, _dp, _dp, _dp, C.c_int, _dp]
    L.orc_llh2xyz.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
    L.orc_ell2lp.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                             C.c_double, C.c_int, C.POINTER(OrbitGeom), _dp, _dp]
    L.orc_geocoreg.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                               C.c_double, C.c_int, C.POINTER(OrbitGeom), C.POINTER(OrbitGeom),
                               C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp,
                               C.POINTER(C.c_int64)]
    L.orc_geocoreg.restype = C.c_int
    L.orc_deramp_params.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                    C.c_double, C.c_int, C.POINTER(Poly), C.POINTER(Poly),
                                    C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
    L.orc_resample.argtypes = [C.c_void_p, C.POINTER(ResampleGeom), _fp, _fp, C.c_int, _dp, _dp,
                               _dp, C.c_int, C.c_int, C.c_int, _dp]
    L.orc_coherence.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.c_int, _dp]
    L.orc_esd.argtypes = [C.c_void_p, C.c_void_p, _dp, _dp, C.POINTER(C.c_int), C.c_int,
                          C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp,
                          _dp]
    L.orc_overlap_sizes.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_double,
                                    C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.orc_spectral_sep.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_double]
    L.orc_spectral_sep.restype = C.c_double
    L.orc_deburst_line.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_int, C.c_int]
    L.orc_deburst_line.restype = C.c_int
    L.orc_deburst_extent.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.orc_deburst.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                              C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                              C.POINTER(C.c_int)]
    L.orc_deburst.restype = C.c_int
    L.orc_xyz2aztime.restype = C.c_double
    L.orc_xyz2aztime.argtypes = [_dp, C.POINTER(MatchOrbit), C.c_double]
    L.orc_lp2xyz.restype = C.c_int
    L.orc_lp2xyz.argtypes = [C.c_double, C.c_double, C.POINTER(MatchSwath),
                             C.POINTER(MatchOrbit), C.c_int, _dp]
    L.orc_burst_overlap.restype = C.c_int
    L.orc_burst_overlap.argtypes = [C.POINTER(MatchSwath), C.POINTER(MatchSwath),
                                    C.POINTER(MatchOrbit), C.POINTER(MatchOrbit), C.c_int,
                                    C.c_int, C.POINTER(C.c_int)]
    _lib = L
    return L


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def lut(npoints=12, prf=1 / 2.055556e-3, abw=327.0, rsr2x=2 * 64.34523076923077e6):
    kaz = np.zeros((2048, npoints), np.float32)
    krg = np.zeros((2048, npoints), np.float32)
    load().orc_lut_init(npoints, prf, abw, rsr2x, _p(kaz, C.c_float), _p(krg, C.c_float))
    return kaz, krg


def polyfit_orbit(t, x, y, z):
    cx, cy, cz = np.zeros(6), np.zeros(6), np.zeros(6)
    n = load().orc_polyfit_orbit(len(t), _p(np.ascontiguousarray(t)), _p(np.ascontiguousarray(x)),
                                 _p(np.ascontiguousarray(y)), _p(np.ascontiguousarray(z)),
                                 _p(cx), _p(cy), _p(cz))
    return cx[:n].copy(), cy[:n].copy(), cz[:n].copy()


def orbit_eval(which, tn, cx, cy, cz):
    out = np.zeros(3)
    cx, cy, cz = (np.ascontiguousarray(a, dtype=np.float64) for a in (cx, cy, cz))
    getattr(load(), "orc_orbit_" + which)(tn, _p(cx), _p(cy), _p(cz), len(cx), _p(out))
    return out


def llh2xyz(lat_rad, lon_rad, h):
    out = np.zeros(3)
    load().orc_llh2xyz(lat_rad, lon_rad, h, _p(out))
    return out


def orbit_geom(coefs, mid_time, dt, first_line_time, slr0, rs, az_init, line_offset):
    cx, cy, cz = (np.ascontiguousarray(c, dtype=np.float64) for c in coefs)
    g = OrbitGeom(_p(cx), _p(cy), _p(cz), len(cx), mid_time, dt, first_line_time, slr0, rs,
                  az_init, line_offset)
    g._keep = (cx, cy, cz)
    return g


def ell2lp(dem, lat0, lon0, dlat, dlon, nodata, g):
    dem = np.ascontiguousarray(dem, dtype=np.int32)
    line = np.zeros(dem.shape)
    pix = np.zeros(dem.shape)
    load().orc_ell2lp(dem.ctypes.data, dem.shape[0], dem.shape[1], lat0, lon0, dlat, dlon,
                      nodata, C.byref(g), _p(line), _p(pix))
    return line, pix


def geocoreg(dem, lat0, lon0, dlat, dlon, nodata, gm, gs, box):
    dem = np.ascontiguousarray(dem, dtype=np.int32)
    ca, cr = np.zeros(6), np.zeros(6)
    used = C.c_int64(0)
    rc = load().orc_geocoreg(dem.ctypes.data, dem.shape[0], dem.shape[1], lat0, lon0, dlat, dlon,
                             nodata, C.byref(gm), C.byref(gs), box[0], box[1], box[2], box[3],
                             _p(ca), _p(cr), C.byref(used))
    return rc, ca, cr, used.value


def deramp_params(nb, spb, slr0, rs, lpb, dt, fvp, dc, fm, steer, wl, speed):
    dcs = (Poly * nb)(*[Poly(*map(float, r)) for r in dc])
    fms = (Poly * nb)(*[Poly(*map(float, r)) for r in fm])
    fdc, ka, tref, kt = (np.zeros((nb, spb)) for _ in range(4))
    load().orc_deramp_params(nb, spb, slr0, rs, lpb, dt, fvp, dcs, fms, steer, wl, speed,
                             _p(fdc), _p(ka), _p(tref), _p(kt))
    return kt, tref, fdc, ka


def set_threads(n):
    """Threads of the oracle's parallel loops (timing only)."""
    load().orc_set_threads(int(n))


def max_threads():
    return int(load().orc_max_threads())


def resample(slave, mbox, sbox, first_line, dt, az_shift, cpm_az, cpm_rg, kaz, krg, kt, tref,
             fdc, row0=0, row1=None, ref_float_phase=False):
    """complex128 [row1-row0, mP] resampled slave (FP64)."""
    slave = np.ascontiguousarray(slave, dtype=np.int16)
    g = ResampleGeom((C.c_int * 4)(*mbox), (C.c_int * 4)(*sbox), first_line, dt, az_shift,
                     (C.c_double * 6)(*cpm_az), (C.c_double * 6)(*cpm_rg))
    mp = mbox[1] - mbox[0] + 1
    ml = mbox[3] - mbox[2] + 1
    row1 = ml if row1 is None else row1
    out = np.zeros((row1 - row0, mp), dtype=np.complex128)
    kt, tref, fdc = (np.ascontiguousarray(a, dtype=np.float64) for a in (kt, tref, fdc))
    load().orc_resample(slave.ctypes.data, C.byref(g), _p(np.ascontiguousarray(kaz), C.c_float),
                        _p(np.ascontiguousarray(krg), C.c_float), kaz.shape[1], _p(kt), _p(tref),
                        _p(fdc), row0, row1, int(ref_float_phase),
                        out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def coherence(master, slave, dx, dy, win_rg, win_az):
    """master int16 [h,w,2], slave complex [h,w] -> coherence float64 [h,w]"""
    master = np.ascontiguousarray(master, dtype=np.int16)
    slave = np.ascontiguousarray(slave, dtype=np.complex128)
    h, w = slave.shape
    out = np.zeros((h, w))
    load().orc_coherence(master.ctypes.data, slave.ctypes.data_as(_dp), w, h, dx, dy, win_rg,
                         win_az, _p(out))
    return out


def esd(m_for, m_back, s_for, s_back, sizes, pixels, win_az, win_rg, thr, spectral_sep, dt):
    m_for, m_back = (np.ascontiguousarray(a, dtype=np.int16) for a in (m_for, m_back))
    s_for, s_back = (np.ascontiguousarray(a, dtype=np.complex128) for a in (s_for, s_back))
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    n = len(sizes)
    shifts = np.zeros(n)
    sums = np.zeros(2 * n)
    load().orc_esd(m_for.ctypes.data, m_back.ctypes.data, s_for.ctypes.data_as(_dp),
                   s_back.ctypes.data_as(_dp), sizes.ctypes.data_as(C.POINTER(C.c_int)), n,
                   pixels, win_az, win_rg, thr, spectral_sep, dt, _p(shifts), _p(sums))
    return shifts, sums.reshape(n, 2)


def overlap_sizes(first, last, b0, bn, dt):
    n = bn - b0
    sizes = np.zeros(max(n, 1), dtype=np.int32)
    total, mx = C.c_int(0), C.c_int(0)
    load().orc_overlap_sizes(_p(np.ascontiguousarray(first)), _p(np.ascontiguousarray(last)), b0,
                             bn, dt, sizes.ctypes.data_as(C.POINTER(C.c_int)), C.byref(total),
                             C.byref(mx))
    return sizes[:n], total.value, mx.value


def spectral_sep(kt, lpb, dt):
    kt = np.ascontiguousarray(kt, dtype=np.float64)
    return load().orc_spectral_sep(_p(kt), kt.shape[0], kt.shape[1], lpb, dt)


def _times(a):
    return _p(np.ascontiguousarray(a, dtype=np.float64))


def deburst_line(first, last, lpb, num_lines, dt, t, b0, bn):
    """getLineIdxInMetaData (Func.cpp:2766-2819)"""
    return load().orc_deburst_line(_times(first), _times(last), len(first), lpb, num_lines, dt, t,
                                   b0, bn)


def deburst_extent(first_valid_b0, last_valid_bn, first_line_time, dt):
    y0, h = C.c_int(0), C.c_int(0)
    load().orc_deburst_extent(first_valid_b0, last_valid_bn, first_line_time, dt, C.byref(y0),
                              C.byref(h))
    return y0.value, h.value


def deburst(first, last, first_valid, last_valid, lpb, num_lines, dt, first_line_time, b0, bn,
            gapped=None):
    """Stitch_OneSubswath (Func.cpp:2821-2933) over the burst-stacked raster
    `gapped` [(bn-b0+1)*lpb, P] of 4- or 8-byte elements (or None: rows
    only).  Returns (stitched raster or None, gapped row of each output row),
    or (None, None) where a line time falls in no burst."""
    first, last = (np.ascontiguousarray(a, dtype=np.float64) for a in (first, last))
    fv, lv = (np.ascontiguousarray(a, dtype=np.float64) for a in (first_valid, last_valid))
    _, h = deburst_extent(fv[b0], lv[bn], first_line_time, dt)
    rows = np.zeros(max(h, 1), np.int32)
    if gapped is not None:
        gapped = np.ascontiguousarray(gapped)
        assert gapped.shape[0] == (bn - b0 + 1) * lpb
        out = np.zeros((max(h, 0),) + gapped.shape[1:], gapped.dtype)
        pixels, eb = gapped.shape[1], gapped.strides[1]
        gp, op = gapped.ctypes.data, out.ctypes.data
    else:
        out, pixels, eb, gp, op = None, 0, 4, None, None
    rc = load().orc_deburst(_p(first), _p(last), _p(fv), _p(lv), len(first), lpb, pixels,
                            num_lines, dt, first_line_time, b0, bn, gp, eb, op,
                            rows.ctypes.data_as(C.POINTER(C.c_int)))
    if rc != 0:
        return None, None
    return out, rows[:h]
