"""Burst matching (ComputerBurstOverlap, Func.cpp:1610-1746, with lp2xyz
1492-1584 and xyz2aztime 1349-1377): the oracle pinned by constructed
geometry (a ground point built at zero Doppler of the true orbit at a chosen
time; the range, Doppler and ellipsoid equations at lp2xyz's answer) and by
burst ranges worked by hand for pairs whose slave burst grid is offset by
-2, 0 and +2 bursts; then the library (tops_xyz2aztime, tops_lp2xyz,
tops_burst_overlap, host code) against the oracle on the same inputs.
"""
import ctypes as C

import numpy as np
import pytest

import s1_files as S1

LINES, PIXELS, NB = 400, 2048, 6


def _orbit_fit(tops, orbit):
    N = tops.native
    t, x, y, z = tops.synth.orbit_polyfit_inputs(orbit)
    cx, cy, cz = N.polyfit_orbit(t, x, y, z)
    return dict(t=np.ascontiguousarray(t), cx=cx, cy=cy, cz=cz)


def _lib_orbit(tops, o):
    N = tops.native
    return N.Orbit(len(o["t"]), N.ptr(o["t"]), len(o["cx"]), N.ptr(o["cx"]), N.ptr(o["cy"]),
                   N.ptr(o["cz"]))


def _orc_orbit(orc, o):
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    return orc.MatchOrbit(len(o["t"]), p(o["t"]), len(o["cx"]), p(o["cx"]), p(o["cy"]),
                          p(o["cz"]))


def _orc_swath(orc, a):
    """oracle view of a library annotation (arrays shared)"""
    return orc.MatchSwath(a.n_bursts, a.lines_per_burst, a.samples_per_burst, a.az_time_interval,
                          a.prf, a.range_sampling_rate, a.slr_time_first_pixel,
                          a.slr_time_last_pixel, a.first_line_time, a.last_line_time,
                          a.burst_first_line_time, a.geo_lines, a.geo_points_per_line,
                          a.geo_az_time, a.geo_slant_range_time, a.geo_latitude,
                          a.geo_longitude)


@pytest.fixture(scope="module")
def pairs(tops, tmp_path_factory):
    """Pairs with the slave's burst grid offset by k bursts, their
    annotations (grid geolocated on the true orbits) and fitted orbits."""
    N = tops.native
    L = N.load()
    d = tmp_path_factory.mktemp("bm")
    out = {}
    keep = []
    for k in (-2, 0, 2, 7, -7):
        pair = tops.synth.make_pair(f"bm{k}", LINES, PIXELS, n_bursts=NB, overlap=100,
                                    slave_burst_offset=k)
        ann = []
        for which, sw, orb in (("m", pair.master, pair.morbit), ("s", pair.slave, pair.sorbit)):
            guess = [pair.scene_center]

            def grid_ll(t, r, orb=orb, guess=guess):
                ll = tops.synth.geolocate(orb, t, r, guess[0])
                guess[0] = ll
                return ll
            p = d / f"{which}{k}.xml"
            p.write_text(S1.annotation_xml(sw, grid_ll=grid_ll))
            a = N.Annotation()
            N.check(L.tops_annotation_read(str(p).encode(), C.byref(a)))
            ann.append(a)
        om, os_ = _orbit_fit(tops, pair.morbit), _orbit_fit(tops, pair.sorbit)
        keep.append((om, os_))
        out[k] = dict(pair=pair, m=ann[0], s=ann[1], om=om, os=os_)
    yield out
    for v in out.values():
        L.tops_annotation_free(C.byref(v["m"]))
        L.tops_annotation_free(C.byref(v["s"]))


def test_oracle_xyz2aztime_constructed(tops, orc, pairs):
    """A ground point at zero Doppler of the true slave orbit at t0 (synth
    geolocate) returns t0; the fitted polynomial deviates from the true
    orbit by millimetres, so 1e-6 s."""
    pr = pairs[0]
    pair, so = pr["pair"], _orc_orbit(orc, pr["os"])
    for dt_off, r in ((0.3, 850e3), (1.7, 870e3), (2.9, 880e3)):
        t0 = pair.slave.first_line_time + dt_off
        lat, lon = tops.synth.geolocate(pair.sorbit, t0, r, pair.scene_center)
        xyz = np.ascontiguousarray(tops.synth.llh2xyz(lat, lon, 0.0))
        t = orc.load().orc_xyz2aztime(xyz.ctypes.data_as(C.POINTER(C.c_double)), C.byref(so),
                                      t0 + 0.4)
        assert abs(t - t0) < 1e-6, (t - t0)


def test_oracle_lp2xyz_equations(tops, orc, pairs):
    """lp2xyz's answer lies on the WGS84 ellipsoid, at the slant range of
    the pixel and at zero Doppler of the line's time (the equations it
    solves), and on the true orbit's geolocation of that line / pixel."""
    pr = pairs[0]
    pair, a = pr["pair"], pr["m"]
    sw, mo = _orc_swath(orc, a), _orc_orbit(orc, pr["om"])
    for b, line, pixel in ((0, 200, 1024), (3, 1, 1), (5, 400, 2048)):
        lat, lon = pair.scene_center
        xyz = np.ascontiguousarray(tops.synth.llh2xyz(lat, lon, 0.0))
        orc.load().orc_lp2xyz(line, pixel, C.byref(sw), C.byref(mo), b,
                              xyz.ctypes.data_as(C.POINTER(C.c_double)))
        t = a.burst_first_line_time[b] + (line - 1.0) / a.prf
        rt = a.slr_time_first_pixel + (pixel - 1.0) / (a.range_sampling_rate * 2)
        A, B = 6378137.0, 6356752.3142451794975639665996337
        assert abs((xyz[0] ** 2 + xyz[1] ** 2) / A ** 2 + (xyz[2] / B) ** 2 - 1) < 1e-12
        s, v = pair.morbit.pos(t), pair.morbit.vel(t)
        assert abs(np.linalg.norm(xyz - s) - S1.SOL * rt) < 1e-2
        assert abs((xyz - s) @ v / np.linalg.norm(v)) < 1e-2
        tl, tn = tops.synth.geolocate(pair.morbit, t, S1.SOL * rt, pair.scene_center)
        assert np.linalg.norm(xyz - tops.synth.llh2xyz(tl, tn, 0.0)) < 1e-2


# (offset k, 1-based burst0, burstN) -> (m_burst0, m_burstN, s_burst0, s_burstN)
HAND = [
    # aligned: master burst b is slave burst b
    ((0, 1, 6), (0, 5, 0, 5)),
    ((0, 2, 4), (1, 3, 1, 3)),
    # slave starts 2 bursts later: its burst 0 is master burst 2; the slave's
    # 6 bursts allow master bursts 2 .. 5 of those asked for
    ((2, 1, 6), (2, 5, 0, 3)),
    # slave starts 2 bursts earlier: master burst 0 is slave burst 2; slave
    # bursts 2 .. 5 cover master bursts 0 .. 3
    ((-2, 1, 6), (0, 3, 2, 5)),
    ((-2, 1, 3), (0, 2, 2, 4)),
]


@pytest.mark.parametrize("case,expect", HAND)
def test_oracle_burst_overlap_hand(orc, pairs, case, expect):
    k, b0, bn = case
    pr = pairs[k]
    out = (C.c_int * 4)()
    rc = orc.load().orc_burst_overlap(C.byref(_orc_swath(orc, pr["m"])),
                                      C.byref(_orc_swath(orc, pr["s"])),
                                      C.byref(_orc_orbit(orc, pr["om"])),
                                      C.byref(_orc_orbit(orc, pr["os"])), b0, bn, out)
    assert rc == 0 and tuple(out) == expect


@pytest.mark.parametrize("k", [7, -7])
def test_oracle_no_overlap(orc, pairs, k):
    pr = pairs[k]
    out = (C.c_int * 4)()
    rc = orc.load().orc_burst_overlap(C.byref(_orc_swath(orc, pr["m"])),
                                      C.byref(_orc_swath(orc, pr["s"])),
                                      C.byref(_orc_orbit(orc, pr["om"])),
                                      C.byref(_orc_orbit(orc, pr["os"])), 1, 6, out)
    assert rc == -2


def test_library_matches_oracle(tops, orc, pairs):
    N = tops.native
    L = N.load()
    for k, pr in pairs.items():
        lm, ls = _lib_orbit(tops, pr["om"]), _lib_orbit(tops, pr["os"])
        om, os_ = _orc_orbit(orc, pr["om"]), _orc_orbit(orc, pr["os"])
        a = pr["m"]
        # lp2xyz from the same start point
        x0 = np.ascontiguousarray(tops.synth.llh2xyz(*pr["pair"].scene_center, 0.0))
        xl, xo = x0.copy(), x0.copy()
        it = L.tops_lp2xyz(200, 1024, C.byref(a), C.byref(lm), 1, N.ptr(xl), 10, 1e-6)
        orc.load().orc_lp2xyz(200, 1024, C.byref(_orc_swath(orc, a)), C.byref(om), 1,
                              xo.ctypes.data_as(C.POINTER(C.c_double)))
        assert 0 <= it <= 10 and np.max(np.abs(xl - xo)) < 1e-6, k
        tl = L.tops_xyz2aztime(N.ptr(xl), C.byref(ls), a.burst_first_line_time[1])
        to = orc.load().orc_xyz2aztime(xl.ctypes.data_as(C.POINTER(C.c_double)), C.byref(os_),
                                       a.burst_first_line_time[1])
        assert abs(tl - to) < 1e-9, k
        for b0, bn in ((1, 6), (2, 4), (1, 3)):
            out = [C.c_int() for _ in range(4)]
            rc = L.tops_burst_overlap(C.byref(pr["m"]), C.byref(pr["s"]), C.byref(lm),
                                      C.byref(ls), b0, bn, *[C.byref(v) for v in out])
            ref = (C.c_int * 4)()
            orc_rc = orc.load().orc_burst_overlap(C.byref(_orc_swath(orc, pr["m"])),
                                                  C.byref(_orc_swath(orc, pr["s"])),
                                                  C.byref(om), C.byref(os_), b0, bn, ref)
            if orc_rc == 0:
                assert rc == 0 and tuple(v.value for v in out) == tuple(ref), (k, b0, bn)
            else:
                assert rc == N.TOPS_ERR_ARG and "overlap" in N.last_error()[1], (k, b0, bn)
    assert L.tops_burst_overlap(C.byref(pairs[0]["m"]), C.byref(pairs[0]["s"]),
                                C.byref(_lib_orbit(tops, pairs[0]["om"])),
                                C.byref(_lib_orbit(tops, pairs[0]["os"])), 0, 6,
                                *[C.byref(C.c_int()) for _ in range(4)]) == N.TOPS_ERR_ARG
    L.tops_clear_error()
