"""C-ABI checks that need no GPU: libtops_mi355x.so loads, exports every
function include/tops_abi.h declares, and its host stage-layer functions
(InKernelInit, Polyfit_Orbit, ComputeDerampPara formulas, EstOverlapSize,
GetSpectralSep) agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np

from conftest import ROOT


def _header_functions():
    text = open(os.path.join(ROOT, "include", "tops_abi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "sizeof", "defined"}
    return sorted({n for n in names if n not in skip})


def test_library_exports_header(tops):
    L = tops.native.load()
    names = _header_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(tops.native.EXPORTS) <= set(names)


def test_single_hip_runtime(tops):
    """The library shares torch's HIP runtime: one libamdhip64 and one HSA
    runtime mapped in the process (a second instance finds no device)."""
    tops.native.load()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64" in ln or "libhsa-runtime64" in ln}
    hip = [p for p in paths if "libamdhip64" in p]
    hsa = [p for p in paths if "libhsa-runtime64" in p]
    assert len(hip) == 1 and len(hsa) == 1, sorted(paths)


def test_status_api(tops):
    L = tops.native.load()
    L.tops_clear_error()
    assert L.tops_last_error() == 0
    # argument errors are reported without touching a device
    rc = L.tops_lut_init(0, 1.0, 1.0, 1.0, None, None)
    assert rc == tops.native.TOPS_ERR_ARG
    assert L.tops_last_error() == tops.native.TOPS_ERR_ARG
    assert b"tops_lut_init" in L.tops_last_error_message()
    L.tops_clear_error()


def test_lut_matches_oracle(tops, orc):
    for args in ((12, 1 / 2.055556e-3, 327.0, 2 * 64.34523076923077e6),
                 (12, 1 / 2.758e-3, 313.0, 2 * 64.34523076923077e6)):
        kaz, krg = tops.native.lut_init(*args)
        oaz, org = orc.lut(*args)
        # same formula in FP64, stored as float: equal to the last float bit
        assert np.max(np.abs(kaz.view(np.int32) - oaz.view(np.int32))) <= 1
        assert np.max(np.abs(krg.view(np.int32) - org.view(np.int32))) <= 1


def test_polyfit_matches_oracle(tops, orc):
    orbit = tops.synth.make_orbit(30.0, 100.0, 43200.0)
    t, x, y, z = tops.synth.orbit_polyfit_inputs(orbit)
    a = tops.native.polyfit_orbit(t, x, y, z)
    b = orc.polyfit_orbit(t, x, y, z)
    for ca, cb in zip(a, b):
        assert len(ca) == len(cb) == 6
        # evaluate both fits over the orbit span: same track to 1e-6 m
        for tn in np.linspace(-4.5, 4.5, 11):
            pa = sum(ca[j] * tn ** j for j in range(6))
            pb = sum(cb[j] * tn ** j for j in range(6))
            assert abs(pa - pb) < 1e-6


def test_deramp_params_match_oracle(tops, orc):
    pair = tops.synth.subswath_iw1(n_bursts=3, lines=200, pixels=300)
    m = pair.master
    m.dc[1, 3] = 1e9      # a quadratic Doppler term in burst 1
    m.fm[2, 2] = -3e5     # a linear FM-rate term in burst 2
    got = tops.native.deramp_params(m, 7421.3)
    ref = orc.deramp_params(m.n_bursts, m.samples_per_burst, m.slr_time_first_pixel,
                            m.range_spacing, m.lines_per_burst, m.az_time_interval,
                            m.first_valid_pixel, m.dc, m.fm, m.steering_rate_deg, m.wavelength,
                            7421.3)
    for g, r in zip(got, ref):
        assert np.allclose(g, r, rtol=1e-13, atol=1e-13)


def test_overlap_and_spectral_sep_match_oracle(tops, orc):
    pair = tops.synth.subswath_iw1(n_bursts=4, lines=300, pixels=64)
    m = pair.master
    L = tops.native.load()
    sizes = np.zeros(8, np.int32)
    tot, mx = C.c_int(0), C.c_int(0)
    n = L.tops_overlap_sizes(tops.native.ptr(m.burst_first_line_time),
                             tops.native.ptr(m.burst_last_line_time), 0, 3, m.az_time_interval,
                             sizes.ctypes.data_as(C.POINTER(C.c_int)), C.byref(tot), C.byref(mx))
    osz, otot, omx = orc.overlap_sizes(m.burst_first_line_time, m.burst_last_line_time, 0, 3,
                                       m.az_time_interval)
    assert n == 3 and list(sizes[:3]) == list(osz) and tot.value == otot and mx.value == omx
    assert list(osz) == [99, 99, 99]
    kt = np.ascontiguousarray(pair.params()[0])
    a = L.tops_spectral_sep(tops.native.ptr(kt), kt.shape[0], kt.shape[1], m.lines_per_burst,
                            m.az_time_interval)
    b = orc.spectral_sep(kt, m.lines_per_burst, m.az_time_interval)
    assert abs(a - b) <= 1e-9 * abs(b)


def test_read_config(tops, tmp_path):
    """config.txt semantics of ReadInput (Func.cpp:7-130)."""
    p = tmp_path / "config.txt"
    p.write_text("process_dir= /data/run one\nmasterpath=/m.SAFE\nslavepath\t/s.SAFE\n"
                 "firstsubswath=1\nlastsubswath=3\nburst0=2\nburstN=7\npolarisation=vv\n"
                 "multi_az=4\nmulti_rg=1\nunknown=5\n")
    N = tops.native
    cfg = N.Config()
    N.check(N.load().tops_read_config(str(p).encode(), C.byref(cfg)))
    assert cfg.process_dir == b"/data/run one" and cfg.masterpath == b"/m.SAFE"
    assert cfg.slavepath == b"/s.SAFE" and cfg.polarisation == b"vv"
    assert (cfg.firstsubswath, cfg.lastsubswath, cfg.burst0, cfg.burstN) == (1, 3, 2, 7)
    assert N.load().tops_read_config(b"/nonexistent/config.txt", C.byref(cfg)) == N.TOPS_ERR_ARG
    N.load().tops_clear_error()


def test_reference_entry_points_clear_the_status(tops):
    """The void reference entry points report failure only through the
    status: each clears it on entry, so a check after a successful call does
    not see an earlier call's error.  (EstESDShifts with no overlaps returns
    before any device work.)"""
    N = tops.native
    L = N.load()
    assert L.tops_read_config(b"/nonexistent/config.txt", C.byref(N.Config())) == N.TOPS_ERR_ARG
    assert L.tops_last_error() == N.TOPS_ERR_ARG
    L.EstESDShifts(None, None, None, None, None, None, 0, 0, 0, 6, 24, 0, 0.3, 1.0, 1.0)
    assert L.tops_last_error() == N.TOPS_OK


def test_esd_workspace_covers_empty_bands(tops):
    """The subswath workspace holds the ESD partials of the launch grid also
    when every overlap band is empty (one tile row per band is still
    launched and written)."""
    N = tops.native
    L = N.load()
    for ov in ([0], [0, 0], [0, 100]):
        o = np.ascontiguousarray(ov, np.int32)
        d = N.SubswathDesc(len(ov) + 1, 400, 2048, 400, 2048, 2e-3, 2e-3, 10.0, N.ptr(o, C.c_int),
                           6, 24, 0.3, 3, 10)
        ws = L.tops_subswath_workspace_size(C.byref(d))
        tiles_x = (2048 + 63) // 64
        rows = max(1, (max(ov) + 15) // 16)
        assert ws >= len(ov) * tiles_x * rows * 16 + len(ov) * 16, (ov, ws)


def test_burst_geometry_rejected_before_device_work(tops):
    """tops_resample_burst and tops_plan_create refuse a burst geometry the
    kernels cannot address, before any device work (the context and buffer
    pointers are only checked for null): empty boxes, slave columns outside
    the per-range arrays, slave lines outside [0, 2^20 - 64), master lines
    beyond +-(2^20 - 64), a slave box over 2^31 bytes of CInt16, an unknown
    mode."""
    N = tops.native
    L = N.load()
    fake = C.cast(C.create_string_buffer(64), C.c_void_p)   # never dereferenced
    good = dict(mb=[0, 99, 0, 49], sb=[0, 99, 0, 49], spb=100)
    cases = {
        "empty master box": dict(mb=[10, 9, 0, 49]),
        "empty slave box": dict(sb=[0, 99, 5, 4]),
        "slave column past the per-range arrays": dict(sb=[0, 100, 0, 49]),
        "negative slave column": dict(sb=[-1, 99, 0, 49]),
        "negative slave line": dict(sb=[0, 99, -3, 49]),
        "slave line past 2^20 - 64": dict(sb=[0, 99, 0, (1 << 20) - 64]),
        "master line past 2^20 - 64": dict(mb=[0, 99, 0, (1 << 20) - 64]),
        "slave box over 2^31 bytes": dict(sb=[0, 99999, 0, 5400], spb=100000),
    }
    for label, over in cases.items():
        d = dict(good, **over)
        g = N.BurstGeom((C.c_int * 4)(*d["mb"]), (C.c_int * 4)(*d["sb"]), 0, d["spb"], 2e-3, 0.0,
                        (C.c_double * 6)(), (C.c_double * 6)())
        L.tops_clear_error()
        rc = L.tops_resample_burst(fake, C.byref(g), fake, fake, fake, fake, fake, 0, None)
        assert rc == N.TOPS_ERR_ARG, label
        msg = L.tops_last_error_message()
        items = (N.BurstItem * 1)()
        items[0] = N.BurstItem(g, fake, fake, fake, fake, fake, None, None)
        assert not L.tops_plan_create(fake, items, 1, 3, 10, 64), label
        assert L.tops_last_error() == N.TOPS_ERR_ARG, label
        print(f"  {label}: {msg.decode()}")
    g = N.BurstGeom((C.c_int * 4)(*good["mb"]), (C.c_int * 4)(*good["sb"]), 0, 100, 2e-3, 0.0,
                    (C.c_double * 6)(), (C.c_double * 6)())
    assert L.tops_resample_burst(fake, C.byref(g), fake, fake, fake, fake, fake, 7,
                                 None) == N.TOPS_ERR_ARG
    L.tops_clear_error()
