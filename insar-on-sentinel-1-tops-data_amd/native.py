"""ctypes bindings of libtops_mi355x.so (the C-ABI in include/tops_abi.h).

The library is built in-tree by the package Makefile.  Loading fails loudly
when it is missing: there is no CPU fallback for any entry point.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtops_mi355x.so")
SYNTH_PATH = os.path.join(HERE, "libtops_synth.so")

TOPS_OK = 0
TOPS_ERR_ARG = 1
TOPS_ERR_HIP = 2
TOPS_ERR_NOT_PD = 3
TOPS_ERR_NODEVICE = 4
TOPS_ERR_ALLOC = 5
MODE_FAST = 0
MODE_FP64 = 1
GDT_FLOAT32 = 6
GDT_CINT16 = 8
GDT_FLOAT64 = 7
GDT_CFLOAT32 = 10

# every symbol declared in include/tops_abi.h
EXPORTS = [
    "tops_last_error", "tops_last_error_message", "tops_clear_error",
    "GeoCoreg_warp_cuda", "DerampDemodResample", "DerampDemodResample_ESD",
    "ResampleFirstBurst", "EstESDShifts", "CoherenceEst", "CoherenceEst_ESD",
    "tops_ctx_create", "tops_ctx_destroy", "tops_resample_burst",
    "tops_coherence", "tops_esd_shifts", "tops_esd_shifts_noalloc", "tops_geocoreg",
    "tops_lut_init", "tops_polyfit_orbit", "tops_orbit_speed", "tops_deramp_params",
    "tops_overlap_sizes", "tops_spectral_sep", "tops_geometric_coreg",
    "tops_deramp_and_resample", "tops_esd_and_coh", "tops_read_config",
    "tops_plan_create", "tops_plan_run", "tops_plan_capture_graph", "tops_plan_tune",
    "tops_plan_strip_rows", "tops_plan_time_launches", "tops_plan_kernel_name",
    "tops_plan_coh_rows", "tops_plan_tune_count", "tops_plan_tune_entry",
    "tops_plan_set_groups", "tops_plan_groups",
    "tops_plan_destroy",
    "tops_subswath_workspace_size", "tops_subswath_esd_coh",
    "tops_deburst_extent", "tops_deburst_rows", "tops_deburst", "tops_stitch_subswath",
    "tops_utc2seconds", "tops_annotation_read", "tops_annotation_free", "tops_eof_read",
    "tops_eof_free", "tops_tiff_open", "tops_tiff_info", "tops_tiff_read",
    "tops_tiff_read_overlaps", "tops_tiff_read_device", "tops_tiff_create", "tops_tiff_write",
    "tops_tiff_close", "tops_xyz2aztime", "tops_lp2xyz", "tops_burst_overlap",
]

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class BurstGeom(C.Structure):
    _fields_ = [("master_box", C.c_int * 4), ("slave_box", C.c_int * 4),
                ("first_line_in_burst", C.c_int), ("samples_per_burst", C.c_int),
                ("az_time_interval", C.c_double), ("az_shift", C.c_double),
                ("cpm_az", C.c_double * 6), ("cpm_rg", C.c_double * 6)]


class OrbitGeom(C.Structure):
    _fields_ = [("coef_x", _dp), ("coef_y", _dp), ("coef_z", _dp), ("ncoef", C.c_int),
                ("mid_time", C.c_double), ("az_time_interval", C.c_double),
                ("first_line_time", C.c_double), ("slr_time_first_pixel", C.c_double),
                ("range_spacing", C.c_double), ("az_init", C.c_double),
                ("line_offset", C.c_int)]


class BurstItem(C.Structure):
    _fields_ = [("geom", BurstGeom), ("slave", _vp), ("kt", _vp), ("tref", _vp), ("fdc", _vp),
                ("out", _vp), ("master", _vp), ("coh", _vp)]


class SubswathDesc(C.Structure):
    _fields_ = [("n_bursts", C.c_int), ("m_lines_per_burst", C.c_int),
                ("m_samples_per_burst", C.c_int), ("s_lines_per_burst", C.c_int),
                ("s_samples_per_burst", C.c_int), ("m_az_time_interval", C.c_double),
                ("s_az_time_interval", C.c_double), ("spectral_sep", C.c_double),
                ("overlap_lines", C.POINTER(C.c_int)), ("esd_win_az", C.c_int),
                ("esd_win_rg", C.c_int), ("threshold", C.c_float), ("coh_win_az", C.c_int),
                ("coh_win_rg", C.c_int)]


class DeburstSwath(C.Structure):
    _fields_ = [("lines_per_burst", C.c_int), ("samples_per_burst", C.c_int),
                ("n_bursts", C.c_int), ("num_lines", C.c_int), ("az_time_interval", C.c_double),
                ("first_line_time", C.c_double), ("burst_first_line_time", _dp),
                ("burst_last_line_time", _dp), ("burst_first_valid_line_time", _dp),
                ("burst_last_valid_line_time", _dp)]


class Annotation(C.Structure):
    _fields_ = [("polarisation", C.c_char * 8), ("num_lines", C.c_int), ("num_samples", C.c_int),
                ("first_line_time", C.c_double), ("last_line_time", C.c_double),
                ("az_time_interval", C.c_double), ("prf", C.c_double),
                ("range_pixel_spacing", C.c_double), ("azimuth_pixel_spacing", C.c_double),
                ("radar_frequency", C.c_double), ("wavelength", C.c_double),
                ("range_sampling_rate", C.c_double), ("azimuth_steering_rate", C.c_double),
                ("heading", C.c_double), ("ascending_node_time", C.c_double),
                ("azimuth_bandwidth", C.c_double), ("range_bandwidth", C.c_double),
                ("slr_time_first_pixel", C.c_double), ("slr_time_last_pixel", C.c_double),
                ("n_bursts", C.c_int), ("lines_per_burst", C.c_int),
                ("samples_per_burst", C.c_int),
                ("burst_first_line_time", _dp), ("burst_last_line_time", _dp),
                ("burst_first_valid_line_time", _dp), ("burst_last_valid_line_time", _dp),
                ("first_valid_line", _ip), ("last_valid_line", _ip),
                ("first_valid_sample", _ip), ("last_valid_sample", _ip),
                ("first_valid_pixel", C.c_int), ("last_valid_pixel", C.c_int),
                ("slr_time_first_valid_pixel", C.c_double),
                ("slr_time_last_valid_pixel", C.c_double),
                ("geo_lines", C.c_int), ("geo_points_per_line", C.c_int),
                ("geo_az_time", _dp), ("geo_slant_range_time", _dp), ("geo_latitude", _dp),
                ("geo_longitude", _dp), ("geo_incidence_angle", _dp),
                ("near_range_on_left", C.c_int), ("burst_ll", _dp), ("dc", _dp), ("fm", _dp),
                ("first_line_utc", C.c_char * 32), ("last_line_utc", C.c_char * 32)]


class EofOrbit(C.Structure):
    _fields_ = [("n", C.c_int), ("t", _dp), ("x", _dp), ("y", _dp), ("z", _dp), ("vx", _dp),
                ("vy", _dp), ("vz", _dp), ("time_interval", C.c_double)]


class Swath(C.Structure):
    _fields_ = [("lines_per_burst", C.c_int), ("samples_per_burst", C.c_int),
                ("n_bursts", C.c_int), ("az_time_interval", C.c_double),
                ("range_spacing", C.c_double), ("slr_time_first_pixel", C.c_double),
                ("wavelength", C.c_double), ("burst_first_line_time", _dp),
                ("burst_last_line_time", _dp), ("burst_ll", _dp), ("doppler_rate", _dp),
                ("reference_time", _dp), ("doppler_centroid", _dp), ("slc", _vp)]


class Orbit(C.Structure):
    _fields_ = [("num_points", C.c_int), ("az_time", _dp), ("ncoef", C.c_int),
                ("coef_x", _dp), ("coef_y", _dp), ("coef_z", _dp)]


class Dem(C.Structure):
    _fields_ = [("data", _vp), ("lines", C.c_int), ("pixels", C.c_int), ("lat_max", C.c_double),
                ("lon_min", C.c_double), ("dlat", C.c_double), ("dlon", C.c_double),
                ("nodata", C.c_int)]


class Config(C.Structure):
    _fields_ = [("process_dir", C.c_char * 512), ("masterpath", C.c_char * 512),
                ("slavepath", C.c_char * 512), ("polarisation", C.c_char * 32),
                ("precise_orbit_master", C.c_char * 512), ("precise_orbit_slave", C.c_char * 512),
                ("specific_dem_path", C.c_char * 512), ("firstsubswath", C.c_int),
                ("lastsubswath", C.c_int), ("multi_az", C.c_int), ("multi_rg", C.c_int),
                ("burst0", C.c_int), ("burstN", C.c_int)]


_lib = None
_synth = None


def load():
    """Load libtops_mi355x.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (__graft_entry__.build())")
    # The torch wheel carries its own HIP runtime.  Loaded first, it is the
    # runtime this library binds to (same soname), so device pointers and
    # streams from torch are valid here.  Loaded after this library, it would
    # be a second runtime instance in the process, and whichever initialises
    # second finds no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    L.tops_last_error.restype = C.c_int
    L.tops_last_error_message.restype = C.c_char_p
    L.tops_clear_error.restype = None
    L.tops_ctx_create.restype = _vp
    L.tops_ctx_create.argtypes = [_fp, _fp, C.c_int]
    L.tops_ctx_destroy.argtypes = [_vp]
    L.tops_ctx_destroy.restype = None
    L.tops_resample_burst.restype = C.c_int
    L.tops_resample_burst.argtypes = [_vp, C.POINTER(BurstGeom), _vp, _vp, _vp, _vp, _vp,
                                      C.c_int, _vp]
    L.tops_coherence.restype = C.c_int
    L.tops_coherence.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_int, _vp, C.c_int, _vp]
    L.tops_esd_shifts.restype = C.c_int
    L.tops_esd_shifts.argtypes = [_vp, _vp, _vp, _vp, _ip, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_float, C.c_double, C.c_double, _dp, _dp, _vp]
    L.tops_esd_shifts_noalloc.restype = C.c_int
    L.tops_esd_shifts_noalloc.argtypes = L.tops_esd_shifts.argtypes
    L.tops_geocoreg.restype = C.c_int
    L.tops_geocoreg.argtypes = [_vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_int, C.POINTER(OrbitGeom),
                                C.POINTER(OrbitGeom), C.c_int, C.c_int, C.c_int, C.c_int,
                                _dp, _dp, C.POINTER(C.c_int64), _vp, _vp]
    # reference entry points
    L.DerampDemodResample.restype = None
    L.DerampDemodResample.argtypes = [_vp, _dp, _dp, C.c_double, _vp, _fp, _fp, C.c_int,
                                      C.c_int, C.c_int, _ip, _ip, C.c_int, C.c_int,
                                      C.c_double, _dp, _dp, _dp, C.c_int]
    L.DerampDemodResample_ESD.restype = _vp
    L.DerampDemodResample_ESD.argtypes = L.DerampDemodResample.argtypes
    L.ResampleFirstBurst.restype = _vp
    L.ResampleFirstBurst.argtypes = [_vp, C.c_int, C.c_int]
    L.CoherenceEst.restype = None
    L.CoherenceEst.argtypes = [C.c_int] * 8 + [_vp, _vp, _fp]
    L.CoherenceEst_ESD.restype = None
    L.CoherenceEst_ESD.argtypes = L.CoherenceEst.argtypes
    L.EstESDShifts.restype = None
    L.EstESDShifts.argtypes = [_vp, _vp, _vp, _vp, _ip, _dp, C.c_int, C.c_int, C.c_int,
                               C.c_int, C.c_int, C.c_int, C.c_float, C.c_double, C.c_double]
    L.GeoCoreg_warp_cuda.restype = None
    L.GeoCoreg_warp_cuda.argtypes = (
        [C.c_int] * 6 + [C.c_double, C.c_double, C.c_int, C.c_int, _dp, _dp, C.c_double,
                         C.c_double, C.c_int] + [C.c_double] * 7 + [C.c_int, C.c_int, _dp, _dp]
        + [C.c_int] * 4 + [C.c_double] * 4 + [C.c_bool, C.c_bool, _ip] + [_dp] * 6
        + [C.c_double, C.c_double])
    # host stage layer
    L.tops_lut_init.restype = C.c_int
    L.tops_lut_init.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _fp, _fp]
    L.tops_polyfit_orbit.restype = C.c_int
    L.tops_polyfit_orbit.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.tops_orbit_speed.restype = C.c_double
    L.tops_orbit_speed.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp, C.c_int]
    L.tops_deramp_params.restype = C.c_int
    L.tops_deramp_params.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                     C.c_double, C.c_int, _dp, _dp, C.c_double, C.c_double,
                                     C.c_double, _dp, _dp, _dp, _dp]
    L.tops_overlap_sizes.restype = C.c_int
    L.tops_overlap_sizes.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_double, _ip, _ip, _ip]
    L.tops_spectral_sep.restype = C.c_double
    L.tops_spectral_sep.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_double]
    L.tops_geometric_coreg.restype = C.c_int
    L.tops_geometric_coreg.argtypes = [C.POINTER(Swath), C.POINTER(Swath), C.POINTER(Orbit),
                                       C.POINTER(Orbit), C.POINTER(Dem), C.c_int, C.c_int,
                                       C.c_int, C.c_int, _dp, _dp]
    L.tops_deramp_and_resample.restype = C.c_int
    L.tops_deramp_and_resample.argtypes = [C.POINTER(Swath), C.POINTER(Swath), _dp, _dp,
                                           C.c_int, C.c_int, C.c_int, C.c_int, _fp, _fp, _vp,
                                           _vp]
    L.tops_esd_and_coh.restype = C.c_int
    L.tops_esd_and_coh.argtypes = [C.POINTER(Swath), C.POINTER(Swath), _dp, _dp, C.c_int,
                                   C.c_int, C.c_int, C.c_int, _fp, _fp, _vp, C.c_int, C.c_int,
                                   C.c_float, C.c_int, C.c_int, _vp, _fp, _dp]
    L.tops_plan_create.restype = _vp
    L.tops_plan_create.argtypes = [_vp, C.POINTER(BurstItem), C.c_int, C.c_int, C.c_int, C.c_int]
    L.tops_plan_run.restype = C.c_int
    L.tops_plan_run.argtypes = [_vp, _vp]
    L.tops_plan_capture_graph.restype = C.c_int
    L.tops_plan_capture_graph.argtypes = [_vp, _vp]
    L.tops_plan_tune.restype = C.c_int
    L.tops_plan_tune.argtypes = [_vp, _vp, _ip]
    L.tops_plan_strip_rows.restype = C.c_int
    L.tops_plan_strip_rows.argtypes = [_vp]
    L.tops_plan_time_launches.restype = C.c_int
    L.tops_plan_time_launches.argtypes = [_vp, _vp, C.c_int, _fp, _ip]
    L.tops_plan_kernel_name.restype = C.c_char_p
    L.tops_plan_kernel_name.argtypes = [_vp, C.c_int]
    L.tops_plan_coh_rows.restype = C.c_int
    L.tops_plan_coh_rows.argtypes = [_vp]
    L.tops_plan_tune_count.restype = C.c_int
    L.tops_plan_tune_count.argtypes = [_vp]
    L.tops_plan_tune_entry.restype = C.c_int
    L.tops_plan_tune_entry.argtypes = [_vp, C.c_int, _ip, _ip, _ip, _ip, _fp]
    L.tops_plan_set_groups.restype = C.c_int
    L.tops_plan_set_groups.argtypes = [_vp, C.c_int]
    L.tops_plan_groups.restype = C.c_int
    L.tops_plan_groups.argtypes = [_vp]
    L.tops_plan_destroy.restype = None
    L.tops_plan_destroy.argtypes = [_vp]
    L.tops_subswath_workspace_size.restype = C.c_size_t
    L.tops_subswath_workspace_size.argtypes = [C.POINTER(SubswathDesc)]
    L.tops_subswath_esd_coh.restype = C.c_int
    L.tops_subswath_esd_coh.argtypes = [_vp, C.POINTER(SubswathDesc), _vp, _vp, _vp, _vp, _vp,
                                        _dp, _dp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]
    _dsw = C.POINTER(DeburstSwath)
    L.tops_deburst_extent.restype = C.c_int
    L.tops_deburst_extent.argtypes = [_dsw, C.c_int, C.c_int, _ip, _ip]
    L.tops_deburst_rows.restype = C.c_int
    L.tops_deburst_rows.argtypes = [_dsw, C.c_int, C.c_int, _ip]
    L.tops_deburst.restype = C.c_int
    L.tops_deburst.argtypes = [_dsw, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]
    L.tops_stitch_subswath.restype = C.c_int
    L.tops_stitch_subswath.argtypes = [_dsw, C.c_int, C.c_int, C.c_int, _vp, _vp]
    L.tops_utc2seconds.restype = C.c_double
    L.tops_utc2seconds.argtypes = [C.c_char_p]
    L.tops_annotation_read.restype = C.c_int
    L.tops_annotation_read.argtypes = [C.c_char_p, C.POINTER(Annotation)]
    L.tops_annotation_free.restype = None
    L.tops_annotation_free.argtypes = [C.POINTER(Annotation)]
    L.tops_eof_read.restype = C.c_int
    L.tops_eof_read.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(EofOrbit)]
    L.tops_eof_free.restype = None
    L.tops_eof_free.argtypes = [C.POINTER(EofOrbit)]
    L.tops_tiff_open.restype = _vp
    L.tops_tiff_open.argtypes = [C.c_char_p]
    L.tops_tiff_info.restype = C.c_int
    L.tops_tiff_info.argtypes = [_vp, _ip, _ip, _ip]
    L.tops_tiff_read.restype = C.c_int
    L.tops_tiff_read.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp]
    L.tops_tiff_read_overlaps.restype = C.c_int
    L.tops_tiff_read_overlaps.argtypes = [_vp, _vp, _vp, C.c_int, _ip, C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_int]
    L.tops_tiff_read_device.restype = C.c_int
    L.tops_tiff_read_device.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp,
                                        _vp]
    L.tops_tiff_create.restype = _vp
    L.tops_tiff_create.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int]
    L.tops_tiff_write.restype = C.c_int
    L.tops_tiff_write.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp]
    L.tops_tiff_close.restype = C.c_int
    L.tops_tiff_close.argtypes = [_vp]
    _op = C.POINTER(Orbit)
    _ap = C.POINTER(Annotation)
    L.tops_xyz2aztime.restype = C.c_double
    L.tops_xyz2aztime.argtypes = [_dp, _op, C.c_double]
    L.tops_lp2xyz.restype = C.c_long
    L.tops_lp2xyz.argtypes = [C.c_double, C.c_double, _ap, _op, C.c_int, _dp, C.c_int,
                              C.c_double]
    L.tops_burst_overlap.restype = C.c_int
    L.tops_burst_overlap.argtypes = [_ap, _ap, _op, _op, C.c_int, C.c_int, _ip, _ip, _ip, _ip]
    L.tops_read_config.restype = C.c_int
    L.tops_read_config.argtypes = [C.c_char_p, C.POINTER(Config)]
    _lib = L
    return L


def lut_init(npoints, prf, abw, rsr2x):
    """InKernelInit (Func.cpp:2105-2158) through the library."""
    import numpy as np
    kaz = np.zeros((2048, npoints), np.float32)
    krg = np.zeros((2048, npoints), np.float32)
    check(load().tops_lut_init(npoints, prf, abw, rsr2x, ptr(kaz, C.c_float), ptr(krg, C.c_float)),
          "tops_lut_init")
    return kaz, krg


def polyfit_orbit(t, x, y, z):
    """Polyfit_Orbit (Func.cpp:1050-1126) through the library."""
    import numpy as np
    cx, cy, cz = np.zeros(6), np.zeros(6), np.zeros(6)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (t, x, y, z)]
    n = load().tops_polyfit_orbit(len(t), *[ptr(a) for a in arrs], ptr(cx), ptr(cy), ptr(cz))
    if n == 0:
        check(load().tops_last_error(), "tops_polyfit_orbit")
    return cx[:n].copy(), cy[:n].copy(), cz[:n].copy()


def deramp_params(sw, speed):
    """ComputeDerampPara formulas through the library: kt, tref, fdc, ka
    [n_bursts x samples_per_burst] for a synth.Swath."""
    import numpy as np
    nb, spb = sw.n_bursts, sw.samples_per_burst
    out = [np.zeros((nb, spb)) for _ in range(4)]
    dc = np.ascontiguousarray(sw.dc, dtype=np.float64)
    fm = np.ascontiguousarray(sw.fm, dtype=np.float64)
    check(load().tops_deramp_params(nb, spb, sw.slr_time_first_pixel, sw.range_spacing,
                                    sw.lines_per_burst, sw.az_time_interval,
                                    sw.first_valid_pixel, ptr(dc), ptr(fm), sw.steering_rate_deg,
                                    sw.wavelength, speed, *[ptr(a) for a in out]),
          "tops_deramp_params")
    return tuple(out)


def check(rc, what="tops"):
    """Raise RuntimeError with the library's message on a non-zero status."""
    if rc != TOPS_OK:
        msg = load().tops_last_error_message().decode(errors="replace")
        raise RuntimeError(f"{what}: status {rc}: {msg}")


def last_error():
    L = load()
    return L.tops_last_error(), L.tops_last_error_message().decode(errors="replace")


def load_synth():
    global _synth
    if _synth is not None:
        return _synth
    if not os.path.exists(SYNTH_PATH):
        raise RuntimeError(f"{SYNTH_PATH} is missing: build it first")
    S = C.CDLL(SYNTH_PATH)
    S.tsyn_burst.restype = None
    S.tsyn_burst.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_double,
                             C.c_int64, C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_double,
                             _dp, _dp, _dp, C.c_double, C.c_uint64, _vp]
    S.tsyn_dem.restype = None
    S.tsyn_dem.argtypes = [C.c_uint64, C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_double,
                           C.c_double, C.c_int, C.c_int, _vp]
    _synth = S
    return S


def ptr(a, ctype=C.c_double):
    """ctypes pointer to a contiguous numpy array."""
    return a.ctypes.data_as(C.POINTER(ctype))
