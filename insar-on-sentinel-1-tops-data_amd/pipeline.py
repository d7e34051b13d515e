"""Stage-layer driver over a synthetic pair: what Main.cpp:84-134 does for one
subswath (GeometricCoreg -> DerampAndResample -> ESDAndCoh), through the
library's tops_geometric_coreg / tops_deramp_and_resample /
tops_esd_and_coh, with the metadata steps of Main.cpp:57-76 (Polyfit_Orbit,
ComputeDerampPara, InKernelInit) done by the library's host functions.
"""
import ctypes as C

import numpy as np

from . import native


class SubswathRun:
    """Inputs of one subswath pair, laid out as the stage functions read
    them (host memory), plus the three stage calls."""

    def __init__(self, pair, dem_margin_deg=0.02):
        N = native
        self.pair = pair
        m, s = pair.master, pair.slave
        self.m, self.s = m, s
        self.nb = m.n_bursts
        self.lpb, self.spb = m.lines_per_burst, m.samples_per_burst
        # SLC rasters (burst-stacked)
        self.m_slc = np.concatenate([pair.burst(b, "master") for b in range(self.nb)])
        self.s_slc = np.concatenate([pair.burst(b, "slave") for b in range(self.nb)])
        # orbits: Polyfit_Orbit
        from .synth import orbit_polyfit_inputs
        self.orbits = {}
        for key, orbit, sw in (("m", pair.morbit, m), ("s", pair.sorbit, s)):
            t, x, y, z = orbit_polyfit_inputs(orbit)
            cx, cy, cz = N.polyfit_orbit(t, x, y, z)
            speed = N.load().tops_orbit_speed(0.5 * (sw.first_line_time + sw.last_line_time),
                                              t[len(t) // 2], N.ptr(cx), N.ptr(cy), N.ptr(cz),
                                              len(cx))
            kt, tref, fdc, ka = N.deramp_params(sw, speed)
            self.orbits[key] = dict(t=np.ascontiguousarray(t), cx=cx, cy=cy, cz=cz, kt=kt,
                                    tref=tref, fdc=fdc, speed=speed)
        # InKernelInit with the master's PRF / ABW / RSR (Main.cpp:75-76)
        self.kaz, self.krg = N.lut_init(12, m.prf, m.azimuth_bandwidth,
                                        2 * m.range_sampling_rate)
        # DEM covering every burst
        ll = m.burst_ll
        dem, lat0, lon0, dlat, dlon = pair.dem_window(ll[:, 0].min() - dem_margin_deg,
                                                      ll[:, 1].max() + dem_margin_deg,
                                                      ll[:, 2].min() - dem_margin_deg,
                                                      ll[:, 3].max() + dem_margin_deg)
        self.dem = np.ascontiguousarray(dem)
        self.dem_geo = (lat0, lon0, dlat, dlon)
        self._keep = []

    @classmethod
    def from_files(cls, m_xml, s_xml, m_eof, s_eof, m_tiff, s_tiff, dem, dem_geo):
        """A pair read from Sentinel-1 product files the way Main.cpp:26-82
        reads it: each subswath's annotation XML (S1_Initialize and the DC /
        FM lists of ComputeDerampPara), each orbit's EOF state vectors around
        the acquisition (S1_OrbitInitialize) fitted with Polyfit_Orbit, the
        CInt16 measurement TIFFs (burst-stacked), the LUT of InKernelInit.
        The DEM is an in-memory grid (dem int32 [lines, pixels]; dem_geo =
        lat0, lon0, dlat, dlon of post (0, 0))."""
        from .synth import Swath
        N = native
        L = N.load()
        self = cls.__new__(cls)
        self.pair = None
        self._keep = []
        sws, slcs, self.orbits = {}, {}, {}
        for key, xml, eof, tif in (("m", m_xml, m_eof, m_tiff), ("s", s_xml, s_eof, s_tiff)):
            a = N.Annotation()
            N.check(L.tops_annotation_read(str(xml).encode(), C.byref(a)), "tops_annotation_read")
            try:
                utcs = (a.first_line_utc.decode(), a.last_line_utc.decode())
                nb, lpb = a.n_bursts, a.lines_per_burst
                arr = lambda q, n: np.ctypeslib.as_array(q, shape=(n,)).copy()
                sw = Swath(lines_per_burst=lpb, samples_per_burst=a.samples_per_burst,
                           n_bursts=nb, az_time_interval=a.az_time_interval,
                           range_spacing=a.range_pixel_spacing,
                           slr_time_first_pixel=a.slr_time_first_pixel,
                           wavelength=a.wavelength, prf=a.prf,
                           azimuth_bandwidth=a.azimuth_bandwidth,
                           range_sampling_rate=a.range_sampling_rate,
                           steering_rate_deg=a.azimuth_steering_rate,
                           first_valid_pixel=a.first_valid_pixel,
                           burst_first_line_time=arr(a.burst_first_line_time, nb),
                           burst_last_line_time=arr(a.burst_last_line_time, nb),
                           dc=arr(a.dc, nb * 4).reshape(nb, 4), fm=arr(a.fm, nb * 4).reshape(nb, 4),
                           burst_ll=arr(a.burst_ll, nb * 4).reshape(nb, 4),
                           scene_line0=np.arange(nb, dtype=np.int64) * lpb,
                           first_line_time=a.first_line_time, last_line_time=a.last_line_time,
                           first_valid_line=arr(a.first_valid_line, nb),
                           last_valid_line=arr(a.last_valid_line, nb), num_lines=a.num_lines)
            finally:
                L.tops_annotation_free(C.byref(a))
            sws[key] = sw
            first_utc, last_utc = utcs
            # state vectors around the acquisition (the product's first and
            # last line UTC times, Main.cpp:39-42)
            o = N.EofOrbit()
            N.check(L.tops_eof_read(str(eof).encode(), first_utc.encode(), last_utc.encode(),
                                    C.byref(o)), "tops_eof_read")
            try:
                arr = lambda q: np.ctypeslib.as_array(q, shape=(o.n,)).copy()
                t, x, y, z = arr(o.t), arr(o.x), arr(o.y), arr(o.z)
            finally:
                L.tops_eof_free(C.byref(o))
            cx, cy, cz = N.polyfit_orbit(t, x, y, z)
            speed = L.tops_orbit_speed(0.5 * (sw.first_line_time + sw.last_line_time),
                                       t[len(t) // 2], N.ptr(cx), N.ptr(cy), N.ptr(cz), len(cx))
            kt, tref, fdc, ka = N.deramp_params(sw, speed)
            self.orbits[key] = dict(t=np.ascontiguousarray(t), cx=cx, cy=cy, cz=cz, kt=kt,
                                    tref=tref, fdc=fdc, speed=speed)
            # measurement raster: the burst-stacked CInt16 subswath
            tf = L.tops_tiff_open(str(tif).encode())
            if not tf:
                N.check(L.tops_last_error(), "tops_tiff_open")
            try:
                w, h, dt_ = C.c_int(), C.c_int(), C.c_int()
                N.check(L.tops_tiff_info(tf, C.byref(w), C.byref(h), C.byref(dt_)))
                if dt_.value != N.GDT_CINT16 or w.value != sw.samples_per_burst or \
                        h.value < sw.n_bursts * sw.lines_per_burst:
                    raise ValueError(f"{tif}: {w.value} x {h.value} of type {dt_.value}")
                rows = sw.n_bursts * sw.lines_per_burst
                slc = np.empty((rows, w.value, 2), np.int16)
                N.check(L.tops_tiff_read(tf, 0, 0, rows, w.value, N.GDT_CINT16,
                                         slc.ctypes.data), "tops_tiff_read")
            finally:
                L.tops_tiff_close(tf)
            slcs[key] = slc
        self.m, self.s = sws["m"], sws["s"]
        self.m_slc, self.s_slc = slcs["m"], slcs["s"]
        self.nb = self.m.n_bursts
        self.lpb, self.spb = self.m.lines_per_burst, self.m.samples_per_burst
        self.kaz, self.krg = N.lut_init(12, self.m.prf, self.m.azimuth_bandwidth,
                                        2 * self.m.range_sampling_rate)
        self.dem = np.ascontiguousarray(dem, dtype=np.int32)
        self.dem_geo = tuple(dem_geo)
        return self

    # ---- ctypes structures ---------------------------------------------
    def swath(self, which):
        N = native
        sw = self.m if which == "m" else self.s
        o = self.orbits[which]
        slc = self.m_slc if which == "m" else self.s_slc
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (sw.burst_first_line_time, sw.burst_last_line_time, sw.burst_ll, o["kt"],
                 o["tref"], o["fdc"])]
        self._keep.append(arrs)
        return N.Swath(sw.lines_per_burst, sw.samples_per_burst, sw.n_bursts,
                       sw.az_time_interval, sw.range_spacing, sw.slr_time_first_pixel,
                       sw.wavelength, *[N.ptr(a) for a in arrs], slc.ctypes.data)

    def orbit(self, which):
        N = native
        o = self.orbits[which]
        return N.Orbit(len(o["t"]), N.ptr(o["t"]), len(o["cx"]), N.ptr(o["cx"]), N.ptr(o["cy"]),
                       N.ptr(o["cz"]))

    def deburst_swath(self, which="m"):
        """tops_deburst_swath of the master (or slave) subswath."""
        sw = self.m if which == "m" else self.s
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (sw.burst_first_line_time, sw.burst_last_line_time,
                 sw.burst_first_valid_line_time, sw.burst_last_valid_line_time)]
        self._keep.append(arrs)
        return native.DeburstSwath(sw.lines_per_burst, sw.samples_per_burst, sw.n_bursts,
                                   sw.num_lines, sw.az_time_interval, sw.first_line_time,
                                   *[native.ptr(a) for a in arrs])

    def dem_struct(self):
        lat0, lon0, dlat, dlon = self.dem_geo
        return native.Dem(self.dem.ctypes.data, self.dem.shape[0], self.dem.shape[1], lat0, lon0,
                          dlat, dlon, -32768)

    # ---- stages ------------------------------------------------------------
    def geometric_coreg(self, b0=0, bn=None):
        N = native
        bn = self.nb - 1 if bn is None else bn
        n = bn - b0 + 1
        ca, cr = np.zeros((n, 6)), np.zeros((n, 6))
        ms, ss = self.swath("m"), self.swath("s")
        mo, so = self.orbit("m"), self.orbit("s")
        dem = self.dem_struct()
        rc = N.load().tops_geometric_coreg(C.byref(ms), C.byref(ss), C.byref(mo), C.byref(so),
                                           C.byref(dem), b0, bn, b0, bn, N.ptr(ca), N.ptr(cr))
        N.check(rc, "tops_geometric_coreg")
        return ca, cr

    def deramp_and_resample(self, ca, cr, b0=0, bn=None):
        N = native
        bn = self.nb - 1 if bn is None else bn
        n = bn - b0 + 1
        out = np.zeros((n * self.lpb, self.spb), np.complex64)
        ms, ss = self.swath("m"), self.swath("s")
        ca, cr = np.ascontiguousarray(ca), np.ascontiguousarray(cr)
        rc = N.load().tops_deramp_and_resample(C.byref(ms), C.byref(ss), N.ptr(ca), N.ptr(cr), b0,
                                               bn, b0, bn, N.ptr(self.kaz, C.c_float),
                                               N.ptr(self.krg, C.c_float), out.ctypes.data, None)
        N.check(rc, "tops_deramp_and_resample")
        return out

    def esd_and_coh(self, ca, cr, reslave, b0=0, bn=None, esd_win=(6, 24), threshold=0.3,
                    coh_win=(6, 24)):
        N = native
        bn = self.nb - 1 if bn is None else bn
        n = bn - b0 + 1
        res = np.zeros((n * self.lpb, self.spb), np.complex64)
        coh = np.zeros((n * self.lpb, self.spb), np.float32)
        shifts = np.zeros(max(n - 1, 1))
        ms, ss = self.swath("m"), self.swath("s")
        ca, cr = np.ascontiguousarray(ca), np.ascontiguousarray(cr)
        reslave = np.ascontiguousarray(reslave, dtype=np.complex64)
        rc = N.load().tops_esd_and_coh(C.byref(ms), C.byref(ss), N.ptr(ca), N.ptr(cr), b0, bn, b0,
                                       bn, N.ptr(self.kaz, C.c_float), N.ptr(self.krg, C.c_float),
                                       reslave.ctypes.data, esd_win[0], esd_win[1], threshold,
                                       coh_win[0], coh_win[1], res.ctypes.data,
                                       N.ptr(coh, C.c_float), N.ptr(shifts))
        N.check(rc, "tops_esd_and_coh")
        return res, coh, shifts[:n - 1]

    def stitch(self, raster, b0=0, bn=None):
        """Stitch_OneSubswath (Func.cpp:2821-2933) of a burst-stacked host
        raster of bursts b0..bn (CInt16 [.., P, 2] int16, complex64 or
        float32) onto the master's continuous line grid."""
        N = native
        L = N.load()
        bn = self.nb - 1 if bn is None else bn
        raster = np.ascontiguousarray(raster)
        if raster.dtype == np.int16:
            dtype = N.GDT_CINT16
        elif raster.dtype == np.complex64:
            dtype = N.GDT_CFLOAT32
        elif raster.dtype == np.float32:
            dtype = N.GDT_FLOAT32
        else:
            raise TypeError(f"unsupported raster type {raster.dtype}")
        sw = self.deburst_swath()
        y0, h = C.c_int(0), C.c_int(0)
        N.check(L.tops_deburst_extent(C.byref(sw), b0, bn, C.byref(y0), C.byref(h)),
                "tops_deburst_extent")
        out = np.empty((h.value,) + raster.shape[1:], raster.dtype)
        N.check(L.tops_stitch_subswath(C.byref(sw), b0, bn, dtype, raster.ctypes.data,
                                       out.ctypes.data), "tops_stitch_subswath")
        return out


class DeviceSubswath:
    """One subswath pair resident in HBM for tops_subswath_esd_coh: the whole
    ESDAndCoh flow (Func.cpp:2496-2763, pass-1 resample included) as one
    sequence of launches on one stream, with no host round trip between the
    ESD shifts and the pass-2 resample.  Device buffers are torch tensors
    (plumbing only); the work is the library's."""

    def __init__(self, run, ca, cr, esd_win=(6, 24), threshold=0.3, coh_win=(6, 24)):
        import torch
        N = native
        L = N.load()
        self.torch = torch
        self.run = run
        m, s = run.m, run.s
        nb, lpb, spb = run.nb, run.lpb, run.spb
        dev = "cuda"
        put = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        self.d_master = put(run.m_slc)
        self.d_slave = put(run.s_slc)
        so = run.orbits["s"]
        self.d_kt, self.d_tref, self.d_fdc = (put(np.asarray(so[k], np.float64))
                                              for k in ("kt", "tref", "fdc"))
        self.ca = np.ascontiguousarray(ca, dtype=np.float64)
        self.cr = np.ascontiguousarray(cr, dtype=np.float64)
        nov = nb - 1
        self.overlap = np.zeros(max(nov, 1), np.int32)
        if nov > 0:
            total, mx = C.c_int(0), C.c_int(0)
            got = L.tops_overlap_sizes(N.ptr(m.burst_first_line_time),
                                       N.ptr(m.burst_last_line_time), 0, nb - 1,
                                       m.az_time_interval, N.ptr(self.overlap, C.c_int),
                                       C.byref(total), C.byref(mx))
            assert got == nov, (got, nov)
        df = L.tops_spectral_sep(N.ptr(run.orbits["m"]["kt"]), m.n_bursts, spb, lpb,
                                 m.az_time_interval)
        self.desc = N.SubswathDesc(nb, lpb, spb, s.lines_per_burst, s.samples_per_burst,
                                   m.az_time_interval, s.az_time_interval, df,
                                   N.ptr(self.overlap, C.c_int), esd_win[0], esd_win[1],
                                   threshold, coh_win[0], coh_win[1])
        ws = L.tops_subswath_workspace_size(C.byref(self.desc))
        if ws == 0:
            N.check(-1 if L.tops_last_error() == 0 else L.tops_last_error(),
                    "tops_subswath_workspace_size")
        self.workspace = torch.empty(ws, dtype=torch.uint8, device=dev)
        npx = nb * lpb * spb
        self.d_res = torch.empty(npx * 8, dtype=torch.uint8, device=dev)
        self.d_coh = torch.empty(npx * 4, dtype=torch.uint8, device=dev)
        self.d_shifts = torch.zeros(max(nov, 1), dtype=torch.float64, device=dev)
        self.ctx = L.tops_ctx_create(N.ptr(run.kaz, C.c_float), N.ptr(run.krg, C.c_float), 12)
        if not self.ctx:
            N.check(L.tops_last_error(), "tops_ctx_create")

    def enqueue(self, stream=None):
        """Launch the flow on `stream` (a torch stream, or the current one)."""
        N = native
        st = stream if stream is not None else self.torch.cuda.current_stream()
        rc = N.load().tops_subswath_esd_coh(
            self.ctx, C.byref(self.desc), self.d_master.data_ptr(), self.d_slave.data_ptr(),
            self.d_kt.data_ptr(), self.d_tref.data_ptr(), self.d_fdc.data_ptr(), N.ptr(self.ca),
            N.ptr(self.cr), self.d_res.data_ptr(), self.d_coh.data_ptr(),
            self.d_shifts.data_ptr(), self.workspace.data_ptr(), self.workspace.numel(),
            C.c_void_p(st.cuda_stream))
        N.check(rc, "tops_subswath_esd_coh")

    def results(self):
        """(resampled slave [n*L, P] complex64, coherence [n*L, P] float32,
        ESD shifts [n-1]) copied back to host memory."""
        self.torch.cuda.synchronize()
        run = self.run
        shape = (run.nb * run.lpb, run.spb)
        res = self.d_res.cpu().numpy().view(np.complex64).reshape(shape)
        coh = self.d_coh.cpu().numpy().view(np.float32).reshape(shape)
        return res, coh, self.d_shifts.cpu().numpy()[:run.nb - 1]

    def deburst_extent(self, b0=0, bn=None):
        bn = self.run.nb - 1 if bn is None else bn
        y0, h = C.c_int(0), C.c_int(0)
        sw = self.run.deburst_swath()
        native.check(native.load().tops_deburst_extent(C.byref(sw), b0, bn, C.byref(y0),
                                                       C.byref(h)), "tops_deburst_extent")
        return y0.value, h.value

    def deburst(self, which, out, stream=None):
        """Stitch the resident resampled slave ("res"), coherence ("coh") or
        master ("master") of every burst into `out` (a device tensor of
        height x P elements) on `stream`."""
        N = native
        src, dtype = {"res": (self.d_res, N.GDT_CFLOAT32), "coh": (self.d_coh, N.GDT_FLOAT32),
                      "master": (self.d_master, N.GDT_CINT16)}[which]
        st = stream if stream is not None else self.torch.cuda.current_stream()
        sw = self.run.deburst_swath()
        N.check(N.load().tops_deburst(C.byref(sw), 0, self.run.nb - 1, dtype, src.data_ptr(),
                                      out.data_ptr(), C.c_void_p(st.cuda_stream)),
                "tops_deburst")

    def close(self):
        if self.ctx:
            native.load().tops_ctx_destroy(self.ctx)
            self.ctx = None
