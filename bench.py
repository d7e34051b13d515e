#!/usr/bin/env python3
"""Throughput of the TOPS burst hot path on one MI355X.

A step is one pass of deramp -> 12x12 sinc resample -> reramp -> coherence
over one synthetic burst already resident in HBM (default workload: a full
IW1 burst, 1500 az x 21000 rg CInt16, 3 az x 10 rg coherence window).  The
metric counts master-grid output samples per second.

Printed JSON line (one): metric, value (Msamples/s), roofline of the
dominant kernel (the fused resample kernel; HIP-event durations on the
stream it runs on) in HBM bytes and, under "valu", in the reference form's
FP32 arithmetic per sample, the plain all-FP64 GPU form measured in the same run
(vs_plain_fp64), the error of the measured output against the FP64 oracle,
and the CPU baseline (the FP64 oracle on the host cores).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "burst Msamples/s (deramp+resample+coh) vs plain FP64 on MI355X; % HBM peak"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per output sample: the resample kernel reads the CInt16
# slave once (4 B) and writes the CF32 result (8 B); the whole path adds the
# CInt16 master read (4 B) and the float coherence write (4 B): 20 B.
RESAMPLE_BYTES = 12
PATH_BYTES = 20
# the reference form's arithmetic per output sample: the 12x12 separable
# complex-by-real sinc (12*12*2 + 12*2 FMA = 312 FMA, 624 flop) and the 12x12
# interpolation of the deramp phase (156 FMA, 312 flop).  The kernel executes
# the azimuth-first moment form, which needs fewer; this is the work it
# replaces, priced against the FP32 vector peak (157.3 TFLOP/s counting packed
# FP32; the build issues no packed FP32, which halves its own issue ceiling).
REFERENCE_FLOP = 624 + 312
FP32_VECTOR_PEAK_TFLOPS = 157.3


def load_package():
    import __graft_entry__ as ge
    return ge._import_package()


def workloads(tops, name):
    s = tops.synth
    if name == "single_burst_iw1":
        return s.single_burst_iw1()
    if name == "single_burst_small":
        return s.single_burst_small()
    raise SystemExit(f"unknown workload {name}")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def read_traffic(kernel, workload):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), from the committed
    profile summary, or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.json")
    try:
        with open(path) as f:
            return json.load(f).get("traffic_bytes")
    except (OSError, ValueError):
        return None


def _surface(ca, cr):
    gl, gp = np.meshgrid(np.linspace(-2, 2, 17), np.linspace(-2, 2, 17))
    A = np.stack([np.ones_like(gl), gl, gp, gl * gl, gl * gp, gp * gp], axis=-1)
    return A @ ca, A @ cr


def bench_geocoreg(tops, torch, pair, L, N, steps, check):
    """DEM-assisted coregistration of the burst: FP64 Newton posts/s over the
    1-arcsec DEM window of the burst footprint (24-post margin), wall clock
    per call (the call ends with the 34-sum readback and the 6x6 solve)."""
    m, s = pair.master, pair.slave
    keep = []

    def geom(sw, orbit):
        t, x, y, z = tops.synth.orbit_polyfit_inputs(orbit)
        cx, cy, cz = N.polyfit_orbit(t, x, y, z)
        keep.append((cx, cy, cz))
        first, last = sw.burst_first_line_time[0], sw.burst_last_line_time[0]
        return (cx, cy, cz), t[len(t) // 2], sw.az_time_interval, first, \
            sw.slr_time_first_pixel, sw.range_spacing, 0.5 * (first + last)

    gm, gs = geom(m, pair.morbit), geom(s, pair.sorbit)

    def ng(g):
        (cx, cy, cz), mid, dt, first, slr, rs, init = g
        return N.OrbitGeom(N.ptr(cx), N.ptr(cy), N.ptr(cz), len(cx), mid, dt, first, slr, rs,
                           init, 0)

    dem, lat0, lon0, dlat, dlon = pair.dem_window(*m.burst_ll[0])
    ddem = torch.from_numpy(np.ascontiguousarray(dem)).to("cuda:0")
    box = [0, m.samples_per_burst - 1, 0, m.lines_per_burst - 1]
    ca, cr = np.zeros(6), np.zeros(6)
    used = C.c_int64(0)
    ngm, ngs = ng(gm), ng(gs)

    def call():
        N.check(L.tops_geocoreg(ddem.data_ptr(), dem.shape[0], dem.shape[1], lat0, lon0, dlat,
                                dlon, -32768, C.byref(ngm), C.byref(ngs), *box, N.ptr(ca),
                                N.ptr(cr), C.byref(used), None, None), "geocoreg")

    call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        call()
    t = (time.perf_counter() - t0) / steps
    out = {"posts": int(dem.size), "posts_used": int(used.value),
           "mposts_per_s": round(dem.size / t / 1e6, 2), "ms_per_call": round(t * 1e3, 3),
           "dem": list(dem.shape), "precision": "FP64 orbit, Newton and fit"}
    if check:
        from oracle import oracle as orc

        def og(g):
            (cx, cy, cz), mid, dt, first, slr, rs, init = g
            return orc.orbit_geom((cx, cy, cz), mid, dt, first, slr, rs, init, 0)
        t0 = time.perf_counter()
        rc, oa, orr, _ = orc.geocoreg(dem, lat0, lon0, dlat, dlon, -32768, og(gm), og(gs), box)
        out["oracle_s"] = round(time.perf_counter() - t0, 3)
        ga, gr = _surface(ca, cr)
        ra, rr = _surface(oa, orr)
        out["offset_err_px"] = float(max(np.max(np.abs(ga - ra)), np.max(np.abs(gr - rr))))
        out["tol_px"] = 1e-3
    return out


def bench_esd(torch, M, S_res, L, N, pair, kt, steps, check):
    """Double-difference ESD over 8 overlap bands of 100 lines x the burst
    width (the band shapes of a 9-burst subswath), band samples/s; the bands
    are cut from the resampled burst and its master."""
    lines, pixels = S_res.shape
    nb, n = 8, 100
    rows = [(i * (lines - n)) // (nb - 1) for i in range(nb)]
    mf = np.concatenate([M[r:r + n] for r in rows])
    mb = np.concatenate([M[r + 1:r + 1 + n] if r + 1 + n <= lines else M[r:r + n] for r in rows])
    sf = np.concatenate([S_res[r:r + n] for r in rows])
    sb = np.concatenate([S_res[r + 1:r + 1 + n] if r + 1 + n <= lines else S_res[r:r + n]
                         for r in rows])
    sizes = np.full(nb, n, np.int32)
    d = [torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0") for a in (mf, mb, sf, sb)]
    shifts, sums = np.zeros(nb), np.zeros(2 * nb)
    m = pair.master
    dt = m.az_time_interval
    df = float(np.mean(kt)) * m.lines_per_burst * dt   # GetSpectralSep, Func.cpp:2962-2976

    def call():
        N.check(L.tops_esd_shifts(d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(),
                                  d[3].data_ptr(), sizes.ctypes.data_as(C.POINTER(C.c_int)), nb,
                                  pixels, 6, 24, 0.3, df, dt, N.ptr(shifts), N.ptr(sums), None),
                "esd")

    call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        call()
    t = (time.perf_counter() - t0) / steps
    out = {"bands": nb, "band_lines": n, "band_samples": int(nb * n * pixels),
           "msamples_per_s": round(nb * n * pixels / t / 1e6, 2), "ms_per_call": round(t * 1e3, 3),
           "window": [6, 24], "precision": "FP32 window sums, Kahan FP32 phasor sums"}
    if check:
        from oracle import oracle as orc
        ref, _ = orc.esd(mf, mb, sf.astype(np.complex128), sb.astype(np.complex128), sizes,
                         pixels, 6, 24, 0.3, df, dt)
        out["shift_err_px"] = float(np.max(np.abs(shifts - ref)))
        out["tol_px"] = 1e-4
    return out


def run_iw_pair(args, tops, torch):
    """27 uneven bursts (3 subswaths x 9; 21000 / 25000 / 24500 samples,
    1480-1520 lines each) in one batched plan: one resample launch and one
    coherence launch over all bursts, tuned strip height, replayed as one
    hipGraph.  Compared with per-burst plain FP64 launches on the same GPU."""
    N = tops.native
    L = N.load()
    rng = np.random.default_rng(2026)
    widths = (21000, 25000, 24500)
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(dev)
    sp = C.c_void_p(stream.cuda_stream)
    keep, specs = [], []
    kaz = krg = None
    for sw, width in enumerate(widths):
        lines = [int(v) for v in rng.integers(1480, 1521, size=9)]
        pair = tops.synth.make_pair(f"iw{sw + 1}", max(lines), width, seed=300 + sw,
                                    dem_spacing=1.0, dc_slope_total=10.0 + 5.0 * sw)
        m = pair.master
        if kaz is None:
            kaz, krg = N.lut_init(12, m.prf, m.azimuth_bandwidth, 2 * m.range_sampling_rate)
        kt, tref, fdc, _ = pair.params()
        M = pair.burst(0, "master")
        S = pair.burst(0, "slave")
        d = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev)
             for k, v in dict(S=S, M=M, kt=kt[0], tref=tref[0], fdc=fdc[0]).items()}
        keep.append(d)
        for b, nl in enumerate(lines):
            ca, cr = tops.synth.cpm_from_field(pair.az_off, pair.rg_off, nl, width, 0)
            specs.append((d, nl, width, ca, cr, m.az_time_interval))
        print(f"[bench] subswath {sw + 1}: {width} samples, lines {lines}", file=sys.stderr,
              flush=True)
    ctx = L.tops_ctx_create(N.ptr(kaz, C.c_float), N.ptr(krg, C.c_float), 12)
    items = (N.BurstItem * len(specs))()
    outs = []
    samples = 0
    geoms = []
    for i, (d, nl, width, ca, cr, dt) in enumerate(specs):
        box = [0, width - 1, 0, nl - 1]
        g = N.BurstGeom((C.c_int * 4)(*box), (C.c_int * 4)(*box), 0, width, dt, 0.0,
                        (C.c_double * 6)(*ca), (C.c_double * 6)(*cr))
        out = torch.empty(nl * width * 2, dtype=torch.float32, device=dev)
        coh = torch.empty(nl * width, dtype=torch.float32, device=dev)
        outs.append((out, coh))
        geoms.append(g)
        items[i] = N.BurstItem(g, d["S"].data_ptr(), d["kt"].data_ptr(), d["tref"].data_ptr(),
                               d["fdc"].data_ptr(), out.data_ptr(), d["M"].data_ptr(),
                               coh.data_ptr())
        samples += nl * width
    plan = L.tops_plan_create(ctx, items, len(specs), 3, 10, 0)
    if not plan:
        raise RuntimeError(N.last_error())
    best = C.c_int(0)
    N.check(L.tops_plan_tune(plan, sp, C.byref(best)), "tune")
    N.check(L.tops_plan_capture_graph(plan, sp), "graph")
    for _ in range(args.warmup):
        N.check(L.tops_plan_run(plan, sp))
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        N.check(L.tops_plan_run(plan, sp))
    stream.synchronize()
    step_s = (time.perf_counter() - t0) / args.steps
    # per-burst plain FP64 launches (same inputs)
    nfp = max(1, args.fp64_steps)
    for _ in range(1):
        for (d, nl, width, *_r), g, (out, coh) in zip(specs, geoms, outs):
            N.check(L.tops_resample_burst(ctx, C.byref(g), d["S"].data_ptr(), d["kt"].data_ptr(),
                                          d["tref"].data_ptr(), d["fdc"].data_ptr(),
                                          out.data_ptr(), N.MODE_FP64, sp))
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(nfp):
        for (d, nl, width, *_r), g, (out, coh) in zip(specs, geoms, outs):
            N.check(L.tops_resample_burst(ctx, C.byref(g), d["S"].data_ptr(), d["kt"].data_ptr(),
                                          d["tref"].data_ptr(), d["fdc"].data_ptr(),
                                          out.data_ptr(), N.MODE_FP64, sp))
            N.check(L.tops_coherence(d["M"].data_ptr(), out.data_ptr(), width, nl, 4, 1, 10, 3,
                                     coh.data_ptr(), N.MODE_FP64, sp))
    stream.synchronize()
    fp64_s = (time.perf_counter() - t0) / nfp
    # the same bursts as per-burst fast launches (no plan, no graph)
    def per_burst():
        for (d, nl, width, *_r), g, (out, coh) in zip(specs, geoms, outs):
            N.check(L.tops_resample_burst(ctx, C.byref(g), d["S"].data_ptr(), d["kt"].data_ptr(),
                                          d["tref"].data_ptr(), d["fdc"].data_ptr(),
                                          out.data_ptr(), N.MODE_FAST, sp))
            N.check(L.tops_coherence(d["M"].data_ptr(), out.data_ptr(), width, nl, 4, 1, 10, 3,
                                     coh.data_ptr(), N.MODE_FAST, sp))
    per_burst()
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        per_burst()
    stream.synchronize()
    per_burst_s = (time.perf_counter() - t0) / args.steps
    L.tops_plan_destroy(plan)
    L.tops_ctx_destroy(ctx)
    value = samples / step_s / 1e6
    print(json.dumps({
        "metric": METRIC, "value": round(value, 2), "unit": "Msamples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(step_s * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "iw_pair", "bursts": len(specs), "samples": samples,
                   "coh_window": [3, 10], "strip_rows": best.value, "hip_graph": True,
                   "parallelism": "single GPU"},
        "vs_plain_fp64": round(fp64_s / step_s, 3),
        "plain_fp64_msamples_per_s": round(samples / fp64_s / 1e6, 2),
        "per_burst_launches_msamples_per_s": round(samples / per_burst_s / 1e6, 2),
        "path_hbm_frac": round(PATH_BYTES * samples / step_s / 1e9 / HBM_PEAK_GBS, 4)}))


def run_subswath(args, tops, torch):
    """The host stage layer on one subswath of 9 bursts of 1500 x 21000 with
    ~100-line overlaps: GeometricCoreg, DerampAndResample and ESDAndCoh
    (Main.cpp:107-121) over host-resident SLCs, as the reference's stages
    run them (host buffers in and out, so PCIe transfers are inside the
    timing).  value = master-grid samples of the subswath / the three
    stages' wall time."""
    import tops_amd.pipeline as pl
    t0 = time.perf_counter()
    run = pl.SubswathRun(tops.synth.subswath_iw1())
    print(f"[bench] subswath generated in {time.perf_counter() - t0:.0f} s", file=sys.stderr,
          flush=True)
    samples = run.nb * run.lpb * run.spb
    torch.cuda.synchronize()
    times = {"geometric_coreg": [], "deramp_and_resample": [], "esd_and_coh": []}
    out = None
    for k in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        ca, cr = run.geometric_coreg()
        t1 = time.perf_counter()
        res1 = run.deramp_and_resample(ca, cr)
        t2 = time.perf_counter()
        res2, coh, shifts = run.esd_and_coh(ca, cr, res1, coh_win=(3, 10))
        t3 = time.perf_counter()
        if k >= args.warmup:
            times["geometric_coreg"].append(t1 - t0)
            times["deramp_and_resample"].append(t2 - t1)
            times["esd_and_coh"].append(t3 - t2)
        out = shifts
    ms = {k: round(float(np.median(v)) * 1e3, 2) for k, v in times.items()}
    total_s = sum(ms.values()) / 1e3
    # the same ESDAndCoh flow with the subswath resident in HBM
    # (tops_subswath_esd_coh: pass 1, ESD in place, device shifts, pass 2,
    # coherence), timed with events on its stream
    ds = pl.DeviceSubswath(run, ca, cr, coh_win=(3, 10))
    st = torch.cuda.current_stream()
    for _ in range(args.warmup):
        ds.enqueue(st)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(args.steps):
        ds.enqueue(st)
    e1.record(st)
    torch.cuda.synchronize()
    dev_ms = e0.elapsed_time(e1) / args.steps
    dev_shifts = ds.results()[2]
    # deburst (Stitch_OneSubswath) of the resident pass-2 slave, CFloat32:
    # 8 B read + 8 B written per output sample
    _, height = ds.deburst_extent()
    stitched = torch.empty(height * run.spb * 8, dtype=torch.uint8, device="cuda")
    for _ in range(args.warmup):
        ds.deburst("res", stitched, st)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(args.steps):
        ds.deburst("res", stitched, st)
    e1.record(st)
    torch.cuda.synchronize()
    db_ms = e0.elapsed_time(e1) / args.steps
    db_bytes = 16 * height * run.spb
    del stitched
    ds.close()
    # ingest: the master SLC as a stripped CInt16 TIFF (the measurement
    # file's layout), read into host memory and straight into HBM
    import tempfile
    N = tops.native
    L = N.load()
    tmpdir = tempfile.mkdtemp(prefix="tops_bench_")
    tif = os.path.join(tmpdir, "slc.tif")
    rows, width = run.m_slc.shape[0], run.m_slc.shape[1]
    t = L.tops_tiff_create(tif.encode(), N.GDT_CINT16, width, rows)
    N.check(L.tops_tiff_write(t, 0, 0, rows, width, N.GDT_CINT16, run.m_slc.ctypes.data))
    N.check(L.tops_tiff_close(t))
    t = L.tops_tiff_open(tif.encode())
    host = np.empty_like(run.m_slc)
    d_slc = torch.empty(rows * width * 4, dtype=torch.uint8, device="cuda")
    ingest = {}
    for name, fn in (("host", lambda: L.tops_tiff_read(t, 0, 0, rows, width, N.GDT_CINT16,
                                                       host.ctypes.data)),
                     ("device", lambda: L.tops_tiff_read_device(t, 0, 0, rows, width,
                                                                N.GDT_CINT16,
                                                                d_slc.data_ptr(), None))):
        N.check(fn())
        t0 = time.perf_counter()
        for _ in range(3):
            N.check(fn())
        torch.cuda.synchronize()
        ingest[name] = (time.perf_counter() - t0) / 3
    L.tops_tiff_close(t)
    ok = bool(np.array_equal(host, run.m_slc)) and bool(torch.equal(
        d_slc.cpu(), torch.from_numpy(run.m_slc.view(np.uint8).reshape(-1))))
    os.remove(tif)
    os.rmdir(tmpdir)
    del d_slc, host
    slc_bytes = run.m_slc.nbytes
    print(json.dumps({
        "metric": METRIC, "value": round(samples / total_s / 1e6, 2), "unit": "Msamples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total_s * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "subswath", "bursts": run.nb, "burst": [run.lpb, run.spb],
                   "samples": samples, "coh_window": [3, 10], "esd_window": [6, 24],
                   "host_buffers": True, "parallelism": "single GPU"},
        "stage_ms": ms, "esd_shifts": [round(float(v), 6) for v in out],
        "device_esd_and_coh": {"ms": round(dev_ms, 3),
                               "msamples_per_s": round(samples / dev_ms / 1e3, 1),
                               "esd_shifts": [round(float(v), 6) for v in dev_shifts]},
        "tiff_ingest_cint16": {"bytes": slc_bytes, "page_cached": True, "exact": ok,
                               "host_gb_per_s": round(slc_bytes / ingest["host"] / 1e9, 2),
                               "to_hbm_gb_per_s": round(slc_bytes / ingest["device"] / 1e9, 2)},
        "device_deburst_cf32": {"ms": round(db_ms, 3), "rows": height,
                                "gb_per_s": round(db_bytes / db_ms / 1e6, 1),
                                "hbm_frac": round(db_bytes / db_ms / 1e6 / 8000.0, 3)}}))


def run_stack(args, tops, torch):
    """One master (3 IW2 bursts of 1500 x 25000) against 16 slaves, each with
    its own offset field (az 0-0.5 px, rg 0-3 px) and Doppler centroid
    (0-60 Hz): 48 slave bursts resampled and given coherence against the
    master in one tuned batched plan replayed as a hipGraph.  The batch is
    the number of slave bursts whose buffers (slave in, CF32 out, coherence
    out) fit in 90% of the free HBM; it is reported as admitted_batch."""
    N = tops.native
    L = N.load()
    rng = np.random.default_rng(2027)
    n_slaves, n_bursts, lines, width = 16, 3, 1500, 25000
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(dev)
    sp = C.c_void_p(stream.cuda_stream)
    per_item = lines * width * (4 + 8 + 4)
    free, total = torch.cuda.mem_get_info(dev)
    admitted = int(min(n_slaves * n_bursts, (0.9 * free - 3 * lines * width * 4) // per_item))
    if admitted < n_slaves * n_bursts:
        raise SystemExit(f"only {admitted} slave bursts fit in HBM; the stack needs "
                         f"{n_slaves * n_bursts}")
    t0 = time.perf_counter()
    pairs = []
    for k in range(n_slaves):
        az = np.array([rng.uniform(0, 0.25), rng.uniform(0, 0.25) * 0.5, rng.uniform(0, 0.25) * 0.3,
                       0.0, rng.uniform(0, 0.25) * 0.2, 0.0])
        rg = np.array([rng.uniform(0, 1.5), rng.uniform(0, 1.5) * 0.5, rng.uniform(0, 1.5) * 0.3,
                       0.0, rng.uniform(0, 1.5) * 0.2, 0.0])
        pairs.append(tops.synth.make_pair("stack_iw2", lines, width, n_bursts=n_bursts,
                                          overlap=100, seed=4100, stream=k, az_off=az, rg_off=rg,
                                          dem_spacing=1.0, dc_slope_total=10.0,
                                          slave_dc0=float(rng.uniform(0, 60))))
    m = pairs[0].master
    kaz, krg = N.lut_init(12, m.prf, m.azimuth_bandwidth, 2 * m.range_sampling_rate)
    masters = [torch.from_numpy(pairs[0].burst(b, "master")).to(dev) for b in range(n_bursts)]
    keep, specs = [], []
    for k, pr in enumerate(pairs):
        kt, tref, fdc, _ = pr.params()
        for b in range(n_bursts):
            d = {"S": torch.from_numpy(pr.burst(b, "slave")).to(dev)}
            for name, a in (("kt", kt[b]), ("tref", tref[b]), ("fdc", fdc[b])):
                d[name] = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
            ca, cr = pr.cpm(b)
            box = [0, width - 1, b * lines, (b + 1) * lines - 1]
            g = N.BurstGeom((C.c_int * 4)(*box), (C.c_int * 4)(*box), b * lines, width,
                            m.az_time_interval, 0.0, (C.c_double * 6)(*ca),
                            (C.c_double * 6)(*cr))
            d["out"] = torch.empty(lines * width * 2, dtype=torch.float32, device=dev)
            d["coh"] = torch.empty(lines * width, dtype=torch.float32, device=dev)
            keep.append(d)
            specs.append((k, b, g, d))
        print(f"[bench] slave {k + 1}/{n_slaves} generated ({time.perf_counter() - t0:.0f} s)",
              file=sys.stderr, flush=True)
    ctx = L.tops_ctx_create(N.ptr(kaz, C.c_float), N.ptr(krg, C.c_float), 12)
    items = (N.BurstItem * len(specs))()
    for i, (k, b, g, d) in enumerate(specs):
        items[i] = N.BurstItem(g, d["S"].data_ptr(), d["kt"].data_ptr(), d["tref"].data_ptr(),
                               d["fdc"].data_ptr(), d["out"].data_ptr(), masters[b].data_ptr(),
                               d["coh"].data_ptr())
    samples = len(specs) * lines * width
    plan = L.tops_plan_create(ctx, items, len(specs), 3, 10, 0)
    if not plan:
        raise RuntimeError(N.last_error())
    best = C.c_int(0)
    N.check(L.tops_plan_tune(plan, sp, C.byref(best)), "tune")
    N.check(L.tops_plan_capture_graph(plan, sp), "graph")
    for _ in range(args.warmup):
        N.check(L.tops_plan_run(plan, sp))
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        N.check(L.tops_plan_run(plan, sp))
    stream.synchronize()
    step_s = (time.perf_counter() - t0) / args.steps
    # plain FP64 per-burst launches over the same 48 bursts
    t0 = time.perf_counter()
    for k, b, g, d in specs:
        N.check(L.tops_resample_burst(ctx, C.byref(g), d["S"].data_ptr(), d["kt"].data_ptr(),
                                      d["tref"].data_ptr(), d["fdc"].data_ptr(),
                                      d["out"].data_ptr(), N.MODE_FP64, sp))
        N.check(L.tops_coherence(masters[b].data_ptr(), d["out"].data_ptr(), width, lines, 4, 1,
                                 10, 3, d["coh"].data_ptr(), N.MODE_FP64, sp))
    stream.synchronize()
    fp64_s = time.perf_counter() - t0
    # error of the fast path on slabs of the last slave's bursts
    N.check(L.tops_plan_run(plan, sp))
    stream.synchronize()
    err = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        k, b, g, d = specs[-1]
        pr = pairs[k]
        kt, tref, fdc, _ = pr.params()
        ca, cr = pr.cpm(b)
        box = [0, width - 1, b * lines, (b + 1) * lines - 1]
        S = pr.burst(b, "slave")
        got = d["out"].cpu().numpy().view(np.complex64).reshape(lines, width)
        err = 0.0
        for r0 in (0, 744, 1488):
            ref = orc.resample(S, box, box, b * lines, m.az_time_interval, 0.0, ca, cr, kaz, krg,
                               kt[b], tref[b], fdc[b], r0, r0 + 12)
            rms = np.sqrt(np.mean(np.abs(ref) ** 2))
            err = max(err, float(np.max(np.abs(got[r0:r0 + 12] - ref)) / rms))
    L.tops_plan_destroy(plan)
    L.tops_ctx_destroy(ctx)
    print(json.dumps({
        "metric": METRIC, "value": round(samples / step_s / 1e6, 2), "unit": "Msamples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "stack", "slaves": n_slaves, "bursts_per_slave": n_bursts,
                   "burst": [lines, width], "samples": samples, "coh_window": [3, 10],
                   "strip_rows": best.value, "hip_graph": True, "admitted_batch": admitted,
                   "hbm_free_gb": round(free / 1e9, 1), "parallelism": "single GPU"},
        "vs_plain_fp64": round(fp64_s / step_s, 3),
        "plain_fp64_msamples_per_s": round(samples / fp64_s / 1e6, 2),
        "path_hbm_frac": round(PATH_BYTES * samples / step_s / 1e9 / HBM_PEAK_GBS, 4),
        "error_vs_oracle": {"resample_max_abs_over_rms_slabs": err, "tol_resample": 1e-5}}))


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="single_burst_iw1")
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (twice) and exit: "
                         "the reference kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "launches": 2}))
        return

    tops = load_package()
    if args.workload == "iw_pair":
        return run_iw_pair(args, tops, torch)
    if args.workload == "stack":
        return run_stack(args, tops, torch)
    if args.workload == "subswath":
        return run_subswath(args, tops, torch)
    N = tops.native
    L = N.load()
    pair = workloads(tops, args.workload)
    m = pair.master
    lines, pixels = m.lines_per_burst, m.samples_per_burst
    samples = lines * pixels
    print(f"[bench] workload {args.workload}: {lines} x {pixels}", file=sys.stderr, flush=True)

    # host stage layer: LUT, orbit fit, deramp parameters (InKernelInit,
    # Polyfit_Orbit, ComputeDerampPara) through the library
    kaz, krg = N.lut_init(12, m.prf, m.azimuth_bandwidth, 2 * m.range_sampling_rate)
    t, x, y, z = tops.synth.orbit_polyfit_inputs(pair.sorbit)
    cx, cy, cz = N.polyfit_orbit(t, x, y, z)
    speed = L.tops_orbit_speed(0.5 * (pair.slave.first_line_time + pair.slave.last_line_time),
                               t[len(t) // 2], N.ptr(cx), N.ptr(cy), N.ptr(cz), len(cx))
    kt, tref, fdc, _ = N.deramp_params(pair.slave, speed)
    ca, cr = pair.cpm(0)
    M = pair.burst(0, "master")
    S = pair.burst(0, "slave")
    print("[bench] inputs generated", file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    sp = C.c_void_p(stream.cuda_stream)
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev)
         for k, v in dict(S=S, M=M, kt=kt[0], tref=tref[0], fdc=fdc[0]).items()}
    out = torch.empty(samples * 2, dtype=torch.float32, device=dev)
    coh = torch.empty(samples, dtype=torch.float32, device=dev)
    ctx = L.tops_ctx_create(N.ptr(kaz, C.c_float), N.ptr(krg, C.c_float), 12)
    if not ctx:
        raise RuntimeError(N.last_error())
    box = [0, pixels - 1, 0, lines - 1]
    g = N.BurstGeom((C.c_int * 4)(*box), (C.c_int * 4)(*box), 0, pixels, m.az_time_interval, 0.0,
                    (C.c_double * 6)(*ca), (C.c_double * 6)(*cr))
    win_az, win_rg = pair.coh_win
    dx, dy = (win_rg - 1) // 2, (win_az - 1) // 2

    def step(mode, ev=None):
        if ev:
            ev[0].record(stream)
        N.check(L.tops_resample_burst(ctx, C.byref(g), d["S"].data_ptr(), d["kt"].data_ptr(),
                                      d["tref"].data_ptr(), d["fdc"].data_ptr(), out.data_ptr(),
                                      mode, sp), "resample")
        if ev:
            ev[1].record(stream)
        N.check(L.tops_coherence(d["M"].data_ptr(), out.data_ptr(), pixels, lines, dx, dy,
                                 win_rg, win_az, coh.data_ptr(), mode, sp), "coherence")
        if ev:
            ev[2].record(stream)

    def timed(mode, warmup, steps):
        for _ in range(warmup):
            step(mode)
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            step(mode, evs[k])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rs = [e[0].elapsed_time(e[1]) for e in evs]
        ch = [e[1].elapsed_time(e[2]) for e in evs]
        return (t1 - t0) / steps, float(np.mean(rs)) * 1e-3, float(np.mean(ch)) * 1e-3

    # plain all-FP64 GPU form, same inputs, same stream
    fp64_step, fp64_rs, fp64_coh = timed(N.MODE_FP64, 1, max(1, args.fp64_steps))
    # the measured fast path
    step_s, rs_s, coh_s = timed(N.MODE_FAST, args.warmup, args.steps)
    value = samples / step_s / 1e6
    got = out.cpu().numpy().view(np.complex64).reshape(lines, pixels)
    got_coh = coh.cpu().numpy().reshape(lines, pixels)
    print(f"[bench] fast {step_s * 1e3:.3f} ms/step (resample {rs_s * 1e3:.3f} ms, coherence "
          f"{coh_s * 1e3:.3f} ms); fp64 {fp64_step * 1e3:.3f} ms/step", file=sys.stderr, flush=True)

    achieved = RESAMPLE_BYTES * samples / rs_s / 1e9
    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "burst_lines": lines, "burst_samples": pixels,
                   "taps": 12, "coh_window": [win_az, win_rg],
                   "precision": "FP32 sinc MACs and window sums, FP64 phase reduced mod 2pi",
                   "parallelism": "single GPU"},
        "vs_plain_fp64": round(fp64_step / step_s, 3),
        "plain_fp64_msamples_per_s": round(samples / fp64_step / 1e6, 2),
        "kernels_ms": {"resample_fast_kernel": round(rs_s * 1e3, 4),
                       "coherence_kernel": round(coh_s * 1e3, 4),
                       "plain_fp64_resample": round(fp64_rs * 1e3, 4),
                       "plain_fp64_coherence": round(fp64_coh * 1e3, 4)},
        "path_hbm_frac": round(PATH_BYTES * samples / step_s / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"kernel": "resample_fast_kernel", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic("resample_fast_kernel", args.workload),
                     "algorithmic_bytes_per_launch": RESAMPLE_BYTES * samples},
        "valu": {"kernel": "resample_fast_kernel", "reference_flop_per_sample": REFERENCE_FLOP,
                 "achieved": round(REFERENCE_FLOP * samples / rs_s / 1e12, 2),
                 "peak": FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": round(REFERENCE_FLOP * samples / rs_s / 1e12 / FP32_VECTOR_PEAK_TFLOPS, 4)},
    }

    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        threads = orc.max_threads()
        box_l = [0, pixels - 1, 0, lines - 1]

        def cpu_pass(r1):
            # deramp+resample+reramp of rows 0..r1-1, then their coherence
            ref = orc.resample(S, box_l, box_l, 0, m.az_time_interval, 0.0, ca, cr, kaz, krg,
                               kt[0], tref[0], fdc[0], 0, r1)
            return ref, orc.coherence(M[:r1], ref, dx, dy, win_rg, win_az)

        cpu_pass(64)   # warm-up (page faults, thread pool)
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            ref, cref = cpu_pass(lines)
            times.append(time.perf_counter() - t0)
        t_cpu = float(np.median(times))
        # one core on a bounded slice (the first rows of the same burst)
        r1 = min(lines, 150)
        orc.set_threads(1)
        t1 = []
        for _ in range(3):
            t0 = time.perf_counter()
            cpu_pass(r1)
            t1.append(time.perf_counter() - t0)
        orc.set_threads(threads)
        t_one = float(np.median(t1))
        rms = np.sqrt(np.mean(np.abs(ref) ** 2))
        err = float(np.max(np.abs(got - ref)) / rms)
        cerr = float(np.max(np.abs(got_coh - cref)))
        result["cpu_baseline"] = {"value": round(samples / t_cpu / 1e6, 3), "unit": "Msamples/s",
                                  "cores": threads, "kind": "port",
                                  "sample": f"the full {lines}x{pixels} burst, deramp+resample+"
                                            f"reramp+coherence in FP64 (oracle/tops_oracle.c, "
                                            f"OpenMP), median of 5 passes ({t_cpu:.2f} s each)",
                                  "cpu": cpu_model(),
                                  "one_core": {"value": round(r1 * pixels / t_one / 1e6, 3),
                                               "cores": 1,
                                               "sample": f"rows 0..{r1 - 1} of the burst, "
                                                         f"median of 3 passes ({t_one:.2f} s "
                                                         f"each)"}}
        result["error_vs_oracle"] = {"resample_max_abs_over_rms": err, "tol_resample": 1e-5,
                                     "coherence_max_abs": cerr, "tol_coherence": 1e-4}
    print("[bench] geocoreg and ESD", file=sys.stderr, flush=True)
    result["geocoreg"] = bench_geocoreg(tops, torch, pair, L, N, 5, not args.no_cpu_baseline)
    result["esd"] = bench_esd(torch, M, got, L, N, pair, kt[0], 10, not args.no_cpu_baseline)
    L.tops_ctx_destroy(ctx)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
