#!/usr/bin/env python3
"""Throughput of the TOPS burst hot path on one MI355X.

A step is one pass of deramp -> 12x12 sinc resample -> reramp -> coherence
over bursts already resident in HBM.  The default workload is the batched
IW pair: 3 subswaths x 9 bursts (21000 / 25000 / 24500 range samples,
1480-1520 lines each, every burst its own seeded scene, offsets and
Doppler), 3 az x 10 rg coherence window, one tuned plan replayed as a
hipGraph.  The metric counts master-grid output samples per second.

Printed JSON line (one): metric, value (Msamples/s), the roofline of the
dominant kernel (its algorithmic bytes over its HIP-event launch duration on
the plan's stream; measured VALU issue utilisation beside it), the plain
all-FP64 GPU form on the same bursts (vs_plain_fp64), the error of the
measured output against the FP64 oracle, and the CPU baseline (the FP64
oracle on the host cores, on a sample of the bursts).  Other workloads:
--workload stack (48 slave bursts against one master), single_burst_iw1,
single_burst_small, subswath.
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
# the azimuth-first moment form, which needs fewer; the rate of the work it
# replaces is reported as context, the measured issue utilisation is the
# arithmetic bound.
REFERENCE_FLOP = 624 + 312


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
    per call (the call ends with the 27-sum readback and the 6x6 solve)."""
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
           "dem": list(dem.shape),
           "precision": "FP32 DEM projection (relative to an FP64 reference post per 4 x 64 "
                        "posts), FP64 orbit, zero-Doppler Newton and fit"}
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


# algorithmic bytes per output sample of each launch kind of a plan:
# resample (slave in, CF32 out), coherence (master, resampled slave, coherence)
# and the fused launch (the whole path)
KIND_BYTES = (RESAMPLE_BYTES, 16, PATH_BYTES)
KIND_LABEL = ("resample", "coherence", "fused resample + coherence")
SIMDS = 256 * 4          # MI355X: 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9         # peak engine clock


def read_sq(kernel, workload):
    """SQ counters per dispatch of a kernel (rocprofv3 --pmc, committed
    profile summary), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_sq.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def tuner_record(L, plan, chosen):
    """The candidates tops_plan_tune timed ([strip_rows, coherence rows per
    block, burst groups, ms per run]; strip_rows 0 = per-burst launches,
    > 0 = batched strips of that height, < 0 = the fused launch over row
    bands), the choice, and the untuned default (per-burst launches, 32
    coherence rows, one group)."""
    timed = []
    for i in range(L.tops_plan_tune_count(plan)):
        r, cr, g, st, ms = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_float()
        if L.tops_plan_tune_entry(plan, i, C.byref(r), C.byref(cr), C.byref(g), C.byref(st),
                                  C.byref(ms)) == 0:
            timed.append([r.value, cr.value, g.value, st.value, round(ms.value, 4)])
    pick = [chosen, L.tops_plan_coh_rows(plan), L.tops_plan_groups(plan)]
    # eager timings (stages 1 and 2) of the default and of the chosen schedule
    # with one group
    eager = {(r, c, g): m for r, c, g, st, m in timed if st in (1, 2)}
    default = eager.get((0, 32, 1))
    tuned = eager.get((pick[0], pick[1], 1))
    return {"timed": timed, "chosen": pick, "chosen_ms_one_group": tuned, "default": [0, 32, 1],
            "default_ms": default,
            "speedup_vs_default": round(default / tuned, 3) if default and tuned else None,
            "note": "entries [strip_rows, coh_rows, groups, stage, ms]; stages 1 and 2 eager "
                    "runs, stage 3 (burst groups) replayed graphs"}


def issue_fields(roof, kernel, workload, launch_s, samples_per_launch):
    """The measured VALU issue utilisation of a launch of `kernel`, from the
    committed SQ counters of that kernel on that workload: SQ_INSTS_VALU x 4
    cycles per wave64 instruction / (1024 SIMDs x 2.4 GHz x launch_s).  Sets
    roof["bound"] to "valu_issue" when that fraction exceeds the HBM one.
    The reference form's FP32 work per sample is context only."""
    sq = read_sq(kernel, workload)
    valu = {"kernel": kernel}
    if sq and sq.get("SQ_INSTS_VALU"):
        issue = sq["SQ_INSTS_VALU"] * 4.0 / (SIMDS * CLOCK_HZ * launch_s)
        valu.update({"sq_insts_valu_per_launch": sq["SQ_INSTS_VALU"],
                     "sq_insts_salu_per_launch": sq.get("SQ_INSTS_SALU"),
                     "valu_lane_ops_per_sample": round(sq["SQ_INSTS_VALU"] * 64.0
                                                       / samples_per_launch, 1),
                     "issue_frac": round(issue, 4),
                     "note": "VALU issue utilisation: SQ_INSTS_VALU x 4 cycles per wave64 "
                             "instruction / (1024 SIMDs x 2.4 GHz x the launch duration)"})
        if issue > roof["frac"]:
            roof["bound"] = "valu_issue"
    roof["limiter"] = roof["bound"]
    valu["reference_form_flop_per_sample"] = REFERENCE_FLOP
    valu["reference_form_tflops"] = round(REFERENCE_FLOP * samples_per_launch / launch_s / 1e12, 2)
    return valu


def run_batched(args, tops, torch):
    """A batched workload through one tuned plan replayed as a hipGraph:
      iw_pair: 27 bursts (3 subswaths x 9; 21000 / 25000 / 24500 samples,
               1480-1520 lines), every burst its own scene, offsets and
               Doppler (synth.iw_pair_bursts);
      stack:   one 3-burst IW2 master (1500 x 25000) against 16 slaves, 48
               slave bursts (synth.stack_bursts); the batch is the number of
               slave bursts whose buffers fit in 90% of the free HBM.
    A step is one replay: deramp -> resample -> reramp -> coherence of every
    burst.  Compared with per-burst plain FP64 launches and per-burst fast
    launches on the same GPU; the CPU baseline and the error are taken on a
    sample of the bursts (one per burst width)."""
    N = tops.native
    L = N.load()
    workload = args.workload
    S_ = tops.synth
    bursts = S_.iw_pair_bursts() if workload == "iw_pair" else S_.stack_bursts()
    dev = torch.device("cuda:0")
    free, _ = torch.cuda.mem_get_info(dev)
    per_burst = max(b.lines * b.width for b in bursts) * (4 + 8 + 4 + 4)
    admitted = int(min(len(bursts), 0.9 * free // per_burst))
    if admitted < len(bursts):
        raise SystemExit(f"only {admitted} of {len(bursts)} bursts fit in HBM")
    # bursts kept on the host for the CPU baseline and the error: the first
    # burst of each width
    sampled, seen = [], set()
    for i, bb in enumerate(bursts):
        if bb.width not in seen:
            seen.add(bb.width)
            sampled.append(i)
    if len(sampled) == 1:
        sampled.append(len(bursts) - 1)
    m0 = bursts[0].pair.master
    kaz, krg = N.lut_init(12, m0.prf, m0.azimuth_bandwidth, 2 * m0.range_sampling_rate)
    stream = torch.cuda.Stream(dev)
    sp = C.c_void_p(stream.cuda_stream)
    keep, host, masters = [], {}, {}
    items = (N.BurstItem * len(bursts))()
    samples = 0
    t0 = time.perf_counter()
    for i, bb in enumerate(bursts):
        S = bb.slave()
        mkey = ("m", bb.index) if workload == "stack" else i
        if mkey not in masters:
            M = bb.master()
            masters[mkey] = (M, torch.from_numpy(M).to(dev))
        M, dM = masters[mkey]
        kt, tref, fdc = bb.params()
        ca, cr = bb.cpm()
        d = {"S": torch.from_numpy(S).to(dev), "M": dM,
             "kt": torch.from_numpy(kt).to(dev), "tref": torch.from_numpy(tref).to(dev),
             "fdc": torch.from_numpy(fdc).to(dev),
             "out": torch.empty(bb.lines * bb.width * 2, dtype=torch.float32, device=dev),
             "coh": torch.empty(bb.lines * bb.width, dtype=torch.float32, device=dev)}
        g = N.BurstGeom((C.c_int * 4)(*bb.box), (C.c_int * 4)(*bb.box), bb.y0, bb.width, bb.dt,
                        0.0, (C.c_double * 6)(*ca), (C.c_double * 6)(*cr))
        d["geom"] = g
        items[i] = N.BurstItem(g, d["S"].data_ptr(), d["kt"].data_ptr(), d["tref"].data_ptr(),
                               d["fdc"].data_ptr(), d["out"].data_ptr(), dM.data_ptr(),
                               d["coh"].data_ptr())
        keep.append(d)
        if i in sampled:
            host[i] = dict(S=S, M=M, kt=kt, tref=tref, fdc=fdc, ca=ca, cr=cr)
        samples += bb.lines * bb.width
        if i % 9 == 8 or i == len(bursts) - 1:
            print(f"[bench] {workload}: {i + 1}/{len(bursts)} bursts generated "
                  f"({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)
        if workload == "stack":
            del S
    ctx = L.tops_ctx_create(N.ptr(kaz, C.c_float), N.ptr(krg, C.c_float), 12)
    plan = L.tops_plan_create(ctx, items, len(bursts), 3, 10, 0)
    if not plan:
        raise RuntimeError(N.last_error())
    best = C.c_int(0)
    N.check(L.tops_plan_tune(plan, sp, C.byref(best)), "tune")
    tuner = tuner_record(L, plan, best.value)
    if args.groups > 0:
        N.check(L.tops_plan_set_groups(plan, args.groups), "groups")
    print(f"[bench] {workload}: tuner " + ", ".join(f"{r}/{c}/{g}/s{st}: {m}"
                                                     for r, c, g, st, m in tuner["timed"]),
          file=sys.stderr, flush=True)
    names = [L.tops_plan_kernel_name(plan, k).decode() for k in range(3)]
    N.check(L.tops_plan_capture_graph(plan, sp), "graph")
    print(f"[bench] {workload}: schedule {best.value} ({', '.join(n for n in names if n)})",
          file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        N.check(L.tops_plan_run(plan, sp))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        N.check(L.tops_plan_run(plan, sp))
    torch.cuda.synchronize()
    step_s = (time.perf_counter() - t0) / args.steps
    # the kernels of the schedule, with HIP events on the plan's stream.  With
    # burst groups a group's coherence runs beside the next group's
    # resampling, which stretches both launches: the launch durations (and
    # the roofline) are taken with one group, each kernel on its own.
    groups = L.tops_plan_groups(plan)
    if groups > 1:
        N.check(L.tops_plan_set_groups(plan, 1), "groups")
    ms = (C.c_float * 3)()
    nl = (C.c_int * 3)()
    N.check(L.tops_plan_time_launches(plan, sp, max(3, min(args.steps, 10)), ms, nl), "timing")
    kernels_ms = {names[k]: round(ms[k], 4) for k in range(3) if nl[k] > 0}
    dom = max((k for k in range(3) if nl[k] > 0), key=lambda k: ms[k])
    dom_s = ms[dom] * 1e-3 / nl[dom]            # one launch of the dominant kind
    dom_bytes = KIND_BYTES[dom] * samples / nl[dom]
    achieved = dom_bytes / dom_s / 1e9
    roof = {"kernel": names[dom], "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": read_traffic(names[dom], workload),
            "algorithmic_bytes_per_launch": int(dom_bytes),
            "algorithmic_bytes_per_sample": KIND_BYTES[dom], "launches_per_step": nl[dom],
            "launch_ms": round(dom_s * 1e3, 4)}
    valu = issue_fields(roof, names[dom], workload, dom_s, samples / nl[dom])
    # per-burst plain FP64 launches (same inputs)
    def fp64_pass():
        for bb, d in zip(bursts, keep):
            N.check(L.tops_resample_burst(ctx, C.byref(d["geom"]), d["S"].data_ptr(),
                                          d["kt"].data_ptr(), d["tref"].data_ptr(),
                                          d["fdc"].data_ptr(), d["out"].data_ptr(), N.MODE_FP64,
                                          sp))
            N.check(L.tops_coherence(d["M"].data_ptr(), d["out"].data_ptr(), bb.width, bb.lines,
                                     4, 1, 10, 3, d["coh"].data_ptr(), N.MODE_FP64, sp))
    fp64_pass()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(1, args.fp64_steps)):
        fp64_pass()
    torch.cuda.synchronize()
    fp64_s = (time.perf_counter() - t0) / max(1, args.fp64_steps)
    # per-burst fast launches (no plan, no graph)
    def per_burst():
        for bb, d in zip(bursts, keep):
            N.check(L.tops_resample_burst(ctx, C.byref(d["geom"]), d["S"].data_ptr(),
                                          d["kt"].data_ptr(), d["tref"].data_ptr(),
                                          d["fdc"].data_ptr(), d["out"].data_ptr(), N.MODE_FAST,
                                          sp))
            N.check(L.tops_coherence(d["M"].data_ptr(), d["out"].data_ptr(), bb.width, bb.lines,
                                     4, 1, 10, 3, d["coh"].data_ptr(), N.MODE_FAST, sp))
    per_burst()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        per_burst()
    torch.cuda.synchronize()
    per_burst_s = (time.perf_counter() - t0) / args.steps
    # the measured outputs: one more replay of the plan
    N.check(L.tops_plan_run(plan, sp))
    torch.cuda.synchronize()
    result = {
        "metric": METRIC, "value": round(samples / step_s / 1e6, 2), "unit": "Msamples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload, "bursts": len(bursts), "samples": samples,
                   "widths": sorted({b.width for b in bursts}),
                   "lines": [min(b.lines for b in bursts), max(b.lines for b in bursts)],
                   "taps": 12, "coh_window": [3, 10], "schedule_strip_rows": best.value,
                   "coh_rows": L.tops_plan_coh_rows(plan), "burst_groups": groups,
                   "tuner": tuner,
                   "kernels": [n for n in names if n], "hip_graph": True,
                   "admitted_batch": admitted, "hbm_free_gb": round(free / 1e9, 1),
                   "precision": "FP32 sinc MACs and window sums, FP64 phase reduced mod 2pi",
                   "parallelism": "single GPU"},
        "vs_plain_fp64": round(fp64_s / step_s, 3),
        "plain_fp64_msamples_per_s": round(samples / fp64_s / 1e6, 2),
        "per_burst_launches_msamples_per_s": round(samples / per_burst_s / 1e6, 2),
        "kernels_ms": kernels_ms,
        "path_hbm_frac": round(PATH_BYTES * samples / step_s / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": roof, "valu": valu,
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        threads = orc.max_threads()
        t_cpu, n_cpu, err, cerr = 0.0, 0, 0.0, 0.0
        for i in sampled:
            bb, h, d = bursts[i], host[i], keep[i]
            def cpu_pass(r1):
                ref = orc.resample(h["S"], bb.box, bb.box, bb.y0, bb.dt, 0.0, h["ca"], h["cr"],
                                   kaz, krg, h["kt"], h["tref"], h["fdc"], 0, r1)
                return ref, orc.coherence(h["M"][:r1], ref, 4, 1, 10, 3)
            cpu_pass(32)   # warm-up
            times = []
            for _ in range(3):
                t0 = time.perf_counter()
                ref, cref = cpu_pass(bb.lines)
                times.append(time.perf_counter() - t0)
            t_cpu += float(np.median(times))
            n_cpu += bb.lines * bb.width
            got = d["out"].cpu().numpy().view(np.complex64).reshape(bb.lines, bb.width)
            gcoh = d["coh"].cpu().numpy().reshape(bb.lines, bb.width)
            err = max(err, float(np.max(np.abs(got - ref)) / np.sqrt(np.mean(np.abs(ref) ** 2))))
            cerr = max(cerr, float(np.max(np.abs(gcoh - cref))))
            print(f"[bench] oracle burst {i} ({bb.lines} x {bb.width}): {times[-1]:.2f} s",
                  file=sys.stderr, flush=True)
        bb, h = bursts[sampled[0]], host[sampled[0]]
        r1 = 150
        orc.set_threads(1)
        t1 = []
        for _ in range(3):
            t0 = time.perf_counter()
            orc.resample(h["S"], bb.box, bb.box, bb.y0, bb.dt, 0.0, h["ca"], h["cr"], kaz, krg,
                         h["kt"], h["tref"], h["fdc"], 0, r1)
            t1.append(time.perf_counter() - t0)
        orc.set_threads(threads)
        sample_txt = ", ".join(f"{bursts[i].lines}x{bursts[i].width}" for i in sampled)
        result["cpu_baseline"] = {
            "value": round(n_cpu / t_cpu / 1e6, 3), "unit": "Msamples/s", "cores": threads,
            "kind": "port",
            "sample": f"{len(sampled)} of the {len(bursts)} bursts ({sample_txt}, one per burst "
                      f"width), full deramp+resample+reramp+coherence in FP64 "
                      f"(oracle/tops_oracle.c, OpenMP), median of 3 passes each; the rate "
                      f"extrapolates to the whole workload by sample count "
                      f"({samples / (n_cpu / t_cpu):.1f} s)",
            "extrapolated": True, "cpu": cpu_model(),
            "one_core": {"value": round(r1 * bb.width / float(np.median(t1)) / 1e6, 3),
                         "cores": 1, "sample": f"rows 0..{r1 - 1} of burst {sampled[0]}, "
                                               f"resample only, median of 3 passes"}}
        result["error_vs_oracle"] = {
            "bursts": sampled, "resample_max_abs_over_rms": err, "tol_resample": 1e-5,
            "coherence_max_abs": cerr, "tol_coherence": 1e-4}
    if workload == "iw_pair" and not args.no_side_stages:
        print("[bench] geocoreg and ESD", file=sys.stderr, flush=True)
        pair = bursts[0].pair
        kt, tref, fdc, _ = pair.params()
        got = keep[0]["out"].cpu().numpy().view(np.complex64).reshape(bursts[0].lines,
                                                                      bursts[0].width)
        result["geocoreg"] = bench_geocoreg(tops, torch, pair, L, N, 5, not args.no_cpu_baseline)
        result["esd"] = bench_esd(torch, masters[0][0], got, L, N, pair, kt[0], 10,
                                  not args.no_cpu_baseline)
    L.tops_plan_destroy(plan)
    L.tops_ctx_destroy(ctx)
    print(json.dumps(result))


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


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="iw_pair",
                    help="iw_pair (default: 27 uneven bursts, one tuned plan), stack, "
                         "single_burst_iw1, single_burst_small, subswath")
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-side-stages", action="store_true",
                    help="iw_pair: skip the geocoreg and ESD figures")
    ap.add_argument("--groups", type=int, default=0,
                    help="iw_pair / stack: burst groups of the plan (0: the tuner's choice)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (twice) and exit: "
                         "the reference kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU only; --gpus must be 1",
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
    if args.workload in ("iw_pair", "stack"):
        return run_batched(args, tops, torch)
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
                     "algorithmic_bytes_per_launch": RESAMPLE_BYTES * samples,
                     "algorithmic_bytes_per_sample": RESAMPLE_BYTES,
                     "launch_ms": round(rs_s * 1e3, 4)},
    }
    result["valu"] = issue_fields(result["roofline"], "resample_fast_kernel", args.workload,
                                  rs_s, samples)

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
