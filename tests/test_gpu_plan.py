"""Plans over many bursts (batched or per-burst schedule, hipGraph replay, strip-height
tuning) against the single-burst path and the FP64 oracle."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch(tops, dev):
    """5 uneven bursts (lines 300-340, samples 700-900), each with its own
    offsets and Doppler slope, master and slave from the generator."""
    N = tops.native
    L = N.load()
    rng = np.random.default_rng(5)
    bursts = []
    kaz = krg = None
    for i in range(5):
        lines, pixels = int(rng.integers(300, 341)), int(rng.integers(700, 901))
        az = np.array([rng.uniform(0, 0.3), 0.1, 0.05, 0, 0.05, 0])
        rg = np.array([rng.uniform(0, 2.0), 0.4, 0.3, 0, 0.2, 0])
        pair = tops.synth.make_pair(f"b{i}", lines, pixels, seed=100 + i, az_off=az, rg_off=rg,
                                    dc_slope_total=float(rng.uniform(0, 60)))
        m = pair.master
        if kaz is None:
            kaz, krg = N.lut_init(12, m.prf, m.azimuth_bandwidth, 2 * m.range_sampling_rate)
        kt, tref, fdc, _ = pair.params()
        ca, cr = pair.cpm(0)
        bursts.append(dict(pair=pair, M=pair.burst(0, "master"), S=pair.burst(0, "slave"),
                           kt=kt[0].copy(), tref=tref[0].copy(), fdc=fdc[0].copy(), ca=ca, cr=cr,
                           L=lines, P=pixels, dt=m.az_time_interval))
    ctx = L.tops_ctx_create(N.ptr(kaz, C.c_float), N.ptr(krg, C.c_float), 12)
    keep = []
    items = (N.BurstItem * len(bursts))()
    for i, b in enumerate(bursts):
        box = [0, b["P"] - 1, 0, b["L"] - 1]
        g = N.BurstGeom((C.c_int * 4)(*box), (C.c_int * 4)(*box), 0, b["P"], b["dt"], 0.0,
                        (C.c_double * 6)(*b["ca"]), (C.c_double * 6)(*b["cr"]))
        d = dict(S=dev.put(b["S"]), M=dev.put(b["M"]), kt=dev.put(b["kt"]),
                 tref=dev.put(b["tref"]), fdc=dev.put(b["fdc"]),
                 out=dev.empty(b["L"] * b["P"] * 8), coh=dev.empty(b["L"] * b["P"] * 4))
        keep.append(d)
        items[i] = N.BurstItem(g, d["S"].data_ptr(), d["kt"].data_ptr(), d["tref"].data_ptr(),
                               d["fdc"].data_ptr(), d["out"].data_ptr(), d["M"].data_ptr(),
                               d["coh"].data_ptr())
        b["geom"] = g
    yield dict(bursts=bursts, ctx=ctx, items=items, dev=keep, kaz=kaz, krg=krg)
    L.tops_ctx_destroy(ctx)


def _single(tops, dev, ctx, b, d):
    N = tops.native
    L = N.load()
    out = dev.empty(b["L"] * b["P"] * 8)
    coh = dev.empty(b["L"] * b["P"] * 4)
    N.check(L.tops_resample_burst(ctx, C.byref(b["geom"]), d["S"].data_ptr(), d["kt"].data_ptr(),
                                  d["tref"].data_ptr(), d["fdc"].data_ptr(), out.data_ptr(), 0,
                                  None))
    N.check(L.tops_coherence(d["M"].data_ptr(), out.data_ptr(), b["P"], b["L"], 4, 1, 10, 3,
                             coh.data_ptr(), 0, None))
    return (dev.get(out, np.complex64, (b["L"], b["P"])),
            dev.get(coh, np.float32, (b["L"], b["P"])))


def test_plan_matches_single_and_oracle(tops, orc, dev, batch):
    N = tops.native
    L = N.load()
    print("  plan: 5 uneven bursts", flush=True)
    plan = L.tops_plan_create(batch["ctx"], batch["items"], 5, 3, 10, 64)
    assert plan, N.last_error()
    N.check(L.tops_plan_run(plan, None), "tops_plan_run")
    for i, (b, d) in enumerate(zip(batch["bursts"], batch["dev"])):
        got = dev.get(d["out"], np.complex64, (b["L"], b["P"]))
        coh = dev.get(d["coh"], np.float32, (b["L"], b["P"]))
        s_out, s_coh = _single(tops, dev, batch["ctx"], b, d)
        # a different strip height re-anchors the deramp recurrence at other
        # rows: the same result up to FP32 rounding
        assert rel_err(got, s_out) <= 1e-6, i
        assert float(np.max(np.abs(coh - s_coh))) <= 1e-6, i
        if True:
            box = [0, b["P"] - 1, 0, b["L"] - 1]
            ref = orc.resample(b["S"], box, box, 0, b["dt"], 0.0, b["ca"], b["cr"], batch["kaz"],
                               batch["krg"], b["kt"], b["tref"], b["fdc"])
            e = rel_err(got, ref)
            cref = orc.coherence(b["M"], got.astype(np.complex128), 4, 1, 10, 3)
            ec = float(np.max(np.abs(coh - cref)))
            print(f"  burst {i} {b['L']}x{b['P']}: resample {e:.2e}, coherence {ec:.2e}")
            assert e <= 1e-5 and ec <= 1e-4
    L.tops_plan_destroy(plan)


def test_plan_per_burst_schedule_is_the_single_burst_path(tops, dev, batch):
    """strip_rows 0: the plan issues the single-burst launches, so its output
    is bit-identical to tops_resample_burst + tops_coherence."""
    N = tops.native
    L = N.load()
    plan = L.tops_plan_create(batch["ctx"], batch["items"], 5, 3, 10, 0)
    assert plan, N.last_error()
    assert L.tops_plan_strip_rows(plan) == 0
    N.check(L.tops_plan_run(plan, None), "tops_plan_run")
    for i, (b, d) in enumerate(zip(batch["bursts"], batch["dev"])):
        got = dev.get(d["out"], np.complex64, (b["L"], b["P"])).copy()
        coh = dev.get(d["coh"], np.float32, (b["L"], b["P"])).copy()
        s_out, s_coh = _single(tops, dev, batch["ctx"], b, d)
        assert np.array_equal(got, s_out) and np.array_equal(coh, s_coh), i
    L.tops_plan_destroy(plan)


def test_plan_tune_and_graph_replay(tops, dev, batch):
    N = tops.native
    L = N.load()
    torch = dev.torch
    plan = L.tops_plan_create(batch["ctx"], batch["items"], 5, 3, 10, 0)
    assert plan, N.last_error()
    s = torch.cuda.Stream()
    sp = C.c_void_p(s.cuda_stream)
    best = C.c_int(0)
    print("  plan: tuning", flush=True)
    N.check(L.tops_plan_tune(plan, sp, C.byref(best)), "tops_plan_tune")
    # per-burst launches (0), a fixed height or a height that fills whole
    # rounds of wave slots (batched), or a row band of the fused launch
    # (negative)
    assert best.value == 0 or 16 <= abs(best.value) <= 4096
    assert L.tops_plan_strip_rows(plan) == best.value
    # the tuner's record: every candidate timed; the choice is the fastest
    n = L.tops_plan_tune_count(plan)
    assert n >= 10
    log = []
    for i in range(n):
        r, cr, g, st, ms = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_float()
        N.check(L.tops_plan_tune_entry(plan, i, C.byref(r), C.byref(cr), C.byref(g), C.byref(st),
                                       C.byref(ms)))
        log.append((r.value, cr.value, g.value, st.value, ms.value))
    print("  plan: tuner " + ", ".join(f"{r}/{cr}/{g}/s{st}: {ms:.3f}" for r, cr, g, st, ms in log),
          flush=True)
    assert all(e[4] > 0 for e in log)
    coh, groups = L.tops_plan_coh_rows(plan), L.tops_plan_groups(plan)
    assert coh in (16, 32, 64, 128, 256) and 1 <= groups <= 5
    ms = lambda e: e[4]
    # each stage keeps its fastest candidate: the schedule among stage 1,
    # the coherence rows among stage 2 and the stage-1 best, the groups
    # among stage 3 (replayed graphs, one group included)
    s1 = [e for e in log if e[3] == 1]
    b1 = min(s1, key=ms)
    assert best.value == b1[0], log
    if best.value > 0:
        s2 = [b1] + [e for e in log if e[3] == 2]
        assert coh == min(s2, key=ms)[1], log
        s3 = [e for e in log if e[3] == 3]
        assert s3 and all(e[:2] == (best.value, coh) for e in s3), log
        assert groups == min(s3, key=ms)[2], log
    assert L.tops_plan_tune_entry(plan, n, None, None, None, None, None) != 0
    L.tops_clear_error()
    N.check(L.tops_plan_run(plan, sp))
    s.synchronize()
    eager = [(dev.get(d["out"], np.complex64, (b["L"], b["P"])).copy(),
              dev.get(d["coh"], np.float32, (b["L"], b["P"])).copy())
             for b, d in zip(batch["bursts"], batch["dev"])]
    print(f"  plan: graph capture (strip rows {best.value})", flush=True)
    N.check(L.tops_plan_capture_graph(plan, sp), "tops_plan_capture_graph")
    for rep in range(3):
        for d in batch["dev"]:
            d["out"].zero_()
            d["coh"].zero_()
        torch.cuda.synchronize()
        N.check(L.tops_plan_run(plan, sp))
        s.synchronize()
        for (b, d), (eo, ec) in zip(zip(batch["bursts"], batch["dev"]), eager):
            assert np.array_equal(dev.get(d["out"], np.complex64, (b["L"], b["P"])), eo), rep
            assert np.array_equal(dev.get(d["coh"], np.float32, (b["L"], b["P"])), ec), rep
    L.tops_plan_destroy(plan)


@pytest.mark.parametrize("groups", [2, 3, 5, 9])
def test_plan_burst_groups(tops, dev, batch, groups):
    """The batched schedule in burst groups (each group's coherence on the
    plan's second stream, waiting for that group's resample launch only):
    bit-identical to one group, eagerly on a caller stream and as a captured
    hipGraph replayed 3 times; more groups than bursts give one per burst."""
    N = tops.native
    L = N.load()
    torch = dev.torch
    s = torch.cuda.Stream()
    sp = C.c_void_p(s.cuda_stream)
    plan = L.tops_plan_create(batch["ctx"], batch["items"], 5, 3, 10, 96)
    assert plan, N.last_error()
    N.check(L.tops_plan_run(plan, sp))
    s.synchronize()
    shp = [(b["L"], b["P"]) for b in batch["bursts"]]
    one = [(dev.get(d["out"], np.complex64, sh).copy(), dev.get(d["coh"], np.float32, sh).copy())
           for d, sh in zip(batch["dev"], shp)]
    N.check(L.tops_plan_set_groups(plan, groups), "tops_plan_set_groups")
    assert L.tops_plan_groups(plan) == min(groups, 5)
    print(f"  plan: {L.tops_plan_groups(plan)} burst groups", flush=True)

    def check(label):
        for (d, sh), (eo, ec) in zip(zip(batch["dev"], shp), one):
            assert np.array_equal(dev.get(d["out"], np.complex64, sh), eo), label
            assert np.array_equal(dev.get(d["coh"], np.float32, sh), ec), label

    for d in batch["dev"]:
        d["out"].zero_()
        d["coh"].zero_()
    torch.cuda.synchronize()
    N.check(L.tops_plan_run(plan, sp))
    s.synchronize()
    check("eager")
    N.check(L.tops_plan_capture_graph(plan, sp), "tops_plan_capture_graph")
    for rep in range(3):
        for d in batch["dev"]:
            d["out"].zero_()
            d["coh"].zero_()
        torch.cuda.synchronize()
        N.check(L.tops_plan_run(plan, sp))
        s.synchronize()
        check(f"replay {rep}")
        # per-launch timing between replays (eager launches of both streams)
        ms = (C.c_float * 3)()
        nl = (C.c_int * 3)()
        N.check(L.tops_plan_time_launches(plan, sp, 1, ms, nl))
        assert nl[0] == min(groups, 5) and nl[1] == min(groups, 5)
    L.tops_plan_destroy(plan)


def test_plan_per_burst_graph_replay(tops, dev, batch):
    """The per-burst schedule (a resample and a coherence launch per burst)
    captured into a hipGraph and replayed gives the eager result."""
    N = tops.native
    L = N.load()
    torch = dev.torch
    plan = L.tops_plan_create(batch["ctx"], batch["items"], 5, 3, 10, 0)
    assert plan, N.last_error()
    s = torch.cuda.Stream()
    sp = C.c_void_p(s.cuda_stream)
    N.check(L.tops_plan_run(plan, sp))
    s.synchronize()
    eager = [(dev.get(d["out"], np.complex64, (b["L"], b["P"])).copy(),
              dev.get(d["coh"], np.float32, (b["L"], b["P"])).copy())
             for b, d in zip(batch["bursts"], batch["dev"])]
    print("  plan: per-burst schedule graph capture", flush=True)
    N.check(L.tops_plan_capture_graph(plan, sp), "tops_plan_capture_graph")
    for rep in range(3):
        for d in batch["dev"]:
            d["out"].zero_()
            d["coh"].zero_()
        torch.cuda.synchronize()
        N.check(L.tops_plan_run(plan, sp))
        s.synchronize()
        for (b, d), (eo, ec) in zip(zip(batch["bursts"], batch["dev"]), eager):
            assert np.array_equal(dev.get(d["out"], np.complex64, (b["L"], b["P"])), eo), rep
            assert np.array_equal(dev.get(d["coh"], np.float32, (b["L"], b["P"])), ec), rep
    L.tops_plan_destroy(plan)


def test_plan_reference_window_6x24(tops, orc, dev, batch):
    """The reference's own ESDAndCoh window (6 az x 24 rg, dx = 11, dy = 2,
    Func.cpp:2583-2584, 2643-2644) through the batched schedule (the tiled
    coherence kernel) and its hipGraph replay, against the oracle."""
    N = tops.native
    L = N.load()
    torch = dev.torch
    plan = L.tops_plan_create(batch["ctx"], batch["items"], 5, 6, 24, 64)
    assert plan, N.last_error()
    s = torch.cuda.Stream()
    sp = C.c_void_p(s.cuda_stream)
    N.check(L.tops_plan_run(plan, sp))
    s.synchronize()
    eager = []
    for i, (b, d) in enumerate(zip(batch["bursts"], batch["dev"])):
        got = dev.get(d["out"], np.complex64, (b["L"], b["P"])).copy()
        coh = dev.get(d["coh"], np.float32, (b["L"], b["P"])).copy()
        cref = orc.coherence(b["M"], got.astype(np.complex128), 11, 2, 24, 6)
        ec = float(np.max(np.abs(coh - cref)))
        print(f"  burst {i}: coherence 6x24 {ec:.2e}")
        assert ec <= 1e-4
        eager.append(coh)
    N.check(L.tops_plan_capture_graph(plan, sp), "tops_plan_capture_graph")
    for rep in range(2):
        for d in batch["dev"]:
            d["coh"].zero_()
        torch.cuda.synchronize()
        N.check(L.tops_plan_run(plan, sp))
        s.synchronize()
        for (b, d), ec in zip(zip(batch["bursts"], batch["dev"]), eager):
            assert np.array_equal(dev.get(d["coh"], np.float32, (b["L"], b["P"])), ec), rep
    # the same window's tiled coherence in 3 burst groups (second stream)
    N.check(L.tops_plan_set_groups(plan, 3), "tops_plan_set_groups")
    for d in batch["dev"]:
        d["coh"].zero_()
    torch.cuda.synchronize()
    N.check(L.tops_plan_run(plan, sp))
    s.synchronize()
    for (b, d), ec in zip(zip(batch["bursts"], batch["dev"]), eager):
        assert np.array_equal(dev.get(d["coh"], np.float32, (b["L"], b["P"])), ec), "groups"
    L.tops_plan_destroy(plan)


def test_plan_window_beyond_lds_tile(tops, orc, dev, batch):
    """A window whose tile does not fit in LDS (21 x 81): the plan is
    accepted and forms coherence per burst (the FP64 kernel), also on the
    batched schedule, instead of failing at run or capture time."""
    N = tops.native
    L = N.load()
    plan = L.tops_plan_create(batch["ctx"], batch["items"], 5, 21, 81, 64)
    assert plan, N.last_error()
    N.check(L.tops_plan_run(plan, None), "tops_plan_run")
    b, d = batch["bursts"][0], batch["dev"][0]
    got = dev.get(d["out"], np.complex64, (b["L"], b["P"])).copy()
    coh = dev.get(d["coh"], np.float32, (b["L"], b["P"])).copy()
    cref = orc.coherence(b["M"], got.astype(np.complex128), 40, 10, 81, 21)
    ec = float(np.max(np.abs(coh - cref)))
    print(f"  coherence 21x81: {ec:.2e}")
    assert ec <= 1e-4
    L.tops_plan_destroy(plan)


def test_plan_shifted_slave_origin(tops, orc, dev):
    """A slave box far down the raster (first line 700000: the walker's
    packed (line, LUT row) keys near the top of their range): the single-
    burst path and a plan agree with the oracle.  Negative slave lines
    (GetIndexes truncates toward zero, so an interior position there would
    take a negative LUT row) and lines beyond 2^20 are refused by both."""
    N = tops.native
    L = N.load()
    pair = tops.synth.make_pair("origin", 96, 300, seed=11)
    m = pair.master
    kaz, krg = N.lut_init(12, m.prf, m.azimuth_bandwidth, 2 * m.range_sampling_rate)
    kt, tref, fdc, _ = pair.params()
    S = pair.burst(0, "slave")
    ctx = L.tops_ctx_create(N.ptr(kaz, C.c_float), N.ptr(krg, C.c_float), 12)
    dS, dk, dtr, dfd = (dev.put(a) for a in (S, kt[0], tref[0], fdc[0]))
    for y0 in (700000,):
        ca, cr = pair.cpm(0)
        ca = ca.copy()
        ca[0] += y0
        mbox = [0, 299, 0, 95]
        sbox = [0, 299, y0, y0 + 95]
        g = N.BurstGeom((C.c_int * 4)(*mbox), (C.c_int * 4)(*sbox), y0, 300, m.az_time_interval,
                        0.0, (C.c_double * 6)(*ca), (C.c_double * 6)(*cr))
        out = dev.empty(96 * 300 * 8)
        N.check(L.tops_resample_burst(ctx, C.byref(g), dS.data_ptr(), dk.data_ptr(),
                                      dtr.data_ptr(), dfd.data_ptr(), out.data_ptr(), 0, None))
        got = dev.get(out, np.complex64, (96, 300)).copy()
        ref = orc.resample(S, mbox, sbox, y0, m.az_time_interval, 0.0, ca, cr, kaz, krg, kt[0],
                           tref[0], fdc[0])
        e = rel_err(got, ref)
        item = (N.BurstItem * 1)()
        out2 = dev.empty(96 * 300 * 8)
        item[0] = N.BurstItem(g, dS.data_ptr(), dk.data_ptr(), dtr.data_ptr(), dfd.data_ptr(),
                              out2.data_ptr(), None, None)
        plan = L.tops_plan_create(ctx, item, 1, 3, 10, 32)
        assert plan, N.last_error()
        N.check(L.tops_plan_run(plan, None))
        got2 = dev.get(out2, np.complex64, (96, 300))
        L.tops_plan_destroy(plan)
        e2 = rel_err(got2, ref)
        print(f"  slave origin {y0}: single {e:.2e}, plan {e2:.2e}, interior "
              f"{int(np.count_nonzero(ref))} of {ref.size}")
        assert e <= 1e-5 and e2 <= 1e-5 and np.count_nonzero(ref) > ref.size // 2
    for lo in (-40, 1 << 20):
        g.slave_box[2], g.slave_box[3] = lo, lo + 95
        assert L.tops_resample_burst(ctx, C.byref(g), dS.data_ptr(), dk.data_ptr(),
                                     dtr.data_ptr(), dfd.data_ptr(), out.data_ptr(), 0,
                                     None) == N.TOPS_ERR_ARG
        item[0] = N.BurstItem(g, dS.data_ptr(), dk.data_ptr(), dtr.data_ptr(), dfd.data_ptr(),
                              out2.data_ptr(), None, None)
        assert not L.tops_plan_create(ctx, item, 1, 3, 10, 32)
        L.tops_clear_error()
    L.tops_ctx_destroy(ctx)


@pytest.mark.parametrize("strip_rows", [0, 64, -64])
def test_plan_azimuth_shifts(tops, orc, dev, strip_rows):
    """Per-burst azimuth shifts (the ESD pass-2 shift of DerampDemodResample_ESD,
    Func.cpp:2702-2705) through a plan's per-burst, batched and fused
    schedules: three bursts shifted by 0, +0.37 and -1.6 lines, the last two
    also moved by -1.6 and +1.6 pixels in range (positions in (-1, 0) at the
    top and left edges, past the box at the right edge), against the oracle
    resampled with the same offsets, and the coherence of each."""
    N = tops.native
    L = N.load()
    shifts = (0.0, 0.37, -1.6)
    rg_moves = (0.0, -1.6, 1.6)
    bursts, keep = [], []
    kaz = krg = None
    for i, sh in enumerate(shifts):
        lines, pixels = 200 + 17 * i, 400 + 33 * i
        pair = tops.synth.make_pair(f"sh{i}", lines, pixels, seed=300 + i, dc_slope_total=20.0 * i)
        m = pair.master
        if kaz is None:
            kaz, krg = N.lut_init(12, m.prf, m.azimuth_bandwidth, 2 * m.range_sampling_rate)
        kt, tref, fdc, _ = pair.params()
        ca, cr = pair.cpm(0)
        cr = cr.copy()
        cr[0] += rg_moves[i]
        bursts.append(dict(M=pair.burst(0, "master"), S=pair.burst(0, "slave"), kt=kt[0].copy(),
                           tref=tref[0].copy(), fdc=fdc[0].copy(), ca=ca, cr=cr, L=lines,
                           P=pixels, dt=m.az_time_interval, sh=sh, rg=rg_moves[i]))
    ctx = L.tops_ctx_create(N.ptr(kaz, C.c_float), N.ptr(krg, C.c_float), 12)
    items = (N.BurstItem * len(bursts))()
    for i, b in enumerate(bursts):
        box = [0, b["P"] - 1, 0, b["L"] - 1]
        g = N.BurstGeom((C.c_int * 4)(*box), (C.c_int * 4)(*box), 0, b["P"], b["dt"], b["sh"],
                        (C.c_double * 6)(*b["ca"]), (C.c_double * 6)(*b["cr"]))
        d = dict(S=dev.put(b["S"]), M=dev.put(b["M"]), kt=dev.put(b["kt"]),
                 tref=dev.put(b["tref"]), fdc=dev.put(b["fdc"]),
                 out=dev.empty(b["L"] * b["P"] * 8), coh=dev.empty(b["L"] * b["P"] * 4), g=g)
        keep.append(d)
        items[i] = N.BurstItem(g, d["S"].data_ptr(), d["kt"].data_ptr(), d["tref"].data_ptr(),
                               d["fdc"].data_ptr(), d["out"].data_ptr(), d["M"].data_ptr(),
                               d["coh"].data_ptr())
    plan = L.tops_plan_create(ctx, items, len(bursts), 3, 10, strip_rows)
    assert plan, N.last_error()
    N.check(L.tops_plan_run(plan, None))
    for b, d in zip(bursts, keep):
        box = [0, b["P"] - 1, 0, b["L"] - 1]
        ref = orc.resample(b["S"], box, box, 0, b["dt"], b["sh"], b["ca"], b["cr"], kaz, krg,
                           b["kt"], b["tref"], b["fdc"])
        got = dev.get(d["out"], np.complex64, (b["L"], b["P"]))
        e = rel_err(got, ref)
        cref = orc.coherence(b["M"], ref, 4, 1, 10, 3)
        cgot = dev.get(d["coh"], np.float32, (b["L"], b["P"]))
        ec = float(np.max(np.abs(cgot - cref)))
        print(f"  schedule {strip_rows}, shift {b['sh']:+.2f} lines, range {b['rg']:+.2f} px: "
              f"resample {e:.2e}, coherence {ec:.2e}")
        assert e <= 1e-5 and ec <= 1e-4
    L.tops_plan_destroy(plan)
    L.tops_ctx_destroy(ctx)


def test_plan_tune_cache(tops, dev, batch, tmp_path, monkeypatch):
    """TOPS_TUNE_CACHE: a tuned plan appends "<key> strip coh groups"; a
    plan over the same burst shapes and window takes that choice without
    timing anything; other shapes get their own key and line; a cached
    per-burst choice (strip_rows 0) is taken as it is."""
    N = tops.native
    L = N.load()
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("TOPS_TUNE_CACHE", str(cache))
    items, n = batch["items"], len(batch["bursts"])
    best = C.c_int(0)

    def tuned(k):
        plan = L.tops_plan_create(batch["ctx"], items, k, 3, 10, 0)
        assert plan, N.last_error()
        N.check(L.tops_plan_tune(plan, None, C.byref(best)), "tune")
        return plan, (L.tops_plan_strip_rows(plan), L.tops_plan_coh_rows(plan),
                      L.tops_plan_groups(plan)), L.tops_plan_tune_count(plan)

    p1, choice1, timed1 = tuned(n)
    lines = cache.read_text().split("\n")
    assert timed1 > 0 and len([x for x in lines if x]) == 1
    key, *vals = lines[0].split()
    assert tuple(int(v) for v in vals) == choice1 and best.value == choice1[0]
    p2, choice2, timed2 = tuned(n)
    assert timed2 == 0 and choice2 == choice1
    p3, choice3, timed3 = tuned(n - 1)
    assert timed3 > 0
    keys = [x.split()[0] for x in cache.read_text().split("\n") if x]
    assert len(keys) == 2 and keys[0] != keys[1]
    # a cached per-burst schedule is reused (not re-tuned)
    cache.write_text(f"{key} 0 32 1\n")
    p4, choice4, timed4 = tuned(n)
    assert timed4 == 0 and choice4 == (0, 32, 1)
    print(f"  tuned {choice1} ({timed1} runs timed), cached {choice2}, per-burst {choice4}")
    for p in (p1, p2, p3, p4):
        L.tops_plan_destroy(p)
