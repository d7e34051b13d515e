"""The fused persistent schedule of a plan (strip_rows < 0: one launch of
deramp -> resample -> reramp -> coherence over (burst, column strip, row
band) items) against the FP64 oracle (DerampAndResample.cu:108-183, 348-520;
EstCoh.cu:101-139, 413-456) and against the single-burst path.

Tolerances: resample max|dz| / RMS <= 1e-5, coherence max|d| <= 1e-4
absolute (against the oracle's coherence of the GPU's own resampled burst).
The fused kernel sums the 3 x 10 window azimuth-first and re-anchors the
deramp recurrence at each row band, so against the single-burst kernels it
agrees to FP32 rounding, not bit for bit."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _make(tops, dev, shapes, seed, az_scale=1.0, coh=None):
    """Bursts of the given (lines, pixels), each with its own offsets and
    Doppler; coh[i] False gives burst i no master (no coherence)."""
    N = tops.native
    L = N.load()
    rng = np.random.default_rng(seed)
    bursts, kaz, krg = [], None, None
    for i, (lines, pixels) in enumerate(shapes):
        az = np.array([rng.uniform(0, 0.3), 0.1, 0.05 * az_scale, 0, 0.05 * az_scale, 0])
        rg = np.array([rng.uniform(0, 2.0), 0.4, 0.3, 0, 0.2, 0])
        pair = tops.synth.make_pair(f"f{seed}_{i}", lines, pixels, seed=seed * 100 + i,
                                    az_off=az, rg_off=rg, dc_slope_total=float(rng.uniform(0, 60)))
        m = pair.master
        if kaz is None:
            kaz, krg = N.lut_init(12, m.prf, m.azimuth_bandwidth, 2 * m.range_sampling_rate)
        kt, tref, fdc, _ = pair.params()
        ca, cr = pair.cpm(0)
        bursts.append(dict(M=pair.burst(0, "master"), S=pair.burst(0, "slave"), kt=kt[0].copy(),
                           tref=tref[0].copy(), fdc=fdc[0].copy(), ca=ca, cr=cr, L=lines,
                           P=pixels, dt=m.az_time_interval))
    ctx = L.tops_ctx_create(N.ptr(kaz, C.c_float), N.ptr(krg, C.c_float), 12)
    items = (N.BurstItem * len(bursts))()
    keep = []
    for i, b in enumerate(bursts):
        box = [0, b["P"] - 1, 0, b["L"] - 1]
        g = N.BurstGeom((C.c_int * 4)(*box), (C.c_int * 4)(*box), 0, b["P"], b["dt"], 0.0,
                        (C.c_double * 6)(*b["ca"]), (C.c_double * 6)(*b["cr"]))
        d = dict(S=dev.put(b["S"]), M=dev.put(b["M"]), kt=dev.put(b["kt"]),
                 tref=dev.put(b["tref"]), fdc=dev.put(b["fdc"]),
                 out=dev.empty(b["L"] * b["P"] * 8), coh=dev.empty(b["L"] * b["P"] * 4))
        d["coh"].fill_(7)   # a burst without coherence must keep this
        with_coh = coh is None or coh[i]
        items[i] = N.BurstItem(g, d["S"].data_ptr(), d["kt"].data_ptr(), d["tref"].data_ptr(),
                               d["fdc"].data_ptr(), d["out"].data_ptr(),
                               d["M"].data_ptr() if with_coh else None,
                               d["coh"].data_ptr() if with_coh else None)
        b["geom"] = g
        b["with_coh"] = with_coh
        keep.append(d)
    return dict(bursts=bursts, ctx=ctx, items=items, dev=keep, kaz=kaz, krg=krg)


def _check_oracle(orc, dev, w, label):
    er = ec = 0.0
    for i, (b, d) in enumerate(zip(w["bursts"], w["dev"])):
        got = dev.get(d["out"], np.complex64, (b["L"], b["P"])).copy()
        coh = dev.get(d["coh"], np.float32, (b["L"], b["P"])).copy()
        box = [0, b["P"] - 1, 0, b["L"] - 1]
        ref = orc.resample(b["S"], box, box, 0, b["dt"], 0.0, b["ca"], b["cr"], w["kaz"],
                           w["krg"], b["kt"], b["tref"], b["fdc"])
        e = rel_err(got, ref) if np.any(ref) else float(np.max(np.abs(got)))
        er = max(er, e)
        if b["with_coh"]:
            cref = orc.coherence(b["M"], got.astype(np.complex128), 4, 1, 10, 3)
            c = float(np.max(np.abs(coh - cref)))
            ec = max(ec, c)
        else:
            c = float("nan")
            assert np.all(coh == np.frombuffer(np.full(4, 7, np.uint8).tobytes(), np.float32)[0])
        print(f"  {label} burst {i} {b['L']}x{b['P']}: resample {e:.2e}, coherence {c:.2e}",
              flush=True)
    assert er <= 1e-5 and ec <= 1e-4, (label, er, ec)
    return er, ec


def _run(tops, w, rows, stream=None):
    N = tops.native
    L = N.load()
    n = len(w["bursts"])
    plan = L.tops_plan_create(w["ctx"], w["items"], n, 3, 10, -rows)
    assert plan, N.last_error()
    assert L.tops_plan_strip_rows(plan) == -rows
    assert L.tops_plan_kernel_name(plan, 2) == b"burst_fused_kernel"
    N.check(L.tops_plan_run(plan, stream), "tops_plan_run")
    return plan


@pytest.fixture(scope="module")
def uneven(tops, dev):
    rng = np.random.default_rng(8)
    shapes = [(int(rng.integers(300, 341)), int(rng.integers(700, 901))) for _ in range(5)]
    w = _make(tops, dev, shapes, 8)
    yield w
    tops.native.load().tops_ctx_destroy(w["ctx"])


@pytest.mark.parametrize("rows", [1, 7, 64, 1000])
def test_fused_matches_oracle(tops, orc, dev, uneven, rows):
    """Row bands of 1 line (every row is next to a band edge), 7, 64 and
    the whole burst."""
    print(f"  fused: 5 uneven bursts, row bands of {rows}", flush=True)
    plan = _run(tops, uneven, rows)
    _check_oracle(orc, dev, uneven, f"rows {rows}")
    tops.native.load().tops_plan_destroy(plan)


def test_fused_matches_single_burst_path(tops, dev, uneven):
    N = tops.native
    L = N.load()
    plan = _run(tops, uneven, 128)
    for i, (b, d) in enumerate(zip(uneven["bursts"], uneven["dev"])):
        got = dev.get(d["out"], np.complex64, (b["L"], b["P"])).copy()
        coh = dev.get(d["coh"], np.float32, (b["L"], b["P"])).copy()
        out = dev.empty(b["L"] * b["P"] * 8)
        sc = dev.empty(b["L"] * b["P"] * 4)
        N.check(L.tops_resample_burst(uneven["ctx"], C.byref(b["geom"]), d["S"].data_ptr(),
                                      d["kt"].data_ptr(), d["tref"].data_ptr(),
                                      d["fdc"].data_ptr(), out.data_ptr(), 0, None))
        N.check(L.tops_coherence(d["M"].data_ptr(), out.data_ptr(), b["P"], b["L"], 4, 1, 10, 3,
                                 sc.data_ptr(), 0, None))
        s_out = dev.get(out, np.complex64, (b["L"], b["P"]))
        s_coh = dev.get(sc, np.float32, (b["L"], b["P"]))
        e, c = rel_err(got, s_out), float(np.max(np.abs(coh - s_coh)))
        print(f"  burst {i}: fused vs single-burst kernels: resample {e:.2e}, coherence {c:.2e}")
        # both are about 1e-6 of the RMS from the oracle, in different
        # FP32 rounding
        assert e <= 3e-6 and c <= 1e-5
    L.tops_plan_destroy(plan)


def test_fused_graph_replay(tops, dev, uneven):
    """Captured once, replayed 3 times with the outputs cleared and the
    program's own work in between: every replay equals the eager run (the
    work queue resets itself)."""
    N = tops.native
    L = N.load()
    torch = dev.torch
    s = torch.cuda.Stream()
    sp = C.c_void_p(s.cuda_stream)
    plan = _run(tops, uneven, 96, sp)
    s.synchronize()
    shp = [(b["L"], b["P"]) for b in uneven["bursts"]]
    eager = [(dev.get(d["out"], np.complex64, sh).copy(), dev.get(d["coh"], np.float32, sh).copy())
             for d, sh in zip(uneven["dev"], shp)]
    N.check(L.tops_plan_capture_graph(plan, sp), "tops_plan_capture_graph")
    for rep in range(3):
        for d in uneven["dev"]:
            d["out"].zero_()
            d["coh"].zero_()
        torch.cuda.synchronize()
        N.check(L.tops_plan_run(plan, sp))
        s.synchronize()
        for (d, sh), (eo, ec) in zip(zip(uneven["dev"], shp), eager):
            assert np.array_equal(dev.get(d["out"], np.complex64, sh), eo), rep
            assert np.array_equal(dev.get(d["coh"], np.float32, sh), ec), rep
        # eager launches of the same plan's kernel between replays
        N.check(L.tops_plan_time_launches(plan, sp, 1, (C.c_float * 3)(), (C.c_int * 3)()))
    L.tops_plan_destroy(plan)


def test_fused_small_and_odd_boxes(tops, orc, dev):
    """Bursts narrower than a strip, one and two columns past a strip's 234
    owned columns, shorter than the coherence window, and a single line."""
    shapes = [(2, 2), (5, 3), (17, 250), (40, 235), (33, 236), (2, 300), (3, 469), (12, 700)]
    w = _make(tops, dev, shapes, 21)
    try:
        plan = _run(tops, w, 16)
        _check_oracle(orc, dev, w, "small")
        tops.native.load().tops_plan_destroy(plan)
    finally:
        tops.native.load().tops_ctx_destroy(w["ctx"])


@pytest.mark.parametrize("rows,groups", [(0, 1), (16, 1), (64, 1), (16, 3)])
def test_batched_small_and_odd_boxes(tops, orc, dev, rows, groups):
    """The same small and odd bursts through the per-burst (0) and batched
    schedules (strips of 16 and 64 rows; 16 rows also in 3 burst groups)."""
    shapes = [(2, 2), (5, 3), (17, 250), (40, 235), (33, 236), (2, 300), (3, 469), (12, 700)]
    w = _make(tops, dev, shapes, 21)
    N = tops.native
    L = N.load()
    try:
        plan = L.tops_plan_create(w["ctx"], w["items"], len(shapes), 3, 10, rows)
        assert plan, N.last_error()
        assert L.tops_plan_strip_rows(plan) == rows
        if groups > 1:
            N.check(L.tops_plan_set_groups(plan, groups))
        N.check(L.tops_plan_run(plan, None), "tops_plan_run")
        _check_oracle(orc, dev, w, f"schedule {rows} / {groups} groups")
        L.tops_plan_destroy(plan)
    finally:
        L.tops_ctx_destroy(w["ctx"])


def test_fused_bursts_without_coherence(tops, orc, dev):
    """Bursts whose item has no master: resampled, coherence left alone."""
    shapes = [(60, 500), (61, 480), (59, 520)]
    w = _make(tops, dev, shapes, 22, coh=[True, False, True])
    try:
        plan = _run(tops, w, 32)
        _check_oracle(orc, dev, w, "no-coherence")
        tops.native.load().tops_plan_destroy(plan)
    finally:
        tops.native.load().tops_ctx_destroy(w["ctx"])


def test_batched_groups_with_bursts_without_coherence(tops, orc, dev):
    """The batched schedule in burst groups where some groups' bursts have
    no coherence item (their coherence launch is skipped): outputs equal the
    oracle, and the coherence buffers of those bursts are left alone."""
    N = tops.native
    L = N.load()
    shapes = [(60, 500), (61, 480), (59, 520), (62, 510), (58, 490)]
    w = _make(tops, dev, shapes, 24, coh=[False, True, False, False, True])
    try:
        n = len(shapes)
        plan = L.tops_plan_create(w["ctx"], w["items"], n, 3, 10, 32)
        assert plan, N.last_error()
        N.check(L.tops_plan_set_groups(plan, 4), "tops_plan_set_groups")
        assert L.tops_plan_groups(plan) == 4
        N.check(L.tops_plan_run(plan, None), "tops_plan_run")
        _check_oracle(orc, dev, w, "groups, some without coherence")
        L.tops_plan_destroy(plan)
    finally:
        L.tops_ctx_destroy(w["ctx"])


@pytest.mark.parametrize("az_scale", [40.0, 400.0])
def test_fused_azimuth_shear(tops, orc, dev, az_scale):
    """Strong azimuth shear along the range: many azimuth keys per row, so
    key slots collide and keys fall beyond the window's reach; those outputs
    take the direct path."""
    w = _make(tops, dev, [(120, 900), (100, 700)], 23, az_scale=az_scale)
    try:
        plan = _run(tops, w, 40)
        _check_oracle(orc, dev, w, f"shear x{az_scale:g}")
        tops.native.load().tops_plan_destroy(plan)
    finally:
        tops.native.load().tops_ctx_destroy(w["ctx"])


@pytest.mark.parametrize("rows", [0, 64, -48])
def test_plans_on_absolute_burst_boxes(tops, orc, dev, rows):
    """Bursts of a subswath at their own boxes (master and slave lines
    b*lpb .. (b+1)*lpb - 1, first line b*lpb, as DerampAndResample passes
    them, Func.cpp:2356-2371) through the per-burst, batched and fused
    schedules: resample and coherence of every burst against the oracle."""
    N = tops.native
    L = N.load()
    pair = tops.synth.subswath_iw1(n_bursts=3, lines=200, pixels=600)
    m = pair.master
    lpb, spb = m.lines_per_burst, m.samples_per_burst
    kaz, krg = N.lut_init(12, m.prf, m.azimuth_bandwidth, 2 * m.range_sampling_rate)
    kt, tref, fdc, _ = pair.params()
    ctx = L.tops_ctx_create(N.ptr(kaz, C.c_float), N.ptr(krg, C.c_float), 12)
    items = (N.BurstItem * m.n_bursts)()
    keep, host = [], []
    for b in range(m.n_bursts):
        ca, cr = pair.cpm(b)
        box = [0, spb - 1, b * lpb, (b + 1) * lpb - 1]
        g = N.BurstGeom((C.c_int * 4)(*box), (C.c_int * 4)(*box), b * lpb, spb,
                        m.az_time_interval, 0.0, (C.c_double * 6)(*ca), (C.c_double * 6)(*cr))
        S, M = pair.burst(b, "slave"), pair.burst(b, "master")
        d = dict(g=g, S=dev.put(S), M=dev.put(M), kt=dev.put(kt[b].copy()),
                 tref=dev.put(tref[b].copy()), fdc=dev.put(fdc[b].copy()),
                 out=dev.empty(lpb * spb * 8), coh=dev.empty(lpb * spb * 4))
        items[b] = N.BurstItem(g, d["S"].data_ptr(), d["kt"].data_ptr(), d["tref"].data_ptr(),
                               d["fdc"].data_ptr(), d["out"].data_ptr(), d["M"].data_ptr(),
                               d["coh"].data_ptr())
        keep.append(d)
        host.append(dict(S=S, M=M, ca=ca, cr=cr, box=box, kt=kt[b], tref=tref[b], fdc=fdc[b]))
    try:
        plan = L.tops_plan_create(ctx, items, m.n_bursts, 3, 10, rows)
        assert plan, N.last_error()
        N.check(L.tops_plan_run(plan, None), "tops_plan_run")
        for b, (d, h) in enumerate(zip(keep, host)):
            ref = orc.resample(h["S"], h["box"], h["box"], b * lpb, m.az_time_interval, 0.0,
                               h["ca"], h["cr"], kaz, krg, h["kt"], h["tref"], h["fdc"])
            got = dev.get(d["out"], np.complex64, (lpb, spb))
            e = rel_err(got, ref)
            cref = orc.coherence(h["M"], got.astype(np.complex128), 4, 1, 10, 3)
            ec = float(np.max(np.abs(dev.get(d["coh"], np.float32, (lpb, spb)) - cref)))
            print(f"  schedule {rows}, burst {b} (lines {h['box'][2]}..{h['box'][3]}): "
                  f"resample {e:.2e}, coherence {ec:.2e}")
            assert e <= 1e-5 and ec <= 1e-4
        L.tops_plan_destroy(plan)
    finally:
        L.tops_ctx_destroy(ctx)
