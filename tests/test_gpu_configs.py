"""The batched workloads at their full shapes, through the C-ABI plan API:

  * the IW pair: 27 bursts (3 subswaths x 9, 21000 / 25000 / 24500 samples,
    1480-1520 lines), every burst its own scene, offsets and Doppler;
  * the stack: 48 slave bursts (16 slaves x 3 bursts of 1500 x 25000)
    against one master.

Each plan is tuned (tops_plan_tune), captured as a hipGraph and replayed;
rows at the top, middle and bottom of every burst are compared with the FP64
oracle (deramp + resample + reramp, DerampAndResample.cu:108-183, 348-520;
coherence, EstCoh.cu:101-139, 413-456).  Sizes are those of the benchmark,
so the generation of the inputs takes most of each test's time."""
import ctypes as C
import time

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

SLAB = 12   # rows per checked slab


def _slabs(lines):
    mid = lines // 2 - SLAB // 2
    return [(0, SLAB), (mid, mid + SLAB), (lines - SLAB, lines)]


def _build(tops, dev, bursts, label):
    """Inputs of every burst on the device (plus the host rasters the oracle
    reads), one tops_burst_item each."""
    N = tops.native
    L = N.load()
    m0 = bursts[0].pair.master
    kaz, krg = N.lut_init(12, m0.prf, m0.azimuth_bandwidth, 2 * m0.range_sampling_rate)
    ctx = L.tops_ctx_create(N.ptr(kaz, C.c_float), N.ptr(krg, C.c_float), 12)
    assert ctx, N.last_error()
    items = (N.BurstItem * len(bursts))()
    host, keep, masters = [], [], {}
    t0 = time.perf_counter()
    for i, bb in enumerate(bursts):
        S = bb.slave()
        # the stack's slaves share one master: its bursts are generated once
        key = ("master", bb.index) if label == "stack" else i
        if key not in masters:
            M = bb.master()
            masters[key] = (M, dev.put(M))
        M, dM = masters[key]
        kt, tref, fdc = bb.params()
        ca, cr = bb.cpm()
        d = dict(S=dev.put(S), M=dM, kt=dev.put(kt), tref=dev.put(tref), fdc=dev.put(fdc),
                 out=dev.empty(bb.lines * bb.width * 8), coh=dev.empty(bb.lines * bb.width * 4))
        g = N.BurstGeom((C.c_int * 4)(*bb.box), (C.c_int * 4)(*bb.box), bb.y0, bb.width, bb.dt,
                        0.0, (C.c_double * 6)(*ca), (C.c_double * 6)(*cr))
        items[i] = N.BurstItem(g, d["S"].data_ptr(), d["kt"].data_ptr(), d["tref"].data_ptr(),
                               d["fdc"].data_ptr(), d["out"].data_ptr(), d["M"].data_ptr(),
                               d["coh"].data_ptr())
        keep.append(d)
        host.append(dict(S=S, M=M, kt=kt, tref=tref, fdc=fdc, ca=ca, cr=cr))
        if i % 8 == 7 or i == len(bursts) - 1:
            print(f"  {label}: {i + 1}/{len(bursts)} bursts generated and uploaded "
                  f"({time.perf_counter() - t0:.0f} s)", flush=True)
    return dict(ctx=ctx, items=items, dev=keep, host=host, kaz=kaz, krg=krg, bursts=bursts)


def _free(tops, dev, w):
    tops.native.load().tops_ctx_destroy(w["ctx"])
    w["dev"].clear()
    w["host"].clear()
    dev.torch.cuda.empty_cache()


def _outputs(dev, w, i):
    bb = w["bursts"][i]
    d = w["dev"][i]
    return (dev.get(d["out"], np.complex64, (bb.lines, bb.width)).copy(),
            dev.get(d["coh"], np.float32, (bb.lines, bb.width)).copy())


def _check_against_oracle(orc, dev, w, label):
    """Slabs of every burst against the oracle; returns the worst errors."""
    worst_r = worst_c = 0.0
    for i, (bb, h) in enumerate(zip(w["bursts"], w["host"])):
        got, coh = _outputs(dev, w, i)
        for r0, r1 in _slabs(bb.lines):
            ref = orc.resample(h["S"], bb.box, bb.box, bb.y0, bb.dt, 0.0, h["ca"], h["cr"],
                               w["kaz"], w["krg"], h["kt"], h["tref"], h["fdc"], r0, r1)
            e = rel_err(got[r0:r1], ref)
            # coherence of rows r0..r1-1 reads rows r0-1 .. r1 (3-row window)
            a, b = max(0, r0 - 1), min(bb.lines, r1 + 1)
            cref = orc.coherence(h["M"][a:b], got[a:b].astype(np.complex128), 4, 1, 10, 3)
            ec = float(np.max(np.abs(coh[r0:r1] - cref[r0 - a:r1 - a])))
            assert e <= 1e-5, (label, i, r0, e)
            assert ec <= 1e-4, (label, i, r0, ec)
            worst_r, worst_c = max(worst_r, e), max(worst_c, ec)
    return worst_r, worst_c


def _run_tuned_graph(tops, dev, w, label):
    """tune -> capture -> replay twice (outputs cleared before each replay,
    the program's own work and synchronisations in between); both replays
    must give identical outputs.  Returns the chosen schedule."""
    N = tops.native
    L = N.load()
    torch = dev.torch
    n = len(w["bursts"])
    plan = L.tops_plan_create(w["ctx"], w["items"], n, 3, 10, 0)
    assert plan, N.last_error()
    s = torch.cuda.Stream()
    sp = C.c_void_p(s.cuda_stream)
    best = C.c_int(0)
    t0 = time.perf_counter()
    N.check(L.tops_plan_tune(plan, sp, C.byref(best)), "tops_plan_tune")
    print(f"  {label}: tuned schedule {best.value} ({time.perf_counter() - t0:.1f} s); "
          f"graph capture", flush=True)
    N.check(L.tops_plan_capture_graph(plan, sp), "tops_plan_capture_graph")
    first = None
    for rep in range(2):
        for d in w["dev"]:
            d["out"].zero_()
            d["coh"].zero_()
        torch.cuda.synchronize()
        N.check(L.tops_plan_run(plan, sp), "tops_plan_run")
        s.synchronize()
        probe = [_outputs(dev, w, i) for i in (0, n // 2, n - 1)]
        if first is None:
            first = probe
        else:
            for (a0, a1), (b0, b1) in zip(first, probe):
                assert np.array_equal(a0, b0) and np.array_equal(a1, b1), label
    L.tops_plan_destroy(plan)
    return best.value


def _batched_matches(tops, dev, w, label, strip_rows):
    """The batched schedule at a fixed strip height gives the tuned plan's
    outputs up to FP32 rounding (other heights re-anchor the deramp
    recurrence at other rows; the fused schedule sums the coherence window
    azimuth-first, the batched one range-first)."""
    N = tops.native
    L = N.load()
    n = len(w["bursts"])
    tuned = [_outputs(dev, w, i) for i in (0, n // 2, n - 1)]
    plan = L.tops_plan_create(w["ctx"], w["items"], n, 3, 10, strip_rows)
    assert plan, N.last_error()
    N.check(L.tops_plan_run(plan, None), "tops_plan_run")
    dev.torch.cuda.synchronize()
    for (o, c), i in zip(tuned, (0, n // 2, n - 1)):
        o2, c2 = _outputs(dev, w, i)
        assert rel_err(o2, o) <= 1e-6, (label, i)
        assert float(np.max(np.abs(c2 - c))) <= 1e-5, (label, i)
    L.tops_plan_destroy(plan)


@pytest.mark.timeout(900)
def test_iw_pair_27_bursts(tops, orc, dev):
    bursts = tops.synth.iw_pair_bursts()
    assert len(bursts) == 27 and len({(b.lines, b.width) for b in bursts}) > 9
    assert {b.width for b in bursts} == {21000, 25000, 24500}
    assert all(1480 <= b.lines <= 1520 for b in bursts)
    w = _build(tops, dev, bursts, "iw_pair")
    try:
        best = _run_tuned_graph(tops, dev, w, "iw_pair")
        er, ec = _check_against_oracle(orc, dev, w, "iw_pair")
        print(f"  iw_pair: 27 bursts, schedule {best}: resample max|dz|/rms {er:.2e} "
              f"(tol 1e-5), coherence max|dc| {ec:.2e} (tol 1e-4)", flush=True)
        _batched_matches(tops, dev, w, "iw_pair", 256)
    finally:
        _free(tops, dev, w)


@pytest.mark.timeout(900)
def test_stack_48_bursts(tops, orc, dev):
    bursts = tops.synth.stack_bursts()
    assert len(bursts) == 48
    torch = dev.torch
    free, _ = torch.cuda.mem_get_info()
    per_burst = 1500 * 25000 * (4 + 8 + 4 + 3 * 8 / 1500)
    admitted = int((0.9 * free - 3 * 1500 * 25000 * 4) // per_burst)
    print(f"  stack: {admitted} slave bursts fit in 90% of the free HBM", flush=True)
    assert admitted >= 48
    w = _build(tops, dev, bursts, "stack")
    try:
        best = _run_tuned_graph(tops, dev, w, "stack")
        er, ec = _check_against_oracle(orc, dev, w, "stack")
        print(f"  stack: 48 bursts, schedule {best}: resample max|dz|/rms {er:.2e} "
              f"(tol 1e-5), coherence max|dc| {ec:.2e} (tol 1e-4)", flush=True)
    finally:
        _free(tops, dev, w)
