"""Concurrent calls (the C-ABI's reentrancy: per-call parameter blocks
instead of the reference's module-global __constant__ symbols and texture
references, cuConstants.cuh:10-50, and a status per host thread):

  * the single-burst resample and coherence of two bursts enqueued on two
    streams in turn, with no synchronisation between them, give the same
    bits as the same calls run one after the other;
  * two host threads, each running its own plan on its own stream over a
    shared context, get the same bits as sequential runs, while one of them
    keeps failing calls: each thread's status reports its own calls only.
"""
import ctypes as C
import threading

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair_set(tops, dev):
    """Four bursts (lines, pixels) with their own offsets and Doppler, one
    context for all."""
    N = tops.native
    L = N.load()
    shapes = [(300, 900), (280, 700), (260, 820), (310, 640)]
    bursts, kaz, krg = [], None, None
    for i, (lines, pixels) in enumerate(shapes):
        pair = tops.synth.make_pair(f"cc{i}", lines, pixels, seed=400 + i,
                                    dc_slope_total=15.0 * i)
        m = pair.master
        if kaz is None:
            kaz, krg = N.lut_init(12, m.prf, m.azimuth_bandwidth, 2 * m.range_sampling_rate)
        kt, tref, fdc, _ = pair.params()
        ca, cr = pair.cpm(0)
        box = [0, pixels - 1, 0, lines - 1]
        g = N.BurstGeom((C.c_int * 4)(*box), (C.c_int * 4)(*box), 0, pixels,
                        m.az_time_interval, 0.0, (C.c_double * 6)(*ca), (C.c_double * 6)(*cr))
        bursts.append(dict(L=lines, P=pixels, g=g, S=dev.put(pair.burst(0, "slave")),
                           M=dev.put(pair.burst(0, "master")), kt=dev.put(kt[0].copy()),
                           tref=dev.put(tref[0].copy()), fdc=dev.put(fdc[0].copy())))
    ctx = L.tops_ctx_create(N.ptr(kaz, C.c_float), N.ptr(krg, C.c_float), 12)
    yield dict(bursts=bursts, ctx=ctx)
    L.tops_ctx_destroy(ctx)


def _enqueue_burst(L, N, ctx, b, out, coh, stream):
    N.check(L.tops_resample_burst(ctx, C.byref(b["g"]), b["S"].data_ptr(), b["kt"].data_ptr(),
                                  b["tref"].data_ptr(), b["fdc"].data_ptr(), out.data_ptr(), 0,
                                  stream))
    N.check(L.tops_coherence(b["M"].data_ptr(), out.data_ptr(), b["P"], b["L"], 4, 1, 10, 3,
                             coh.data_ptr(), 0, stream))


def test_two_streams_interleaved(tops, dev, pair_set):
    N = tops.native
    L = N.load()
    ctx, bs = pair_set["ctx"], pair_set["bursts"][:2]
    refs = []
    for b in bs:
        out, coh = dev.empty(b["L"] * b["P"] * 8), dev.empty(b["L"] * b["P"] * 4)
        _enqueue_burst(L, N, ctx, b, out, coh, None)
        torch.cuda.synchronize()
        refs.append((out.cpu().numpy().copy(), coh.cpu().numpy().copy()))
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    bufs = [(dev.empty(b["L"] * b["P"] * 8), dev.empty(b["L"] * b["P"] * 4)) for b in bs]
    for rep in range(4):
        for k, b in enumerate(bs):
            _enqueue_burst(L, N, ctx, b, bufs[k][0], bufs[k][1],
                           C.c_void_p(streams[k].cuda_stream))
    for s in streams:
        s.synchronize()
    for k in range(2):
        same = (np.array_equal(bufs[k][0].cpu().numpy(), refs[k][0]) and
                np.array_equal(bufs[k][1].cpu().numpy(), refs[k][1]))
        print(f"  stream {k}: outputs equal to the sequential run: {same}")
        assert same


def test_threads_plans_and_status(tops, dev, pair_set):
    N = tops.native
    L = N.load()
    ctx, bs = pair_set["ctx"], pair_set["bursts"]

    def make_plan(group):
        items = (N.BurstItem * len(group))()
        outs = []
        for i, b in enumerate(group):
            out, coh = dev.empty(b["L"] * b["P"] * 8), dev.empty(b["L"] * b["P"] * 4)
            items[i] = N.BurstItem(b["g"], b["S"].data_ptr(), b["kt"].data_ptr(),
                                   b["tref"].data_ptr(), b["fdc"].data_ptr(), out.data_ptr(),
                                   b["M"].data_ptr(), coh.data_ptr())
            outs.append((out, coh))
        plan = L.tops_plan_create(ctx, items, len(group), 3, 10, 64)
        assert plan, N.last_error()
        return plan, outs

    groups = [bs[:2], bs[2:]]
    plans = [make_plan(g) for g in groups]
    # sequential reference runs
    refs = []
    for plan, outs in plans:
        N.check(L.tops_plan_run(plan, None))
        torch.cuda.synchronize()
        refs.append([(o.cpu().numpy().copy(), c.cpu().numpy().copy()) for o, c in outs])
        for o, c in outs:
            o.zero_()
            c.zero_()
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    errors = []

    def worker(k):
        try:
            plan, _ = plans[k]
            sp = C.c_void_p(streams[k].cuda_stream)
            for rep in range(6):
                rc = L.tops_plan_run(plan, sp)
                assert rc == 0, (k, rep, rc)
                if k == 0:
                    # a failing call in this thread sets this thread's status
                    assert not L.tops_plan_create(None, None, 0, 3, 10, 0)
                    assert L.tops_last_error() == N.TOPS_ERR_ARG
                    L.tops_clear_error()
                else:
                    assert L.tops_last_error() == N.TOPS_OK
        except Exception as e:   # reported by the main thread
            errors.append((k, repr(e)))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for s in streams:
        s.synchronize()
    assert not errors, errors
    for k, (plan, outs) in enumerate(plans):
        for j, (o, c) in enumerate(outs):
            same = (np.array_equal(o.cpu().numpy(), refs[k][j][0]) and
                    np.array_equal(c.cpu().numpy(), refs[k][j][1]))
            print(f"  thread {k} burst {j}: outputs equal to the sequential run: {same}")
            assert same
        L.tops_plan_destroy(plan)
