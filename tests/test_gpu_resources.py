"""Device resources over repeated calls: plans (with burst groups, so a
second stream and events), contexts, the ESD entry that allocates its tile
partials per call, and the reference's DerampDemodResample_ESD ->
CoherenceEst hand-over of a device buffer (EstCoh.cu:708, 900 free it) are
created and released many times; the device's free memory must come back
to within a small margin (hipMalloc'd memory outside torch's caching
allocator shows in hipMemGetInfo)."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

MARGIN = 64 << 20   # allocation granularity and runtime bookkeeping


def _free():
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info()[0]


@pytest.fixture(scope="module")
def small(tops, dev):
    N = tops.native
    L = N.load()
    pairs = [tops.synth.make_pair(f"rs{i}", 120 + 10 * i, 300 + 20 * i, seed=500 + i)
             for i in range(3)]
    m = pairs[0].master
    kaz, krg = N.lut_init(12, m.prf, m.azimuth_bandwidth, 2 * m.range_sampling_rate)
    bursts = []
    for p in pairs:
        kt, tref, fdc, _ = p.params()
        ca, cr = p.cpm(0)
        lines, pixels = p.master.lines_per_burst, p.master.samples_per_burst
        box = [0, pixels - 1, 0, lines - 1]
        g = N.BurstGeom((C.c_int * 4)(*box), (C.c_int * 4)(*box), 0, pixels,
                        p.master.az_time_interval, 0.0, (C.c_double * 6)(*ca),
                        (C.c_double * 6)(*cr))
        S, M = p.burst(0, "slave"), p.burst(0, "master")
        bursts.append(dict(g=g, L=lines, P=pixels, S=S, M=M, ca=ca, cr=cr, dt=p.master.az_time_interval,
                           kt=kt[0].copy(), tref=tref[0].copy(), fdc=fdc[0].copy(),
                           dS=dev.put(S), dM=dev.put(M), dkt=dev.put(kt[0].copy()),
                           dtref=dev.put(tref[0].copy()), dfdc=dev.put(fdc[0].copy()),
                           out=dev.empty(lines * pixels * 8), coh=dev.empty(lines * pixels * 4)))
    return dict(bursts=bursts, kaz=kaz, krg=krg)


def test_plans_and_contexts_release_what_they_take(tops, small):
    N = tops.native
    L = N.load()
    bs = small["bursts"]
    items = (N.BurstItem * len(bs))()
    for i, b in enumerate(bs):
        items[i] = N.BurstItem(b["g"], b["dS"].data_ptr(), b["dkt"].data_ptr(),
                               b["dtref"].data_ptr(), b["dfdc"].data_ptr(), b["out"].data_ptr(),
                               b["dM"].data_ptr(), b["coh"].data_ptr())
    stream = torch.cuda.Stream()
    sp = C.c_void_p(stream.cuda_stream)

    def cycle():
        ctx = L.tops_ctx_create(N.ptr(small["kaz"], C.c_float), N.ptr(small["krg"], C.c_float), 12)
        assert ctx, N.last_error()
        for rows in (0, 32, -32):
            plan = L.tops_plan_create(ctx, items, len(bs), 3, 10, rows)
            assert plan, N.last_error()
            if rows > 0:
                N.check(L.tops_plan_set_groups(plan, 3))
            N.check(L.tops_plan_capture_graph(plan, sp))
            N.check(L.tops_plan_run(plan, sp))
            stream.synchronize()
            L.tops_plan_destroy(plan)
        L.tops_ctx_destroy(ctx)

    for _ in range(3):   # first uses: the runtime's own pools grow, then stay
        cycle()
    before = _free()
    for _ in range(20):
        cycle()
    after = _free()
    print(f"  60 plans and 20 contexts: free memory {before >> 20} -> {after >> 20} MiB")
    assert before - after < MARGIN


def test_allocating_calls_release_what_they_take(tops, small):
    N = tops.native
    L = N.load()
    b = small["bursts"][0]
    n, P = 40, b["P"]
    rng = np.random.default_rng(9)
    mf = np.ascontiguousarray(rng.integers(-300, 300, (2 * n, P, 2)).astype(np.int16))
    sf = np.ascontiguousarray((rng.normal(size=(2 * n, P)) +
                               1j * rng.normal(size=(2 * n, P))).astype(np.complex64))
    d = [torch.from_numpy(a).cuda() for a in (mf, mf, sf, sf)]
    sizes = np.array([n, n], np.int32)
    shifts, sums = np.zeros(2), np.zeros(4)
    L_, P_ = b["L"], b["P"]
    box = (C.c_int * 4)(0, P_ - 1, 0, L_ - 1)
    out = np.zeros((L_, P_), np.complex64)
    coh = np.zeros((L_, P_), np.float32)
    S = np.ascontiguousarray(b["S"])
    M = np.ascontiguousarray(b["M"])

    def cycle():
        N.check(L.tops_esd_shifts(d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(),
                                  d[3].data_ptr(), sizes.ctypes.data_as(C.POINTER(C.c_int)), 2,
                                  P, 6, 24, 0.3, 1.0, 1.0, N.ptr(shifts), N.ptr(sums), None))
        L.tops_clear_error()
        dptr = L.DerampDemodResample_ESD(
            S.ctypes.data, N.ptr(b["ca"]), N.ptr(b["cr"]), 0.0, out.ctypes.data,
            N.ptr(small["kaz"], C.c_float), N.ptr(small["krg"], C.c_float), 0, P_, L_, box, box,
            L_, P_, b["dt"], N.ptr(b["kt"]), N.ptr(b["tref"]), N.ptr(b["fdc"]), 12)
        assert dptr and L.tops_last_error() == 0, N.last_error()
        L.CoherenceEst(0, 0, P_, L_, 4, 1, 10, 3, M.ctypes.data, dptr, N.ptr(coh, C.c_float))
        assert L.tops_last_error() == 0, N.last_error()

    cycle()
    before = _free()
    for _ in range(30):
        cycle()
    after = _free()
    print(f"  30 ESD calls and 30 ESD-resample + coherence hand-overs: free memory "
          f"{before >> 20} -> {after >> 20} MiB")
    assert before - after < MARGIN
