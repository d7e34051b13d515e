"""ESDAndCoh (Func.cpp:2496-2763) through the reference's own entry points,
and the |shift| > 100 -> 0 reset (Func.cpp:2612-2621).

* The burst loop of Func.cpp:2646-2723 replayed with the C-ABI symbols it
  calls: EstESDShifts on the gathered overlap bands, then burst 0 through
  ResampleFirstBurst -> CoherenceEst_ESD (the pass-1 slave, no shift) and
  every later burst through DerampDemodResample_ESD (its overlap's shift) ->
  CoherenceEst, with the reference's 6 x 24 coherence window.
* A spectral separation a million times too small turns the 0.002-line
  misregistration into shifts of thousands of lines: the host stage
  (tops_esd_and_coh) and the device flow (tops_subswath_esd_coh) must reset
  them to 0 and resample pass 2 without a shift.

Tolerances: shifts 1e-4 px, resampled samples 1e-5 of the RMS, coherence
1e-4."""
import ctypes as C

import numpy as np
import pytest

from conftest import oracle_resample_burst, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(tops):
    import tops_amd.pipeline as pl
    pair = tops.synth.subswath_iw1(n_bursts=3, lines=400, pixels=2048, delta=0.002)
    run = pl.SubswathRun(pair)
    print("  esd flow: geometric coreg and pass 1", flush=True)
    ca, cr = run.geometric_coreg()
    res1 = run.deramp_and_resample(ca, cr)
    return dict(run=run, ca=ca, cr=cr, res1=res1)


def _bands(orc, run, res1):
    m = run.m
    lpb = run.lpb
    sizes, total, mx = orc.overlap_sizes(m.burst_first_line_time, m.burst_last_line_time, 0,
                                         run.nb - 1, m.az_time_interval)
    M = run.m_slc
    cut = lambda a, i, n, fwd: a[(i + 1) * lpb - n:(i + 1) * lpb] if fwd else \
        a[(i + 1) * lpb:(i + 1) * lpb + n]
    mf = np.concatenate([cut(M, i, n, True) for i, n in enumerate(sizes)])
    mb = np.concatenate([cut(M, i, n, False) for i, n in enumerate(sizes)])
    sf = np.concatenate([cut(res1, i, n, True) for i, n in enumerate(sizes)])
    sb = np.concatenate([cut(res1, i, n, False) for i, n in enumerate(sizes)])
    return np.ascontiguousarray(sizes, dtype=np.int32), mf, mb, sf, sb


def test_reference_burst_loop(tops, orc, case):
    N = tops.native
    L = N.load()
    run, ca, cr, res1 = case["run"], case["ca"], case["cr"], case["res1"]
    m, s = run.m, run.s
    lpb, spb, nb = run.lpb, run.spb, run.nb
    dt = m.az_time_interval
    sizes, mf, mb, sf, sb = _bands(orc, run, res1)
    n = len(sizes)
    df = orc.spectral_sep(run.orbits["m"]["kt"], lpb, dt)
    shifts = np.zeros(n)
    L.tops_clear_error()
    L.EstESDShifts(mf.ctypes.data, mb.ctypes.data, sf.ctypes.data, sb.ctypes.data,
                   sizes.ctypes.data_as(C.POINTER(C.c_int)), N.ptr(shifts), n, int(sizes.sum()),
                   spb, 6, 24, int(sizes.max()), 0.3, df, dt)
    assert L.tops_last_error() == 0, N.last_error()
    ref_sh, _ = orc.esd(mf, mb, sf.astype(np.complex128), sb.astype(np.complex128), sizes, spb,
                        6, 24, 0.3, df, dt)
    print(f"  EstESDShifts {shifts} oracle {ref_sh}")
    assert np.max(np.abs(shifts - ref_sh)) <= 1e-4
    shifts = np.where(np.abs(shifts) > 100, 0.0, shifts)   # the caller's clamp
    dx, dy, wrg, waz = 11, 2, 24, 6                          # Func.cpp:2583-2584, 2643-2644
    so = run.orbits["s"]
    for b in range(nb):
        M = np.ascontiguousarray(run.m_slc[b * lpb:(b + 1) * lpb])
        out = np.zeros((lpb, spb), np.complex64)
        coh = np.zeros((lpb, spb), np.float32)
        if b == 0:
            # burst 0: the pass-1 slave uploaded as it is, prefix-sum coherence
            out[:] = res1[:lpb]
            dptr = L.ResampleFirstBurst(out.ctypes.data, spb, lpb)
            assert dptr, N.last_error()
            L.CoherenceEst_ESD(0, 0, spb, lpb, dx, dy, wrg, waz, M.ctypes.data, dptr,
                               N.ptr(coh, C.c_float))
            ref = oracle_resample_burst(orc, run, 0, ca[0], cr[0], 0.0, s.az_time_interval)
        else:
            S = np.ascontiguousarray(run.s_slc[b * lpb:(b + 1) * lpb])
            mbox = (C.c_int * 4)(0, spb - 1, b * lpb, (b + 1) * lpb - 1)
            sbox = (C.c_int * 4)(0, spb - 1, b * lpb, (b + 1) * lpb - 1)
            kt, tref, fdc = (np.ascontiguousarray(so[k][b]) for k in ("kt", "tref", "fdc"))
            dptr = L.DerampDemodResample_ESD(
                S.ctypes.data, N.ptr(np.ascontiguousarray(ca[b])),
                N.ptr(np.ascontiguousarray(cr[b])), float(shifts[b - 1]), out.ctypes.data,
                N.ptr(run.kaz, C.c_float), N.ptr(run.krg, C.c_float), b, spb, lpb, mbox, sbox,
                lpb, spb, dt, N.ptr(kt), N.ptr(tref), N.ptr(fdc), 12)
            assert dptr, N.last_error()
            L.CoherenceEst(0, b * lpb, spb, lpb, dx, dy, wrg, waz, M.ctypes.data, dptr,
                           N.ptr(coh, C.c_float))
            ref = oracle_resample_burst(orc, run, b, ca[b], cr[b], float(shifts[b - 1]), dt)
        assert L.tops_last_error() == 0, N.last_error()
        er = rel_err(out, ref)
        cref = orc.coherence(M, out.astype(np.complex128), dx, dy, wrg, waz)
        ec = float(np.max(np.abs(coh - cref)))
        print(f"  burst {b}: resample {er:.2e}, coherence 6x24 {ec:.2e}")
        assert er <= 1e-5 and ec <= 1e-4


def _unclamped(orc, run, res1, df):
    sizes, mf, mb, sf, sb = _bands(orc, run, res1)
    sh, _ = orc.esd(mf, mb, sf.astype(np.complex128), sb.astype(np.complex128), sizes, run.spb,
                    6, 24, 0.3, df, run.m.az_time_interval)
    return sh


def _check_pass2_unshifted(orc, run, ca, cr, res, coh, win):
    lpb = run.lpb
    waz, wrg = win
    for b in range(run.nb):
        rows = slice(b * lpb, (b + 1) * lpb)
        dt = run.s.az_time_interval if b == 0 else run.m.az_time_interval
        ref = oracle_resample_burst(orc, run, b, ca[b], cr[b], 0.0, dt)
        er = rel_err(res[rows], ref)
        cref = orc.coherence(run.m_slc[rows], res[rows].astype(np.complex128), (wrg - 1) // 2,
                             (waz - 1) // 2, wrg, waz)
        ec = float(np.max(np.abs(coh[rows] - cref)))
        print(f"  burst {b}: resample (no shift) {er:.2e}, coherence {ec:.2e}")
        assert er <= 1e-5 and ec <= 1e-4


def test_host_stage_resets_large_shifts(orc, case):
    run, ca, cr, res1 = case["run"], case["ca"], case["cr"], case["res1"]
    kt = run.orbits["m"]["kt"]
    df = orc.spectral_sep(kt, run.lpb, run.m.az_time_interval)
    raw = _unclamped(orc, run, res1, df * 1e-6)
    print(f"  unclamped shifts with df x 1e-6: {raw}")
    assert np.all(np.abs(raw) > 100)
    # GetSpectralSep reads the master's Doppler rate: scaled, df is 1e-6 x
    run.orbits["m"]["kt"] = kt * 1e-6
    try:
        res, coh, shifts = run.esd_and_coh(ca, cr, res1, coh_win=(3, 10))
    finally:
        run.orbits["m"]["kt"] = kt
    print(f"  tops_esd_and_coh shifts: {shifts}")
    assert np.all(shifts == 0.0)
    _check_pass2_unshifted(orc, run, ca, cr, res, coh, (3, 10))


def test_device_flow_resets_large_shifts(orc, case):
    import tops_amd.pipeline as pl
    run, ca, cr, res1 = case["run"], case["ca"], case["cr"], case["res1"]
    ds = pl.DeviceSubswath(run, ca, cr, coh_win=(3, 10))
    try:
        df = ds.desc.spectral_sep
        raw = _unclamped(orc, run, res1, df * 1e-6)
        assert np.all(np.abs(raw) > 100)
        ds.desc.spectral_sep = df * 1e-6
        ds.d_shifts.fill_(123.0)
        ds.enqueue()
        res, coh, shifts = (a.copy() for a in ds.results())
    finally:
        ds.close()
    print(f"  tops_subswath_esd_coh shifts: {shifts} (unclamped {raw})")
    assert np.all(shifts == 0.0)
    _check_pass2_unshifted(orc, run, ca, cr, res, coh, (3, 10))
