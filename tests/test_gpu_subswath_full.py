"""Full-size subswath through the stage layer: 9 bursts of 1500 x 21000
CInt16 with ~100-line overlaps and a 0.002-line injected misregistration
(GeometricCoreg -> DerampAndResample -> ESDAndCoh, Main.cpp:107-121).

Against the FP64 oracle:
  * offsets of every burst (fitted slave line / pixel over the burst grid)
    within 1e-3 px;
  * ESD shifts within 1e-4 px, from the oracle's ESD over the overlap bands
    of the pass-1 output;
  * pass-1 and pass-2 resampled samples within 1e-5 of the RMS, and the 3x10
    coherence within 1e-4, on slabs of rows at the top, middle and bottom of
    every burst (the FP64 oracle over all 283 M samples would take minutes).
"""
import numpy as np
import pytest

from conftest import offset_surface, oracle_geocoreg_burst, oracle_resample_burst

pytestmark = pytest.mark.gpu

SLABS = ((0, 12), (744, 756), (1488, 1500))


@pytest.fixture(scope="module")
def full(tops):
    import tops_amd.pipeline as pl
    print("  full subswath: generating 9 x 1500 x 21000", flush=True)
    run = pl.SubswathRun(tops.synth.subswath_iw1())
    print("  full subswath: geometric coreg", flush=True)
    ca, cr = run.geometric_coreg()
    print("  full subswath: deramp and resample", flush=True)
    res1 = run.deramp_and_resample(ca, cr)
    print("  full subswath: ESD and coherence", flush=True)
    res2, coh, shifts = run.esd_and_coh(ca, cr, res1, coh_win=(3, 10))
    return dict(run=run, ca=ca, cr=cr, res1=res1, res2=res2, coh=coh, shifts=shifts)


def test_full_offsets(orc, full):
    run = full["run"]
    worst = 0.0
    for b in range(run.nb):
        rc, oa, orr, used = oracle_geocoreg_burst(orc, run, b)
        assert rc == 0
        ga, gr = offset_surface(full["ca"][b], full["cr"][b])
        ra, rr = offset_surface(oa, orr)
        e = max(np.max(np.abs(ga - ra)), np.max(np.abs(gr - rr)))
        worst = max(worst, e)
        print(f"  burst {b}: {used} posts, offset |d| = {e:.2e} px", flush=True)
    assert worst <= 1e-3


def test_full_esd(orc, full):
    run = full["run"]
    m, lpb, spb, nb = run.m, run.lpb, run.spb, run.nb
    sizes, _, _ = orc.overlap_sizes(m.burst_first_line_time, m.burst_last_line_time, 0, nb - 1,
                                    m.az_time_interval)
    M, r1 = run.m_slc, full["res1"]
    mf = np.concatenate([M[(i + 1) * lpb - n:(i + 1) * lpb] for i, n in enumerate(sizes)])
    mb = np.concatenate([M[(i + 1) * lpb:(i + 1) * lpb + n] for i, n in enumerate(sizes)])
    sf = np.concatenate([r1[(i + 1) * lpb - n:(i + 1) * lpb] for i, n in enumerate(sizes)])
    sb = np.concatenate([r1[(i + 1) * lpb:(i + 1) * lpb + n] for i, n in enumerate(sizes)])
    df = orc.spectral_sep(run.orbits["m"]["kt"], lpb, m.az_time_interval)
    ref, _ = orc.esd(mf, mb, sf.astype(np.complex128), sb.astype(np.complex128), sizes, spb, 6,
                     24, 0.3, df, m.az_time_interval)
    ref = np.where(np.abs(ref) > 100, 0.0, ref)
    e = float(np.max(np.abs(full["shifts"] - ref)))
    print(f"  {len(sizes)} overlaps {list(sizes)}: shifts {np.round(full['shifts'], 6)}, "
          f"|d| = {e:.2e} px", flush=True)
    assert e <= 1e-4


def _slab_errors(orc, full, b, which):
    run = full["run"]
    lpb = run.lpb
    if which == 1:
        out, shift, dt = full["res1"], 0.0, run.s.az_time_interval
    else:
        out, shift, dt = full["res2"], full["shifts"][b - 1], run.m.az_time_interval
    ers, ecs = [], []
    for r0, r1 in SLABS:
        ref = oracle_resample_burst(orc, run, b, full["ca"][b], full["cr"][b], shift, dt, r0, r1)
        got = out[b * lpb + r0:b * lpb + r1]
        rms = np.sqrt(np.mean(np.abs(ref) ** 2))
        ers.append(float(np.max(np.abs(got - ref)) / rms))
        if which == 2:
            # coherence of the inner rows of the slab from the GPU's own
            # resampled rows (3-line window: one row of context each side)
            lo, hi = max(r0 - 1, 0), min(r1 + 1, lpb)
            mm = run.m_slc[b * lpb + lo:b * lpb + hi]
            ss = out[b * lpb + lo:b * lpb + hi].astype(np.complex128)
            cref = orc.coherence(mm, ss, 4, 1, 10, 3)
            # rows r0..r1-1 see only rows lo..hi-1 (or the burst's zero
            # padding, which the slab shares)
            inner = slice(r0 - lo, r1 - lo)
            rows = slice(b * lpb + lo, b * lpb + hi)
            ecs.append(float(np.max(np.abs(full["coh"][rows][inner] - cref[inner]))))
    return max(ers), (max(ecs) if ecs else 0.0)


def test_full_pass1_resample(orc, full):
    run = full["run"]
    worst = 0.0
    for b in range(run.nb):
        er, _ = _slab_errors(orc, full, b, 1)
        worst = max(worst, er)
    print(f"  pass 1: max|dz|/rms over {run.nb} bursts x {len(SLABS)} slabs = {worst:.2e}")
    assert worst <= 1e-5


def test_full_pass2_resample_and_coherence(orc, full):
    run = full["run"]
    worst_r, worst_c = 0.0, 0.0
    for b in range(1, run.nb):
        er, ec = _slab_errors(orc, full, b, 2)
        worst_r, worst_c = max(worst_r, er), max(worst_c, ec)
        print(f"  burst {b}: pass-2 {er:.2e}, coherence {ec:.2e}", flush=True)
    assert np.array_equal(full["res2"][:run.lpb], full["res1"][:run.lpb])
    assert worst_r <= 1e-5 and worst_c <= 1e-4


def test_full_deburst(orc, full):
    """Stitch_OneSubswath of the pass-2 slave (CFloat32), the coherence
    (Float32) and the master (CInt16) of all 9 bursts, bit-exact against the
    oracle's stitch."""
    run = full["run"]
    m = run.m
    arrs = (m.burst_first_line_time, m.burst_last_line_time, m.burst_first_valid_line_time,
            m.burst_last_valid_line_time)
    for name, src in (("res2", full["res2"]), ("coh", full["coh"]), ("master", run.m_slc)):
        got = run.stitch(src)
        ref, _ = orc.deburst(*arrs, run.lpb, m.num_lines, m.az_time_interval, m.first_line_time,
                             0, run.nb - 1, gapped=src)
        print(f"  stitch {name}: {got.shape}", flush=True)
        assert np.array_equal(got, ref)
        del got, ref
