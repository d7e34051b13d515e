"""GPU parity of the device-resident ESDAndCoh flow (tops_subswath_esd_coh:
pass-1 resample, ESD over the overlap bands read in place, shifts formed on
the device, pass-2 resample, coherence; Func.cpp:2496-2763) on a 3-burst
subswath, against the oracle and against the host-memory stage
(tops_esd_and_coh) on the same inputs, eager and replayed from a hipGraph.

Tolerances: ESD shift 1e-4 px, resampled samples 1e-5 of the RMS, coherence
1e-4.
"""
import numpy as np
import pytest

from conftest import oracle_resample_burst, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def run(tops):
    import tops_amd.pipeline as pl
    pair = tops.synth.subswath_iw1(n_bursts=3, lines=400, pixels=2048, delta=0.002)
    return pl.SubswathRun(pair)


@pytest.fixture(scope="module")
def flow(run):
    import tops_amd.pipeline as pl
    print("  subswath: geometric coreg", flush=True)
    ca, cr = run.geometric_coreg()
    ds = pl.DeviceSubswath(run, ca, cr, coh_win=(3, 10))
    print("  subswath: device ESD and coherence", flush=True)
    ds.enqueue()
    res, coh, shifts = (a.copy() for a in ds.results())
    yield dict(ds=ds, ca=ca, cr=cr, res=res, coh=coh, shifts=shifts)
    ds.close()


def test_subswath_device_matches_oracle(orc, run, flow):
    m = run.m
    lpb, spb, nb = run.lpb, run.spb, run.nb
    ca, cr = flow["ca"], flow["cr"]
    # pass 1 by the oracle (slave azimuth interval, no shift), for the ESD
    # bands and for burst 0
    r1 = np.concatenate([oracle_resample_burst(orc, run, b, ca[b], cr[b], 0.0,
                                               run.s.az_time_interval) for b in range(nb)])
    sizes, total, mx = orc.overlap_sizes(m.burst_first_line_time, m.burst_last_line_time, 0,
                                         nb - 1, m.az_time_interval)
    M = run.m_slc
    mf = np.concatenate([M[(i + 1) * lpb - n:(i + 1) * lpb] for i, n in enumerate(sizes)])
    mb = np.concatenate([M[(i + 1) * lpb:(i + 1) * lpb + n] for i, n in enumerate(sizes)])
    sf = np.concatenate([r1[(i + 1) * lpb - n:(i + 1) * lpb] for i, n in enumerate(sizes)])
    sb = np.concatenate([r1[(i + 1) * lpb:(i + 1) * lpb + n] for i, n in enumerate(sizes)])
    df = orc.spectral_sep(run.orbits["m"]["kt"], lpb, m.az_time_interval)
    shifts, _ = orc.esd(mf, mb, sf, sb, sizes, spb, 6, 24, 0.3, df, m.az_time_interval)
    shifts = np.where(np.abs(shifts) > 100, 0.0, shifts)
    e = float(np.max(np.abs(flow["shifts"] - shifts)))
    print(f"  shifts gpu={flow['shifts']} oracle={shifts} |d|={e:.2e}")
    assert e <= 1e-4
    for b in range(nb):
        rows = slice(b * lpb, (b + 1) * lpb)
        if b == 0:
            ref = r1[rows]
        else:
            # with the shift the device applied (checked above): a 1e-9-line
            # change of the shift can move an azimuth position across a
            # LUT-row boundary, a step of the resample itself
            ref = oracle_resample_burst(orc, run, b, ca[b], cr[b], float(flow["shifts"][b - 1]),
                                        m.az_time_interval)
        er = rel_err(flow["res"][rows], ref)
        cref = orc.coherence(M[rows], flow["res"][rows].astype(np.complex128), 4, 1, 10, 3)
        ec = float(np.max(np.abs(flow["coh"][rows] - cref)))
        print(f"  burst {b}: resample {er:.2e}, coherence {ec:.2e}")
        assert er <= 1e-5 and ec <= 1e-4


def test_subswath_device_matches_host_stages(run, flow):
    """The host-memory stages (tops_deramp_and_resample then
    tops_esd_and_coh) give the same shifts and the same samples: both run
    the same kernels; the host stage uploads the gathered bands."""
    res1 = run.deramp_and_resample(flow["ca"], flow["cr"])
    res2, coh, shifts = run.esd_and_coh(flow["ca"], flow["cr"], res1, coh_win=(3, 10))
    e = float(np.max(np.abs(shifts - flow["shifts"])))
    print(f"  shifts host={shifts} device={flow['shifts']} |d|={e:.2e}")
    assert e <= 1e-9
    er = rel_err(flow["res"], res2)
    ec = float(np.max(np.abs(flow["coh"] - coh)))
    print(f"  resample {er:.2e}, coherence {ec:.2e}")
    assert er <= 1e-6 and ec <= 1e-6


def test_subswath_device_graph_replay(dev, flow):
    torch = dev.torch
    ds = flow["ds"]
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    print("  subswath: graph capture", flush=True)
    with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
        ds.enqueue(s)
    for rep in range(3):
        ds.d_res.zero_()
        ds.d_coh.zero_()
        ds.d_shifts.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        res, coh, shifts = ds.results()
        assert np.array_equal(shifts, flow["shifts"]), rep
        assert np.array_equal(res, flow["res"]), rep
        assert np.array_equal(coh, flow["coh"]), rep
    del g


def test_subswath_device_deburst(orc, dev, run, flow):
    """The resident coherence, resampled slave and master stitched on the
    device (tops_deburst) equal the oracle's stitch of the same rasters."""
    ds = flow["ds"]
    y0, h = ds.deburst_extent()
    m = run.m
    arrs = (m.burst_first_line_time, m.burst_last_line_time, m.burst_first_valid_line_time,
            m.burst_last_valid_line_time)
    for which, src, eb in (("coh", flow["coh"], 4), ("res", flow["res"], 8),
                           ("master", run.m_slc, 4)):
        out = dev.empty(h * run.spb * eb)
        ds.deburst(which, out)
        got = dev.get(out, src.dtype, (h,) + src.shape[1:])
        ref, _ = orc.deburst(*arrs, run.lpb, m.num_lines, m.az_time_interval, m.first_line_time,
                             0, run.nb - 1, gapped=src)
        print(f"  deburst {which}: {h} rows from line {y0}", flush=True)
        assert np.array_equal(got, ref)
