"""A subswath pair from product files, end to end: annotation XML, EOF and
CInt16 measurement TIFFs written for a 3-burst synthetic pair
(tests/s1_files.py), read back by SubswathRun.from_files (the library's
annotation / EOF / TIFF readers, orbit fit, deramp parameters, LUT), then
GeometricCoreg, the device-resident ESDAndCoh flow and the deburst, checked
against the oracle composed on the same parsed inputs.

Tolerances: offsets 1e-3 px, resampled samples 1e-5 of the RMS, coherence
1e-4, ESD shift 1e-4 px.
"""

import numpy as np
import pytest

import s1_files as S1
from conftest import offset_surface, oracle_geocoreg_burst, oracle_resample_burst, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def files(tops, tmp_path_factory):
    import tops_amd.pipeline as pl
    pair = tops.synth.subswath_iw1(n_bursts=3, lines=400, pixels=2048, delta=0.002)
    mem = pl.SubswathRun(pair)
    d = tmp_path_factory.mktemp("s1")
    paths = {}
    for key, sw, orbit, slc in (("m", pair.master, pair.morbit, mem.m_slc),
                                ("s", pair.slave, pair.sorbit, mem.s_slc)):
        guess = [pair.scene_center]

        def grid_ll(t, r, orbit=orbit, guess=guess):
            guess[0] = tops.synth.geolocate(orbit, t, r, guess[0])
            return guess[0]
        xml = d / f"s1a-iw1-slc-vv-{key}.xml"
        xml.write_text(S1.annotation_xml(sw, grid_ll=grid_ll))
        t0 = np.floor(sw.first_line_time / 10.0) * 10.0
        text, _ = S1.eof_xml(t0 - 120, t0 + 600, orbit.pos, orbit.vel)
        eof = d / f"S1A_OPER_AUX_POEORB_{key}.EOF"
        eof.write_text(text)
        tif = d / f"s1a-iw1-slc-vv-{key}.tiff"
        S1.write_tiff(tif, slc.view(S1.TIFF_TYPES[8][2]).reshape(slc.shape[:2]), 8,
                      rows_per_strip=3)
        paths[key] = (xml, eof, tif)
    print("  files: written; reading back", flush=True)
    run = pl.SubswathRun.from_files(paths["m"][0], paths["s"][0], paths["m"][1], paths["s"][1],
                                    paths["m"][2], paths["s"][2], mem.dem, mem.dem_geo)
    return dict(pair=pair, mem=mem, run=run, dir=d)


def test_files_metadata(files):
    pair, mem, run = files["pair"], files["mem"], files["run"]
    m = pair.master
    first = np.array([S1.seconds(S1.utc(t)) for t in m.burst_first_line_time])
    assert np.array_equal(run.m.burst_first_line_time, first)
    assert np.array_equal(run.m_slc, mem.m_slc) and np.array_equal(run.s_slc, mem.s_slc)
    assert (run.nb, run.lpb, run.spb) == (mem.nb, mem.lpb, mem.spb)
    # the EOF window: 10 vectors from 40 s before the acquisition's 10 s slot
    t = run.orbits["m"]["t"]
    assert len(t) == 10 and np.all(np.diff(t) == 10.0)
    assert t[0] <= m.first_line_time - 40 and t[-1] >= m.last_line_time + 30
    # the fitted orbit reproduces the true one at the burst times (cm)
    for key, orb in (("m", pair.morbit), ("s", pair.sorbit)):
        o = run.orbits[key]
        for tt in (m.first_line_time, m.last_line_time):
            tn = (tt - o["t"][5]) / 10.0
            p = np.array([np.polyval(o[c][::-1], tn) for c in ("cx", "cy", "cz")])
            assert np.linalg.norm(p - orb.pos(tt)) < 1e-2
    assert np.array_equal(run.kaz, mem.kaz) and np.array_equal(run.krg, mem.krg)


@pytest.fixture(scope="module")
def flow(files):
    import tops_amd.pipeline as pl
    run = files["run"]
    print("  files: geometric coreg", flush=True)
    ca, cr = run.geometric_coreg()
    ds = pl.DeviceSubswath(run, ca, cr, coh_win=(3, 10))
    print("  files: device ESD and coherence", flush=True)
    ds.enqueue()
    res, coh, shifts = (a.copy() for a in ds.results())
    yield dict(ds=ds, ca=ca, cr=cr, res=res, coh=coh, shifts=shifts)
    ds.close()


def test_files_pipeline_matches_oracle(orc, files, flow):
    run = files["run"]
    m = run.m
    lpb, spb, nb = run.lpb, run.spb, run.nb
    ca, cr = flow["ca"], flow["cr"]
    for b in range(nb):
        rc, oa, orr, used = oracle_geocoreg_burst(orc, run, b)
        assert rc == 0
        ga, gr = offset_surface(ca[b], cr[b])
        ra, rr = offset_surface(oa, orr)
        e = max(np.max(np.abs(ga - ra)), np.max(np.abs(gr - rr)))
        print(f"  burst {b}: offsets |d| = {e:.2e} px")
        assert e <= 1e-3
    r1 = np.concatenate([oracle_resample_burst(orc, run, b, ca[b], cr[b], 0.0,
                                               run.s.az_time_interval) for b in range(nb)])
    sizes, _, _ = orc.overlap_sizes(m.burst_first_line_time, m.burst_last_line_time, 0, nb - 1,
                                    m.az_time_interval)
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
        ref = r1[rows] if b == 0 else oracle_resample_burst(
            orc, run, b, ca[b], cr[b], float(flow["shifts"][b - 1]), m.az_time_interval)
        er = rel_err(flow["res"][rows], ref)
        cref = orc.coherence(M[rows], flow["res"][rows].astype(np.complex128), 4, 1, 10, 3)
        ec = float(np.max(np.abs(flow["coh"][rows] - cref)))
        print(f"  burst {b}: resample {er:.2e}, coherence {ec:.2e}")
        assert er <= 1e-5 and ec <= 1e-4


def test_files_stitched_coherence_to_tiff(tops, orc, dev, files, flow):
    """The coherence stitched on the device, written as a Float32 TIFF and
    read back, equals the oracle's stitch of the same coherence."""
    N = tops.native
    L = N.load()
    run, ds = files["run"], flow["ds"]
    y0, h = ds.deburst_extent()
    out = dev.empty(h * run.spb * 4)
    ds.deburst("coh", out)
    stitched = dev.get(out, np.float32, (h, run.spb))
    m = run.m
    ref, _ = orc.deburst(m.burst_first_line_time, m.burst_last_line_time,
                         m.burst_first_valid_line_time, m.burst_last_valid_line_time, run.lpb,
                         m.num_lines, m.az_time_interval, m.first_line_time, 0, run.nb - 1,
                         gapped=flow["coh"])
    assert np.array_equal(stitched, ref)
    path = str(files["dir"] / "coh_stitched.tif").encode()
    t = L.tops_tiff_create(path, N.GDT_FLOAT32, run.spb, h)
    assert t, N.last_error()
    N.check(L.tops_tiff_write(t, 0, 0, h, run.spb, N.GDT_FLOAT32, stitched.ctypes.data))
    N.check(L.tops_tiff_close(t))
    t = L.tops_tiff_open(path)
    back = np.empty_like(stitched)
    try:
        N.check(L.tops_tiff_read(t, 0, 0, h, run.spb, N.GDT_FLOAT32, back.ctypes.data))
    finally:
        L.tops_tiff_close(t)
    assert np.array_equal(back, stitched)
