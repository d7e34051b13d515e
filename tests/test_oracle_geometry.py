"""Pins of the oracle's zero-Doppler projection and coregistration fit
(Geocoreg_cuda.cu:211-323, 548-605, 812-855), CPU only.

  * forward geometry: a target built at zero Doppler of the true orbit at
    time t0 and slant range R must project to line (t0 - t_first)/dt and
    pixel (R - tau0 c)/dr (polyfit error of the orbit < 1e-3 m);
  * exact offsets: a slave that is the master delayed by k lines and shifted
    by m pixels must fit Cpm = identity + (k, m) to 1e-6 px;
  * the normal-equation Cholesky solve equals an independent numpy lstsq on
    the projected posts;
  * an all-nodata DEM gives the singular-fit status, not garbage.
"""
import numpy as np
import pytest

SOL = 299792458.0


@pytest.fixture(scope="module")
def scene(tops, orc):
    pair = tops.synth.single_burst_small()
    m = pair.master
    t, x, y, z = tops.synth.orbit_polyfit_inputs(pair.morbit)
    coefs = orc.polyfit_orbit(t, x, y, z)
    return pair, m, t, coefs


def _geom(orc, coefs, t, m, b=0, slr_shift=0.0, t_shift=0.0):
    first = m.burst_first_line_time[b] + t_shift
    last = m.burst_last_line_time[b] + t_shift
    return orc.orbit_geom(coefs, t[len(t) // 2], m.az_time_interval, first,
                          m.slr_time_first_pixel + slr_shift, m.range_spacing,
                          0.5 * (first + last), b * m.lines_per_burst)


def test_projection_forward_geometry(tops, orc, scene):
    pair, m, t, coefs = scene
    g = _geom(orc, coefs, t, m)
    rng = np.random.default_rng(11)
    for _ in range(6):
        t0 = m.burst_first_line_time[0] + rng.uniform(0.1, 0.9) * (m.lines_per_burst - 1) * m.az_time_interval
        r = (m.slr_time_first_pixel + rng.uniform(0.1, 0.9) * 2047 * m.range_spacing / SOL) * SOL
        lat, lon = tops.synth.geolocate(pair.morbit, t0, r, pair.scene_center)
        dem = np.zeros((1, 1), np.int32)
        line, pix = orc.ell2lp(dem, lat, lon, 0.0, 0.0, -32768, g)
        assert abs(line[0, 0] - (t0 - m.burst_first_line_time[0]) / m.az_time_interval) < 1e-3
        assert abs(pix[0, 0] - (r - m.slr_time_first_pixel * SOL) / m.range_spacing) < 1e-3


def test_projection_nodata(orc, scene):
    pair, m, t, coefs = scene
    g = _geom(orc, coefs, t, m)
    dem = np.full((2, 3), -32768, np.int32)
    line, pix = orc.ell2lp(dem, 30.0, 100.0, 1e-3, 1e-3, -32768, g)
    assert np.all(line == -99999999) and np.all(pix == -99999999)


def test_geocoreg_exact_offsets(orc, scene):
    """slave = master orbit, burst delayed by k lines, near range earlier by
    m pixels: slave line = master line - k, slave pixel = master pixel + m."""
    pair, m, t, coefs = scene
    k, mpx = 3.0, 2.0
    gm = _geom(orc, coefs, t, m)
    gs = _geom(orc, coefs, t, m, slr_shift=-mpx * m.range_spacing / SOL,
               t_shift=k * m.az_time_interval)
    ll = m.burst_ll[0]
    dem, lat0, lon0, dlat, dlon = pair.dem_window(*ll)
    L, P = m.lines_per_burst, m.samples_per_burst
    box = [0, P - 1, 0, L - 1]
    rc, ca, cr, used = orc.geocoreg(dem, lat0, lon0, dlat, dlon, -32768, gm, gs, box)
    assert rc == 0 and used > 1000
    mid_l, q_l = 0.5 * (L - 1), 0.25 * (L - 1)
    mid_p, q_p = 0.5 * (P - 1), 0.25 * (P - 1)
    assert np.allclose(ca, [mid_l - k, q_l, 0, 0, 0, 0], atol=1e-6)
    assert np.allclose(cr, [mid_p + mpx, 0, q_p, 0, 0, 0], atol=1e-6)


def test_geocoreg_matches_lstsq(tops, orc, scene):
    pair, m, t, coefs = scene
    ts, xs, ys, zs = tops.synth.orbit_polyfit_inputs(pair.sorbit)
    scoefs = orc.polyfit_orbit(ts, xs, ys, zs)
    gm = _geom(orc, coefs, t, m)
    sw = pair.slave
    gs = orc.orbit_geom(scoefs, ts[len(ts) // 2], sw.az_time_interval,
                        sw.burst_first_line_time[0], sw.slr_time_first_pixel, sw.range_spacing,
                        0.5 * (sw.burst_first_line_time[0] + sw.burst_last_line_time[0]), 0)
    dem, lat0, lon0, dlat, dlon = pair.dem_window(*m.burst_ll[0], nodata_every=97)
    L, P = m.lines_per_burst, m.samples_per_burst
    box = [0, P - 1, 0, L - 1]
    rc, ca, cr, used = orc.geocoreg(dem, lat0, lon0, dlat, dlon, -32768, gm, gs, box)
    assert rc == 0
    ml, mp = orc.ell2lp(dem, lat0, lon0, dlat, dlon, -32768, gm)
    sl, sp = orc.ell2lp(dem, lat0, lon0, dlat, dlon, -32768, gs)
    w = (dem != -32768) & (ml >= 0) & (ml <= L - 1) & (mp >= 0) & (mp <= P - 1)
    assert w.sum() == used
    ln = (ml[w] - 0.5 * (L - 1)) / (0.25 * (L - 1))
    pn = (mp[w] - 0.5 * (P - 1)) / (0.25 * (P - 1))
    a = np.stack([np.ones_like(ln), ln, pn, ln * ln, ln * pn, pn * pn], axis=1)
    xa = np.linalg.lstsq(a, sl[w], rcond=None)[0]
    xr = np.linalg.lstsq(a, sp[w], rcond=None)[0]
    # compare the fitted offset surfaces over the burst grid
    gl, gp = np.meshgrid(np.linspace(-2, 2, 9), np.linspace(-2, 2, 9))
    A = np.stack([np.ones_like(gl), gl, gp, gl * gl, gl * gp, gp * gp], axis=-1)
    assert np.max(np.abs(A @ ca - A @ xa)) < 1e-6
    assert np.max(np.abs(A @ cr - A @ xr)) < 1e-6


def test_geocoreg_all_nodata(orc, scene):
    pair, m, t, coefs = scene
    g = _geom(orc, coefs, t, m)
    dem = np.full((20, 20), -32768, np.int32)
    rc, ca, cr, used = orc.geocoreg(dem, 30.0, 100.0, 1e-3, 1e-3, -32768, g, g, [0, 10, 0, 10])
    assert rc == 1 and used == 0
    assert np.all(ca == 0) and np.all(cr == 0)


def test_deramp_params_closed_form(orc, tops):
    """constant FM rate and linear Doppler centroid: tref is linear in range
    sample, kt constant = ka krot/(ka - krot)."""
    pair = tops.synth.single_burst_small()
    m = pair.master
    kt, tref, fdc, ka = orc.deramp_params(1, m.samples_per_burst, m.slr_time_first_pixel,
                                          m.range_spacing, m.lines_per_burst,
                                          m.az_time_interval, 0, m.dc, m.fm,
                                          m.steering_rate_deg, m.wavelength, 7500.0)
    krot = 2 * 7500.0 * np.radians(m.steering_rate_deg) / m.wavelength
    assert np.allclose(ka, -2300.0)
    assert np.allclose(kt, -2300.0 * krot / (-2300.0 - krot), rtol=1e-13)
    x = np.arange(m.samples_per_burst)
    assert np.allclose(fdc[0], 20.0 + 5e-3 * x, atol=1e-9)
    half = m.lines_per_burst * m.az_time_interval / 2
    assert np.allclose(tref[0], half + 20.0 / -2300.0 - fdc[0] / -2300.0, atol=1e-12)


def test_overlap_and_spectral_sep(orc):
    dt = 2e-3
    first = 100.0 + np.arange(4) * 1400 * dt - np.arange(4) * 1e-9
    last = first + 1499 * dt
    sizes, total, mx = orc.overlap_sizes(first, last, 0, 3, dt)
    assert list(sizes) == [99, 99, 99] and total == 297 and mx == 99
    kt = np.full((3, 50), 1700.0)
    assert abs(orc.spectral_sep(kt, 1500, dt) - 1700.0 * 1500 * dt) < 1e-9
