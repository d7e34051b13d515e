"""Pins of the oracle's deramp/resample/reramp (DerampAndResample.cu:108-183,
348-520), coherence (EstCoh.cu:101-139, 413-456) and ESD (EstESD.cu:62-161,
243-546, 1073-1217), CPU only.

Resample:
  * identity geometry, no TOPS ramp -> output == slave (LUT row 0 is a delta);
  * integer offsets -> exact shift;
  * identity geometry WITH the ramp -> output == slave: deramp and reramp use
    the same phase, so a wrong sign, dt or tap index in either breaks it;
  * a half-pixel range shift of a band-limited ramp lands at +0.5 px;
  * the reramp phase equals the FP64 moment identity
    sum_j kL_j phi(y0+j-5) = M0 a + (y0 M0 + M1) b + (y0^2 M0 + 2 y0 M1 + M2) c
    (phi quadratic in line) evaluated independently in numpy;
  * zero / edge-band / interior regions follow the reference's box rules
    (brute force classification of every output of a shifted burst).
Coherence: M = S and S = M exp(j th) give 1, |S| < 1e-3 gives 0, and a
brute-force python window sum agrees on a small random case.
ESD: a constructed double-difference phase th is recovered as
th / (2 pi df dt); incoherent bands are masked to a zero sum.
"""
import numpy as np

from conftest import zero_master_band
import pytest

DT = 2.055556e-3


@pytest.fixture(scope="module")
def lut(orc):
    return orc.lut(12, 1 / DT, 327.0, 2 * 64.34523076923077e6)


def _identity_cpm(L, P, y0=0):
    """slave = master position (+1e-9 so that rounding in the normalised
    coordinates cannot truncate an exact integer position one sample low)"""
    return (np.array([y0 + 0.5 * (L - 1) + 1e-9, 0.25 * (L - 1), 0, 0, 0, 0.0]),
            np.array([0.5 * (P - 1) + 1e-9, 0, 0.25 * (P - 1), 0, 0, 0.0]))


def _rand_slave(L, P, seed=1):
    rng = np.random.default_rng(seed)
    return rng.integers(-3000, 3000, size=(L, P, 2)).astype(np.int16)


def _ramp(P, kt=1760.0, tref=0.2, fdc0=20.0, slope=5e-3):
    x = np.arange(P)
    return np.full(P, kt), np.full(P, tref) + 1e-6 * x, fdc0 + slope * x


def _c(s):
    return s[..., 0].astype(np.float64) + 1j * s[..., 1]


def test_identity_no_ramp(orc, lut):
    L, P = 40, 64
    s = _rand_slave(L, P)
    ca, cr = _identity_cpm(L, P)
    z = np.zeros(P)
    out = orc.resample(s, [0, P - 1, 0, L - 1], [0, P - 1, 0, L - 1], 0, DT, 0.0, ca, cr,
                       *lut, z, z, z)
    assert np.max(np.abs(out - _c(s))) < 1e-9


def test_integer_shift(orc, lut):
    L, P = 48, 80
    s = _rand_slave(L, P, 2)
    ca, cr = _identity_cpm(L, P)
    ca[0] += 3
    cr[0] += 2
    z = np.zeros(P)
    out = orc.resample(s, [0, P - 1, 0, L - 1], [0, P - 1, 0, L - 1], 0, DT, 0.0, ca, cr,
                       *lut, z, z, z)
    ref = _c(s)
    assert np.max(np.abs(out[:L - 3, :P - 2] - ref[3:, 2:])) < 1e-9
    assert np.all(out[L - 2:, :] == 0)        # slave line >= L: outside the box


def test_identity_with_ramp(orc, lut):
    L, P = 40, 64
    s = _rand_slave(L, P, 3)
    ca, cr = _identity_cpm(L, P, y0=200)
    kt, tref, fdc = _ramp(P)
    out = orc.resample(s, [0, P - 1, 200, 200 + L - 1], [0, P - 1, 200, 200 + L - 1], 180, DT,
                       0.0, ca, cr, *lut, kt, tref, fdc)
    assert np.max(np.abs(out - _c(s))) < 1e-8


def test_half_pixel_shift_direction(orc, lut):
    """slave = exp(j w x) (no ramp): resampled at x + 0.5 -> exp(j w (x+0.5))"""
    L, P = 24, 96
    w = 0.02 * np.pi
    x = np.arange(P)
    amp = 3000.0
    sig = amp * np.exp(1j * w * x)[None, :].repeat(L, 0)
    s = np.stack([np.round(sig.real), np.round(sig.imag)], -1).astype(np.int16)
    ca, cr = _identity_cpm(L, P)
    cr[0] += 0.5
    z = np.zeros(P)
    out = orc.resample(s, [0, P - 1, 0, L - 1], [0, P - 1, 0, L - 1], 0, DT, 0.0, ca, cr,
                       *lut, z, z, z)
    inner = out[6:-7, 6:-8]
    ref = amp * np.exp(1j * w * (x[6:-8] + 0.5))[None, :]
    assert np.max(np.abs(inner - ref)) / amp < 3e-3
    # the same shift the other way is far off: the sign is pinned
    bad = amp * np.exp(1j * w * (x[6:-8] - 0.5))[None, :]
    assert np.max(np.abs(inner - bad)) / amp > 2e-2


def test_reramp_phase_moment_identity(orc, lut):
    """constant deramped scene C: slave = C exp(-j phi); the output phase is
    -Phi with Phi the LUT-weighted tap sum of phi, which the quadratic
    moment identity evaluates independently."""
    L, P = 30, 60
    y0 = 600
    kaz, krg = lut
    kt, tref, fdc = _ramp(P)
    yy = np.arange(L)[:, None] + y0 - 500         # burst-relative line
    ta = yy * DT
    phi = -np.pi * kt[None, :] * (ta - tref[None, :]) ** 2 - 2 * np.pi * fdc[None, :] * ta
    sig = 2000.0 * np.exp(-1j * phi)
    s = np.stack([np.round(sig.real), np.round(sig.imag)], -1).astype(np.int16)
    ca, cr = _identity_cpm(L, P, y0=y0)
    ca[0] += 0.3137
    cr[0] += 0.4412
    out = orc.resample(s, [0, P - 1, y0, y0 + L - 1], [0, P - 1, y0, y0 + L - 1], 500, DT, 0.0,
                       ca, cr, kaz, krg, kt, tref, fdc)
    a = -np.pi * kt * tref ** 2
    b = 2 * np.pi * DT * (kt * tref - fdc)
    c = -np.pi * kt * DT ** 2
    u = np.arange(12) - 5
    for r in (8, 15, 20):
        for col in (10, 31, 44):
            ys = y0 + r + 0.3137
            xs = col + 0.4412
            iy, ix = int(ys), int(xs)
            fy = int((ys - iy) * 2047 + 0.5)
            fx = int((xs - ix) * 2047 + 0.5)
            kl = kaz[fy].astype(np.float64)
            kp = krg[fx].astype(np.float64)
            m0, m1, m2 = kl.sum(), (kl * u).sum(), (kl * u * u).sum()
            ycen = iy - 500
            cols = ix - 5 + np.arange(12)
            Phi = (kp * (m0 * a[cols] + (ycen * m0 + m1) * b[cols] +
                         (ycen * ycen * m0 + 2 * ycen * m1 + m2) * c[cols])).sum()
            # deramped scene is 2000 up to the int16 rounding of the slave
            # (+-0.5 per component): phase within ~3e-4 rad, amplitude 1e-3
            ref = 2000.0 * kl.sum() * kp.sum() * np.exp(-1j * Phi)
            assert abs(np.angle(out[r, col] / ref)) < 5e-4, (r, col, out[r, col], ref)
            assert abs(abs(out[r, col]) / abs(ref) - 1) < 1e-3


def test_region_rules_brute_force(orc, lut):
    L, P = 40, 50
    s = _rand_slave(L, P, 4)
    ca, cr = _identity_cpm(L, P)
    ca[0] += 2.6        # push the bottom rows into the edge band and beyond
    cr[0] -= 3.3
    z = np.zeros(P)
    out = orc.resample(s, [0, P - 1, 0, L - 1], [0, P - 1, 0, L - 1], 0, DT, 0.0, ca, cr,
                       *lut, z, z, z)
    sc = _c(s)
    n_zero = n_edge = 0
    for r in range(L):
        for col in range(P):
            ys, xs = r + 2.6, col - 3.3
            iy, ix = int(np.trunc(ys)), int(np.trunc(xs))
            if iy > L - 1 or iy < 0 or ix > P - 1 or ix < 0:
                assert out[r, col] == 0
                n_zero += 1
            elif iy > L - 1 - 7 or iy < 5 or ix > P - 1 - 7 or ix < 5:
                assert abs(out[r, col] - sc[iy, ix]) < 1e-9
                n_edge += 1
    assert n_zero > 50 and n_edge > 50


def test_float_phase_rounding_is_small(orc, lut):
    """the reference stores phi as float (DerampAndResample.cu:163-164); the
    oracle keeps FP64.  The difference is bounded by float spacing of phi."""
    L, P = 30, 40
    s = _rand_slave(L, P, 5)
    ca, cr = _identity_cpm(L, P, y0=1400)
    ca[0] += 0.25
    kt, tref, fdc = _ramp(P)
    args = ([0, P - 1, 1400, 1400 + L - 1], [0, P - 1, 1400, 1400 + L - 1], 0, DT, 0.0, ca, cr,
            *lut, kt, tref, fdc)
    a = orc.resample(s, *args)
    b = orc.resample(s, *args, ref_float_phase=True)
    nz = np.abs(a) > 0
    d = np.abs(np.angle(a[nz] / b[nz]))
    assert d.max() < 5e-3 and d.max() > 1e-6


# ---------------------------------------------------------------- coherence
def test_coherence_identities(orc):
    rng = np.random.default_rng(6)
    h, w = 20, 30
    m = rng.integers(-500, 500, size=(h, w, 2)).astype(np.int16)
    mc = _c(m)
    c1 = orc.coherence(m, mc, 4, 1, 10, 3)
    assert np.max(np.abs(c1 - 1.0)) < 1e-12
    c2 = orc.coherence(m, mc * np.exp(1j * 0.7), 4, 1, 10, 3)
    assert np.max(np.abs(c2 - 1.0)) < 1e-12
    s = mc.copy()
    s[5, 7] = 0.0005
    c3 = orc.coherence(m, s, 4, 1, 10, 3)
    assert c3[5, 7] == 0.0


def test_coherence_brute_force(orc):
    rng = np.random.default_rng(7)
    h, w = 9, 13
    m = rng.integers(-50, 50, size=(h, w, 2)).astype(np.int16)
    s = rng.normal(size=(h, w)) + 1j * rng.normal(size=(h, w))
    mc = _c(m)
    for (waz, wrg, dy, dx) in ((3, 5, 1, 2), (6, 4, 2, 1), (2, 7, 0, 3)):
        got = orc.coherence(m, s, dx, dy, wrg, waz)
        for r in range(h):
            for c in range(w):
                num, ps, pm = 0j, 0.0, 0.0
                for j in range(waz):
                    for i in range(wrg):
                        rr, cc = r - dy + j, c - dx + i
                        if 0 <= rr < h and 0 <= cc < w:
                            num += mc[rr, cc] * np.conj(s[rr, cc])
                            ps += abs(s[rr, cc]) ** 2
                            pm += abs(mc[rr, cc]) ** 2
                ref = abs(num) / np.sqrt(ps * pm) if ps * pm > 0 else 0.0
                assert abs(got[r, c] - ref) < 1e-12


# ---------------------------------------------------------------- ESD
def _bands(n_lines, P, theta, noise, seed):
    rng = np.random.default_rng(seed)
    sf = (rng.normal(size=(n_lines, P)) + 1j * rng.normal(size=(n_lines, P))) * 300
    sb = (rng.normal(size=(n_lines, P)) + 1j * rng.normal(size=(n_lines, P))) * 300
    nf = (rng.normal(size=(n_lines, P)) + 1j * rng.normal(size=(n_lines, P))) * noise
    nb = (rng.normal(size=(n_lines, P)) + 1j * rng.normal(size=(n_lines, P))) * noise
    mf = sf * np.exp(1j * theta) + nf
    mb = sb + nb
    q = lambda z: np.stack([np.round(z.real), np.round(z.imag)], -1).astype(np.int16)
    return q(mf), q(mb), sf, sb


def test_esd_recovers_constructed_phase(orc):
    P = 200
    sizes = [30, 25]
    df, dt = 5400.0, DT
    thetas = [0.21, -0.37]
    parts = [_bands(n, P, th, 60.0, 10 + i) for i, (n, th) in enumerate(zip(sizes, thetas))]
    mf = np.concatenate([p[0] for p in parts])
    mb = np.concatenate([p[1] for p in parts])
    sf = np.concatenate([p[2] for p in parts])
    sb = np.concatenate([p[3] for p in parts])
    shifts, sums = orc.esd(mf, mb, sf, sb, sizes, P, 6, 24, 0.3, df, dt)
    for i, th in enumerate(thetas):
        assert abs(shifts[i] * 2 * np.pi * df * dt - th) < 0.02
        assert np.hypot(*sums[i]) > 0.5 * sizes[i] * P


def test_esd_incoherent_masked(orc):
    P, n = 120, 20
    rng = np.random.default_rng(21)
    rnd = lambda: (rng.normal(size=(n, P)) + 1j * rng.normal(size=(n, P))) * 300
    q = lambda z: np.stack([np.round(z.real), np.round(z.imag)], -1).astype(np.int16)
    shifts, sums = orc.esd(q(rnd()), q(rnd()), rnd(), rnd(), [n], P, 6, 24, 0.9, 5400.0, DT)
    assert sums[0, 0] == 0 and sums[0, 1] == 0 and shifts[0] == 0


def test_esd_zero_products_take_the_sign_of_zero(orc):
    n, P = 12, 64
    mf, mb, sf, sb, expect = zero_master_band(n, P, 5)
    shifts, sums = orc.esd(mf, mb, sf, sb, [n], P, 6, 24, 0.3, 5400.0, DT)
    assert sums[0, 0] == expect
    assert abs(sums[0, 1]) < 1e-12   # sin(pi) = 1.2e-16 per -1 sample
    assert 0 < np.sum(sf.real > 0) < n * P


def test_esd_responds_to_injected_misregistration(tops, orc):
    """End to end on the synthetic TOPS scene (3 bursts, 100-line overlaps):
    resample with the closed-form offsets, gather the bands, estimate.  The
    estimate moves with the injected misregistration d by d * df_cycle/df,
    where df_cycle = kt * (burst cycle) is the true spectral separation and
    df = kt * lpb * dt the one GetSpectralSep uses (Func.cpp:2962-2976).  The
    estimate at d = 0 is not 0: the reference's 12-tap LUT is not exact on
    linear phase (first moment off by ~1e-3 line), which shows up as a
    constant bias of a few 1e-3 lines."""
    est = {}
    for delta in (0.0, 0.004):
        pair = tops.synth.subswath_iw1(n_bursts=3, lines=240, pixels=512, delta=delta)
        m = pair.master
        lpb, spb = m.lines_per_burst, m.samples_per_burst
        kaz, krg = orc.lut(prf=m.prf, abw=m.azimuth_bandwidth, rsr2x=2 * m.range_sampling_rate)
        kt, tref, fdc, _ = pair.params()
        Ms, Rs = [], []
        for b in range(3):
            ca, cr = pair.cpm(b)
            box = [0, spb - 1, b * lpb, (b + 1) * lpb - 1]
            Rs.append(orc.resample(pair.burst(b, "slave"), box, box, b * lpb, m.az_time_interval,
                                   0.0, ca, cr, kaz, krg, kt[b], tref[b], fdc[b]))
            Ms.append(pair.burst(b, "master"))
        sizes, _, _ = orc.overlap_sizes(m.burst_first_line_time, m.burst_last_line_time, 0, 2,
                                        m.az_time_interval)
        cat = lambda X, back: np.concatenate(
            [X[i + 1][:n] if back else X[i][lpb - n:] for i, n in enumerate(sizes)])
        df = orc.spectral_sep(kt, lpb, m.az_time_interval)
        sh, _ = orc.esd(cat(Ms, 0), cat(Ms, 1), cat(Rs, 0), cat(Rs, 1), sizes, spb, 6, 24, 0.3,
                        df, m.az_time_interval)
        est[delta] = sh
        ratio = (lpb - 100) / lpb
    slope = (est[0.004] - est[0.0]) / 0.004
    assert np.all(np.abs(slope - ratio) < 0.05), (slope, ratio)
    assert np.all(np.abs(est[0.0]) < 0.01)
