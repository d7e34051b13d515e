"""Pins of the oracle's resample LUT and orbit functions (CPU only).

LUT (InKernelInit, Func.cpp:2105-2158):
  * row 0 is the unit impulse at tap 5 and row 2047 the impulse at tap 6
    (the kernel is a sinc, zero at non-zero integers);
  * every row sums to 1 within float rounding;
  * k[i][j] = k[2047-i][11-j] (the kernel is even in x);
  * interpolating a band-limited complex exponential at the row's fractional
    position reproduces it (a sign or index error in x fails this).
Orbit (Polyfit_Orbit, Func.cpp:1050-1126; getPos/getVelo/getAcc,
Geocoreg_cuda.cu:123-205):
  * an exact degree-5 polynomial track is recovered to rounding;
  * velocity and acceleration equal central finite differences of position;
  * a circular orbit is fitted to sub-millimetre over +-50 s.
"""
import numpy as np
import pytest

PRF = 1 / 2.055556e-3
ABW = 327.0
RSR2X = 2 * 64.34523076923077e6


@pytest.fixture(scope="module")
def lut(orc):
    return orc.lut(12, PRF, ABW, RSR2X)


def test_lut_impulse_rows(lut):
    for k in lut:
        e0 = np.zeros(12)
        e0[5] = 1.0
        e1 = np.zeros(12)
        e1[6] = 1.0
        assert np.max(np.abs(k[0] - e0)) < 1e-15
        assert np.max(np.abs(k[2047] - e1)) < 1e-15


def test_lut_rows_sum_to_one(lut):
    for k in lut:
        s = k.astype(np.float64).sum(axis=1)
        assert np.max(np.abs(s - 1.0)) < 6e-7


def test_lut_symmetry(lut):
    for k in lut:
        a = k.astype(np.float64)
        b = a[::-1, ::-1]
        assert np.max(np.abs(a - b)) < 2e-7


@pytest.mark.parametrize("which", [0, 1])
def test_lut_interpolates_band_limited_signal(lut, which):
    """A slowly varying exp(j w n) sampled at the taps and weighted by row i
    must land at the row's fractional position i/2047: phase w*i/2047.  The
    raised-cosine kernel is a coarse interpolator (the range kernel has v ~ 1
    because chi_rg uses the azimuth bandwidth), so only a low frequency is
    checked; a sign or offset error in x (e.g. j-5+i/2047) misplaces the
    phase by ~2w = 0.13 rad and fails by two orders of magnitude."""
    k = lut[which].astype(np.float64)
    frac = np.arange(2048) / 2047.0
    taps = np.arange(12) - 5            # sample offsets of the taps
    omega = 0.02 * np.pi
    est = (k * np.exp(1j * omega * taps)[None, :]).sum(axis=1)
    exact = np.exp(1j * omega * frac)
    assert np.max(np.abs(est - exact)) < 2e-3
    assert np.max(np.abs(np.angle(est / exact))) < 2e-3


def test_polyfit_exact_polynomial(orc):
    t = 1000.0 + 10.0 * np.arange(10)
    tn = (t - t[5]) / 10.0
    cx = np.array([4.0e6, 7.1e3, -3.2, 0.51, -0.02, 0.003])
    cy = np.array([-1.0e6, 2.0e2, 11.0, -0.4, 0.01, -0.002])
    cz = np.array([5.5e6, -6.9e3, 1.5, 0.3, 0.07, 0.0011])
    pv = lambda c: sum(c[j] * tn ** j for j in range(6))
    fx, fy, fz = orc.polyfit_orbit(t, pv(cx), pv(cy), pv(cz))
    assert len(fx) == 6
    for got, ref in ((fx, cx), (fy, cy), (fz, cz)):
        assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)) < 1e-7


def test_polyfit_degree_rule(orc):
    # degree = min(5, N-2): 5 points -> 4 coefficients
    t = 10.0 * np.arange(5)
    x = 1.0 + 2.0 * ((t - t[2]) / 10.0)
    fx, _, _ = orc.polyfit_orbit(t, x, x, x)
    assert len(fx) == 4
    assert abs(fx[0] - 1.0) < 1e-12 and abs(fx[1] - 2.0) < 1e-12


def test_velocity_acceleration_finite_difference(orc):
    rng = np.random.default_rng(3)
    c = [rng.normal(size=6) * 10.0 ** -np.arange(6) for _ in range(3)]
    for tn in (-2.3, 0.0, 0.7, 4.1):
        h = 1e-4
        p = lambda u: orc.orbit_eval("pos", u, *c)
        fd_v = (p(tn + h) - p(tn - h)) / (2 * h) / 10.0          # d/dt, t = 10 tn
        fd_a = (p(tn + h) - 2 * p(tn) + p(tn - h)) / (h * h) / 100.0
        v = orc.orbit_eval("vel", tn, *c)
        a = orc.orbit_eval("acc", tn, *c)
        assert np.max(np.abs(v - fd_v)) < 1e-6
        assert np.max(np.abs(a - fd_a)) < 1e-5


def test_circular_orbit_fit(orc, tops):
    orbit = tops.synth.make_orbit(30.0, 100.0, 43200.0)
    t, x, y, z = tops.synth.orbit_polyfit_inputs(orbit)
    cx, cy, cz = orc.polyfit_orbit(t, x, y, z)
    for tt in np.linspace(t[1], t[-2], 17):
        tn = (tt - t[5]) / 10.0
        p = orc.orbit_eval("pos", tn, cx, cy, cz)
        assert np.max(np.abs(p - orbit.pos(tt))) < 1e-3
        v = orc.orbit_eval("vel", tn, cx, cy, cz)
        assert np.max(np.abs(v - orbit.vel(tt))) < 1e-4


def test_llh2xyz_closed_forms(orc):
    a, b = 6378137.0, 6356752.3142451794975639665996337
    p = orc.llh2xyz(0.0, 0.0, 0.0)
    assert np.allclose(p, [a, 0, 0], atol=1e-9)
    p = orc.llh2xyz(np.pi / 2, 0.3, 100.0)
    assert np.allclose(p, [0, 0, b + 100.0], atol=1e-6)
    p = orc.llh2xyz(0.0, np.pi / 2, 10.0)
    assert np.allclose(p, [0, a + 10.0, 0], atol=1e-8)
