"""Seeded synthetic Sentinel-1 IW scenarios (metadata, orbit, DEM, bursts).

The metadata fields are the subset of the reference's SubSwathInfo
(DataTypes.h:51-140) that the coregistration, deramp/resample, ESD and
coherence stages read.  Heavy arrays (bursts, DEM windows) come from
libtops_synth.so (csrc/synth.c); everything here is small numpy.

Workloads (sizes of the benchmark configurations):
  single_burst_small : one 512 x 2048 burst, 3" DEM over 0.1 x 0.1 deg
  single_burst_iw1   : one 1500 x 21000 IW1 burst, 1" DEM
  subswath_iw1       : 9 bursts of 1500 x 21000 with ~100-line overlaps and
                       a 0.002-line injected azimuth misregistration
  iw_pair            : 3 subswaths x 9 bursts, 21000/25000/24500 samples,
                       1480-1520 lines per burst
  stack_iw2          : one 3-burst IW2 master (1500 x 25000) and 16 slaves
"""
from dataclasses import dataclass, field
import ctypes as C

import numpy as np

from . import native

SOL = 299792458.0
PI = 3.1415926535897931
ELL_A = 6378137.0
ELL_B = 6356752.3142451794975639665996337
ELL_E2 = 1.0 - (ELL_B * ELL_B / ELL_A / ELL_A)
DEM_NODATA = -32768

# IW1-like constants (annotation values of a typical S1 IW1 product)
AZ_TIME_INTERVAL = 2.055556e-3
RANGE_SAMPLING_RATE = 64.34523076923077e6
RANGE_SPACING = SOL / (2.0 * RANGE_SAMPLING_RATE)
WAVELENGTH = 0.05546576
AZ_BANDWIDTH = 327.0
STEERING_RATE = 1.590368784   # deg/s


# ---------------------------------------------------------------------------
# geometry helpers
# ---------------------------------------------------------------------------
def llh2xyz(lat_deg, lon_deg, h):
    phi = np.radians(lat_deg)
    lam = np.radians(lon_deg)
    n = ELL_A / np.sqrt(1.0 - ELL_E2 * np.sin(phi) ** 2)
    return np.array([(n + h) * np.cos(phi) * np.cos(lam),
                     (n + h) * np.cos(phi) * np.sin(lam),
                     (n * (1 - ELL_E2) + h) * np.sin(phi)])


def xyz2ll(p):
    """geodetic lat/lon (deg) of an ECEF point near the ellipsoid"""
    lon = np.degrees(np.arctan2(p[1], p[0]))
    r = np.hypot(p[0], p[1])
    lat = np.arctan2(p[2], r * (1 - ELL_E2))
    for _ in range(6):
        n = ELL_A / np.sqrt(1 - ELL_E2 * np.sin(lat) ** 2)
        h = r / np.cos(lat) - n
        lat = np.arctan2(p[2], r * (1 - ELL_E2 * n / (n + h)))
    return np.degrees(lat), lon


@dataclass
class Orbit:
    """Circular orbit: S(t) = R (cos w(t-tc) e1 + sin w(t-tc) e2)."""
    tc: float
    radius: float
    omega: float
    e1: np.ndarray
    e2: np.ndarray
    offset: np.ndarray = field(default_factory=lambda: np.zeros(3))

    def pos(self, t):
        a = self.omega * (np.asarray(t, dtype=np.float64) - self.tc)
        return (self.radius * (np.multiply.outer(np.cos(a), self.e1) +
                               np.multiply.outer(np.sin(a), self.e2)) + self.offset)

    def vel(self, t):
        a = self.omega * (np.asarray(t, dtype=np.float64) - self.tc)
        return self.radius * self.omega * (np.multiply.outer(-np.sin(a), self.e1) +
                                           np.multiply.outer(np.cos(a), self.e2))

    def state_vectors(self, n=10, spacing=10.0):
        """n state vectors every `spacing` s with tc at index n//2 (the
        reference normalises time at orbitAzTime[N/2], Func.cpp:1060)."""
        t = self.tc + (np.arange(n) - n // 2) * spacing
        return t, self.pos(t), self.vel(t)


def make_orbit(lat0, lon0, tc, altitude=693e3, speed=7500.0, alpha=0.045, offset=None):
    """Right-looking circular orbit whose zero-Doppler point at tc is
    (lat0, lon0) at Earth-central angle alpha from the sub-satellite point."""
    pc = llh2xyz(lat0, lon0, 0.0)
    up = pc / np.linalg.norm(pc)
    north = np.array([0.0, 0.0, 1.0]) - up[2] * up
    north /= np.linalg.norm(north)
    east = np.cross(north, up)
    e1 = np.cos(alpha) * up - np.sin(alpha) * east
    radius = ELL_A + altitude
    return Orbit(tc=tc, radius=radius, omega=speed / radius, e1=e1, e2=north,
                 offset=np.zeros(3) if offset is None else np.asarray(offset, float))


def zero_doppler(orbit, p, t0):
    t = t0
    for _ in range(20):
        s = orbit.pos(t)
        v = orbit.vel(t)
        a = -orbit.omega ** 2 * (s - orbit.offset)
        d = p - s
        dt = -(v @ d) / (a @ d - v @ v)
        t += dt
        if abs(dt) < 1e-12:
            break
    return t, float(np.linalg.norm(p - orbit.pos(t)))


def geolocate(orbit, t, r, guess_ll):
    """Point on the ellipsoid (h=0) at zero Doppler time t and slant range r."""
    lat, lon = guess_ll
    for _ in range(30):
        p = llh2xyz(lat, lon, 0.0)
        s = orbit.pos(t)
        v = orbit.vel(t)
        f = np.array([np.linalg.norm(p - s) - r, (p - s) @ v / np.linalg.norm(v)])
        eps = 1e-6
        j = np.zeros((2, 2))
        for k, (dl, dn) in enumerate(((eps, 0.0), (0.0, eps))):
            q = llh2xyz(lat + dl, lon + dn, 0.0)
            j[:, k] = (np.array([np.linalg.norm(q - s) - r, (q - s) @ v / np.linalg.norm(v)]) - f) / eps
        step = np.linalg.solve(j, -f)
        lat += step[0]
        lon += step[1]
        if np.max(np.abs(step)) < 1e-12:
            break
    return lat, lon


# ---------------------------------------------------------------------------
# metadata
# ---------------------------------------------------------------------------
@dataclass
class Swath:
    """SubSwathInfo subset (DataTypes.h:51-140) of one acquisition."""
    lines_per_burst: int
    samples_per_burst: int
    n_bursts: int
    az_time_interval: float
    range_spacing: float
    slr_time_first_pixel: float      # one way
    wavelength: float
    prf: float
    azimuth_bandwidth: float
    range_sampling_rate: float
    steering_rate_deg: float
    first_valid_pixel: int
    burst_first_line_time: np.ndarray
    burst_last_line_time: np.ndarray
    dc: np.ndarray                   # [n_bursts, 4] (t0, c0, c1, c2)
    fm: np.ndarray                   # [n_bursts, 4]
    burst_ll: np.ndarray             # [n_bursts, 4] lat_min lat_max lon_min lon_max
    scene_line0: np.ndarray          # absolute scene line of each burst's first line
    first_line_time: float = 0.0
    last_line_time: float = 0.0
    # first / last valid line of each burst (the annotation's firstValidSample
    # rows that are not -1, Func.cpp:370-406) and the subswath's line count
    first_valid_line: np.ndarray = None
    last_valid_line: np.ndarray = None
    num_lines: int = 0

    @property
    def burst_first_valid_line_time(self):
        return self.burst_first_line_time + self.first_valid_line * self.az_time_interval

    @property
    def burst_last_valid_line_time(self):
        return self.burst_first_line_time + self.last_valid_line * self.az_time_interval


def deramp_params(sw, speed):
    """kt, tref, fdc per burst and range sample: the formulas of
    ComputeDerampPara (Func.cpp:1955-2090) with the orbit speed given."""
    nb, spb = sw.n_bursts, sw.samples_per_burst
    x = np.arange(spb, dtype=np.float64)
    slrt = (sw.slr_time_first_pixel + x * sw.range_spacing / SOL) * 2
    fdc = np.empty((nb, spb))
    ka = np.empty((nb, spb))
    for b in range(nb):
        t0, c0, c1, c2 = sw.dc[b]
        d = slrt - t0
        fdc[b] = c0 + c1 * d + c2 * (d * d)
        t0, c0, c1, c2 = sw.fm[b]
        d = slrt - t0
        ka[b] = c0 + c1 * d + c2 * d * d
    tref = np.empty((nb, spb))
    temp1 = sw.lines_per_burst * sw.az_time_interval / 2.0
    for b in range(nb):
        fv = sw.first_valid_pixel
        tref[b] = temp1 + fdc[b, fv] / ka[b, fv] - fdc[b] / ka[b]
    krot = 2 * speed * (sw.steering_rate_deg * PI / 180.0) / sw.wavelength
    kt = ka * krot / (ka - krot)
    return kt, tref, fdc, ka


def offset_to_cpm(c, n_minus_1, base):
    """Offset polynomial c in (u, v) in [0,1]^2 -> reference Cpm in
    normalised master coordinates (x = Ln, y = Pn, cuConstants.cuh:57-66),
    plus the identity part base + (n-1)*u (az) handled by the caller."""
    c0, c1, c2, c3, c4, c5 = c
    return np.array([c0 + 0.5 * c1 + 0.5 * c2 + 0.25 * c3 + 0.25 * c4 + 0.25 * c5,
                     0.25 * c1 + 0.25 * c3 + 0.125 * c4,
                     0.25 * c2 + 0.125 * c4 + 0.25 * c5,
                     0.0625 * c3, 0.0625 * c4, 0.0625 * c5])


def cpm_from_field(az_off, rg_off, lines, pixels, slave_line0):
    """CpmAz/CpmRg giving absolute slave line = slave_line0 + l + az(u,v)
    and slave pixel = p + rg(u,v) for master burst-relative (l, p)."""
    ca = offset_to_cpm(az_off, lines - 1, 0.0)
    ca[0] += slave_line0 + 0.5 * (lines - 1)
    ca[1] += 0.25 * (lines - 1)
    cr = offset_to_cpm(rg_off, pixels - 1, 0.0)
    cr[0] += 0.5 * (pixels - 1)
    cr[2] += 0.25 * (pixels - 1)
    return ca, cr


BILINEAR_AZ = np.array([0.05, 0.25 * 0.5, 0.25 * 0.3, 0.0, 0.25 * 0.2, 0.0])
BILINEAR_RG = np.array([0.5, 1.5 * 0.5, 1.5 * 0.3, 0.0, 1.5 * 0.2, 0.0])


@dataclass
class Pair:
    """Master/slave pair over one subswath (burst-aligned, BurstGap = 0)."""
    name: str
    seed: int
    stream: int
    master: Swath
    slave: Swath
    morbit: Orbit
    sorbit: Orbit
    az_off: np.ndarray
    rg_off: np.ndarray
    delta: float                     # injected azimuth misregistration (lines)
    phase_noise: float
    scene_center: tuple
    dem_spacing: float               # arc-seconds
    speed: float = 7500.0
    coh_win: tuple = (3, 10)         # (az, rg)
    noise_stream: int = None         # phase-noise stream of the slave (default stream + 1000)

    def params(self, which="slave"):
        sw = self.slave if which == "slave" else self.master
        return deramp_params(sw, self.speed)

    def cpm(self, b):
        """closed-form CpmAz/CpmRg of burst b (no misregistration)"""
        L, P = self.master.lines_per_burst, self.master.samples_per_burst
        return cpm_from_field(self.az_off, self.rg_off, L, P, b * self.slave.lines_per_burst)

    def burst(self, b, which):
        """CInt16 burst b of 'master' or 'slave' as int16 [L, P, 2], ramped
        with that swath's own Doppler / FM parameters."""
        S = native.load_synth()
        sw = self.slave if which == "slave" else self.master
        L, P = sw.lines_per_burst, sw.samples_per_burst
        kt, tref, fdc, _ = deramp_params(sw, self.speed)
        kt_b = np.ascontiguousarray(kt[b])
        tr_b = np.ascontiguousarray(tref[b])
        fd_b = np.ascontiguousarray(fdc[b])
        out = np.empty((L, P, 2), dtype=np.int16)
        az_band = sw.azimuth_bandwidth / sw.prf
        rg_band = 0.85
        if which == "master":
            S.tsyn_burst(self.seed, self.stream, az_band, rg_band, 300.0, int(sw.scene_line0[b]),
                         L, P, None, None, 0.0, sw.az_time_interval, native.ptr(kt_b),
                         native.ptr(tr_b), native.ptr(fd_b), 0.0, 0, out.ctypes.data)
        else:
            az = np.ascontiguousarray(self.az_off, dtype=np.float64)
            rg = np.ascontiguousarray(self.rg_off, dtype=np.float64)
            noise = self.stream + 1000 if self.noise_stream is None else self.noise_stream
            S.tsyn_burst(self.seed, self.stream, az_band, rg_band, 300.0, int(sw.scene_line0[b]),
                         L, P, native.ptr(az), native.ptr(rg), self.delta, sw.az_time_interval,
                         native.ptr(kt_b), native.ptr(tr_b), native.ptr(fd_b),
                         self.phase_noise, noise, out.ctypes.data)
        return out

    # --- DEM --------------------------------------------------------------
    def dem_grid(self):
        """global DEM grid: (lat_top, lon_left, step_deg)"""
        d = self.dem_spacing / 3600.0
        lat_c, lon_c = self.scene_center
        return lat_c + 1.0, lon_c - 1.0, d

    def dem_window(self, lat_min, lat_max, lon_min, lon_max, margin_posts=24, nodata_every=0):
        """RefDem::getData semantics (DataTypes.cpp:240-300): window around
        the lat/lon box extended by margin_posts; returns (dem int32 [l,p],
        lat0, lon0, dlat, dlon) with post (0,0) at (lat0, lon0)."""
        top, left, d = self.dem_grid()
        la0, la1 = lat_min - margin_posts * d, lat_max + margin_posts * d
        lo0, lo1 = lon_min - margin_posts * d, lon_max + margin_posts * d
        r0 = int(np.floor((top - la1) / d))
        c0 = int(np.floor((lo0 - left) / d))
        r1 = int(np.ceil((top - la0) / d))
        c1 = int(np.ceil((lo1 - left) / d))
        lines, pixels = r1 - r0 + 1, c1 - c0 + 1
        out = np.empty((lines, pixels), dtype=np.int32)
        native.load_synth().tsyn_dem(self.seed, r0, c0, lines, pixels, self.dem_spacing, 1500.0,
                                     nodata_every, DEM_NODATA, out.ctypes.data)
        return out, top - r0 * d, left + c0 * d, d, d


def _swath(name_seed, lines, pixels, n_bursts, t_first, cycle_lines, slr0, dc_slope_total,
           orbit, scene_center, dc0=20.0):
    dt = AZ_TIME_INTERVAL
    first = t_first + np.arange(n_bursts) * cycle_lines * dt
    # a hair less than the cycle so EstOverlapSize's floor is not marginal
    first = first - np.arange(n_bursts) * 1e-9
    last = first + (lines - 1) * dt
    # invalid lines at each end of a burst, as in IW annotations
    margin = min(lines // 30, 40)
    two_way0 = 2 * slr0
    drange = 2 * RANGE_SPACING / SOL
    c1 = dc_slope_total / (pixels - 1) / drange
    dc = np.tile([two_way0, dc0, c1, 0.0], (n_bursts, 1))
    fm = np.tile([two_way0, -2300.0, 0.0, 0.0], (n_bursts, 1))
    ll = np.zeros((n_bursts, 4))
    for b in range(n_bursts):
        lats, lons = [], []
        for tl in (first[b], last[b]):
            for px in (0, pixels - 1):
                la, lo = geolocate(orbit, tl, (slr0 + px * RANGE_SPACING / SOL) * SOL, scene_center)
                lats.append(la)
                lons.append(lo)
        ll[b] = [min(lats), max(lats), min(lons), max(lons)]
    return Swath(lines_per_burst=lines, samples_per_burst=pixels, n_bursts=n_bursts,
                 az_time_interval=dt, range_spacing=RANGE_SPACING, slr_time_first_pixel=slr0,
                 wavelength=WAVELENGTH, prf=1.0 / dt, azimuth_bandwidth=AZ_BANDWIDTH,
                 range_sampling_rate=RANGE_SAMPLING_RATE, steering_rate_deg=STEERING_RATE,
                 first_valid_pixel=0, burst_first_line_time=first, burst_last_line_time=last,
                 dc=dc, fm=fm, burst_ll=ll,
                 scene_line0=np.arange(n_bursts, dtype=np.int64) * cycle_lines,
                 first_line_time=float(first[0]), last_line_time=float(last[-1]),
                 first_valid_line=np.full(n_bursts, margin, dtype=np.int64),
                 last_valid_line=np.full(n_bursts, lines - 1 - margin, dtype=np.int64),
                 num_lines=n_bursts * lines)


def make_pair(name, lines, pixels, n_bursts=1, overlap=100, seed=20260101, stream=0,
              az_off=BILINEAR_AZ, rg_off=BILINEAR_RG, delta=0.0, phase_noise=0.2,
              dem_spacing=3.0, dc_slope_total=5e-3 * 2047, scene_center=(30.0, 100.0),
              slr_shift=0.0, slave_dc0=20.0, slave_burst_offset=0):
    """Burst-aligned master/slave pair; the slave orbit carries a ~120 m
    baseline and a small along-track delay so geocoreg sees real offsets.
    slave_dc0: the slave's Doppler centroid at the first range sample (Hz);
    the master's is 20 Hz.  slave_burst_offset: the slave's bursts start
    that many burst cycles after the master's (metadata only: the burst
    rasters assume aligned bursts)."""
    tc = 43200.0
    morbit = make_orbit(scene_center[0], scene_center[1], tc)
    cycle = lines - overlap if n_bursts > 1 else lines
    total = (n_bursts - 1) * cycle + lines
    t_first = tc - 0.5 * (total - 1) * AZ_TIME_INTERVAL
    pc = llh2xyz(scene_center[0], scene_center[1], 0.0)
    _, rc = zero_doppler(morbit, pc, tc)
    slr0 = (rc - 0.5 * (pixels - 1) * RANGE_SPACING) / SOL + slr_shift
    master = _swath(seed, lines, pixels, n_bursts, t_first, cycle, slr0, dc_slope_total, morbit,
                    scene_center)
    up = pc / np.linalg.norm(pc)
    sorbit = make_orbit(scene_center[0], scene_center[1], tc + 0.2 * AZ_TIME_INTERVAL,
                        offset=80.0 * morbit.e1 + 90.0 * np.cross(morbit.e2, morbit.e1))
    slave = _swath(seed + 1, lines, pixels, n_bursts,
                   t_first + slave_burst_offset * cycle * AZ_TIME_INTERVAL, cycle, slr0,
                   dc_slope_total, sorbit, scene_center, dc0=slave_dc0)
    del up
    return Pair(name=name, seed=seed, stream=stream, master=master, slave=slave, morbit=morbit,
                sorbit=sorbit, az_off=np.asarray(az_off, float), rg_off=np.asarray(rg_off, float),
                delta=delta, phase_noise=phase_noise, scene_center=scene_center,
                dem_spacing=dem_spacing)


def single_burst_small(seed=20260101):
    """512 az x 2048 rg burst, 3" DEM, 3x10 coherence window."""
    return make_pair("single_burst_small", 512, 2048, seed=seed, dem_spacing=3.0,
                     dc_slope_total=5e-3 * 2047)


def single_burst_iw1(seed=20260102):
    """Full IW1 burst 1500 x 21000, offsets and Doppler scaled to the extent."""
    return make_pair("single_burst_iw1", 1500, 21000, seed=seed, dem_spacing=1.0,
                     dc_slope_total=5e-3 * 2047)


def subswath_iw1(seed=20260103, n_bursts=9, lines=1500, pixels=21000, delta=0.002):
    """9 bursts with ~100-line overlaps and an injected misregistration."""
    return make_pair("subswath_iw1", lines, pixels, n_bursts=n_bursts, overlap=100, seed=seed,
                     dem_spacing=1.0, delta=delta, phase_noise=0.2)


def orbit_polyfit_inputs(orbit, n=10, spacing=10.0):
    t, p, v = orbit.state_vectors(n, spacing)
    return t, p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()


# ---------------------------------------------------------------------------
# batched workloads: many bursts of one plan
# ---------------------------------------------------------------------------
@dataclass
class BatchBurst:
    """One burst of a batched workload: burst `index` of `pair`, with its
    master box on the master grid (absolute lines y0 .. y0 + lines - 1)."""
    pair: Pair
    index: int
    group: int       # subswath (IW pair) or slave (stack) it belongs to
    lines: int
    width: int
    y0: int

    @property
    def box(self):
        return [0, self.width - 1, self.y0, self.y0 + self.lines - 1]

    @property
    def dt(self):
        return self.pair.master.az_time_interval

    def cpm(self):
        return self.pair.cpm(self.index)

    def master(self):
        return self.pair.burst(self.index, "master")

    def slave(self):
        return self.pair.burst(self.index, "slave")

    def params(self):
        """kt, tref, fdc of the slave burst (contiguous [width] arrays)"""
        kt, tref, fdc, _ = self.pair.params()
        return tuple(np.ascontiguousarray(a[self.index]) for a in (kt, tref, fdc))


def iw_pair_bursts(seed=2026, widths=(21000, 25000, 24500), n_bursts=9, lines=(1480, 1520)):
    """The batched IW pair: 3 subswaths x 9 bursts of 21000 / 25000 / 24500
    samples with line counts drawn uniformly from 1480-1520.  Every burst is
    its own seeded scene with its own offset field (az 0-0.3 px, rg 0.5-2
    px, bilinear) and Doppler centroid slope (10-40 Hz over the swath)."""
    rng = np.random.default_rng(seed)
    out = []
    for sw, width in enumerate(widths):
        counts = rng.integers(lines[0], lines[1] + 1, size=n_bursts)
        for b in range(n_bursts):
            amp_az, amp_rg = rng.uniform(0.0, 0.3), rng.uniform(0.5, 2.0)
            az = np.array([rng.uniform(0.0, 0.3 - 0.5 * amp_az) + 0.0, amp_az * 0.5,
                           amp_az * 0.3, 0.0, amp_az * 0.2, 0.0])
            rg = np.array([0.5, (amp_rg - 0.5) * 0.5, (amp_rg - 0.5) * 0.3, 0.0,
                           (amp_rg - 0.5) * 0.2, 0.0])
            nl = int(counts[b])
            pair = make_pair(f"iw{sw + 1}_b{b}", nl, width, seed=seed * 1000 + sw * 100 + b,
                             az_off=az, rg_off=rg, dem_spacing=1.0,
                             dc_slope_total=float(rng.uniform(10.0, 40.0)))
            out.append(BatchBurst(pair, 0, sw, nl, width, 0))
    return out


def stack_bursts(seed=4100, n_slaves=16, n_bursts=3, lines=1500, width=25000):
    """The stack: one IW2 master of 3 bursts (1500 x 25000, 100-line
    overlaps) against 16 slaves that image the same scene, each with its own
    offset field (az 0-0.5 px, rg 0-3 px) and Doppler centroid (0-60 Hz) and
    its own phase noise.  Bursts are listed slave by slave."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_slaves):
        a, r = rng.uniform(0.0, 0.25, size=4), rng.uniform(0.0, 1.5, size=4)
        az = np.array([a[0], a[1] * 0.5, a[2] * 0.3, 0.0, a[3] * 0.2, 0.0])
        rg = np.array([r[0], r[1] * 0.5, r[2] * 0.3, 0.0, r[3] * 0.2, 0.0])
        pair = make_pair("stack_iw2", lines, width, n_bursts=n_bursts, overlap=100, seed=seed,
                         stream=0, az_off=az, rg_off=rg, dem_spacing=1.0, dc_slope_total=10.0,
                         slave_dc0=float(rng.uniform(0.0, 60.0)))
        pair.noise_stream = 1000 + k
        for b in range(n_bursts):
            out.append(BatchBurst(pair, b, k, lines, width, b * lines))
    return out
