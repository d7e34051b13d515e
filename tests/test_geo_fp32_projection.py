"""The FP32 DEM projection of the geocoreg kernel (csrc/geocoreg.hip:
llh_offset, tile_reference), restated in numpy float32 step by step, against
the oracle's FP64 llh2xyz (Geocoreg_cuda.cu:87-99), which is pinned by the
equator / pole closed forms in test_oracle_geometry.py.

The split projects each post relative to its tile's reference post (tiles
of 4 x 64 posts): the offsets of sin / cos from small-angle series, the
prime-vertical radius change from a series in u, all in FP32, added to the
FP64 reference position.  The position error must stay below a millimetre
(a range pixel is 2.3 m, the offset tolerance 1e-3 px) for post spacings
from 1 to 30 arcsec, also next to the pole; the height along the reference
post's ellipsoid normal is added in FP64.  The GPU kernel itself is checked
against the oracle's fitted offsets in test_gpu_esd_geocoreg.py."""
import numpy as np
import pytest

F = np.float32
A = 6378137.0
B = 6356752.3142451794975639665996337


def _tile_reference(lat0, lon0, dlat, dlon, rc, cc):
    """llh2xyz constants at the tile's reference post (rc, cc), height 0"""
    d2r = np.pi / 180.0
    e2 = 1.0 - (B * B / A / A)
    sr, cr = np.sin((lat0 - dlat * rc) * d2r), np.cos((lat0 - dlat * rc) * d2r)
    slr, clr = np.sin((lon0 + dlon * cc) * d2r), np.cos((lon0 + dlon * cc) * d2r)
    w2 = 1.0 - e2 * sr * sr
    nr = A / np.sqrt(w2)
    pr = np.stack([nr * cr * clr, nr * cr * slr, nr * (1 - e2) * sr], axis=-1)
    nv = np.stack([cr * clr, cr * slr, sr], axis=-1)
    return dict(sr=sr, cr=cr, slr=slr, clr=clr, nr=nr, inv_w2=1.0 / w2, pr=pr, nv=nv, e2=e2)


def _sin_small(x):
    x2 = x * x
    return x * (F(1) - x2 * F(1.0 / 6.0) * (F(1) - x2 * F(1.0 / 20.0)))


def _omc_small(x):
    x2 = x * x
    return F(0.5) * x2 * (F(1) - x2 * F(1.0 / 12.0) * (F(1) - x2 * F(1.0 / 30.0)))


def _fp32_projection(lat0, lon0, dlat, dlon, rows, cols, h):
    """the kernel's projection: tiles of 4 x 64 posts, reference post
    (tile row 2, tile column 32), FP32 offsets from it"""
    rc = (rows // 4) * 4 + 2
    cc = (cols // 64) * 64 + 32
    t = _tile_reference(lat0, lon0, dlat, dlon, rc, cc)
    d2r = np.pi / 180.0
    dphi = -(rows - rc).astype(F) * F(dlat * d2r)
    dlam = (cols - cc).astype(F) * F(dlon * d2r)
    h = h.astype(F)
    sp, sl = _sin_small(dphi), _sin_small(dlam)
    cp1, cl1 = _omc_small(dphi), _omc_small(dlam)
    sr, cr, slr, clr = (t[k].astype(F) for k in ("sr", "cr", "slr", "clr"))
    e2, one_e2 = F(t["e2"]), F(1.0 - t["e2"])
    ds = cr * sp - sr * cp1
    dc = -cr * cp1 - sr * sp
    dsl = clr * sl - slr * cl1
    dcl = -clr * cl1 - slr * sl
    u = e2 * (F(2) * sr * ds + ds * ds) * t["inv_w2"].astype(F)
    nr = t["nr"].astype(F)
    dn = nr * (u * (F(0.5) + u * (F(0.375) + u * F(0.3125))))
    c = cr + dc
    nh = nr + h
    d0 = dn * c * (clr + dcl) + nh * (cr * dcl + dc * clr + dc * dcl)
    d1 = dn * c * (slr + dsl) + nh * (cr * dsl + dc * slr + dc * dsl)
    d2 = dn * one_e2 * (sr + ds) + (nr * one_e2 + h) * ds
    # the height along the reference normal in FP64 (the kernel's fma)
    base = t["pr"] + h.astype(np.float64)[:, None] * t["nv"]
    return base + np.stack([np.asarray(d, np.float64) for d in (d0, d1, d2)], axis=-1)


@pytest.mark.parametrize("lat0, lon0, spacing, n, tol_m", [
    (30.25, 100.0, 1.0 / 3600, 2000, 1e-3),     # 1-arcsec posts (the configs' DEM)
    (61.0, -150.0, 3.0 / 3600, 2000, 1e-3),     # 3-arcsec posts, high latitude
    (-89.9, 170.0, 30.0 / 3600, 400, 1e-3),     # 30-arcsec posts next to the pole
])
def test_fp32_projection_matches_fp64_llh2xyz(orc, lat0, lon0, spacing, n, tol_m):
    dlat = dlon = spacing
    rng = np.random.default_rng(3)
    rows = rng.integers(0, n, 4000)
    cols = rng.integers(0, n, 4000)
    rows[:4] = [0, 0, n - 1, n - 1]
    cols[:4] = [0, n - 1, 0, n - 1]
    h = rng.integers(-400, 8800, 4000).astype(np.float64)
    got = _fp32_projection(lat0, lon0, dlat, dlon, rows, cols, h)
    d2r = np.pi / 180.0
    ref = np.array([orc.llh2xyz((lat0 - dlat * r) * d2r, (lon0 + dlon * c) * d2r, hh)
                    for r, c, hh in zip(rows, cols, h)])
    err = float(np.max(np.linalg.norm(got - ref, axis=1)))
    print(f"  {spacing * 3600:g}-arcsec posts at {lat0}N: max |P32 - P64| = {err * 1e3:.3f} mm")
    assert err <= tol_m


def test_series_terms_against_libm():
    """The small-angle series against libm in FP64 up to 1e-2 rad (32 posts
    of 60 arcsec): the truncation stays far below FP32 rounding."""
    x = np.linspace(-1e-2, 1e-2, 20001)
    s = x * (1 - x * x / 6 * (1 - x * x / 20))
    o = 0.5 * x * x * (1 - x * x / 12 * (1 - x * x / 30))
    assert np.max(np.abs(s - np.sin(x)) / np.maximum(np.abs(x), 1e-30)) <= 1e-15
    # 1 - cos in FP64 itself loses digits for small x: compare with 2 sin^2(x/2)
    assert np.max(np.abs(o - 2 * np.sin(x / 2) ** 2) / np.maximum(x * x, 1e-30)) <= 1e-15
