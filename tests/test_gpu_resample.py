"""GPU parity of deramp -> resample -> reramp and coherence against the FP64
oracle (DerampAndResample.cu:348-520, EstCoh.cu:413-456 semantics).

Tolerances: resampled samples max|dz|/RMS|z| <= 1e-5, coherence max |dc|
<= 1e-4 absolute.  The error is printed for every case.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL_RESAMPLE = 1e-5
TOL_COH = 1e-4


@pytest.fixture(scope="module")
def c1(tops, orc):
    pair = tops.synth.single_burst_small()
    m = pair.master
    M = pair.burst(0, "master")
    S = pair.burst(0, "slave")
    kt, tref, fdc, _ = pair.params()
    kaz, krg = orc.lut(prf=m.prf, abw=m.azimuth_bandwidth, rsr2x=2 * m.range_sampling_rate)
    ca, cr = pair.cpm(0)
    return dict(pair=pair, M=M, S=S, kt=kt[0].copy(), tref=tref[0].copy(), fdc=fdc[0].copy(),
                kaz=kaz, krg=krg, ca=ca, cr=cr, L=m.lines_per_burst, P=m.samples_per_burst,
                dt=m.az_time_interval)


class Resampler:
    def __init__(self, tops, dev, d):
        self.N = tops.native
        self.L = self.N.load()
        self.dev = dev
        self.ctx = self.L.tops_ctx_create(self.N.ptr(d["kaz"], C.c_float),
                                          self.N.ptr(d["krg"], C.c_float), 12)
        assert self.ctx, self.N.last_error()

    def close(self):
        self.L.tops_ctx_destroy(self.ctx)

    def run(self, slave, mbox, sbox, first_line, dt, az_shift, ca, cr, kt, tref, fdc, mode,
            spb=None):
        N, dev = self.N, self.dev
        g = N.BurstGeom((C.c_int * 4)(*mbox), (C.c_int * 4)(*sbox), first_line,
                        spb if spb is not None else len(kt), dt, az_shift,
                        (C.c_double * 6)(*ca), (C.c_double * 6)(*cr))
        ds = dev.put(slave)
        dk, dtr, dfd = dev.put(kt), dev.put(tref), dev.put(fdc)
        ml, mp = mbox[3] - mbox[2] + 1, mbox[1] - mbox[0] + 1
        do = dev.empty(ml * mp * 8)
        print(f"  resample mode={mode} mbox={mbox} sbox={sbox} shift={az_shift}", flush=True)
        rc = self.L.tops_resample_burst(self.ctx, C.byref(g), ds.data_ptr(), dk.data_ptr(),
                                        dtr.data_ptr(), dfd.data_ptr(), do.data_ptr(), mode, None)
        N.check(rc, "tops_resample_burst")
        out = dev.get(do, np.complex64, (ml, mp))
        return out, do


@pytest.fixture(scope="module")
def rs(tops, dev, c1):
    r = Resampler(tops, dev, c1)
    yield r
    r.close()


def _oracle(orc, d, mbox, sbox, first_line, az_shift, ca, cr, S=None):
    return orc.resample(d["S"] if S is None else S, mbox, sbox, first_line, d["dt"], az_shift, ca,
                        cr, d["kaz"], d["krg"], d["kt"], d["tref"], d["fdc"])


@pytest.mark.parametrize("mode", [0, 1], ids=["fast", "fp64"])
def test_resample_small_burst(orc, c1, rs, mode):
    L, P = c1["L"], c1["P"]
    box = [0, P - 1, 0, L - 1]
    ref = _oracle(orc, c1, box, box, 0, 0.0, c1["ca"], c1["cr"])
    got, _ = rs.run(c1["S"], box, box, 0, c1["dt"], 0.0, c1["ca"], c1["cr"], c1["kt"],
                    c1["tref"], c1["fdc"], mode)
    e = rel_err(got, ref)
    p999 = np.quantile(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30), 0.999)
    print(f"  resample {['fast', 'fp64'][mode]}: max|dz|/rms = {e:.3e}, p99.9 pointwise rel = {p999:.3e}")
    assert e <= TOL_RESAMPLE


def test_resample_shift_and_offset_boxes(orc, c1, rs):
    """Burst index > 0 (absolute slave lines), an ESD azimuth shift and a
    master box not starting at column 0."""
    L, P = c1["L"], c1["P"]
    b = 3
    mbox = [16, P - 1, b * L, (b + 1) * L - 1]
    sbox = [0, P - 1, b * L, (b + 1) * L - 1]
    # Cpm for absolute slave lines of burst b, master columns offset by 16
    ca = c1["ca"].copy()
    ca[0] += b * L
    shift = 0.37
    ref = _oracle(orc, c1, mbox, sbox, b * L, shift, ca, c1["cr"])
    got, _ = rs.run(c1["S"], mbox, sbox, b * L, c1["dt"], shift, ca, c1["cr"], c1["kt"],
                    c1["tref"], c1["fdc"], 0)
    e = rel_err(got, ref)
    print(f"  shifted burst {b}: max|dz|/rms = {e:.3e}")
    assert e <= TOL_RESAMPLE


@pytest.mark.parametrize("shear", [0.02, 0.5, 40.0])
def test_resample_azimuth_shear(orc, c1, rs, shear):
    """An azimuth offset that varies along range: at 0.02 px per unit of Pn
    a 50-column row segment holds a handful of azimuth keys, at 0.5 nearly
    every output has its own key (the shared pass runs in several rounds),
    at 40 the keys span more lines than the window reaches and those outputs
    take the direct 144-tap path."""
    L, P = c1["L"], c1["P"]
    box = [0, P - 1, 0, L - 1]
    ca = np.array(c1["ca"], dtype=float)
    ca[2] += shear
    ref = _oracle(orc, c1, box, box, 0, 0.0, ca, c1["cr"])
    got, _ = rs.run(c1["S"], box, box, 0, c1["dt"], 0.0, ca, c1["cr"], c1["kt"], c1["tref"],
                    c1["fdc"], 0)
    e = rel_err(got, ref)
    print(f"  azimuth shear {shear}: max|dz|/rms = {e:.3e}")
    assert e <= TOL_RESAMPLE


@pytest.mark.parametrize("wh", [(2, 2), (7, 3), (51, 16), (52, 17), (103, 5)])
def test_resample_and_coherence_small_boxes(tops, orc, dev, c1, rs, wh):
    """Master boxes narrower than one strip, shorter than one tile, or one
    column wider than a strip, inside the burst.  The burst's Cpm over a
    box this small spreads the slave positions over the whole burst, so
    keys change at almost every output and most windows leave the strip's
    columns (the direct path); resample and coherence match the oracle."""
    w, h = wh
    L, P = c1["L"], c1["P"]
    x0, y0 = 300, 200
    mbox = [x0, x0 + w - 1, y0, y0 + h - 1]
    sbox = [0, P - 1, 0, L - 1]
    ref = _oracle(orc, c1, mbox, sbox, 0, 0.0, c1["ca"], c1["cr"])
    got, dres = rs.run(c1["S"], mbox, sbox, 0, c1["dt"], 0.0, c1["ca"], c1["cr"], c1["kt"],
                       c1["tref"], c1["fdc"], 0)
    assert np.count_nonzero(ref) > 0
    e = rel_err(got, ref)
    N = tops.native
    Lb = N.load()
    Mb = np.ascontiguousarray(c1["M"][y0:y0 + h, x0:x0 + w])
    dm = dev.put(Mb)
    dc = dev.empty(w * h * 4)
    N.check(Lb.tops_coherence(dm.data_ptr(), dres.data_ptr(), w, h, 4, 1, 10, 3, dc.data_ptr(),
                              0, None), "tops_coherence")
    coh = dev.get(dc, np.float32, (h, w))
    cref = orc.coherence(Mb, got.astype(np.complex128), 4, 1, 10, 3)
    ec = float(np.max(np.abs(coh - cref)))
    print(f"  box {w}x{h}: resample max|dz|/rms = {e:.3e}, coherence max|dc| = {ec:.3e}")
    assert e <= TOL_RESAMPLE and ec <= TOL_COH


def test_resample_out_of_box_is_zero(orc, c1, rs):
    """Offsets pushing part of the burst outside the slave box: zero region
    and edge band must match the oracle exactly in structure."""
    L, P = c1["L"], c1["P"]
    box = [0, P - 1, 0, L - 1]
    ca = c1["ca"].copy()
    cr = c1["cr"].copy()
    ca[0] += 40.3       # 40 lines down: the last lines fall off the box
    cr[0] -= 25.6       # 25 pixels left: the first columns fall off
    ref = _oracle(orc, c1, box, box, 0, 0.0, ca, cr)
    got, _ = rs.run(c1["S"], box, box, 0, c1["dt"], 0.0, ca, cr, c1["kt"], c1["tref"],
                    c1["fdc"], 0)
    zero_ref = np.abs(ref) == 0
    assert zero_ref.sum() > 0
    assert np.all(got[zero_ref] == 0)
    e = rel_err(got, ref)
    print(f"  out-of-box case: {zero_ref.sum()} zeros, max|dz|/rms = {e:.3e}")
    assert e <= TOL_RESAMPLE


def test_coherence_small_burst(tops, orc, dev, c1, rs):
    L, P = c1["L"], c1["P"]
    box = [0, P - 1, 0, L - 1]
    got, dres = rs.run(c1["S"], box, box, 0, c1["dt"], 0.0, c1["ca"], c1["cr"], c1["kt"],
                       c1["tref"], c1["fdc"], 0)
    N = tops.native
    Lb = N.load()
    dm = dev.put(c1["M"])
    for (waz, wrg) in [(3, 10), (6, 24), (1, 1), (5, 7), (17, 31)]:
        dx, dy = (wrg - 1) // 2, (waz - 1) // 2
        ref = orc.coherence(c1["M"], got.astype(np.complex128), dx, dy, wrg, waz)
        for mode in (0, 1):
            dc = dev.empty(L * P * 4)
            rc = Lb.tops_coherence(dm.data_ptr(), dres.data_ptr(), P, L, dx, dy, wrg, waz,
                                   dc.data_ptr(), mode, None)
            N.check(rc, "tops_coherence")
            coh = dev.get(dc, np.float32, (L, P))
            e = float(np.max(np.abs(coh - ref)))
            print(f"  coherence {waz}x{wrg} mode={mode}: max|dc| = {e:.3e}, mean = {coh.mean():.4f}")
            assert e <= TOL_COH


@pytest.mark.parametrize("waz,wrg,dx,dy", [(3, 10, 0, 0), (3, 10, 9, 2), (3, 10, 5, 1),
                                            (6, 24, 0, 5), (4, 9, 1, 3), (1, 10, 9, 0)])
def test_coherence_offset_windows(tops, orc, dev, c1, rs, waz, wrg, dx, dy):
    """Windows that start dx columns and dy rows before the sample other than
    the centred (win - 1) / 2 (CoherenceEst's dx, dy arguments,
    EstCoh.cu:101-139, 430-437): the 3 x 10 window off its walking form's
    offsets takes the tile kernel, the others as their size selects; fast
    and FP64 modes against the oracle's window sums."""
    L, P = c1["L"], c1["P"]
    box = [0, P - 1, 0, L - 1]
    got, dres = rs.run(c1["S"], box, box, 0, c1["dt"], 0.0, c1["ca"], c1["cr"], c1["kt"],
                       c1["tref"], c1["fdc"], 0)
    N = tops.native
    Lb = N.load()
    dm = dev.put(c1["M"])
    ref = orc.coherence(c1["M"], got.astype(np.complex128), dx, dy, wrg, waz)
    for mode in (0, 1):
        dc = dev.empty(L * P * 4)
        N.check(Lb.tops_coherence(dm.data_ptr(), dres.data_ptr(), P, L, dx, dy, wrg, waz,
                                  dc.data_ptr(), mode, None), "tops_coherence")
        coh = dev.get(dc, np.float32, (L, P))
        e = float(np.max(np.abs(coh - ref)))
        print(f"  coherence {waz}x{wrg} from ({dy}, {dx}) mode={mode}: max|dc| = {e:.3e}")
        assert e <= TOL_COH


def test_reference_entry_points(tops, orc, c1):
    """DerampDemodResample / _ESD + CoherenceEst through host buffers, as
    Func.cpp:2382 and 2701-2716 call them."""
    N = tops.native
    Lb = N.load()
    L, P = c1["L"], c1["P"]
    box = (C.c_int * 4)(0, P - 1, 0, L - 1)
    out = np.zeros((L, P), np.complex64)
    S = np.ascontiguousarray(c1["S"])
    args = [S.ctypes.data, N.ptr(c1["ca"]), N.ptr(c1["cr"]), 0.0, out.ctypes.data,
            N.ptr(c1["kaz"], C.c_float), N.ptr(c1["krg"], C.c_float), 0, P, L, box, box, L, P,
            c1["dt"], N.ptr(c1["kt"]), N.ptr(c1["tref"]), N.ptr(c1["fdc"]), 12]
    Lb.tops_clear_error()
    Lb.DerampDemodResample(*args)
    assert Lb.tops_last_error() == 0, N.last_error()
    ref = _oracle(orc, c1, [0, P - 1, 0, L - 1], [0, P - 1, 0, L - 1], 0, 0.0, c1["ca"], c1["cr"])
    assert rel_err(out, ref) <= TOL_RESAMPLE
    out2 = np.zeros((L, P), np.complex64)
    args[4] = out2.ctypes.data
    dptr = Lb.DerampDemodResample_ESD(*args)
    assert dptr, N.last_error()
    assert np.array_equal(out, out2)
    coh = np.zeros((L, P), np.float32)
    M = np.ascontiguousarray(c1["M"])
    Lb.CoherenceEst(0, 0, P, L, 4, 1, 10, 3, M.ctypes.data, dptr, N.ptr(coh, C.c_float))
    assert Lb.tops_last_error() == 0, N.last_error()
    cref = orc.coherence(M, out.astype(np.complex128), 4, 1, 10, 3)
    assert float(np.max(np.abs(coh - cref))) <= TOL_COH
    # npoints other than 12 is refused with a status, not a crash
    args[-1] = 6
    Lb.DerampDemodResample(*args)
    assert Lb.tops_last_error() == N.TOPS_ERR_ARG
    Lb.tops_clear_error()


def test_coherence_wide_raster(tops, orc, dev):
    """A raster wider than 2^23 samples: the tile kernels address their
    product loads as a 32-bit byte offset from the window's first sample up
    to that width, and take their bounds-tested loads beyond it.  Output
    slices (the start, the middle, across column 2^23, the end) against the
    oracle on the same slices widened by 32 columns each side (the windows
    reach at most 12 columns across, so the margins hold every term)."""
    N = tops.native
    Lb = N.load()
    h, w = 20, (1 << 23) + 70
    rng = np.random.default_rng(23)
    M = rng.integers(-500, 500, (h, w, 2), dtype=np.int16)
    S = rng.standard_normal((h, w, 2), dtype=np.float32)
    dm, ds = dev.put(M), dev.put(S)
    dc = dev.empty(h * w * 4)
    for waz, wrg in [(6, 24), (5, 7)]:
        dx, dy = (wrg - 1) // 2, (waz - 1) // 2
        N.check(Lb.tops_coherence(dm.data_ptr(), ds.data_ptr(), w, h, dx, dy, wrg, waz,
                                  dc.data_ptr(), 0, None), "tops_coherence")
        coh = dev.get(dc, np.float32, (h, w))
        for c0, c1 in [(0, 200), (w // 2, w // 2 + 200), ((1 << 23) - 100, (1 << 23) + 40),
                       (w - 200, w)]:
            a, b = max(0, c0 - 32), min(w, c1 + 32)
            sub = S[:, a:b, 0].astype(np.float64) + 1j * S[:, a:b, 1]
            ref = orc.coherence(M[:, a:b], sub, dx, dy, wrg, waz)
            e = float(np.max(np.abs(coh[:, c0:c1] - ref[:, c0 - a:c1 - a])))
            print(f"  {waz}x{wrg} columns {c0}..{c1 - 1} of {w}: max|dc| = {e:.3e}")
            assert e <= TOL_COH
