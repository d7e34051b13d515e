"""GPU parity of the deburst (Stitch_OneSubswath, Func.cpp:2821-2933) against
the oracle: bit-exact output for CInt16, Float32 and CFloat32 rasters, over
several burst ranges, a width that takes the 4-byte path, rows outside the
gapped raster (zero lines), the host-buffer entry point and hipGraph replay.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def run(tops):
    import tops_amd.pipeline as pl
    pair = tops.synth.subswath_iw1(n_bursts=3, lines=400, pixels=2048, delta=0.002)
    return pl.SubswathRun(pair)


def _dtype_code(N, a):
    return {np.dtype(np.int16): N.GDT_CINT16, np.dtype(np.float32): N.GDT_FLOAT32,
            np.dtype(np.complex64): N.GDT_CFLOAT32}[a.dtype]


def _device_deburst(tops, dev, sw, b0, bn, gapped, stream=None):
    N = tops.native
    L = N.load()
    y0, h = C.c_int(0), C.c_int(0)
    N.check(L.tops_deburst_extent(C.byref(sw), b0, bn, C.byref(y0), C.byref(h)))
    d_in = dev.put(gapped.view(np.uint8))
    row_bytes = gapped[0].nbytes
    d_out = dev.empty(h.value * row_bytes)
    d_out.fill_(0xAB)            # every output byte must be written
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    N.check(L.tops_deburst(C.byref(sw), b0, bn, _dtype_code(N, gapped), d_in.data_ptr(),
                           d_out.data_ptr(), sp), "tops_deburst")
    return d_in, d_out, dev.get(d_out, gapped.dtype, (h.value,) + gapped.shape[1:])


def _oracle(orc, sw_arrays, lpb, num_lines, dt, t0, b0, bn, gapped):
    first, last, fv, lv = sw_arrays
    out, rows = orc.deburst(first, last, fv, lv, lpb, num_lines, dt, t0, b0, bn, gapped=gapped)
    assert out is not None
    return out, rows


def _run_arrays(run):
    m = run.m
    return (m.burst_first_line_time, m.burst_last_line_time, m.burst_first_valid_line_time,
            m.burst_last_valid_line_time)


@pytest.mark.parametrize("b0,bn", [(0, 2), (1, 1), (1, 2), (0, 0)])
@pytest.mark.parametrize("kind", ["cint16", "float32", "cfloat32"])
def test_deburst_matches_oracle(tops, orc, dev, run, b0, bn, kind):
    lpb, spb = run.lpb, run.spb
    rows = slice(b0 * lpb, (bn + 1) * lpb)
    rng = np.random.default_rng(7)
    if kind == "cint16":
        g = run.m_slc[rows]
    elif kind == "float32":
        g = rng.random(((bn - b0 + 1) * lpb, spb), dtype=np.float32)
    else:
        g = (rng.standard_normal(((bn - b0 + 1) * lpb, spb)) +
             1j * rng.standard_normal(((bn - b0 + 1) * lpb, spb))).astype(np.complex64)
    g = np.ascontiguousarray(g)
    print(f"  deburst {kind} bursts {b0}..{bn}", flush=True)
    _, _, got = _device_deburst(tops, dev, run.deburst_swath(), b0, bn, g)
    ref, _ = _oracle(orc, _run_arrays(run), lpb, run.m.num_lines, run.m.az_time_interval,
                     run.m.first_line_time, b0, bn, g)
    assert got.shape == ref.shape
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))


def _hand_swath(tops, width, first_valid_shift=1.0):
    """3 bursts of 10 lines, dt = 1 s, first line times 0, 6, 12."""
    N = tops.native
    first = np.array([0.0, 6.0, 12.0])
    arrs = [first, first + 9, first + first_valid_shift, first + 8]
    arrs = [np.ascontiguousarray(a) for a in arrs]
    sw = N.DeburstSwath(10, width, 3, 30, 1.0, 0.0, *[N.ptr(a) for a in arrs])
    return sw, arrs


def test_deburst_narrow_rows_and_zero_lines(tops, orc, dev):
    """Width 2047 of CFloat32 (16376-byte rows: the 4-byte path), and output
    rows whose source line lies in burst 0 while only burst 1 is given:
    zero lines."""
    sw, arrs = _hand_swath(tops, 2047, first_valid_shift=-3.0)
    rng = np.random.default_rng(3)
    g = (rng.standard_normal((10, 2047)) + 1j * rng.standard_normal((10, 2047))).astype(
        np.complex64)
    _, _, got = _device_deburst(tops, dev, sw, 1, 1, g)
    ref, rows = _oracle(orc, arrs, 10, 30, 1.0, 0.0, 1, 1, g)
    assert (rows < 0).any()
    assert np.array_equal(got, ref)
    assert not got[rows < 0].any()


def test_deburst_many_runs(tops, orc, dev):
    """A line map of 263 runs of constant source offset (the rounding of
    half-line burst offsets): the gather spans several launches."""
    from test_oracle_deburst import wobble_swath
    N = tops.native
    c = wobble_swath()
    arrs = [np.ascontiguousarray(c[k], dtype=np.float64)
            for k in ("first", "last", "first_valid", "last_valid")]
    width = 1000
    sw = N.DeburstSwath(c["lpb"], width, 4, c["num_lines"], c["dt"], c["first_line_time"],
                        *[N.ptr(a) for a in arrs])
    rng = np.random.default_rng(11)
    g = (rng.standard_normal((4 * c["lpb"], width)) +
         1j * rng.standard_normal((4 * c["lpb"], width))).astype(np.complex64)
    _, _, got = _device_deburst(tops, dev, sw, 0, 3, g)
    ref, rows = _oracle(orc, arrs, c["lpb"], c["num_lines"], c["dt"], c["first_line_time"], 0, 3,
                        g)
    assert 1 + np.count_nonzero(np.diff(rows - np.arange(len(rows)))) > 64
    assert np.array_equal(got, ref)


def test_stitch_subswath_host_buffers(tops, orc, run):
    """tops_stitch_subswath (host rasters in and out) gives the oracle's
    stitch of the master SLC."""
    got = run.stitch(run.m_slc)
    ref, _ = _oracle(orc, _run_arrays(run), run.lpb, run.m.num_lines, run.m.az_time_interval,
                     run.m.first_line_time, 0, run.nb - 1, run.m_slc)
    assert np.array_equal(got, ref)


def test_deburst_errors(tops, dev, run):
    N = tops.native
    L = N.load()
    sw = run.deburst_swath()
    d = dev.empty(16)
    assert L.tops_deburst(C.byref(sw), 0, 2, 7, d.data_ptr(), d.data_ptr(), None) == \
        N.TOPS_ERR_ARG
    assert "unsupported image type" in N.last_error()[1]
    assert L.tops_deburst(C.byref(sw), 0, 3, N.GDT_FLOAT32, d.data_ptr(), d.data_ptr(),
                          None) == N.TOPS_ERR_ARG
    L.tops_clear_error()


def test_deburst_graph_replay(tops, dev, run):
    torch = dev.torch
    N = tops.native
    L = N.load()
    g = np.ascontiguousarray(run.m_slc)
    sw = run.deburst_swath()
    d_in, d_out, eager = _device_deburst(tops, dev, sw, 0, run.nb - 1, g)
    eager = eager.copy()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s, capture_error_mode="thread_local"):
        N.check(L.tops_deburst(C.byref(sw), 0, run.nb - 1, N.GDT_CINT16, d_in.data_ptr(),
                               d_out.data_ptr(), C.c_void_p(s.cuda_stream)))
    for rep in range(3):
        d_out.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert np.array_equal(dev.get(d_out, np.int16, eager.shape), eager), rep
    del graph
