"""Deburst (Stitch_OneSubswath, Func.cpp:2821-2933; line map
getLineIdxInMetaData, Func.cpp:2766-2819): the oracle pinned by hand-worked
maps and by properties of the rule, then the library's host line map
(tops_deburst_extent / tops_deburst_rows, no GPU work) against the oracle.

Hand-worked swath A: 3 bursts of 10 lines, dt = 1 s, first line times 0, 6,
12 (4 overlapping lines between neighbours), last line time = first + 9,
valid lines 1 .. 8, 30 subswath lines, subswath first line time 0.
"""
import ctypes as C

import numpy as np
import pytest

FIRST_A = np.array([0.0, 6.0, 12.0])
LAST_A = FIRST_A + 9


def _a(b0, bn, fv=1, lv=8):
    return dict(first=FIRST_A, last=LAST_A, first_valid=FIRST_A + fv, last_valid=FIRST_A + lv,
                lpb=10, num_lines=30, dt=1.0, first_line_time=0.0, b0=b0, bn=bn)


def _oracle_rows(orc, c):
    return orc.deburst(c["first"], c["last"], c["first_valid"], c["last_valid"], c["lpb"],
                       c["num_lines"], c["dt"], c["first_line_time"], c["b0"], c["bn"])[1]


# output row -> gapped row, worked by hand from the rule: in an overlap the
# earlier burst's line up to the mid time (last[b] + first[b+1]) / 2, the
# later one's after it; the last line of a burst is outside it (t < last)
HAND = [
    # all three bursts: line times 1 .. 20; seams at 7.5 and 13.5
    ((0, 2, 1, 8), [1, 2, 3, 4, 5, 6, 7, 12, 13, 14, 15, 16, 17, 22, 23, 24, 25, 26, 27, 28]),
    # burst 1 alone: times 7 .. 14, every line from burst 1 (the neighbour
    # outside the range never wins an overlap)
    ((1, 1, 1, 8), [1, 2, 3, 4, 5, 6, 7, 8]),
    # bursts 1 .. 2: times 7 .. 20; the overlap with burst 0 goes to burst 1
    ((1, 2, 1, 8), [1, 2, 3, 4, 5, 6, 7, 12, 13, 14, 15, 16, 17, 18]),
    # whole bursts valid: times 0 .. 21; t = 21 is the last burst's last line
    # time, in no burst, and maps to the subswath's last line (29)
    ((0, 2, 0, 9), [0, 1, 2, 3, 4, 5, 6, 7, 12, 13, 14, 15, 16, 17, 22, 23, 24, 25, 26, 27,
                    28, 29]),
]


@pytest.mark.parametrize("case,expect", HAND)
def test_oracle_deburst_hand_maps(orc, case, expect):
    b0, bn, fv, lv = case
    assert _oracle_rows(orc, _a(b0, bn, fv, lv)).tolist() == expect


def test_oracle_deburst_gap_is_an_error(orc):
    """Bursts that touch without overlapping (first[i+1] = last[i] + dt)
    leave each burst's last line time in no burst: the reference exits
    there."""
    f = np.array([0.0, 10.0, 20.0])
    out, rows = orc.deburst(f, f + 9, f, f + 9, 10, 30, 1.0, 0.0, 0, 2)
    assert out is None and rows is None


def test_oracle_deburst_extent(orc):
    # yMin / yMax truncate toward zero
    assert orc.deburst_extent(1.0, 20.0, 0.0, 1.0) == (1, 20)
    assert orc.deburst_extent(1.7, 20.9, 0.2, 1.0) == (1, 20)


def test_oracle_deburst_copies_rows(orc):
    """The stitched raster is the gapped raster's rows in map order, zero
    where the map leaves the gapped raster."""
    c = _a(1, 1)
    c["first_valid"] = FIRST_A - 3.0        # starts 3 lines before burst 1's first line
    g = np.arange(10 * 5, dtype=np.float32).reshape(10, 5) + 1
    out, rows = orc.deburst(c["first"], c["last"], c["first_valid"], c["last_valid"], 10, 30,
                            1.0, 0.0, 1, 1, gapped=g)
    assert rows[0] < 0                      # a line of burst 0: outside burst 1's raster
    for yy, r in enumerate(rows):
        if 0 <= r < 10:
            assert np.array_equal(out[yy], g[r])
        else:
            assert not out[yy].any()


@pytest.fixture(scope="module")
def iw(tops):
    """IW-like metadata: 9 bursts of 1500 lines, ~100-line overlaps, 40
    invalid lines at each end of a burst, dt = 2.055556 ms, times near
    43200 s (metadata only)."""
    return tops.synth.subswath_iw1(pixels=8).master


@pytest.mark.parametrize("b0,bn", [(0, 8), (2, 5), (4, 4), (0, 0), (8, 8)])
def test_oracle_deburst_rule_properties(orc, iw, b0, bn):
    dt, lpb = iw.az_time_interval, iw.lines_per_burst
    first, last = iw.burst_first_line_time, iw.burst_last_line_time
    fv, lv = iw.burst_first_valid_line_time, iw.burst_last_valid_line_time
    _, rows = orc.deburst(first, last, fv, lv, lpb, iw.num_lines, dt, iw.first_line_time, b0, bn)
    y0, h = orc.deburst_extent(fv[b0], lv[bn], iw.first_line_time, dt)
    assert len(rows) == h
    t = fv[b0] + np.arange(h) * dt
    line = rows + b0 * lpb                       # subswath line
    burst, j = line // lpb, line % lpb
    # every output line comes from a burst of the range, at the line nearest
    # its time
    assert burst.min() >= b0 and burst.max() <= bn
    src_t = first[burst] + j * dt
    assert np.max(np.abs(src_t - t)) <= 0.5 * dt * (1 + 1e-6)
    # first output row: burst0's first valid line; last: burstN's last valid
    # line to within one line (yMin and yMax truncate their own line counts
    # from the subswath start, Func.cpp:2837-2838)
    assert (burst[0], j[0]) == (b0, iw.first_valid_line[b0])
    assert burst[-1] == bn and abs(int(j[-1]) - int(iw.last_valid_line[bn])) <= 1
    # burst switches only forward, at the mid time of each overlap
    assert np.all(np.diff(burst) >= 0)
    for b in range(b0, bn):
        mid = 0.5 * (last[b] + first[b + 1])
        assert np.all(burst[(t < mid - dt) & (burst <= b + 1)] <= b)
        assert np.all(burst[t > mid + dt] >= b + 1)
    # lines increase by one per row inside a burst
    same = np.diff(burst) == 0
    assert np.all(np.diff(j)[same] == 1)


# ---- the library's host line map against the oracle ----------------------
def _lib_rows(tops, c):
    N = tops.native
    L = N.load()
    arrs = [np.ascontiguousarray(c[k], dtype=np.float64)
            for k in ("first", "last", "first_valid", "last_valid")]
    sw = N.DeburstSwath(c["lpb"], 4, len(arrs[0]), c["num_lines"], c["dt"], c["first_line_time"],
                        *[N.ptr(a) for a in arrs])
    y0, h = C.c_int(0), C.c_int(0)
    rc = L.tops_deburst_extent(C.byref(sw), c["b0"], c["bn"], C.byref(y0), C.byref(h))
    if rc != N.TOPS_OK:
        return rc, None, None
    rows = np.zeros(h.value, np.int32)
    N.check(L.tops_deburst_rows(C.byref(sw), c["b0"], c["bn"], N.ptr(rows, C.c_int)))
    return rc, y0.value, rows


@pytest.mark.parametrize("case,expect", HAND)
def test_library_deburst_rows_hand(tops, case, expect):
    rc, _, rows = _lib_rows(tops, _a(*case))
    assert rc == 0 and rows.tolist() == expect


def test_library_deburst_gap_and_bad_args(tops):
    N = tops.native
    f = np.array([0.0, 10.0, 20.0])
    c = dict(first=f, last=f + 9, first_valid=f, last_valid=f + 9, lpb=10, num_lines=30, dt=1.0,
             first_line_time=0.0, b0=0, bn=2)
    rc, _, _ = _lib_rows(tops, c)
    assert rc == N.TOPS_ERR_ARG and "no burst" in N.last_error()[1]
    for b0, bn in ((-1, 1), (2, 1), (0, 3)):
        rc, _, _ = _lib_rows(tops, dict(_a(0, 2), b0=b0, bn=bn))
        assert rc == N.TOPS_ERR_ARG


@pytest.mark.parametrize("b0,bn", [(0, 8), (2, 5), (4, 4), (7, 8)])
def test_library_deburst_rows_match_oracle(tops, orc, iw, b0, bn):
    c = dict(first=iw.burst_first_line_time, last=iw.burst_last_line_time,
             first_valid=iw.burst_first_valid_line_time,
             last_valid=iw.burst_last_valid_line_time, lpb=iw.lines_per_burst,
             num_lines=iw.num_lines, dt=iw.az_time_interval, first_line_time=iw.first_line_time,
             b0=b0, bn=bn)
    rc, y0, rows = _lib_rows(tops, c)
    assert rc == 0
    oy0, _ = orc.deburst_extent(c["first_valid"][b0], c["last_valid"][bn], c["first_line_time"],
                                c["dt"])
    assert y0 == oy0
    assert np.array_equal(rows, _oracle_rows(orc, c))


def wobble_swath():
    """4 bursts of 400 lines, dt = 0.3 s, odd bursts offset by half a line:
    (t - first) / dt sits on .5 in those bursts and the rounding of
    getLineIdxInMetaData alternates between neighbouring lines (263 runs of
    constant source offset)."""
    dt, lpb, nb = 0.3, 400, 4
    first = np.array([43200.0 + i * 300 * dt + (0.15 if i % 2 else 0.0) for i in range(nb)])
    return dict(first=first, last=first + (lpb - 1) * dt, first_valid=first + 2 * dt,
                last_valid=first + (lpb - 3) * dt, lpb=lpb, num_lines=nb * lpb, dt=dt,
                first_line_time=float(first[0]), b0=0, bn=nb - 1)


def test_library_deburst_rows_wobble(tops, orc):
    c = wobble_swath()
    rc, _, rows = _lib_rows(tops, c)
    assert rc == 0
    ref = _oracle_rows(orc, c)
    assert np.array_equal(rows, ref)
    runs = 1 + np.count_nonzero(np.diff(ref - np.arange(len(ref))))
    assert runs > 64
