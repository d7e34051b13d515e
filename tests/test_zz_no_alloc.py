"""tops_esd_shifts_noalloc: the allocation-free ESD entry point (partials and
band sums in a fixed per-GPU device array) against tops_esd_shifts, the
entry point that allocates its partials per call (EstESD.cu:914-1217 flow).

Both run the same kernels in the same launch shapes, so the band sums and
shifts must be bit-identical.  Two calls in a row, the second on fewer and
smaller bands, check that nothing of the first call leaks into the second.
Calls beyond the fixed array are refused before any device work (this part
needs no GPU)."""
import ctypes as C

import numpy as np
import pytest


def _bands(seed, sizes, pixels):
    """Forward / backward master (CInt16) and resampled slave (CF32) bands,
    concatenated as EstESDShifts takes them; slave = master x a band phase
    ramp + noise, so the coherence passes the 0.3 threshold."""
    rng = np.random.default_rng(seed)
    n = int(np.sum(sizes))
    out = []
    for k in range(2):
        m = np.round(rng.normal(0, 300, (n, pixels, 2))).astype(np.int16)
        mc = m[..., 0].astype(np.float64) + 1j * m[..., 1]
        th = 0.3 + 0.2 * k + 1e-4 * np.arange(pixels)[None, :]
        s = mc * np.exp(1j * th) + rng.normal(0, 60, (n, pixels)) + 1j * rng.normal(0, 60, (n, pixels))
        out.append(np.ascontiguousarray(m))
        out.append(np.ascontiguousarray(s.astype(np.complex64)))
    mf, sf, mb, sb = out
    return mf, mb, sf, sb


def _call(L, N, fn, d, sizes, pixels, df, dt):
    n = len(sizes)
    shifts, sums = np.zeros(n), np.zeros(2 * n)
    rc = fn(d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr(),
            sizes.ctypes.data_as(C.POINTER(C.c_int)), n, pixels, 6, 24, 0.3, df, dt,
            N.ptr(shifts), N.ptr(sums), None)
    N.check(rc, fn.__name__)
    return shifts, sums


@pytest.mark.gpu
def test_noalloc_matches_allocating_entry(tops, dev):
    N = tops.native
    L = N.load()
    df, dt = 4.9e3 * 2.055556e-3 / 1.5, 2.055556e-3
    cases = [(11, np.array([100, 97, 103, 1, 100, 64, 17, 100], np.int32), 3001),
             (12, np.array([40, 9, 33], np.int32), 1000)]
    for seed, sizes, pixels in cases:
        print(f"  noalloc esd: seed {seed}, {len(sizes)} bands of {list(sizes)} lines x {pixels}",
              flush=True)
        d = [dev.put(a) for a in _bands(seed, sizes, pixels)]
        ref_sh, ref_sums = _call(L, N, L.tops_esd_shifts, d, sizes, pixels, df, dt)
        got_sh, got_sums = _call(L, N, L.tops_esd_shifts_noalloc, d, sizes, pixels, df, dt)
        print(f"    shifts {got_sh} (allocating entry {ref_sh})", flush=True)
        assert np.all(np.isfinite(ref_sh)) and np.any(ref_sums != 0.0)
        assert np.array_equal(got_sums, ref_sums), (got_sums, ref_sums)
        assert np.array_equal(got_sh, ref_sh)


def test_noalloc_refuses_beyond_its_array(tops):
    """Refused with TOPS_ERR_ARG before any HIP call (no device needed)."""
    N = tops.native
    L = N.load()
    shifts, sums = np.zeros(64), np.zeros(128)
    null = C.c_void_p(0)

    def call(sizes, pixels):
        sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        L.tops_clear_error()
        return L.tops_esd_shifts_noalloc(null, null, null, null,
                                         sizes.ctypes.data_as(C.POINTER(C.c_int)), len(sizes),
                                         pixels, 6, 24, 0.3, 1.0, 1.0, N.ptr(shifts), N.ptr(sums),
                                         None)

    # 8 bands x 2048 tiles of 16 lines x 64 samples x 17 tile rows > 262144 tiles
    rc = call([272] * 8, 2048 * 64)
    assert rc == N.TOPS_ERR_ARG
    assert b"fixed scratch" in L.tops_last_error_message()
    rc = call([10] * 65, 64)
    assert rc == N.TOPS_ERR_ARG
    assert b"exceed the limit" in L.tops_last_error_message()
    rc = call([10, -1], 64)
    assert rc == N.TOPS_ERR_ARG
    L.tops_clear_error()
