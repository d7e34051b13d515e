// coherence.hip -- boxcar interferometric coherence (gfx950).
//
// Semantics of EstCoh.cu:101-139 (CohMatSet, short2 master) and 413-456
// (Cohbasic): per output (r, c) the window covers rows r-dy .. r-dy+winAz-1
// and columns c-dx .. c-dx+winRg-1 of the burst, zero outside it (the
// reference leaves that padding uninitialised, EstCoh.cu:581-583);
//   coh = |sum M conj(S)| / sqrt(sum|S|^2 * sum|M|^2)  (0 when the product
//   is not positive), and coh = 0 where |S(r, c)| < 0.001.
//
// Fast forms: the 3 az x 10 rg window walks down bands of columns
// (coherence_walk.hpp); any other window stages the products (M conj S,
// |S|^2, |M|^2) of a 16x64 output tile's (16+winAz-1) x (64+winRg-1) window
// in LDS and forms plain FP32 range sums, then short azimuth sums (each
// window sum is a direct sum of its own terms, tile_window.hpp).
// FP64 form: one thread per output, direct FP64 window sums from HBM.
#include "tops_internal.hpp"
#include "tile_window.hpp"
#include "coherence_walk.hpp"

#include <algorithm>

namespace tops {

__global__ void __launch_bounds__(256) coherence_fp64_kernel(const tops_ci16* M, const tops_cf32* S, int w,
                                      int h, int dx, int dy, int win_rg, int win_az,
                                      float* coh)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int r = blockIdx.y;
    if (c >= w) return;
    size_t q0 = (size_t)r * w + c;
    tops_cf32 sc = S[q0];
    double mag = sqrt((double)sc.x * sc.x + (double)sc.y * sc.y);
    if (mag < 0.001) { coh[q0] = 0.0f; return; }
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (int j = 0; j < win_az; j++) {
        int rr = r - dy + j;
        if (rr < 0 || rr >= h) continue;
        for (int k = 0; k < win_rg; k++) {
            int cc = c - dx + k;
            if (cc < 0 || cc >= w) continue;
            size_t q = (size_t)rr * w + cc;
            tops_ci16 m = M[q];
            tops_cf32 s = S[q];
            double mr = m.re, mi = m.im, sr = s.x, si = s.y;
            a0 += mr * sr + mi * si;
            a1 += mi * sr - mr * si;
            a2 += sr * sr + si * si;
            a3 += mr * mr + mi * mi;
        }
    }
    double pp = a2 * a3;
    coh[q0] = pp > 0.0 ? (float)(sqrt(a0 * a0 + a1 * a1) / sqrt(pp)) : 0.0f;
}

// CWR x CWA: a window fixed at compile time (0 x 0: win_rg x win_az)
template <int CWR = 0, int CWA = 0>
__device__ __forceinline__ void coherence_tile(const tops_ci16* M, const tops_cf32* S, int w,
                                               int h, int dx, int dy, int win_rg, int win_az,
                                               float* coh, int tx, int ty, float4* smem)
{
    const int r0 = ty * TW_ROWS, c0 = tx * TW_COLS;
    WindowSums ws[4];
    // (uniform branch; the launch's LDS is tile_window_launch_lds for it)
    if (CWR ? true : tile_window_shared_ok(win_rg, win_az))
        tile_window_sums_shared<CWR, CWA>(M, S, w, h, r0, c0, dx, dy, win_rg, win_az, smem, ws,
                                          nullptr);
    else
        tile_window_sums(M, S, w, h, r0, c0, dx, dy, win_rg, win_az, smem, ws, nullptr);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = c0 + lane;
    if (c >= w) return;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int r = r0 + wv * 4 + k;
        if (r >= h) break;
        size_t q = (size_t)r * w + c;
        tops_cf32 s = S[q];
        float out = 0.0f;
        if (amplitude_ok(s.x * s.x + s.y * s.y)) out = ws[k].coherence();
        coh[q] = out;
    }
}

template <int CWR = 0, int CWA = 0>
__global__ void __launch_bounds__(256) coherence_fast_kernel(const tops_ci16* M, const tops_cf32* S, int w,
                                      int h, int dx, int dy, int win_rg, int win_az,
                                      float* coh)
{
    extern __shared__ float4 smem[];
    coherence_tile<CWR, CWA>(M, S, w, h, dx, dy, win_rg, win_az, coh, blockIdx.x, blockIdx.y,
                             smem);
}

// Batched form: work[i] = {item, tile x, tile y, 0}.
template <int CWR = 0, int CWA = 0>
__global__ void __launch_bounds__(256) coherence_batch_kernel(const CohItem* items, const int4* work,
                                                              int dx, int dy, int win_rg, int win_az)
{
    extern __shared__ float4 smem[];
    const int4 wk = work[blockIdx.x];
    const CohItem& it = items[wk.x];
    coherence_tile<CWR, CWA>(it.m, it.s, it.w, it.h, dx, dy, win_rg, win_az, it.coh, wk.y, wk.z,
                             smem);
}

// the reference's ESDAndCoh window (6 az x 24 rg) takes kernels with the
// window fixed at compile time, its loops unrolled
static bool reference_window(int win_rg, int win_az) { return win_rg == 24 && win_az == 6; }

// Walking form for the 3 az x 10 rg window centred as the stage layer
// centres it (dx = 4, dy = 1): 256 threads = 4 waves, 2 columns per
// thread, so a workgroup walks down a band of CW_BAND = 512 columns over
// the rows it is given (CW_ROWS by default in the batched form, a plan may
// tune it).
using CohWalk3x10 = CohWalk<10, 3, 4, 1>;

// Workgroup (bx, by): band bx (CW_BAND columns), output rows by*rows ..
// by*rows + rows - 1.
__device__ __forceinline__ void coherence_walk_block(const tops_ci16* M, const tops_cf32* S,
                                                     int w, int h, float* coh, int bx, int by,
                                                     int vec, int rows, CohWalk3x10::Lds& sm)
{
    const int cb = bx * CW_BAND;
    const int r0 = by * rows;
    if (cb >= w || r0 >= h) return;   // uniform over the workgroup
    CohWalk3x10::walk(M, S, w, h, coh, cb, r0, min(r0 + rows, h), vec != 0, sm);
}

__global__ void __launch_bounds__(CW_THREADS)
coherence_walk_kernel(const tops_ci16* M, const tops_cf32* S, int w, int h, float* coh, int vec,
                      int rows)
{
    __shared__ CohWalk3x10::Lds sm;
    // blocks numbered bands fastest and dealt to the XCDs in contiguous
    // ranges: bands whose halo columns overlap share an L2
    const int t = xcd_remap((int)(blockIdx.x + blockIdx.y * gridDim.x),
                            (int)(gridDim.x * gridDim.y));
    const int by = t / (int)gridDim.x, bx = t - by * (int)gridDim.x;
    coherence_walk_block(M, S, w, h, coh, bx, by, vec, rows, sm);
}

// Batched form: work[i] = {item, band, row block, vec}, `rows` rows per block.
// The table (bands fastest within a row block) is dealt to the XCDs in
// contiguous ranges, so neighbouring blocks, whose halo rows and columns
// overlap, share an L2; table and items are read through the constant
// address space (scalar loads: read-only, wave-uniform indices).
__global__ void __launch_bounds__(CW_THREADS) coherence_walk_batch_kernel(const CohItem* items,
                                                                          const int4* work,
                                                                          int rows)
{
    __shared__ CohWalk3x10::Lds sm;
    typedef __attribute__((address_space(4))) const int CInt;
    typedef __attribute__((address_space(4))) const CohItem CCohItem;
    const CInt* wk = (const CInt*)work + 4 * xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int item = wk[0], band = wk[1], block = wk[2], vec = wk[3];
    CCohItem* it = (CCohItem*)items + item;
    coherence_walk_block(it->m, it->s, it->w, it->h, it->coh, band, block, vec, rows, sm);
}

bool coherence_walk_fits(int dx, int dy, int win_rg, int win_az)
{
    return win_rg == 10 && win_az == 3 && dx == 4 && dy == 1;
}

// the batched coherence launch serves this window (walking form, or a tile
// whose window fits in LDS); otherwise a plan forms coherence per burst
bool coherence_batch_fits(int dx, int dy, int win_rg, int win_az)
{
    return coherence_walk_fits(dx, dy, win_rg, win_az) ||
           tile_window_launch_lds(win_rg, win_az) != 0;
}

// pair loads: w even, master 8-byte, slave 16-byte, coherence 8-byte aligned
int coherence_walk_vec(const void* m, const void* s, const void* coh, int w)
{
    return (w % 2 == 0) && ((uintptr_t)m % 8 == 0) && ((uintptr_t)s % 16 == 0) &&
           ((uintptr_t)coh % 8 == 0);
}

int coherence_walk_count(int w, int h, int rows, int* nx, int* ny)
{
    *nx = (w + CW_BAND - 1) / CW_BAND;
    *ny = (h + rows - 1) / rows;
    return *nx * *ny;
}

// Rows per workgroup for a single burst: 5 workgroups per CU in one round
// (94 VGPRs allow 5 waves per SIMD, the LDS row buffers 9 workgroups per
// CU), at least 16 rows.
static int coherence_walk_rows(int w, int h)
{
    static std::atomic<int> cache[kMaxCachedDevices];
    const int slots = per_device_cached(cache, 256 * 5, [](int dev) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
        return cus * 5;
    });
    const int nx = (w + CW_BAND - 1) / CW_BAND;
    const int ny = std::max(1, slots / nx);
    return std::max(16, (h + ny - 1) / ny);
}

int coherence_tile_count(int w, int h, int* nx, int* ny)
{
    *nx = (w + TW_COLS - 1) / TW_COLS;
    *ny = (h + TW_ROWS - 1) / TW_ROWS;
    return *nx * *ny;
}

int launch_coherence_batch(const CohItem* d_items, const int4* d_work, int n_work, int dx, int dy,
                           int win_rg, int win_az, int walk_rows, hipStream_t st)
{
    if (n_work <= 0) return TOPS_OK;
    if (coherence_walk_fits(dx, dy, win_rg, win_az)) {
        if (walk_rows <= 0)
            return set_error(TOPS_ERR_ARG, "batched coherence: %d rows per block", walk_rows);
        hipLaunchKernelGGL(coherence_walk_batch_kernel, dim3(n_work), dim3(CW_THREADS), 0, st,
                           d_items, d_work, walk_rows);
        TOPS_HIP(hipGetLastError());
        return TOPS_OK;
    }
    const size_t lds = tile_window_launch_lds(win_rg, win_az);
    if (lds == 0)
        return set_error(TOPS_ERR_ARG, "batched coherence: window %dx%d exceeds the LDS tile",
                         win_az, win_rg);
    if (reference_window(win_rg, win_az))
        hipLaunchKernelGGL((coherence_batch_kernel<24, 6>), dim3(n_work), dim3(256), lds, st,
                           d_items, d_work, dx, dy, win_rg, win_az);
    else
        hipLaunchKernelGGL(coherence_batch_kernel, dim3(n_work), dim3(256), lds, st, d_items,
                           d_work, dx, dy, win_rg, win_az);
    TOPS_HIP(hipGetLastError());
    return TOPS_OK;
}

int launch_coherence(const tops_ci16* m, const tops_cf32* s, int w, int h, int dx,
                     int dy, int win_rg, int win_az, float* coh, int mode, hipStream_t st)
{
    if (mode == TOPS_MODE_FP64 || tile_window_launch_lds(win_rg, win_az) == 0) {
        dim3 g((w + 255) / 256, h);
        hipLaunchKernelGGL(coherence_fp64_kernel, g, dim3(256), 0, st, m, s, w, h, dx, dy,
                           win_rg, win_az, coh);
    } else if (coherence_walk_fits(dx, dy, win_rg, win_az)) {
        const int rows = coherence_walk_rows(w, h);
        const int nx = (w + CW_BAND - 1) / CW_BAND;
        const int ny = (h + rows - 1) / rows;
        hipLaunchKernelGGL(coherence_walk_kernel, dim3(nx, ny), dim3(CW_THREADS), 0, st, m, s, w,
                           h, coh, coherence_walk_vec(m, s, coh, w), rows);
    } else {
        dim3 g((w + TW_COLS - 1) / TW_COLS, (h + TW_ROWS - 1) / TW_ROWS);
        if (reference_window(win_rg, win_az))
            hipLaunchKernelGGL((coherence_fast_kernel<24, 6>), g, dim3(256),
                               tile_window_launch_lds(win_rg, win_az), st, m, s, w, h, dx, dy,
                               win_rg, win_az, coh);
        else
            hipLaunchKernelGGL(coherence_fast_kernel, g, dim3(256),
                               tile_window_launch_lds(win_rg, win_az), st, m, s, w, h, dx, dy,
                               win_rg, win_az, coh);
    }
    TOPS_HIP(hipGetLastError());
    return TOPS_OK;
}

}  // namespace tops
