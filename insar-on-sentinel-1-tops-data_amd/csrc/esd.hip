// esd.hip -- enhanced spectral diversity over burst-overlap bands (gfx950).
//
// Semantics of EstESDShifts_NonOverlap (EstESD.cu:914-1217): for overlap i
// with n_i lines, over the forward band (last n_i lines of burst i) and the
// backward band (first n_i lines of burst i+1):
//   coh_f, coh_b : boxcar coherence, window winAz x winRg starting
//                  (winAz-1)/2 rows and (winRg-1)/2 columns before the
//                  sample, zero outside the band, no amplitude mask
//                  (CohMatrixSet, CohFirstLineSet_ESD, CohDiffer_vertical_ESD,
//                  shfl_vertical_ESD, CpCoh2_ESD)
//   coh          : (coh_f + coh_b) / 2                       (Divide2)
//   dd           : (Mf conj Sf) conj(Mb conj Sb)              (DoubleDifference4OneD)
//   phasor       : 0 if coh < thr or coh >= 1, else exp(j arg dd) (MaskOut;
//                  a zero dd gives 1 or -1 by the sign of its real part,
//                  as atan2 of signed zeros does there)
//   sum_i        : sum of phasors over the band               (reduce0 x3)
// The shift itself, atan2(sum_i) / (2 pi df dt), is formed on the host, or
// on the device by esd_shift_kernel for the device-resident two-pass flow.
//
// One fused kernel per call: each 16x64 tile computes both band coherences
// from LDS (tile_window_sums_shared for the reference's 6x24 window and
// others it serves, tile_window_sums otherwise), the double difference and
// the masked phasors, sums them per thread with Kahan FP32, reduces the
// workgroup by wave shuffles and writes one FP64 partial; a second kernel
// adds the partials of each band in FP64 in a fixed order.
#include "tops_internal.hpp"
#include "tile_window.hpp"

#include <mutex>
#include <vector>

namespace tops {

struct EsdBand {
    int f_offset;      // first line of the forward band in mf / sf
    int b_offset;      // first line of the backward band in mb / sb
    int lines;         // n_i
};

constexpr int kMaxEsdBands = 64;

struct EsdParams {
    const tops_ci16* mf;
    const tops_ci16* mb;
    const tops_cf32* sf;
    const tops_cf32* sb;
    int pixels, win_az, win_rg;
    float thr;
    int tiles_x, tiles_y;
    int n_bands;
    EsdBand band[kMaxEsdBands];
    double2* partials;   // [n_bands][tiles_y][tiles_x]
};

// a / |a| for a nonzero a: the squared magnitude scaled by the hardware
// reciprocal square root (about 1 ulp); a tiny a is first scaled by 2^64
// (exactly), so its squared magnitude stays a normal number
__device__ inline float2 unit_phasor(float2 a)
{
    float n2 = a.x * a.x + a.y * a.y;
    if (n2 < 1e-30f) {
        a.x *= 0x1p64f;
        a.y *= 0x1p64f;
        n2 = a.x * a.x + a.y * a.y;
    }
    const float r = __builtin_amdgcn_rsqf(n2);
    return make_float2(a.x * r, a.y * r);
}

__device__ inline float wave_sum(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// kShared: the shared window sums (tile_window_sums_shared); CWR x CWA: a
// window fixed at compile time (0 x 0: the parameter block's)
template <bool kShared, int CWR = 0, int CWA = 0>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) esd_band_kernel(EsdParams p)
{
    extern __shared__ float4 smem[];
    __shared__ float red[4][2];
    // tiles numbered band rows fastest and dealt to the XCDs in contiguous
    // ranges: the neighbours whose halos overlap a tile's own run on its XCD
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int ty = t % p.tiles_y, tx = (t / p.tiles_y) % p.tiles_x;
    const int b = t / (p.tiles_y * p.tiles_x);
    const int n = p.band[b].lines;
    const int r0 = ty * TW_ROWS, c0 = tx * TW_COLS;
    double2* out = p.partials + ((size_t)b * p.tiles_y + ty) * p.tiles_x + tx;
    if (r0 >= n) {   // whole workgroup leaves together: no barrier skipped
        if (threadIdx.x == 0) *out = make_double2(0.0, 0.0);
        return;
    }
    const size_t off_f = (size_t)p.band[b].f_offset * p.pixels;
    const size_t off_b = (size_t)p.band[b].b_offset * p.pixels;
    const int win_az = CWA ? CWA : p.win_az, win_rg = CWR ? CWR : p.win_rg;
    const int dh = (win_az - 1) / 2, dw = (win_rg - 1) / 2;
    // the two bands in turn through one copy of the window code (a loop kept
    // rolled, so both share its registers): coherence and centre product of
    // each output sample
    float coh_f[4], coh_b[4];
    float2 cf[4], cb[4];
#pragma unroll 1
    for (int band = 0; band < 2; band++) {
        const tops_ci16* M = band ? p.mb + off_b : p.mf + off_f;
        const tops_cf32* S = band ? p.sb + off_b : p.sf + off_f;
        WindowSums ws[4];
        float2 c[4];
        if (kShared)
            tile_window_sums_shared<CWR, CWA>(M, S, p.pixels, n, r0, c0, dw, dh, win_rg, win_az,
                                              smem, ws, c);
        else
            tile_window_sums(M, S, p.pixels, n, r0, c0, dw, dh, win_rg, win_az, smem, ws, c);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float co = ws[k].coherence();
            if (band == 0) {
                coh_f[k] = co;
                cf[k] = c[k];
            } else {
                coh_b[k] = co;
                cb[k] = c[k];
            }
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float sx = 0.f, sy = 0.f, ex = 0.f, ey = 0.f;
    if (c0 + lane < p.pixels) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            int r = r0 + wv * 4 + k;
            if (r >= n) break;
            float coh = (coh_f[k] + coh_b[k]) / 2.0f;
            if (coh < p.thr || coh >= 1.0f) continue;
            // exp(j arg(a conj b)), a = Mf Sf*, b = Mb Sb* (MaskOut's
            // (cos, sin) of atan2 of the double difference)
            float px, py;
            if ((cf[k].x != 0.0f || cf[k].y != 0.0f) && (cb[k].x != 0.0f || cb[k].y != 0.0f)) {
                // from the unit phasors of a and b
                const float2 ua = unit_phasor(cf[k]), ub = unit_phasor(cb[k]);
                const float dx = ua.x * ub.x + ua.y * ub.y, dy = ua.y * ub.x - ua.x * ub.y;
                const float id = __builtin_amdgcn_rsqf(dx * dx + dy * dy);
                px = dx * id;
                py = dy * id;
            } else {
                // a or b is zero, so the double difference is a pair of
                // signed zeros: atan2 gives +-0 (phasor 1) or +-pi (phasor
                // -1) by the sign of its real part, which the products give
                // as the reference's complex multiplications do
                const float dr = cf[k].x * cb[k].x + cf[k].y * cb[k].y;
                px = __builtin_signbit(dr) ? -1.0f : 1.0f;
                py = 0.0f;
            }
            kahan_add(sx, ex, px);
            kahan_add(sy, ey, py);
        }
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    if (lane == 0) { red[wv][0] = sx; red[wv][1] = sy; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tx = 0.0, ty = 0.0;
        for (int w = 0; w < 4; w++) { tx += red[w][0]; ty += red[w][1]; }
        *out = make_double2(tx, ty);
    }
}

__global__ void __launch_bounds__(256) esd_final_kernel(const double2* partials, int per_band, double* sums)
{
    __shared__ double sx[256], sy[256];
    const double2* q = partials + (size_t)blockIdx.x * per_band;
    double ax = 0.0, ay = 0.0;
    for (int i = threadIdx.x; i < per_band; i += 256) { ax += q[i].x; ay += q[i].y; }
    sx[threadIdx.x] = ax;
    sy[threadIdx.x] = ay;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sx[threadIdx.x] += sx[threadIdx.x + s];
            sy[threadIdx.x] += sy[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sums[2 * blockIdx.x] = sx[0];
        sums[2 * blockIdx.x + 1] = sy[0];
    }
}

// shift_i = atan2(sum_i) / (2 pi df dt) (EstESD.cu:1213-1217), 0 when its
// magnitude exceeds 100 lines (Func.cpp:2612-2621)
__global__ void esd_shift_kernel(const double* sums, int n, double df, double dt, double* shifts)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sh = atan2(sums[2 * i + 1], sums[2 * i]) / (2 * kPi * df * dt);
    shifts[i] = fabs(sh) > 100 ? 0.0 : sh;
}

int launch_esd_shifts(const double* d_sums, int n, double df, double dt, double* d_shifts,
                      hipStream_t st)
{
    if (n <= 0) return TOPS_OK;
    hipLaunchKernelGGL(esd_shift_kernel, dim3((n + 63) / 64), dim3(64), 0, st, d_sums, n, df, dt,
                       d_shifts);
    TOPS_HIP(hipGetLastError());
    return TOPS_OK;
}

int esd_partial_capacity(const int* overlap_size, int n_overlap, int pixels)
{
    int maxn = 0;
    for (int i = 0; i < n_overlap; i++) maxn = overlap_size[i] > maxn ? overlap_size[i] : maxn;
    // the launch's grid: at least one tile row, also when every band is empty
    int tx = (pixels + TW_COLS - 1) / TW_COLS, ty = maxn > 0 ? (maxn + TW_ROWS - 1) / TW_ROWS : 1;
    return n_overlap * tx * ty;
}

int launch_esd(const tops_ci16* mf, const tops_ci16* mb, const tops_cf32* sf,
               const tops_cf32* sb, const int* overlap_size, int n_overlap, int pixels,
               int win_az, int win_rg, float thr, double* d_sums, double* d_partials,
               int /*max_blocks_per_band*/, hipStream_t st)
{
    // bands concatenated in order in all four arrays
    std::vector<int> off(n_overlap > 0 ? n_overlap : 1, 0);
    for (int i = 1; i < n_overlap; i++) off[i] = off[i - 1] + overlap_size[i - 1];
    return launch_esd_bands(mf, mb, sf, sb, off.data(), off.data(), overlap_size, n_overlap,
                            pixels, win_az, win_rg, thr, d_sums, d_partials, st);
}

int launch_esd_bands(const tops_ci16* mf, const tops_ci16* mb, const tops_cf32* sf,
                     const tops_cf32* sb, const int* f_offset, const int* b_offset,
                     const int* overlap_size, int n_overlap, int pixels, int win_az, int win_rg,
                     float thr, double* d_sums, double* d_partials, hipStream_t st)
{
    if (n_overlap <= 0) return TOPS_OK;
    if (n_overlap > kMaxEsdBands)
        return set_error(TOPS_ERR_ARG, "EstESDShifts: at most %d overlaps per call", kMaxEsdBands);
    const size_t lds = tile_window_launch_lds(win_rg, win_az);
    if (lds == 0)
        return set_error(TOPS_ERR_ARG, "EstESDShifts: window %dx%d exceeds the LDS tile", win_az,
                         win_rg);
    EsdParams p;
    p.mf = mf; p.mb = mb; p.sf = sf; p.sb = sb;
    p.pixels = pixels; p.win_az = win_az; p.win_rg = win_rg; p.thr = thr;
    int maxn = 0;
    for (int i = 0; i < n_overlap; i++) {
        p.band[i].f_offset = f_offset[i];
        p.band[i].b_offset = b_offset[i];
        p.band[i].lines = overlap_size[i];
        if (overlap_size[i] > maxn) maxn = overlap_size[i];
    }
    p.n_bands = n_overlap;
    p.tiles_x = (pixels + TW_COLS - 1) / TW_COLS;
    p.tiles_y = maxn > 0 ? (maxn + TW_ROWS - 1) / TW_ROWS : 1;
    p.partials = (double2*)d_partials;
    const dim3 grid(p.tiles_x * p.tiles_y * n_overlap);
    if (win_rg == 24 && win_az == 6)   // the reference's ESD window, unrolled
        hipLaunchKernelGGL((esd_band_kernel<true, 24, 6>), grid, dim3(256), lds, st, p);
    else if (tile_window_shared_ok(win_rg, win_az))
        hipLaunchKernelGGL(esd_band_kernel<true>, grid, dim3(256), lds, st, p);
    else
        hipLaunchKernelGGL(esd_band_kernel<false>, grid, dim3(256), lds, st, p);
    TOPS_HIP(hipGetLastError());
    hipLaunchKernelGGL(esd_final_kernel, dim3(n_overlap), dim3(256), 0, st,
                       (const double2*)d_partials, p.tiles_x * p.tiles_y, d_sums);
    TOPS_HIP(hipGetLastError());
    return TOPS_OK;
}

// ---- allocation-free ESD -------------------------------------------------
// The tile partials and the band sums of tops_esd_shifts_noalloc live in one
// fixed device array.  Every GPU the library runs on holds its own copy until
// the process ends; its address on each device is looked up once, with that
// device current, and kept here.  The host name `scratch` is only a stand-in
// for the symbol: kernels and copies take the looked-up address.
constexpr int kEsdNoAllocTiles = 262144;   // tile partials: 4 MiB of double2
static __device__ __align__(256) double2 scratch[kEsdNoAllocTiles + kMaxEsdBands];

constexpr int kMaxDevices = 64;
static double2* g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

static int esd_scratch_ptr(double2** out)
{
    int dev = -1;
    TOPS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(TOPS_ERR_NODEVICE, "tops_esd_shifts_noalloc: device %d beyond %d", dev,
                         kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        TOPS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = (double2*)p;
    }
    *out = g_scratch_dev[dev];
    return TOPS_OK;
}

}  // namespace tops

using namespace tops;

extern "C" int tops_esd_shifts_noalloc(const tops_ci16* d_m_for, const tops_ci16* d_m_back,
                                       const tops_cf32* d_s_for, const tops_cf32* d_s_back,
                                       const int* overlap_size, int n_overlap, int pixels,
                                       int win_az, int win_rg, float threshold,
                                       double spectral_sep, double az_time_interval,
                                       double* shifts, double* sums, void* stream)
{
    // every check comes before the first HIP call
    if (n_overlap < 0 || (n_overlap > 0 && (!overlap_size || !shifts || pixels <= 0)))
        return set_error(TOPS_ERR_ARG, "tops_esd_shifts_noalloc: bad arguments");
    if (n_overlap == 0) return TOPS_OK;
    if (n_overlap > kMaxEsdBands)
        return set_error(TOPS_ERR_ARG, "tops_esd_shifts_noalloc: %d overlaps exceed the limit of %d",
                         n_overlap, kMaxEsdBands);
    for (int i = 0; i < n_overlap; i++)
        if (overlap_size[i] < 0)
            return set_error(TOPS_ERR_ARG, "tops_esd_shifts_noalloc: negative overlap");
    int maxn = 0;
    for (int i = 0; i < n_overlap; i++) maxn = overlap_size[i] > maxn ? overlap_size[i] : maxn;
    // the launch's grid (esd_partial_capacity), counted without int overflow
    const long tiles = (long)n_overlap * ((pixels + TW_COLS - 1L) / TW_COLS) *
                       (maxn > 0 ? (maxn + TW_ROWS - 1L) / TW_ROWS : 1L);
    if (tiles > kEsdNoAllocTiles)
        return set_error(TOPS_ERR_ARG,
                         "tops_esd_shifts_noalloc: %ld tiles of 16 x 64 exceed the fixed scratch "
                         "of %d tiles (use tops_esd_shifts)", tiles, kEsdNoAllocTiles);
    if (tile_window_launch_lds(win_rg, win_az) == 0)
        return set_error(TOPS_ERR_ARG, "tops_esd_shifts_noalloc: window %dx%d exceeds the LDS tile",
                         win_az, win_rg);
    double2* base = nullptr;
    int rc = esd_scratch_ptr(&base);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    // bands concatenated in order in all four arrays (EstESDShifts layout);
    // every tile writes its partial and every band its sum, so nothing in the
    // array needs clearing between calls
    int off[kMaxEsdBands];
    off[0] = 0;
    for (int i = 1; i < n_overlap; i++) off[i] = off[i - 1] + overlap_size[i - 1];
    double* d_sums = (double*)(base + kEsdNoAllocTiles);
    rc = launch_esd_bands(d_m_for, d_m_back, d_s_for, d_s_back, off, off, overlap_size, n_overlap,
                          pixels, win_az, win_rg, threshold, d_sums, (double*)base, s);
    if (rc) return rc;
    double h[2 * kMaxEsdBands];
    TOPS_HIP(hipMemcpyAsync(h, d_sums, sizeof(double) * 2 * n_overlap, hipMemcpyDeviceToHost, s));
    TOPS_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n_overlap; i++) {
        // EstESD.cu:1213-1217
        shifts[i] = atan2(h[2 * i + 1], h[2 * i]) / (2 * kPi * spectral_sep * az_time_interval);
        if (sums) {
            sums[2 * i] = h[2 * i];
            sums[2 * i + 1] = h[2 * i + 1];
        }
    }
    return TOPS_OK;
}
