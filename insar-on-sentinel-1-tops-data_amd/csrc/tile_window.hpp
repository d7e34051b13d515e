// tile_window.hpp -- LDS-tiled boxcar window sums shared by the coherence
// kernel (EstCoh.cu Cohbasic semantics) and the ESD band coherence
// (EstESD.cu:94-128, 297-546 semantics).
//
// A 256-thread workgroup owns a 16x64 tile of output samples; lane = column,
// wave w = rows 4w..4w+3.  For window winAz x winRg starting dy rows and dx
// columns before the output sample (zero outside [0,h) x [0,w)):
//   1. the products P = (M conj S, |S|^2, |M|^2) of the (16+winAz-1) x
//      (64+winRg-1) input window go to LDS as float4;
//   2. direct FP32 range sums over winRg go to LDS (error <= (winRg-1) 2^-24
//      of the window's sum of magnitudes, i.e. <= 2e-6 relative for the
//      reference's 24-sample window: no compensation needed for sums this
//      short);
//   3. each output adds winAz range sums.
// tile_window_sums_shared computes the same window sums with the work of
// neighbouring windows shared (runs of 8 range windows, 4 azimuth windows
// per thread); every sum is still a plain sum of the window's own terms.
#pragma once

#include "tops_internal.hpp"

namespace tops {

constexpr int TW_ROWS = 16;
constexpr int TW_COLS = 64;
constexpr size_t TW_LDS_LIMIT = 64 * 1024;

__host__ __device__ inline size_t tile_window_lds_bytes(int win_rg, int win_az)
{
    size_t pr = TW_ROWS + win_az - 1, pc = TW_COLS + win_rg - 1;
    return (pr * pc + pr * TW_COLS) * sizeof(float4);
}

inline bool tile_window_fits(int win_rg, int win_az)
{
    return win_rg >= 1 && win_az >= 1 && tile_window_lds_bytes(win_rg, win_az) <= TW_LDS_LIMIT;
}

// |sum M S*| / sqrt(sum |S|^2 sum |M|^2), 0 unless the product is positive,
// with the hardware square root and reciprocal square root (about 1 ulp
// each; the coherence tolerance is 1e-4 absolute)
__device__ __forceinline__ float coherence_of(float a0, float a1, float a2, float a3)
{
    const float pp = a2 * a3;
    return pp > 0.0f ? __builtin_amdgcn_sqrtf(a0 * a0 + a1 * a1) * __builtin_amdgcn_rsqf(pp)
                     : 0.0f;
}

// |S| >= 0.001 (EstCoh.cu:446-449) from |S|^2
__device__ __forceinline__ bool amplitude_ok(float s2)
{
    return __builtin_amdgcn_sqrtf(s2) >= 0.001f;
}

struct WindowSums {
    float a0, a1, a2, a3;   // sum Re(M S*), sum Im(M S*), sum |S|^2, sum |M|^2
    __device__ float coherence() const { return coherence_of(a0, a1, a2, a3); }
};

__device__ inline void kahan_add(float& s, float& comp, float v)
{
    float y = v - comp;
    float t = s + y;
    comp = (t - s) - y;
    s = t;
}

__device__ __forceinline__ float4 window_product(tops_ci16 m, tops_cf32 v)
{
    const float mr = m.re, mi = m.im;
    return make_float4(mr * v.x + mi * v.y, mi * v.x - mr * v.y, v.x * v.x + v.y * v.y,
                       mr * mr + mi * mi);
}

// Products P = (M conj S, |S|^2, |M|^2) of the PR x PC input window whose
// first sample is (gr0, gc0), zero outside [0,h) x [0,w), to smem[PR*PC]
// (row-major).  Each thread issues the loads of 8 positions before it uses
// any of them; its position advances by 256 with no division in the loop.
// A window inside the raster (most tiles) takes a loop without the bounds
// tests.
template <int CPR = 0, int CPC = 0>
__device__ __forceinline__ void tile_products(const tops_ci16* M, const tops_cf32* S, int w, int h,
                                     int gr0, int gc0, int PR_, int PC_, float4* smem)
{
    // compile-time tile shape when the caller fixes the window
    const int PR = CPR ? CPR : PR_, PC = CPC ? CPC : PC_;
    const int n = PR * PC;
    const int step_r = 256 / PC, step_c = 256 - step_r * PC;
    // the thread index through an empty asm: the positions below are formed
    // on each call instead of being hoisted out of a caller's loop (the ESD
    // kernel's two bands), where they held about 32 VGPRs across it
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
    tid &= 255;   // its range again (256 threads), for the bounds below
    int rr = tid / PC, cc = tid - rr * PC;
    // inside the raster, and its byte offsets from the window's first sample
    // (< PR * w * 8 <= 64 * 2^23 * 8) fit 32 bits: one uniform base plus a
    // 32-bit offset per load
    if (gr0 >= 0 && gr0 + PR <= h && gc0 >= 0 && gc0 + PC <= w && w <= (1 << 23)) {
        const char* Mt = reinterpret_cast<const char*>(M + (size_t)gr0 * w + gc0);
        const char* St = reinterpret_cast<const char*>(S + (size_t)gr0 * w + gc0);
        for (int q0 = 0; q0 < n; q0 += 8 * 256) {
            tops_ci16 m[8];
            tops_cf32 v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (q0 + k * 256 + tid < n) {
                    const unsigned g = (unsigned)(rr * w + cc);   // < PR * w
                    m[k] = *reinterpret_cast<const tops_ci16*>(Mt + g * 4u);
                    v[k] = *reinterpret_cast<const tops_cf32*>(St + g * 8u);
                }
                rr += step_r;
                cc += step_c;
                if (cc >= PC) { cc -= PC; rr++; }
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int q = q0 + k * 256 + tid;
                if (q < n) smem[q] = window_product(m[k], v[k]);
            }
        }
        return;
    }
    for (int q0 = 0; q0 < n; q0 += 8 * 256) {
        tops_ci16 m[8];
        tops_cf32 v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int q = q0 + k * 256 + tid;
            const int gr = gr0 + rr, gc = gc0 + cc;
            m[k].re = 0; m[k].im = 0;
            v[k].x = 0.f; v[k].y = 0.f;
            if (q < n && gr >= 0 && gr < h && gc >= 0 && gc < w) {
                const size_t g = (size_t)gr * w + gc;
                m[k] = M[g];
                v[k] = S[g];
            }
            rr += step_r;
            cc += step_c;
            if (cc >= PC) { cc -= PC; rr++; }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int q = q0 + k * 256 + tid;
            if (q < n) smem[q] = window_product(m[k], v[k]);
        }
    }
}

// Fills ws[k] for output rows r0 + 4*wave + k, column c0 + lane.  When
// center != nullptr it also returns M conj S of each output sample itself.
// Ends with a barrier, so smem may be reused right after.
__device__ inline void tile_window_sums(const tops_ci16* M, const tops_cf32* S, int w,
                                        int h, int r0, int c0, int dx, int dy,
                                        int win_rg, int win_az, float4* smem,
                                        WindowSums ws[4], float2* center)
{
    const int PR = TW_ROWS + win_az - 1, PC = TW_COLS + win_rg - 1;
    float4* sP = smem;
    float4* sH = smem + PR * PC;
    tile_products(M, S, w, h, r0 - dy, c0 - dx, PR, PC, sP);
    __syncthreads();
    // range sums: direct sums of winRg products (no running subtraction, so
    // the error stays relative to the window's own sum of magnitudes)
    for (int q = threadIdx.x; q < PR * TW_COLS; q += 256) {
        int rr = q / TW_COLS, cc = q - rr * TW_COLS;
        const float4* row = sP + rr * PC + cc;
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
        for (int k = 0; k < win_rg; k++) {
            float4 v = row[k];
            s0 += v.x; s1 += v.y; s2 += v.z; s3 += v.w;
        }
        sH[q] = make_float4(s0, s1, s2, s3);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int rr = wv * 4 + k;
        float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
        for (int j = 0; j < win_az; j++) {
            float4 v = sH[(rr + j) * TW_COLS + lane];
            a0 += v.x; a1 += v.y; a2 += v.z; a3 += v.w;
        }
        ws[k].a0 = a0; ws[k].a1 = a1; ws[k].a2 = a2; ws[k].a3 = a3;
        if (center) {
            float4 v = sP[(rr + dy) * PC + lane + dx];
            // read as the whole 16-byte entry: 8-byte reads at this 16-byte
            // lane stride put two lanes of each half-wave on one bank pair
            asm volatile("" :: "v"(v.z), "v"(v.w));
            center[k] = make_float2(v.x, v.y);
        }
    }
    __syncthreads();
}

// row pitch (float4) of the range sums in tile_window_sums_shared: odd, so
// the rows start in different LDS banks; PR x TW_HSTRIDE fits in the
// PR x PC product tile (PC >= 70 for the windows it serves)
constexpr int TW_HSTRIDE = TW_COLS + 1;

__device__ __forceinline__ float4 add4(float4 a, float4 b)
{
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

// Whether tile_window_sums_shared serves a window: its range runs need
// winRg >= 7, its azimuth runs winAz >= 3, and its 8 range runs per product
// row are one pass of the 256 threads (16 + winAz - 1 rows <= 32).
__host__ __device__ inline bool tile_window_shared_ok(int win_rg, int win_az)
{
    return win_rg >= 7 && win_az >= 3 && win_az <= 17 &&
           (size_t)(TW_ROWS + win_az - 1) * (TW_COLS + win_rg - 1) * sizeof(float4) <= TW_LDS_LIMIT;
}

// LDS bytes of tile_window_sums_shared (the range sums overwrite the
// products, so only the product tile is held)
__host__ __device__ inline size_t tile_window_shared_lds_bytes(int win_rg, int win_az)
{
    return (size_t)(TW_ROWS + win_az - 1) * (TW_COLS + win_rg - 1) * sizeof(float4);
}

// LDS bytes a tile launch needs for the window code it will take
// (tile_window_sums_shared where it serves the window, else
// tile_window_sums), or 0 when neither fits the LDS limit
__host__ __device__ inline size_t tile_window_launch_lds(int win_rg, int win_az)
{
    if (tile_window_shared_ok(win_rg, win_az)) return tile_window_shared_lds_bytes(win_rg, win_az);
    return tile_window_fits(win_rg, win_az) ? tile_window_lds_bytes(win_rg, win_az) : 0;
}

// The window sums of tile_window_sums (same arguments and results; only
// for windows tile_window_shared_ok accepts):
//   range: task (row rr, run g) forms the 8 windows starting at product
//     columns 8g .. 8g+7 from the suffix sums s_j = P[j] + .. + P[6], the
//     shared middle P[7] + .. + P[winRg-1] and the prefix sums
//     p_j = P[winRg] + .. + P[winRg+j-1]: window j = s_j + middle + p_j,
//     31 product reads for 8 windows instead of 8 x 24;
//   azimuth: the thread's 4 output rows share the range sums of rows
//     3 .. winAz-1 in the same way (winAz + 3 reads for 4 outputs).
// Ends with a barrier, so smem may be reused right after.  CWR / CWA fix
// the window at compile time (0: the run-time win_rg / win_az): its loops
// then unroll completely.
template <int CWR = 0, int CWA = 0>
__device__ inline void tile_window_sums_shared(const tops_ci16* M, const tops_cf32* S, int w,
                                               int h, int r0, int c0, int dx, int dy,
                                               int win_rg_, int win_az_, float4* smem,
                                               WindowSums ws[4], float2* center)
{
    const int win_rg = CWR ? CWR : win_rg_, win_az = CWA ? CWA : win_az_;
    const int PR = TW_ROWS + win_az - 1, PC = TW_COLS + win_rg - 1;
    tile_products<CWA ? TW_ROWS + CWA - 1 : 0, CWR ? TW_COLS + CWR - 1 : 0>(
        M, S, w, h, r0 - dy, c0 - dx, PR, PC, smem);
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (center) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float4 v = smem[(wv * 4 + k + dy) * PC + lane + dx];
            // the whole 16-byte entry, as in tile_window_sums (8-byte reads
            // here were all of the ESD kernel's 2.4 M conflict cycles per call)
            asm volatile("" :: "v"(v.z), "v"(v.w));
            center[k] = make_float2(v.x, v.y);
        }
    }
    // range sums of row rr = task % PRP, windows 8g .. 8g+7 (g = task / PRP),
    // with the rows padded to PRP, a multiple of 8: the lanes of a wave take
    // different rows, whose starts (PC float4 apart; PC is odd for even
    // windows such as the reference's 24) fall in different LDS banks, and the
    // range sums are stored TW_HSTRIDE (odd) float4 apart per row for the same
    // reason.  With PRP a multiple of 8, a 16-lane LDS phase that spans two
    // runs takes the same 8 row offsets of both, shifted by the 8-float4 run
    // width, so its 16 float4 accesses still fall in 16 different bank
    // groups.  (Unpadded, the 21 rows of the 6 x 24 window put two rows of a
    // phase in one bank group: 16.95 M conflict cycles per 8 x 100 x 21000
    // call, 1.18 M padded.)
    const int PRP = (PR + 7) & ~7;
    const int task = threadIdx.x;
    const int trow = task % PRP, trun = task / PRP;
    const bool has = trow < PR && trun < 8;
    float4 hs[8];
    if (has) {
        const float4* q = smem + trow * PC + trun * 8;
        // the middle first: its reads do not overlap the suffix sums' registers
        float4 mid = make_float4(0.f, 0.f, 0.f, 0.f);
        // (a fixed window's middle unrolled by 4: fully unrolled, its reads
        // all issue at once and the ESD kernel spills at 4 waves per SIMD)
#pragma unroll 4
        for (int i = 7; i < win_rg; i++) mid = add4(mid, q[i]);
        float4 sfx[7];
        sfx[6] = q[6];
#pragma unroll
        for (int j = 5; j >= 0; j--) sfx[j] = add4(q[j], sfx[j + 1]);
        const float4* t = q + win_rg;
        hs[0] = add4(sfx[0], mid);
        float4 pfx = t[0];
#pragma unroll
        for (int j = 1; j < 7; j++) {
            hs[j] = add4(add4(sfx[j], mid), pfx);
            pfx = add4(pfx, t[j]);
        }
        hs[7] = add4(mid, pfx);
    }
    __syncthreads();   // every product read: the range sums overwrite them
    if (has) {
        float4* hrow = smem + trow * TW_HSTRIDE + trun * 8;
#pragma unroll
        for (int j = 0; j < 8; j++) hrow[j] = hs[j];
    }
    __syncthreads();
    // azimuth: output row 4wv+k sums range-sum rows 4wv+k .. 4wv+k+winAz-1
    const float4* hc = smem + wv * 4 * TW_HSTRIDE + lane;
    const float4 h0 = hc[0], h1 = hc[TW_HSTRIDE], h2 = hc[2 * TW_HSTRIDE];
    const float4 s1 = add4(h1, h2), s0 = add4(h0, s1);
    float4 mid = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int j = 3; j < win_az; j++) mid = add4(mid, hc[j * TW_HSTRIDE]);
    const float4* ht = hc + win_az * TW_HSTRIDE;
    const float4 p1 = ht[0], p2 = add4(p1, ht[TW_HSTRIDE]), p3 = add4(p2, ht[2 * TW_HSTRIDE]);
    const float4 o[4] = {add4(s0, mid), add4(add4(s1, mid), p1), add4(add4(h2, mid), p2),
                         add4(mid, p3)};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        ws[k].a0 = o[k].x; ws[k].a1 = o[k].y; ws[k].a2 = o[k].z; ws[k].a3 = o[k].w;
    }
    __syncthreads();
}

}  // namespace tops
