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
    const int gr0 = r0 - dy, gc0 = c0 - dx;
    for (int q = threadIdx.x; q < PR * PC; q += 256) {
        int rr = q / PC, cc = q - rr * PC;
        int gr = gr0 + rr, gc = gc0 + cc;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (gr >= 0 && gr < h && gc >= 0 && gc < w) {
            size_t g = (size_t)gr * w + gc;
            tops_ci16 m = M[g];
            tops_cf32 s = S[g];
            float mr = m.re, mi = m.im;
            v.x = mr * s.x + mi * s.y;
            v.y = mi * s.x - mr * s.y;
            v.z = s.x * s.x + s.y * s.y;
            v.w = mr * mr + mi * mi;
        }
        sP[q] = v;
    }
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
            center[k] = make_float2(v.x, v.y);
        }
    }
    __syncthreads();
}

}  // namespace tops
