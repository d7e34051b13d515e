// coherence_walk.hpp -- boxcar coherence with a compile-time window, one
// thread per 4 adjacent output columns walking down a run of output rows.
//
// Same semantics as the LDS-tiled form (EstCoh.cu Cohbasic, 413-456, with
// the zero padding of CohMatSet, 101-139): output (r, c) sums the products
// M conj S, |S|^2, |M|^2 over rows r-DY .. r-DY+WA-1 and columns
// c-DX .. c-DX+WR-1 (zero outside the burst) and is 0 where |S(r, c)| <
// 0.001.
//
// Per input row the thread loads the WR+3 samples its 4 outputs need with
// 16-byte vector loads (CInt16 master 4 per load, CF32 slave 2 per load),
// forms their products in registers and the 4 range sums.  The sums share
// their middle terms but are still plain sums of WR products each (no
// running subtraction, so the rounding error stays relative to each
// window's own sum of magnitudes).  The WA most recent rows of range sums
// sit in a rolling register buffer; an output row is their direct sum.
#pragma once

#include "tops_internal.hpp"
#include "tile_window.hpp"

namespace tops {

constexpr int CW_COLS = 4;    // output columns per thread
constexpr int CW_ROWS = 32;   // output rows per thread walk

__device__ __forceinline__ float4 cw_add(float4 a, float4 b)
{
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

// products of one sample: Re(M S*), Im(M S*), |S|^2, |M|^2
__device__ __forceinline__ float4 cw_prod(int mraw, float sr, float si)
{
    const float mr = (float)(short)(mraw & 0xffff), mi = (float)(short)(mraw >> 16);
    return make_float4(mr * sr + mi * si, mi * sr - mr * si, sr * sr + si * si, mr * mr + mi * mi);
}

template <int WR, int WA, int DX, int DY>
struct CohWalk {
    static_assert(WR >= 4, "the shared-middle range sums need WR >= 4");
    static constexpr int N = CW_COLS + WR - 1;          // samples a row needs
    static constexpr int RDX = (DX + 3) / 4 * 4;        // DX rounded up to 4
    static constexpr int OFF = RDX - DX;                // first needed sample in the load
    static constexpr int L = (OFF + N + 3) / 4 * 4;     // samples loaded per row

    // raw samples of one input row: L master (packed CInt16) and L slave
    // (CF32) samples from column c0 - RDX on
    struct Raw {
        int4 m[L / 4];
        float4 s[L / 2];
    };

    // the row's raw samples (zero outside the burst).  vec: w % 4 == 0 and
    // M, S, coh 16-byte aligned; rows in the middle of the burst then use
    // 16-byte loads only.
    __device__ static void fetch(const tops_ci16* M, const tops_cf32* S, int w, int h, int gr,
                                 int c0, bool vec, Raw& raw)
    {
        if (gr < 0 || gr >= h) {
#pragma unroll
            for (int q = 0; q < L / 4; q++) raw.m[q] = make_int4(0, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < L / 2; q++) raw.s[q] = make_float4(0.f, 0.f, 0.f, 0.f);
            return;
        }
        const size_t rb = (size_t)gr * w;
        const int lo = c0 - RDX;
        if (vec && lo >= 0 && lo + L <= w) {
            // aligned: rb + lo is a multiple of 4 samples (w % 4 == 0, c0 % 4 == 0)
            const int4* mq = reinterpret_cast<const int4*>(M + rb + lo);
            const float4* sq = reinterpret_cast<const float4*>(S + rb + lo);
#pragma unroll
            for (int q = 0; q < L / 4; q++) raw.m[q] = mq[q];
#pragma unroll
            for (int q = 0; q < L / 2; q++) raw.s[q] = sq[q];
        } else {
            const int* mi = reinterpret_cast<const int*>(M);
            int m[L];
            float sv[2 * L];
#pragma unroll
            for (int k = 0; k < L; k++) {
                const int c = lo + k;
                m[k] = 0;
                sv[2 * k] = sv[2 * k + 1] = 0.f;
                if (c >= 0 && c < w) {
                    const tops_cf32 t = S[rb + c];
                    m[k] = mi[rb + c];
                    sv[2 * k] = t.x;
                    sv[2 * k + 1] = t.y;
                }
            }
#pragma unroll
            for (int q = 0; q < L / 4; q++)
                raw.m[q] = make_int4(m[4 * q], m[4 * q + 1], m[4 * q + 2], m[4 * q + 3]);
#pragma unroll
            for (int q = 0; q < L / 2; q++)
                raw.s[q] = make_float4(sv[4 * q], sv[4 * q + 1], sv[4 * q + 2], sv[4 * q + 3]);
        }
    }

    // products of the N samples the 4 outputs need (columns c0-DX .. c0-DX+N-1)
    __device__ static void products(const Raw& raw, float4 (&v)[N])
    {
#pragma unroll
        for (int k = 0; k < N; k++) {
            const int e = OFF + k;
            const int4 mq = raw.m[e / 4];
            const int mv = (e % 4 == 0) ? mq.x : (e % 4 == 1) ? mq.y : (e % 4 == 2) ? mq.z : mq.w;
            const float4 sq = raw.s[e / 2];
            const float sr = (e % 2 == 0) ? sq.x : sq.z, si = (e % 2 == 0) ? sq.y : sq.w;
            v[k] = cw_prod(mv, sr, si);
        }
    }

    // range sums of the 4 outputs: s_j = sum of v[j .. j+WR-1]
    __device__ static void range_sums(const float4 (&v)[N], float4 (&s)[CW_COLS])
    {
        float4 mid = v[3];
#pragma unroll
        for (int k = 4; k < WR; k++) mid = cw_add(mid, v[k]);
        const float4 a12 = cw_add(v[1], v[2]);
        const float4 c01 = cw_add(v[WR], v[WR + 1]);
        s[0] = cw_add(cw_add(v[0], a12), mid);
        s[1] = cw_add(cw_add(a12, mid), v[WR]);
        s[2] = cw_add(cw_add(v[2], mid), c01);
        s[3] = cw_add(mid, cw_add(c01, v[WR + 2]));
    }

    // output rows r_begin .. r_end-1, columns c0 .. c0+3
    __device__ static void walk(const tops_ci16* M, const tops_cf32* S, int w, int h, float* coh,
                                int c0, int r_begin, int r_end, bool vec)
    {
        float4 H[WA][CW_COLS];     // range sums of the WA latest input rows, by slot
        unsigned ok[WA];           // |S| >= 0.001 flags of those rows' own 4 columns
        const int first = r_begin - DY;            // first input row
        const int last = r_end - 1 - DY + WA - 1;  // last input row
        Raw nxt;                                   // next row, loads in flight
        fetch(M, S, w, h, first, c0, vec, nxt);
        for (int base = first; base <= last; base += WA) {
#pragma unroll
            for (int q = 0; q < WA; q++) {
                const int gr = base + q;
                if (gr > last) break;
                const Raw cur = nxt;
                if (gr < last) fetch(M, S, w, h, gr + 1, c0, vec, nxt);
                float4 v[N];
                products(cur, v);
                range_sums(v, H[q]);
                unsigned f = 0;
#pragma unroll
                for (int j = 0; j < CW_COLS; j++)
                    f |= (amplitude_ok(v[DX + j].z) ? 1u : 0u) << j;
                ok[q] = f;
                if (gr - first < WA - 1) continue;
                // output row r: input rows r-DY .. r-DY+WA-1 are in the buffer
                const int r = gr - (WA - 1) + DY;
                const int slot_r = ((q - (WA - 1) + DY) % WA + WA) % WA;
                float out[CW_COLS];
#pragma unroll
                for (int j = 0; j < CW_COLS; j++) {
                    float4 a = H[0][j];
#pragma unroll
                    for (int t = 1; t < WA; t++) a = cw_add(a, H[t][j]);
                    out[j] = (ok[slot_r] >> j) & 1u ? coherence_of(a.x, a.y, a.z, a.w) : 0.0f;
                }
                float* dst = coh + (size_t)r * w + c0;
                if (vec && c0 + CW_COLS <= w) {
                    *reinterpret_cast<float4*>(dst) = make_float4(out[0], out[1], out[2], out[3]);
                } else {
#pragma unroll
                    for (int j = 0; j < CW_COLS; j++)
                        if (c0 + j < w) dst[j] = out[j];
                }
            }
        }
    }
};

}  // namespace tops
