// coherence_walk.hpp -- boxcar coherence with a compile-time window: a
// 256-thread workgroup walks down a band of 512 output columns, 2 adjacent
// columns per thread.
//
// Same semantics as the LDS-tiled form (EstCoh.cu Cohbasic, 413-456, with
// the zero padding of CohMatSet, 101-139): output (r, c) sums the products
// M conj S, |S|^2, |M|^2 over rows r-DY .. r-DY+WA-1 and columns
// c-DX .. c-DX+WR-1 (zero outside the burst) and is 0 where |S(r, c)| <
// 0.001.
//
// Per input row each thread loads its own 2 samples (one 8-byte CInt16 pair,
// one 16-byte CF32 pair; threads 0 .. WR-2 also load one halo sample of the
// band), writes their products to an LDS row (double-buffered: one barrier
// per row; even and odd columns in separate arrays, so the reads are
// conflict-free) and reads back the WR+1 products its two outputs need, all
// before summing any.  The two
// range sums share their WR-1 middle terms but each is still a plain sum of
// WR products (no running subtraction, so the rounding error stays relative
// to each window's own sum of magnitudes).  The WA latest rows of range sums
// sit in a rolling register buffer; an output row is their direct sum.  The
// next row's samples are requested before the current row is summed.  The
// pair loads and the output stores are nontemporal (each sample is read by
// one workgroup, each output written once).
#pragma once

#include "stream_access.hpp"
#include "tile_window.hpp"
#include "tops_internal.hpp"

namespace tops {

constexpr int CW_THREADS = 256;                       // threads per workgroup
constexpr int CW_PER_THREAD = 2;                      // output columns per thread
constexpr int CW_BAND = CW_THREADS * CW_PER_THREAD;   // columns per workgroup
constexpr int CW_ROWS = 32;                           // output rows per workgroup (batched form)

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
    static_assert(WR >= 2 && DX >= 0 && DX < WR && DY >= 0 && DY < WA, "window");
    static constexpr int HALO_L = DX;               // columns left of the band
    static constexpr int HALO_R = WR - 1 - DX;      // columns right of the band
    static constexpr int ROW = CW_BAND + WR - 1;    // products per LDS row
    static_assert(HALO_L + HALO_R <= CW_THREADS, "halo loads");

    static constexpr int ROWH = (ROW + 1) / 2;    // products per column parity
    // a row's products split by column parity (column k at ev[k/2] or
    // od[k/2]): a thread's window reads, two columns apart per lane in one
    // array, become consecutive 16-byte reads across the lanes of a wave
    struct Lds {
        float4 ev[2][ROWH];
        float4 od[2][ROWH];
    };
    __device__ static float4& at(Lds& sm, int b, int k)
    {
        return (k & 1) ? sm.od[b][k >> 1] : sm.ev[b][k >> 1];
    }

    // this thread's two samples of one row (zero outside the burst), plus
    // its halo sample when it has one
    struct Raw {
        int m0, m1, mh;
        float4 s;        // (re0, im0, re1, im1)
        float2 sh;
    };

    // vec: w even and M, S, coh 8- / 16- / 8-byte aligned
    __device__ static void fetch(const tops_ci16* M, const tops_cf32* S, int w, int h, int gr,
                                 int cb, bool vec, Raw& raw)
    {
        const int t = threadIdx.x;
        const int c = cb + CW_PER_THREAD * t;
        raw.m0 = raw.m1 = raw.mh = 0;
        raw.s = make_float4(0.f, 0.f, 0.f, 0.f);
        raw.sh = make_float2(0.f, 0.f);
        if (gr < 0 || gr >= h) return;
        const size_t rb = (size_t)gr * w;
        const int* mi = reinterpret_cast<const int*>(M);
        if (vec && c + 1 < w) {
            // rb + c is even: the pair loads are aligned
            const int2 m2 = stream_load_int2(mi + rb + c);
            raw.m0 = m2.x;
            raw.m1 = m2.y;
            raw.s = stream_load_float4(S + rb + c);
        } else {
            if (c < w) {
                raw.m0 = mi[rb + c];
                raw.s.x = S[rb + c].x;
                raw.s.y = S[rb + c].y;
            }
            if (c + 1 < w) {
                raw.m1 = mi[rb + c + 1];
                raw.s.z = S[rb + c + 1].x;
                raw.s.w = S[rb + c + 1].y;
            }
        }
        if (t < HALO_L + HALO_R) {
            const int ch = t < HALO_L ? cb - HALO_L + t : cb + CW_BAND + (t - HALO_L);
            if (ch >= 0 && ch < w) {
                raw.mh = mi[rb + ch];
                raw.sh = make_float2(S[rb + ch].x, S[rb + ch].y);
            }
        }
    }

    // products of a fetched row into LDS row buffer b
    __device__ static void store(const Raw& raw, Lds& sm, int b)
    {
        const int t = threadIdx.x;
        at(sm, b, HALO_L + CW_PER_THREAD * t) = cw_prod(raw.m0, raw.s.x, raw.s.y);
        at(sm, b, HALO_L + CW_PER_THREAD * t + 1) = cw_prod(raw.m1, raw.s.z, raw.s.w);
        if (t < HALO_L + HALO_R) {
            const int k = t < HALO_L ? t : HALO_L + CW_BAND + (t - HALO_L);
            at(sm, b, k) = cw_prod(raw.mh, raw.sh.x, raw.sh.y);
        }
    }

    // output rows r_begin .. r_end-1 (the same for every thread of the
    // workgroup) of the band starting at column cb
    __device__ static void walk(const tops_ci16* M, const tops_cf32* S, int w, int h, float* coh,
                                int cb, int r_begin, int r_end, bool vec, Lds& sm)
    {
        const int t = threadIdx.x;
        const int c = cb + CW_PER_THREAD * t;      // this thread's first column
        float4 H[WA][CW_PER_THREAD];               // range sums of the WA latest rows, by slot
        unsigned ok[WA];                           // |S| >= 0.001 flags of those rows' own columns
        const int first = r_begin - DY;            // first input row
        const int last = r_end - 1 - DY + WA - 1;  // last input row
        Raw nxt;
        fetch(M, S, w, h, first, cb, vec, nxt);
        for (int base = first; base <= last; base += WA) {
#pragma unroll
            for (int q = 0; q < WA; q++) {
                const int gr = base + q;
                if (gr > last) break;               // uniform over the workgroup
                const int b = (gr - first) & 1;
                const Raw cur = nxt;
                if (gr < last) fetch(M, S, w, h, gr + 1, cb, vec, nxt);
                store(cur, sm, b);
                ok[q] = (amplitude_ok(cur.s.x * cur.s.x + cur.s.y * cur.s.y) ? 1u : 0u) |
                        (amplitude_ok(cur.s.z * cur.s.z + cur.s.w * cur.s.w) ? 2u : 0u);
                __syncthreads();
                // window of output c: LDS entries 2t .. 2t+WR-1; of c+1: one more
                // the WR + 1 products read together, before any is summed
                // (one LDS wait per row instead of one per pair of reads)
                // (columns 2t + k: the even ones at ev[t + k/2], the odd at
                // od[t + k/2])
                static_assert(CW_PER_THREAD == 2, "window start 2t is even");
                float4 v[WR + 1];
#pragma unroll
                for (int k = 0; k <= WR; k++)
                    v[k] = (k & 1) ? sm.od[b][t + (k >> 1)] : sm.ev[b][t + (k >> 1)];
                __builtin_amdgcn_sched_barrier(0);
                float4 mid = v[1];
#pragma unroll
                for (int k = 2; k < WR; k++) mid = cw_add(mid, v[k]);
                H[q][0] = cw_add(v[0], mid);
                H[q][1] = cw_add(mid, v[WR]);
                if (gr - first < WA - 1) continue;
                // output row r: input rows r-DY .. r-DY+WA-1 are in the buffer
                const int r = gr - (WA - 1) + DY;
                const int slot_r = ((q - (WA - 1) + DY) % WA + WA) % WA;
                float out[CW_PER_THREAD];
#pragma unroll
                for (int j = 0; j < CW_PER_THREAD; j++) {
                    float4 a = H[0][j];
#pragma unroll
                    for (int u = 1; u < WA; u++) a = cw_add(a, H[u][j]);
                    out[j] = (ok[slot_r] >> j) & 1u ? coherence_of(a.x, a.y, a.z, a.w) : 0.0f;
                }
                float* dst = coh + (size_t)r * w + c;
                if (vec && c + 1 < w) {
                    stream_store_float2(dst, out[0], out[1]);
                } else {
                    if (c < w) dst[0] = out[0];
                    if (c + 1 < w) dst[1] = out[1];
                }
            }
        }
    }
};

}  // namespace tops
