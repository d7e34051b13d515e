// fused.hip -- deramp -> resample -> reramp -> coherence of many bursts in
// one persistent launch (gfx950).
//
// Semantics: DerampAndResample.cu:108-183, 348-520 (the resample, as
// resample.hip) followed by EstCoh.cu:101-139, 413-456 (boxcar coherence
// with zero padding and the |S| < 0.001 mask) with the 3 az x 10 rg window
// starting 1 row and 4 columns before the sample, as the burst loop of
// Func.cpp:2646-2716 chains them.
//
// Work items are (burst, column strip, row band).  A 256-lane workgroup (4
// waves) walks one item; lane g is bound to slave column ucol0 + g and to
// output column c_first + g:
//   * every lane keeps the deramped rows of its slave column in a register
//     window (ColumnWalker: FP64 phasor recurrence, FP32 samples);
//   * per output row each wave lists the distinct azimuth keys (Iy, fy) of
//     the outputs whose 12 range taps reach its lanes (its own outputs and the
//     last 12 of the wave to its left), forms for each key the 12-tap azimuth
//     sum U and the moment-identity phase PU of its columns, and writes them
//     to LDS slot key & 3, tagging the slot with the key;
//   * after a barrier every output does its 12 range taps over the U and PU
//     of that slot when every wave its taps touch has served the key, and the
//     direct 144-tap path otherwise, then the reramp;
//   * the same lane forms the coherence products M S*, |S|^2, |M|^2 of its
//     output, adds those of the two rows above (the window's 3 rows), and
//     puts the sum in LDS; in the next row's step each owned column adds the
//     10 sums of its range window and writes the coherence.
// A workgroup forms 243 outputs per row and owns 234 of them: the others are
// the coherence window's range halo (4 columns left, 5 right), also formed by
// the neighbouring strips.  A row band [r0, r1) also resamples rows r0-1 and
// r1 (the window's azimuth halo).  So the resampled slave never goes back to
// HBM to be read: per output sample the launch reads the CInt16 slave and
// master (4 + 4 B) and writes the CF32 sample and the coherence (8 + 4 B).
//
// Items are pulled from a queue (an atomic counter) until it is empty; the
// last workgroup to leave resets the counter, so a captured graph replays
// without clearing it.
#include "stream_access.hpp"
#include "tile_window.hpp"
#include "walker.hpp"

#include <climits>

namespace tops {

constexpr int FK_WAVES = 4;
constexpr int FK_LANES = 64 * FK_WAVES;            // slave columns per workgroup
constexpr int FK_OUTS = FK_LANES - kTaps - 1;      // outputs per row: 12 tap columns + 1 slack
constexpr int FK_WR = 10, FK_WA = 3, FK_DX = 4, FK_DY = 1;   // the coherence window
constexpr int FK_OWN = FK_OUTS - (FK_WR - 1);      // owned output columns per strip
constexpr int FK_SLOTS = 4;                        // azimuth-key slots (slot = key & 3)
constexpr int kFusedWavesPerSimd = 4;
static_assert(FK_WA == 3 && FK_DY == 1, "the register history holds the window's 3 rows");

struct FusedLds {
    float2 u[FK_SLOTS][FK_LANES];     // azimuth sums per key slot and column
    double pu[FK_SLOTS][FK_LANES];    // azimuth-interpolated phase per key slot and column
    float4 arow[FK_LANES];            // 3-row sums of the coherence products per column
    double pos[5][FK_LANES];          // per lane: Pn, c2 Pn and c5 Pn Pn of both polynomials
    double abc[3][FK_LANES];          // per lane: the column's phase coefficients a, b, c
    double rowt[32][6];               // row-uniform position terms of 32 rows
    int key[FK_LANES];                // this row's azimuth key per output (INT_MAX: none)
    int tag[FK_WAVES][FK_SLOTS];      // key each wave served in each slot (INT_MIN: none)
    int item;                         // work item taken from the queue
};

__device__ __forceinline__ float4 f4_add(float4 a, float4 b)
{
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

// products of one sample: Re(M S*), Im(M S*), |S|^2, |M|^2
__device__ __forceinline__ float4 coh_products(int mraw, float sr, float si)
{
    const float mr = (float)(short)(mraw & 0xffff), mi = (float)(short)(mraw >> 16);
    return make_float4(mr * sr + mi * si, mi * sr - mr * si, sr * sr + si * si, mr * mr + mi * mi);
}

typedef __attribute__((address_space(4))) const int CInt;
typedef __attribute__((address_space(4))) const ResampleParams CParams;
typedef __attribute__((address_space(4))) const CohItem CCoh;

// The parameter block again, through a copy of its pointer the compiler
// cannot see through: the fields read after this are loaded here (scalar
// loads) instead of being held in scalar registers across the row loop.
// Used where fields are read rarely (edge and direct outputs, the row-term
// blocks), so the walker's registers are not spent on them.
__device__ __forceinline__ const ResampleParams* reload(CParams* it)
{
    asm volatile("" : "+s"(it));
    return (const ResampleParams*)it;
}

// Byte offset of this lane's element in a row, formed where it is used: the
// row is then addressed as a scalar base plus a 32-bit lane offset, instead
// of a 64-bit pointer per lane held in registers across the row loop.
__device__ __forceinline__ unsigned lane_off(int c, int bytes)
{
    unsigned o = (unsigned)c * (unsigned)bytes;
    asm volatile("" : "+v"(o));
    return o;
}

// One item: output columns [k*FK_OWN, k*FK_OWN + FK_OWN) and rows [r0, r1)
// of burst p (coherence into ci.coh when ci.m is set).
__device__ __forceinline__ void fused_item(CParams* it, const tops_ci16* cm, float* ccoh, int k,
                                           int r0, int r1, FusedLds& sm)
{
    const ResampleParams& p = *(const ResampleParams*)it;
    const int g = threadIdx.x, lane = g & 63, wv = g >> 6;
    const int c_first = k * FK_OWN - FK_DX;
    const int c = c_first + g;
    const bool out_lane = g < FK_OUTS && c >= 0 && c < p.m_pixels;
    const bool own_lane = g >= FK_DX && g < FK_DX + FK_OWN && c < p.m_pixels;
    const bool with_coh = cm != nullptr;

    // Position polynomial (polyval6 order, contraction off), split into
    //   ys = ((((A + c2 Pn) + B) + D Pn) + c5 Pn Pn) + AzShift
    // with A = c0 + c1 Ln, B = c3 Ln Ln, D = c4 Ln uniform over a row and
    // c2 Pn, c5 Pn Pn fixed per lane: the same operations in the same order.
    {
#pragma clang fp contract(off)
        const double pn = norm_warp((double)(p.mb[0] + min(max(c, 0), p.m_pixels - 1)),
                                    (double)p.mb[0], (double)p.mb[1]);
        sm.pos[0][g] = pn;
        sm.pos[1][g] = p.cpm_az[2] * pn;
        sm.pos[2][g] = p.cpm_az[5] * pn * pn;
        sm.pos[3][g] = p.cpm_rg[2] * pn;
        sm.pos[4][g] = p.cpm_rg[5] * pn * pn;
    }
    // rows resampled: the band and its azimuth halo, inside the burst
    const int rs_lo = max(r0 - 1, 0), rs_hi = min(r1, p.m_lines - 1);

    // slave column of lane 0: the first output column inside the burst,
    // less its lane, less the 5 taps left of the output and one slack column
    // (the same value in every lane)
    int iy0, fy0, ix0, fx0;
    const int cv = max(c_first, 0);
    slave_pos(p, rs_lo, cv, iy0, fy0, ix0, fx0);
    // (computed in vector registers; the same in every lane, so kept scalar)
    const int ucol0 = __builtin_amdgcn_readfirstlane(ix0 - kHalf - 1 - (cv - c_first));
    ColumnWalker w;
    w.init(p, ucol0 + g);
    {
        double a = 0.0, b = 0.0, cc = 0.0;
        if (w.in_box) phase_coeffs(p, w.xcol, a, b, cc);
        sm.abc[0][g] = a;
        sm.abc[1][g] = b;
        sm.abc[2][g] = cc;
        double s, co;
        const double d = 2.0 * cc;
        sincos(fma(-rint(d * kInvTwoPi), kTwoPi, d), &s, &co);
        w.qr = co;
        w.qi = s;
    }

    const double azs = az_shift_of(p);
    WinRow win[WK_WIN];
    int wbase = INT_MIN;   // absolute slave line held in win[0]
    float4 pm1 = make_float4(0.f, 0.f, 0.f, 0.f);   // products of row t-1
    float4 pm2 = pm1;                                 // products of row t-2

    // step t: resample row t (if it is one of the band's rows or its halo),
    // then its coherence products; the coherence of row t-2 is written from
    // the 3-row sums of step t-1
    for (int t = r0 - 1; t <= r1 + 1; t++) {
        const bool rs = t >= rs_lo && t <= rs_hi;   // workgroup-uniform
        if (rs && ((t - rs_lo) & 31) == 0) {
            // row-uniform position terms of rows t .. t+31 (the previous
            // block's last reads came before the previous step's barriers)
            const int rl = t + g;
            if (g < 32 && rl <= rs_hi) {
#pragma clang fp contract(off)
                const ResampleParams& q = *reload(it);
                const double ln = norm_warp((double)(q.mb[2] + rl), (double)q.mb[2],
                                            (double)q.mb[3]);
                sm.rowt[g][0] = q.cpm_az[0] + q.cpm_az[1] * ln;
                sm.rowt[g][1] = q.cpm_az[3] * ln * ln;
                sm.rowt[g][2] = q.cpm_az[4] * ln;
                sm.rowt[g][3] = q.cpm_rg[0] + q.cpm_rg[1] * ln;
                sm.rowt[g][4] = q.cpm_rg[3] * ln * ln;
                sm.rowt[g][5] = q.cpm_rg[4] * ln;
            }
            __syncthreads();
        }

        // ---- this lane's output position, region and azimuth key.  Kept
        // across the barriers packed: ky = Iy * 2048 + fy, kxp = Ix * 2048 +
        // fx (regions 1 and 2 lie in the slave box: 0 <= Iy, Ix < 2^20, the
        // plan offering this schedule only for slave boxes inside that
        // range); reg 3 marks an interior output the shared pass can serve
        int ky = 0, kxp = 0, reg = -1, key = INT_MAX, mraw = 0;
        if (rs) {
            int iy, fy, ix, fx;
            {
#pragma clang fp contract(off)
                const int rr = (t - rs_lo) & 31;
                const double2* t2 = reinterpret_cast<const double2*>(sm.rowt[rr]);
                const double2 ta = t2[0], tb = t2[1], tc = t2[2];
                const double pn = sm.pos[0][g], q1 = sm.pos[1][g], q2 = sm.pos[2][g];
                const double q3 = sm.pos[3][g], q4 = sm.pos[4][g];
                double ys = ta.x + q1 + ta.y + tb.x * pn + q2 + azs;
                double xs = tb.y + q3 + tc.x + tc.y * pn + q4;
                ys = fmin(fmax(ys, -1.0e9), 1.0e9);
                xs = fmin(fmax(xs, -1.0e9), 1.0e9);
                iy = (int)trunc(ys);
                fy = (int)((ys - iy) * kInterval + 0.5);
                ix = (int)trunc(xs);
                fx = (int)((xs - ix) * kInterval + 0.5);
            }
            if (out_lane) {
                reg = region_of(p, iy, ix);
                if (with_coh)
                    mraw = __builtin_nontemporal_load(reinterpret_cast<const int*>(
                        reinterpret_cast<const char*>(cm) + (size_t)t * p.m_pixels * 4 +
                        lane_off(c, 4)));
            }
            // packed only for outputs in the slave box.  An edge output
            // (region 1) takes the sample at the truncated index alone; its
            // fraction is dropped, since a position in (-1, 0) truncates to 0
            // with a negative fraction, which would unpack to line or column
            // -1.  Interior positions are at least 5 lines and columns into
            // the box, so their fractions are in 0 .. 2047.
            if (reg >= 1) {
                ky = iy * kLutRows + (reg >= 2 ? fy : 0);
                kxp = ix * kLutRows + (reg >= 2 ? fx : 0);
            }
            const int base = ix - kHalf - ucol0;
            if (reg == 2 && base >= 0 && base <= FK_LANES - kTaps) {
                key = ky;
                reg = 3;
            }
        }
        sm.key[g] = key;
        __syncthreads();   // B1: keys of this row; 3-row sums of step t-1

        // ---- coherence of row t-2 (window rows t-3 .. t-1, columns c-4 .. c+5)
        if (with_coh && own_lane && t - 2 >= r0) {
            // in two groups of 5: each group's LDS reads issued together
            // before its sums (all 10 at once hold 40 registers)
            const float4* a = &sm.arow[g - FK_DX];
            float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int h = 0; h < FK_WR; h += 5) {
                float4 v[5];
#pragma unroll
                for (int j = 0; j < 5; j++) v[j] = a[h + j];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int j = 0; j < 5; j++) s = (h + j == 0) ? v[0] : f4_add(s, v[j]);
            }
            // |S(t-2, c)| >= 0.001 from the products of row t-2
            const float co = amplitude_ok(pm2.z) ? coherence_of(s.x, s.y, s.z, s.w) : 0.0f;
            __builtin_nontemporal_store(co, reinterpret_cast<float*>(
                reinterpret_cast<char*>(ccoh) + (size_t)(t - 2) * p.m_pixels * 4 + lane_off(c, 4)));
        }

        // ---- azimuth sums of the keys this wave's lanes serve
        if (lane < FK_SLOTS) sm.tag[wv][lane] = INT_MIN;
        if (rs) {
            // keys whose range taps reach this wave: its own outputs' and
            // those of the last 12 outputs of the wave to its left; a key is
            // listed where it differs from the output to its left
            const int kl = g > 0 ? sm.key[g - 1] : INT_MAX;
            const int kx = (lane < kTaps && wv > 0) ? sm.key[g - kTaps] : INT_MAX;
            const int kxl = __shfl_up(kx, 1, 64);
            const unsigned long long th = __ballot(kx != INT_MAX && (lane == 0 || kxl != kx));
            const unsigned long long oh = __ballot(key != INT_MAX && key != kl);
            if ((th | oh) != 0ull) {
                // keys are iy * 2048 + fy with 0 <= iy < 2^20 and 0 <= fy < 2048
                int iy_min = INT_MAX;
                for (unsigned long long h = th; h; h &= h - 1)
                    iy_min = min(iy_min, __builtin_amdgcn_readlane(kx, __builtin_ctzll(h)) >> 11);
                for (unsigned long long h = oh; h; h &= h - 1)
                    iy_min = min(iy_min, __builtin_amdgcn_readlane(key, __builtin_ctzll(h)) >> 11);
                // window: rows iy_min-5 .. iy_min+7 must be held
                int o_win = wbase == INT_MIN ? -1 : iy_min - kHalf - wbase;
                if (o_win < 0 || o_win > 64) {
                    wbase = iy_min - kHalf;
                    w.anchor(p, wbase, sm.abc[0][g], sm.abc[1][g], sm.abc[2][g]);
                    int raw[WK_WIN];
#pragma unroll
                    for (int q = 0; q < WK_WIN; q++) raw[q] = w.fetch(p, wbase + q);
#pragma unroll
                    for (int q = 0; q < WK_WIN; q++) win[q] = win_pack(w.deramp(raw[q]));
                    w.prefetch(p);
                    o_win = 0;
                }
                while (o_win > 3) {   // slide by 4 rows
                    slide4(win, w, p);
                    wbase += 4;
                    o_win -= 4;
                }
                // the column's phase coefficients, read once per row
                const double abc0 = sm.abc[0][g], abc1 = sm.abc[1][g], abc2 = sm.abc[2][g];
                unsigned used = 0u;   // slots taken in this row (wave-uniform)
                auto serve = [&](int kq) {
                    const int s = kq & (FK_SLOTS - 1);
                    // a taken slot: the same key again (served already) or a
                    // colliding one, whose outputs then take the direct path
                    if (used & (1u << s)) return;
                    const int iyq = kq >> 11, fyq = kq & (kLutRows - 1);
                    const int oq = o_win + (iyq - iy_min);
                    if (oq > 4) return;   // beyond the window's reach: direct path
                    used |= 1u << s;
                    // the key's 12 taps placed at window rows oq .. oq+11 of a
                    // 16-tap row (zero weights elsewhere add exactly nothing);
                    // uniform addresses read through the constant address space
                    const float2 u = az_taps_shifted(win, p.kaz_sh, oq, fyq);
                    CDouble* mq = (CDouble*)(p.mom + 4 * fyq);
                    const double m0 = mq[0], m1 = mq[1], m2 = mq[2];
                    const double y0 = (double)(iyq - p.first_line);
                    const double bq = fma(y0, m0, m1);
                    const double cq = y0 * y0 * m0 + 2.0 * y0 * m1 + m2;
                    sm.u[s][g] = u;
                    sm.pu[s][g] = fma(m0, abc0, fma(bq, abc1, cq * abc2));
                    if (lane == 0) sm.tag[wv][s] = kq;
                };
                for (unsigned long long h = th; h; h &= h - 1)
                    serve(__builtin_amdgcn_readlane(kx, __builtin_ctzll(h)));
                for (unsigned long long h = oh; h; h &= h - 1)
                    serve(__builtin_amdgcn_readlane(key, __builtin_ctzll(h)));
            }
        }
        __syncthreads();   // B2: azimuth sums and slot tags of this row

        // ---- range taps and reramp of this lane's output
        tops_cf32 o = {0.0f, 0.0f};
        if (rs) {
            const int fx = kxp & (kLutRows - 1), ix = kxp >> 11;
            if (reg == 1) {
                const int iy = ky >> 11;
                const ResampleParams& q = *reload(it);
                const tops_ci16 v = q.slave[(size_t)(iy - q.sb[2]) * q.s_pixels + (ix - q.sb[0])];
                o.x = (float)v.re;
                o.y = (float)v.im;
            } else if (reg >= 2) {
                bool served = false;
                if (reg == 3) {
                    const int s = key & (FK_SLOTS - 1);
                    const int base = ix - kHalf - ucol0;
                    served = sm.tag[base >> 6][s] == key && sm.tag[(base + kTaps - 1) >> 6][s] == key;
                    if (served) {
                        const float2* u = &sm.u[s][base];
                        const double* pu = &sm.pu[s][base];
                        const float* kr = p.krg + fx * kTaps;
                        float sr = 0.f, si = 0.f;
                        double ph = 0.0;
                        // in two groups of 6 taps: each group's taps and LDS
                        // reads issued together before its sums
#pragma unroll
                        for (int h = 0; h < kTaps; h += 6) {
                            const float4 ka = *reinterpret_cast<const float4*>(kr + h);
                            const float2 kb = *reinterpret_cast<const float2*>(kr + h + 4);
                            const float kp[6] = {ka.x, ka.y, ka.z, ka.w, kb.x, kb.y};
                            float2 uu[6];
                            double pp[6];
#pragma unroll
                            for (int i = 0; i < 6; i++) {
                                uu[i] = u[h + i];
                                pp[i] = pu[h + i];
                            }
                            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                            for (int i = 0; i < 6; i++) {
                                sr = fmaf(kp[i], uu[i].x, sr);
                                si = fmaf(kp[i], uu[i].y, si);
                                ph = fma((double)kp[i], pp[i], ph);
                            }
                        }
                        o = reramp(sr, si, ph);
                    }
                }
                if (!served) {
                    const ResampleParams& q = *reload(it);
                    o = interior_direct(q.slave, q.kt, q.tref, q.fdc, q.mom, q.kaz, q.krg,
                                        q.s_pixels, q.sb[0], q.sb[2], q.first_line, q.dt,
                                        ky >> 11, ky & (kLutRows - 1), ix, fx);
                }
            }
            if (own_lane && t >= r0 && t < r1)
                stream_store_float2(reinterpret_cast<char*>(p.out) + (size_t)t * p.m_pixels * 8 +
                                        lane_off(c, 8), o.x, o.y);
        }

        // ---- coherence products of row t and the 3-row sums (rows t-2 .. t)
        if (with_coh) {
            const float4 pr = coh_products(mraw, o.x, o.y);   // zero off the burst
            const float4 a = f4_add(f4_add(pm2, pm1), pr);
            pm2 = pm1;
            pm1 = pr;
            if (g < FK_OUTS) sm.arow[g] = a;
        }
    }
}


// Persistent: each workgroup takes items {burst, column strip, first row,
// rows} from work[] through queue[0] until none is left.
__global__ void __launch_bounds__(FK_LANES) __attribute__((amdgpu_waves_per_eu(kFusedWavesPerSimd)))
burst_fused_kernel(const ResampleParams* items, const CohItem* coh_items, const int* coh_of,
                   const int4* work, int n_work, unsigned* queue)
{
    __shared__ FusedLds sm;
    for (;;) {
        if (threadIdx.x == 0) sm.item = (int)atomicAdd(&queue[0], 1u);
        __syncthreads();
        const int wi = __builtin_amdgcn_readfirstlane(sm.item);
        if (wi >= n_work) break;   // the same value in every thread
        const CInt* wk = (const CInt*)work + 4 * wi;
        const int b = wk[0], k = wk[1], r0 = wk[2], r1 = r0 + wk[3];
        const CParams* it = (const CParams*)items + b;
        const int ci = ((const CInt*)coh_of)[b];
        const tops_ci16* cm = nullptr;
        float* ccoh = nullptr;
        if (ci >= 0) {
            const CCoh* cc = (const CCoh*)coh_items + ci;
            cm = cc->m;
            ccoh = cc->coh;
        }
        fused_item(it, cm, ccoh, k, r0, r1, sm);
        __syncthreads();   // sm (and sm.item) is rewritten by the next item
    }
    // the last workgroup out resets the queue for the next launch
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
        }
    }
}

bool fused_window_fits(int dx, int dy, int win_rg, int win_az)
{
    return win_rg == FK_WR && win_az == FK_WA && dx == FK_DX && dy == FK_DY;
}

int fused_strip_count(int m_pixels) { return (m_pixels + FK_OWN - 1) / FK_OWN; }

// workgroups the GPU holds at once (CUs x resident workgroups per CU),
// queried once per device
int fused_grid()
{
    static std::atomic<int> grid[kMaxCachedDevices];
    return per_device_cached(grid, 256, [](int dev) {
        int cus = 0, per = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, burst_fused_kernel, FK_LANES, 0) !=
                hipSuccess ||
            per <= 0)
            per = 1;
        return cus * per;
    });
}

int launch_fused_batch(const ResampleParams* d_items, const CohItem* d_coh, const int* d_coh_of,
                       const int4* d_work, int n_work, unsigned* d_queue, hipStream_t s)
{
    if (n_work <= 0) return TOPS_OK;
    hipLaunchKernelGGL(burst_fused_kernel, dim3(fused_grid()), dim3(FK_LANES), 0, s, d_items, d_coh,
                       d_coh_of, d_work, n_work, d_queue);
    TOPS_HIP(hipGetLastError());
    return TOPS_OK;
}

}  // namespace tops
