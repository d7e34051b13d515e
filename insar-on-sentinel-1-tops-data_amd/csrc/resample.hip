// resample.hip -- deramp -> 12x12 truncated-sinc resample -> reramp of one
// slave burst onto the master grid (gfx950).
//
// Semantics follow DerampAndResample.cu:108-183 (GetIndexes,
// DerampDemod_Shared) and 348-520 (Interpolation_12p):
//   slave line  y = poly(CpmAz; Ln, Pn) + AzShift, pixel x = poly(CpmRg; Ln, Pn)
//   (Iy, fy) = (trunc(y), int(frac*2047 + 0.5)), same for x
//   outside the slave box            -> 0
//   within 5 / 7 samples of its edge -> the sample at (Iy, Ix) (deramp and
//                                       reramp with the same phase cancel)
//   interior -> S = sum_j kL_j sum_i kP_i D(Iy-5+j, Ix-5+i),
//               Phi = sum_j kL_j sum_i kP_i phi(Iy-5+j, Ix-5+i),
//               out = S * exp(-j Phi)
// with D = slave * exp(j phi), phi(y, x) = -pi kt (ta - tref)^2 - 2 pi fdc ta,
// ta = (y - firstLine) * dt.
//
// Two implementations:
//   resample_fp64_*  plain all-FP64 per-stage form (deramp kernel writing an
//                    FP64 deramped burst + phase to HBM, then one thread per
//                    output reading 144+144 taps from HBM): the baseline.
//   resample_fast    fused column walker: each 64-lane wave walks a strip
//                    of output rows; every lane deramps one slave column on
//                    the fly (FP64 phasor recurrence, FP32 samples), the
//                    wave shares one 12-tap azimuth pass per distinct
//                    azimuth key through LDS, and each output does its 12
//                    range taps over those sums (FP32 MACs) and over the
//                    azimuth-interpolated phase from the moment identity
//                    below (FP64), then the reramp.
//
// Moment identity: phi is quadratic in the burst-relative line y for a fixed
// column, phi(y, x) = a(x) + b(x) y + c(x) y^2 with
//   a = -pi kt tref^2, b = 2 pi dt (kt tref - fdc), c = -pi kt dt^2.
// With y_j = y0 + (j - 5) and Mk = sum_j kL_j (j-5)^k (float LUT values in
// FP64), sum_j kL_j phi(y_j, x) = M0 a + (y0 M0 + M1) b + (y0^2 M0 + 2 y0 M1
// + M2) c exactly, so Phi needs three 12-tap FP64 range dot products.
#include "stream_access.hpp"
#include "tops_internal.hpp"
#include "walker.hpp"

#include <algorithm>
#include <climits>

namespace tops {

// ======================================================================
// Plain all-FP64 form
// ======================================================================
__global__ void __launch_bounds__(256) deramp_fp64_kernel(ResampleParams p, double2* dre, double* ph)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= p.s_pixels) return;
    int xa = p.sb[0] + x, ya = p.sb[2] + y;
    double phi = deramp_phase_direct(p, ya, xa);
    double sn, cs;
    sincos(phi, &sn, &cs);
    size_t q = (size_t)y * p.s_pixels + x;
    tops_ci16 s = p.slave[q];
    dre[q] = make_double2(s.re * cs - s.im * sn, s.re * sn + s.im * cs);
    ph[q] = phi;
}

__global__ void __launch_bounds__(256) resample_fp64_kernel(ResampleParams p, const double2* dre, const double* ph)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int r = blockIdx.y;
    if (c >= p.m_pixels) return;
    int iy, fy, ix, fx;
    slave_pos(p, r, c, iy, fy, ix, fx);
    int reg = region_of(p, iy, ix);
    tops_cf32 o = {0.0f, 0.0f};
    if (reg == 1) {
        size_t q = (size_t)(iy - p.sb[2]) * p.s_pixels + (ix - p.sb[0]);
        double2 d = dre[q];
        double sn, cs;
        sincos(ph[q], &sn, &cs);
        o.x = (float)(d.x * cs + d.y * sn);
        o.y = (float)(d.y * cs - d.x * sn);
    } else if (reg == 2) {
        const float* kl = p.kaz + fy * kTaps;
        const float* kp = p.krg + fx * kTaps;
        size_t base = (size_t)(iy - kHalf - p.sb[2]) * p.s_pixels + (ix - kHalf - p.sb[0]);
        double sr = 0.0, si = 0.0, sp = 0.0;
        for (int j = 0; j < kTaps; j++) {
            double tr = 0.0, ti = 0.0, tp = 0.0;
            const double2* drow = dre + base + (size_t)j * p.s_pixels;
            const double* prow = ph + base + (size_t)j * p.s_pixels;
            for (int i = 0; i < kTaps; i++) {
                double k = (double)kp[i];
                double2 d = drow[i];
                tr = fma(k, d.x, tr);
                ti = fma(k, d.y, ti);
                tp = fma(k, prow[i], tp);
            }
            double l = (double)kl[j];
            sr = fma(l, tr, sr);
            si = fma(l, ti, si);
            sp = fma(l, tp, sp);
        }
        double sn, cs;
        sincos(sp, &sn, &cs);
        o.x = (float)(sr * cs + si * sn);
        o.y = (float)(si * cs - sr * sn);
    }
    p.out[(size_t)r * p.m_pixels + c] = o;
}

int launch_resample_fp64(tops_ctx* ctx, const ResampleParams& p, hipStream_t s)
{
    size_t n = (size_t)p.s_lines * p.s_pixels;
    size_t need = n * (sizeof(double2) + sizeof(double));
    int rc = ctx->scratch.ensure(need);
    if (rc) return rc;
    double2* dre = (double2*)ctx->scratch.ptr;
    double* ph = (double*)(dre + n);
    dim3 g1((p.s_pixels + 255) / 256, p.s_lines);
    hipLaunchKernelGGL(deramp_fp64_kernel, g1, dim3(256), 0, s, p, dre, (double*)ph);
    TOPS_HIP(hipGetLastError());
    dim3 g2((p.m_pixels + 255) / 256, p.m_lines);
    hipLaunchKernelGGL(resample_fp64_kernel, g2, dim3(256), 0, s, p, (const double2*)dre,
                       (const double*)ph);
    TOPS_HIP(hipGetLastError());
    return TOPS_OK;
}

// Fused fast form: column-walking waves
// ======================================================================
// One 64-lane workgroup owns a strip of WK_ROWS output rows x WK_OUT output
// columns.  Lane l is bound to slave column ucol0 + l (ucol0 = the first
// output's truncated slave column - 6) for the whole strip:
//   * it keeps the deramped slave samples of rows wbase..wbase+15 of its
//     column in registers; new rows are deramped on arrival with an FP64
//     phasor recurrence along the column (exp(j phi) is advanced by
//     exp(j (b + c (2y+1))), which is advanced by exp(j 2c)), anchored with
//     an FP64 sincos at the strip start and after a resync;
//   * per output row and per distinct azimuth key (Iy, fy) among the row's
//     interior outputs it forms the 12-tap azimuth sum U of its column and
//     the azimuth-interpolated phase PU = M0 a + (y0 M0 + M1) b +
//     (y0^2 M0 + 2 y0 M1 + M2) c (FP64), and writes both to LDS;
//   * each output then needs only its own 12 range taps over U and PU.
// Outputs whose key or window the shared pass cannot serve (more than
// WK_KEYS keys in one row segment, or a window outside the strip's 64
// columns) are computed by interior_direct.

// One strip of one burst: output columns sx*WK_OUT .. +WK_OUT-1, output rows
// r0 .. r1-1 of the master box.  p may live in global memory (batched
// launches): its fields are wave-uniform.
struct WalkLds {
    float2 u[WK_KEYS][64];      // azimuth sums per key and column
    double pu[WK_KEYS][64];     // azimuth-interpolated phase per key and column
    double rowt[32][6];         // row-uniform position terms of 32 rows
    double pos[5][64];          // per lane: Pn, c2 Pn and c5 Pn Pn of both polynomials
    double abc[3][64];          // per lane: the column's phase coefficients a, b, c
};

__device__ __forceinline__ void resample_strip(const ResampleParams& p, int sx, int r0, int r1,
                                               WalkLds& sm)
{
    const int lane = threadIdx.x;
    const int c0 = sx * WK_OUT;
    const int c = c0 + lane;
    const bool out_lane = lane < WK_OUT && c < p.m_pixels;

    // Position polynomial (polyval6 order, contraction off), split into
    //   ys = ((((A + c2 Pn) + B) + D Pn) + c5 Pn Pn) + AzShift
    // with A = c0 + c1 Ln, B = c3 Ln Ln, D = c4 Ln uniform over a row and
    // c2 Pn, c5 Pn Pn fixed per lane: the same operations in the same order.
    // (kept in LDS, read once per row: registers are the walker's limit)
    {
#pragma clang fp contract(off)
        const double pn = norm_warp((double)(p.mb[0] + min(c, p.m_pixels - 1)),
                                    (double)p.mb[0], (double)p.mb[1]);
        sm.pos[0][lane] = pn;
        sm.pos[1][lane] = p.cpm_az[2] * pn;
        sm.pos[2][lane] = p.cpm_az[5] * pn * pn;
        sm.pos[3][lane] = p.cpm_rg[2] * pn;
        sm.pos[4][lane] = p.cpm_rg[5] * pn * pn;
    }

    // slave column of lane 0.  Output j (lane j) takes its range taps from
    // lanes ix_j - 5 - ucol0 .. +11, which must lie in 0 .. 63.  With
    // rel_j = ix_j - j over the strip within [lo, hi] (its corners: the
    // offset fields are smooth), ucol0 = lo - 5 serves every output when
    // hi - lo <= 1 (the slave column drifts by at most one column along the
    // strip's 52 outputs and rows); any other output takes the direct path.
    int ucol0;
    {
        const int cl = min(c0 + WK_OUT - 1, p.m_pixels - 1);
        const int cc = (lane & 2) ? cl : c0;
        int iyc, fyc, ixc, fxc;
        slave_pos(p, (lane & 1) ? r1 - 1 : r0, cc, iyc, fyc, ixc, fxc);
        const int rel = ixc - (cc - c0);
        int lo = __builtin_amdgcn_readlane(rel, 0);
        for (int k = 1; k < 4; k++) lo = min(lo, __builtin_amdgcn_readlane(rel, k));
        ucol0 = lo - kHalf;
    }
    ColumnWalker w;
    w.init(p, ucol0 + lane);
    {
        double a = 0.0, b = 0.0, cc = 0.0;
        if (w.in_box) phase_coeffs(p, w.xcol, a, b, cc);
        sm.abc[0][lane] = a;
        sm.abc[1][lane] = b;
        sm.abc[2][lane] = cc;
    }
    {
        double s, cc;
        double d = 2.0 * sm.abc[2][lane];
        sincos(fma(-rint(d * kInvTwoPi), kTwoPi, d), &s, &cc);
        w.qr = cc; w.qi = s;
    }

    const double azs = az_shift_of(p);
    // a strip whose outputs are all interior skips the per-output region
    // tests (wave-uniform)
    const bool all_interior = block_interior(p, r0, r1, c0, min(c0 + WK_OUT - 1, p.m_pixels - 1));
    WinRow win[WK_WIN];
    int wbase = INT_MIN;   // absolute slave line held in win[0]

    for (int r = r0; r < r1; r++) {
        // ---- row-uniform position terms, 32 rows at a time (one per lane)
        const int rr = (r - r0) & 31;
        if (rr == 0) {
            __syncthreads();
            const int rl = r + lane;
            if (lane < 32 && rl < r1) {
#pragma clang fp contract(off)
                const double ln = norm_warp((double)(p.mb[2] + rl), (double)p.mb[2],
                                            (double)p.mb[3]);
                sm.rowt[lane][0] = p.cpm_az[0] + p.cpm_az[1] * ln;
                sm.rowt[lane][1] = p.cpm_az[3] * ln * ln;
                sm.rowt[lane][2] = p.cpm_az[4] * ln;
                sm.rowt[lane][3] = p.cpm_rg[0] + p.cpm_rg[1] * ln;
                sm.rowt[lane][4] = p.cpm_rg[3] * ln * ln;
                sm.rowt[lane][5] = p.cpm_rg[4] * ln;
            }
            __syncthreads();
        }

        // ---- this lane's output position
        int iy = 0, fy = 0, ix = 0, fx = 0, reg = -1;
        {
#pragma clang fp contract(off)
            // all LDS operands first (their reads overlap), then the sums in order
            const double2* t2 = reinterpret_cast<const double2*>(sm.rowt[rr]);
            const double2 ta = t2[0], tb = t2[1], tc = t2[2];
            const double pn = sm.pos[0][lane], q1 = sm.pos[1][lane], q2 = sm.pos[2][lane];
            const double q3 = sm.pos[3][lane], q4 = sm.pos[4][lane];
            double ys = ta.x + q1 + ta.y + tb.x * pn + q2 + azs;
            double xs = tb.y + q3 + tc.x + tc.y * pn + q4;
            ys = fmin(fmax(ys, -1.0e9), 1.0e9);
            xs = fmin(fmax(xs, -1.0e9), 1.0e9);
            iy = (int)trunc(ys);
            fy = (int)((ys - iy) * kInterval + 0.5);
            ix = (int)trunc(xs);
            fx = (int)((xs - ix) * kInterval + 0.5);
        }
        if (out_lane) reg = all_interior ? 2 : region_of(p, iy, ix);
        const int base = ix - kHalf - ucol0;
        const bool servable = reg == 2 && base >= 0 && base <= 64 - kTaps;
        const int key = servable ? iy * kLutRows + fy : INT_MAX;

        tops_cf32 o = {0.0f, 0.0f};
        if (reg == 1) {
            tops_ci16 v = p.slave[(size_t)(iy - p.sb[2]) * p.s_pixels + (ix - p.sb[0])];
            o.x = (float)v.re;
            o.y = (float)v.im;
        }
        bool need = reg == 2;   // interior output not formed yet

        // ---- distinct azimuth keys of the row.  Keys change slowly along
        // the row, so a key is listed at each lane where it differs from the
        // lane to its left: one neighbour exchange and a ballot, then a
        // scalar walk over the set bits.
        const int kleft = __shfl_up(key, 1, 64);
        const unsigned long long heads = __ballot(key != INT_MAX && (lane == 0 || kleft != key));
        if (heads != 0ull) {
            // keys are iy * 2048 + fy with 0 <= iy < 2^20 (interior rows of a
            // slave box checked at launch) and 0 <= fy < 2048
            int iy_min = INT_MAX;
            for (unsigned long long h = heads; h; h &= h - 1)
                iy_min = min(iy_min, __builtin_amdgcn_readlane(key, __builtin_ctzll(h)) >> 11);
            // window: rows iy_min-5 .. iy_min+7 must be held
            int o_win = wbase == INT_MIN ? -1 : iy_min - kHalf - wbase;
            if (o_win < 0 || o_win > 64) {
                wbase = iy_min - kHalf;
                w.anchor(p, wbase, sm.abc[0][lane], sm.abc[1][lane], sm.abc[2][lane]);
                int raw[WK_WIN];
#pragma unroll
                for (int k = 0; k < WK_WIN; k++) raw[k] = w.fetch(p, wbase + k);
#pragma unroll
                for (int k = 0; k < WK_WIN; k++) win[k] = win_pack(w.deramp(raw[k]));
                w.prefetch(p);
                o_win = 0;
            }
            while (o_win > 3) {   // slide by 4 rows
                slide4(win, w, p);
                wbase += 4;
                o_win -= 4;
            }
            // rounds of up to WK_KEYS keys (one round unless the row segment
            // holds more keys); a key whose rows lie beyond the window's
            // reach (offset > 4) is left to the direct path.  A key that
            // reappears further along the row takes another slot (its lanes
            // use the later one): extra work in that rare case, same result.
            static_assert(kLutRows == 2048, "key packing below");
            auto serve_round = [&](unsigned long long& h) {
                int nk = 0, mybuf = -1;
                // the column's phase coefficients, read once per round (their
                // LDS reads then overlap the keys' scalar tap loads)
                const double abc0 = sm.abc[0][lane], abc1 = sm.abc[1][lane];
                const double abc2 = sm.abc[2][lane];
                for (; h && nk < WK_KEYS; h &= h - 1) {
                    const int kq = __builtin_amdgcn_readlane(key, __builtin_ctzll(h));
                    const int iyq = kq >> 11, fyq = kq & (kLutRows - 1);
                    const int oq = o_win + (iyq - iy_min);
                    if (oq > 4) continue;
                    // the key's 12 taps placed at window rows oq .. oq+11 (zero
                    // weights elsewhere add exactly nothing, so the sum equals
                    // the 12-term sum in the same order)
                    // uniform addresses: read through the constant address
                    // space so they become scalar loads
                    const float2 u = az_taps_shifted(win, p.kaz_sh, oq, fyq);
                    CDouble* mq = (CDouble*)(p.mom + 4 * fyq);
                    const double4 m = make_double4(mq[0], mq[1], mq[2], 0.0);
                    double y0 = (double)(iyq - p.first_line);
                    double bq = fma(y0, m.x, m.y);
                    double cq = y0 * y0 * m.x + 2.0 * y0 * m.y + m.z;
                    sm.u[nk][lane] = u;
                    sm.pu[nk][lane] = fma(m.x, abc0, fma(bq, abc1, cq * abc2));
                    if (key == kq) mybuf = nk;
                    nk++;
                }
                __syncthreads();
                if (need && mybuf >= 0) {
                    // this output's 12 range taps over U and PU (the phase
                    // sum takes the float taps widened to double, exactly)
                    float kp[kTaps];
                    load_taps(p.krg, fx, kp);
                    const float2* u = &sm.u[mybuf][base];
                    const double* pu = &sm.pu[mybuf][base];
                    // in two groups of 6 taps, each group's LDS reads issued
                    // together before its sums (one wait per group instead of
                    // one per tap pair; all 12 at once would spill)
                    static_assert(kTaps % 6 == 0, "tap groups of 6");
                    float sr = 0.f, si = 0.f;
                    double ph = 0.0;
#pragma unroll
                    for (int h = 0; h < kTaps; h += 6) {
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
                            sr = fmaf(kp[h + i], uu[i].x, sr);
                            si = fmaf(kp[h + i], uu[i].y, si);
                            ph = fma((double)kp[h + i], pp[i], ph);
                        }
                    }
                    o = reramp(sr, si, ph);
                    need = false;
                }
                __syncthreads();   // LDS is rewritten by the next round or row
            };
            unsigned long long h = heads;
            serve_round(h);                  // the common case: one round
            while (h != 0ull) serve_round(h);
        }

        // ---- interior outputs the shared pass did not serve
        if (need) {
            o = interior_direct(p.slave, p.kt, p.tref, p.fdc, p.mom, p.kaz, p.krg, p.s_pixels,
                                p.sb[0], p.sb[2], p.first_line, p.dt, iy, fy, ix, fx);
        }
        // written once: a nontemporal store
        if (out_lane) stream_store_float2(p.out + (size_t)r * p.m_pixels + c, o.x, o.y);
    }
}

// Output rows row0 .. row_end-1 of the box, strip_rows per strip.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kWalkWavesPerSimd))) resample_fast_kernel(ResampleParams p, int strip_rows, int row0, int row_end)
{
    __shared__ WalkLds sm;
    // strips numbered row strips fastest and dealt to the XCDs in
    // contiguous ranges: strips whose slave columns overlap run on one XCD
    const int t = xcd_remap((int)(blockIdx.x + blockIdx.y * gridDim.x),
                            (int)(gridDim.x * gridDim.y));
    const int sx = t / (int)gridDim.y, sy = t - sx * (int)gridDim.y;
    const int r0 = row0 + sy * strip_rows;
    // the parameter block read in place in the kernel-argument segment (it
    // is the first argument): its fields then become scalar loads where they
    // are used, as in the batched kernel, instead of SGPRs held for the whole
    // kernel (which spill to VGPR lanes inside the row loop)
    typedef __attribute__((address_space(4))) const ResampleParams CParams;
    static_assert(offsetof(ResampleParams, mb) == 0, "parameter block layout");
    CParams* pk = (CParams*)__builtin_amdgcn_kernarg_segment_ptr();
    (void)p;
    resample_strip(*(const ResampleParams*)pk, sx, r0, min(r0 + strip_rows, row_end), sm);
}

// Batched form: one workgroup per entry of a strip table covering many
// bursts (work[i] = {burst, strip x, strip y, strip rows}).
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kWalkWavesPerSimd))) resample_batch_kernel(const ResampleParams* items,
                                                            const int4* work)
{
    __shared__ WalkLds sm;
    // the table and the parameter blocks are read-only and read with
    // wave-uniform indices: through the constant address space they become
    // scalar loads, as the single-burst kernel's argument block is
    typedef __attribute__((address_space(4))) const int CInt;
    typedef __attribute__((address_space(4))) const ResampleParams CParams;
    const CInt* w = (const CInt*)work + 4 * xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int item = w[0], sx = w[1], sy = w[2], rows = w[3];
    CParams* it = (CParams*)items + item;
    const int r0 = sy * rows;
    resample_strip(*(const ResampleParams*)it, sx, r0, min(r0 + rows, it->m_lines), sm);
}

int resample_strip_count(int m_lines, int m_pixels, int strip_rows, int* nx, int* ny)
{
    *nx = (m_pixels + WK_OUT - 1) / WK_OUT;
    *ny = (m_lines + strip_rows - 1) / strip_rows;
    return *nx * *ny;
}

// Walker waves the GPU holds at once: CUs x 4 SIMDs x the kernel's waves per
// SIMD (kWalkWavesPerSimd).  Queried once per device.
int resample_wave_slots()
{
    static std::atomic<int> slots[kMaxCachedDevices];
    return per_device_cached(slots, 256 * 4 * kWalkWavesPerSimd, [](int dev) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
        return cus * 4 * kWalkWavesPerSimd;
    });
}

// Strip height whose strips (each one wave) fill `rounds` rounds of the
// GPU's wave slots without a partial extra round: the smallest height with
// sum_i nx_i * ceil(lines_i / h) <= rounds * slots.
int resample_rows_for_rounds(const int* lines, const int* pixels, int n, int rounds)
{
    const long cap = (long)rounds * resample_wave_slots();
    int lo = 8, hi = 8;
    for (int i = 0; i < n; i++) hi = std::max(hi, lines[i]);
    auto strips = [&](int h) {
        long t = 0;
        for (int i = 0; i < n; i++)
            t += (long)((pixels[i] + WK_OUT - 1) / WK_OUT) * ((lines[i] + h - 1) / h);
        return t;
    };
    if (strips(hi) > cap) return hi;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (strips(mid) <= cap) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

int launch_resample_rows(const ResampleParams& p, int row0, int rows, hipStream_t s,
                         int strip_rows)
{
    if (row0 < 0 || rows < 0 || row0 + rows > p.m_lines)
        return set_error(TOPS_ERR_ARG, "resample: rows %d .. %d outside the box's %d lines", row0,
                         row0 + rows - 1, p.m_lines);
    if (rows == 0) return TOPS_OK;
    if ((long long)p.s_lines * p.s_pixels * 4 > kMaxSlaveBytes)
        return set_error(TOPS_ERR_ARG, "resample: slave box of %d x %d samples over 2^31 bytes",
                         p.s_lines, p.s_pixels);
    if (strip_rows <= 0) {
        // strips for about 5 rounds of the GPU's wave slots: on one IW burst
        // (1514 x 21000) 31-34-row strips took 0.303-0.305 ms against 0.331 ms
        // for one round of 169-row strips (the slots finish together); a
        // launch of few rows keeps one round rather than strips under 24 rows
        strip_rows = resample_rows_for_rounds(&rows, &p.m_pixels, 1, 5);
        if (strip_rows < 24) strip_rows = resample_rows_for_rounds(&rows, &p.m_pixels, 1, 1);
    }
    dim3 grid((p.m_pixels + WK_OUT - 1) / WK_OUT, (rows + strip_rows - 1) / strip_rows);
    hipLaunchKernelGGL(resample_fast_kernel, grid, dim3(64), 0, s, p, strip_rows, row0,
                       row0 + rows);
    TOPS_HIP(hipGetLastError());
    return TOPS_OK;
}

int launch_resample_fast(const ResampleParams& p, hipStream_t s, int strip_rows)
{
    return launch_resample_rows(p, 0, p.m_lines, s, strip_rows);
}

int launch_resample_batch(const ResampleParams* d_items, const int4* d_work, int n_work,
                          hipStream_t s)
{
    if (n_work <= 0) return TOPS_OK;
    hipLaunchKernelGGL(resample_batch_kernel, dim3(n_work), dim3(64), 0, s, d_items, d_work);
    TOPS_HIP(hipGetLastError());
    return TOPS_OK;
}

}  // namespace tops
