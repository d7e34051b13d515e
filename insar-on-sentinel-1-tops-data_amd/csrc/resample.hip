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
//   resample_fast    fused: a 16x64 output tile stages its deramped slave
//                    window in LDS (FP64 phase reduced mod 2*pi, FP32
//                    phasor), FP32 sinc MACs, and the reramp phase from the
//                    FP64 moment identity below instead of 144 FP64 taps.
//
// Moment identity: phi is quadratic in the burst-relative line y for a fixed
// column, phi(y, x) = a(x) + b(x) y + c(x) y^2 with
//   a = -pi kt tref^2, b = 2 pi dt (kt tref - fdc), c = -pi kt dt^2.
// With y_j = y0 + (j - 5) and Mk = sum_j kL_j (j-5)^k (float LUT values in
// FP64), sum_j kL_j phi(y_j, x) = M0 a + (y0 M0 + M1) b + (y0^2 M0 + 2 y0 M1
// + M2) c exactly, so Phi needs three 12-tap FP64 range dot products.
#include "tops_internal.hpp"

#include <climits>

namespace tops {

__device__ inline void slave_pos(const ResampleParams& p, int r, int c, int& iy,
                                 int& fy, int& ix, int& fx)
{
#pragma clang fp contract(off)
    double pn = norm_warp((double)(p.mb[0] + c), (double)p.mb[0], (double)p.mb[1]);
    double ln = norm_warp((double)(p.mb[2] + r), (double)p.mb[2], (double)p.mb[3]);
    double ys = polyval6(ln, pn, p.cpm_az) + p.az_shift;
    double xs = polyval6(ln, pn, p.cpm_rg);
    // keep the int conversion defined for wild polynomials; anything this
    // far out lands in the zero region anyway
    ys = fmin(fmax(ys, -1.0e9), 1.0e9);
    xs = fmin(fmax(xs, -1.0e9), 1.0e9);
    iy = (int)trunc(ys);
    fy = (int)((ys - iy) * kInterval + 0.5);
    ix = (int)trunc(xs);
    fx = (int)((xs - ix) * kInterval + 0.5);
}

// 0: outside, 1: edge band, 2: interior (DerampAndResample.cu:383-403)
__device__ inline int region_of(const ResampleParams& p, int iy, int ix)
{
    if (iy > p.sb[3] || iy < p.sb[2] || ix > p.sb[1] || ix < p.sb[0]) return 0;
    if (iy > p.sb[3] - kHalf - 2 || iy < p.sb[2] + kHalf || ix > p.sb[1] - kHalf - 2 ||
        ix < p.sb[0] + kHalf)
        return 1;
    return 2;
}

__device__ inline double deramp_phase_direct(const ResampleParams& p, int ya, int xa)
{
    double ta = (ya - p.first_line) * p.dt;
    double d = ta - p.tref[xa];
    return -kPi * p.kt[xa] * (d * d) - 2.0 * kPi * p.fdc[xa] * ta;
}

__device__ inline void phase_coeffs(const ResampleParams& p, int xa, double& a,
                                    double& b, double& c)
{
    double kt = p.kt[xa], tr = p.tref[xa], fd = p.fdc[xa];
    a = -kPi * kt * (tr * tr);
    b = 2.0 * kPi * p.dt * (kt * tr - fd);
    c = -kPi * kt * (p.dt * p.dt);
}

__device__ inline float reduce_2pi(double ph)
{
    double k = rint(ph * kInvTwoPi);
    return (float)fma(-k, kTwoPi, ph);
}

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

// ======================================================================
// Fused fast form
// ======================================================================
constexpr int FT_ROWS = 16;          // output rows per tile (4 per wave)
constexpr int FT_COLS = 64;          // output cols per tile (one per lane)
constexpr int FT_MAXY = 44;          // LDS slave window capacity (rows)
constexpr int FT_MAXX = 100;         // LDS slave window capacity (cols)

__device__ inline int wave_min(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline int wave_max(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline void load_taps(const float* lut, int row, float k[kTaps])
{
    const float4* q = reinterpret_cast<const float4*>(lut + row * kTaps);
    float4 a = q[0], b = q[1], c = q[2];
    k[0] = a.x; k[1] = a.y; k[2] = a.z; k[3] = a.w;
    k[4] = b.x; k[5] = b.y; k[6] = b.z; k[7] = b.w;
    k[8] = c.x; k[9] = c.y; k[10] = c.z; k[11] = c.w;
}

__device__ inline tops_cf32 reramp(float sr, float si, double phi)
{
    float sn, cs;
    sincosf(reduce_2pi(phi), &sn, &cs);
    tops_cf32 o;
    o.x = sr * cs + si * sn;
    o.y = si * cs - sr * sn;
    return o;
}

// Interior output computed straight from HBM (used when the tile's slave
// window does not fit the LDS capacity: strongly warped offset fields).
__device__ __noinline__ tops_cf32 interior_direct(const ResampleParams& p, int iy, int fy, int ix, int fx)
{
    float kl[kTaps], kp[kTaps];
    load_taps(p.kaz, fy, kl);
    load_taps(p.krg, fx, kp);
    float sr = 0.f, si = 0.f;
    double A = 0.0, B = 0.0, C = 0.0;
    for (int j = 0; j < kTaps; j++) {
        int ya = iy - kHalf + j;
        float tr = 0.f, ti = 0.f;
        for (int i = 0; i < kTaps; i++) {
            int xa = ix - kHalf + i;
            double a, b, c;
            phase_coeffs(p, xa, a, b, c);
            double y = (double)(ya - p.first_line);
            float sn, cs;
            sincosf(reduce_2pi(fma(y, fma(y, c, b), a)), &sn, &cs);
            tops_ci16 s = p.slave[(size_t)(ya - p.sb[2]) * p.s_pixels + (xa - p.sb[0])];
            float dr = (float)s.re * cs - (float)s.im * sn;
            float di = (float)s.re * sn + (float)s.im * cs;
            tr = fmaf(kp[i], dr, tr);
            ti = fmaf(kp[i], di, ti);
            if (j == 0) {
                A = fma((double)kp[i], a, A);
                B = fma((double)kp[i], b, B);
                C = fma((double)kp[i], c, C);
            }
        }
        sr = fmaf(kl[j], tr, sr);
        si = fmaf(kl[j], ti, si);
    }
    double4 m = reinterpret_cast<const double4*>(p.mom)[fy];
    double y0 = (double)(iy - p.first_line);
    double phi = m.x * A + fma(y0, m.x, m.y) * B + (y0 * y0 * m.x + 2.0 * y0 * m.y + m.z) * C;
    return reramp(sr, si, phi);
}

__global__ void __launch_bounds__(256) resample_fast_kernel(ResampleParams p)
{
    __shared__ float2 sD[FT_MAXY * FT_MAXX];
    __shared__ double sA[FT_MAXX], sB[FT_MAXX], sC[FT_MAXX];
    __shared__ int sRed[4][4];
    __shared__ int sBox[4];

    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int c = blockIdx.x * FT_COLS + lane;
    const int rbase = blockIdx.y * FT_ROWS + wv * 4;

    int iy[4], fy[4], ix[4], fx[4], reg[4];
    int ymin = INT_MAX, ymax = INT_MIN, xmin = INT_MAX, xmax = INT_MIN;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int r = rbase + k;
        reg[k] = -1;
        if (r < p.m_lines && c < p.m_pixels) {
            slave_pos(p, r, c, iy[k], fy[k], ix[k], fx[k]);
            reg[k] = region_of(p, iy[k], ix[k]);
            if (reg[k] == 2) {
                ymin = min(ymin, iy[k]);
                ymax = max(ymax, iy[k]);
                xmin = min(xmin, ix[k]);
                xmax = max(xmax, ix[k]);
            }
        }
    }
    ymin = wave_min(ymin);
    ymax = wave_max(ymax);
    xmin = wave_min(xmin);
    xmax = wave_max(xmax);
    if (lane == 0) {
        sRed[wv][0] = ymin; sRed[wv][1] = ymax; sRed[wv][2] = xmin; sRed[wv][3] = xmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = INT_MAX, b = INT_MIN, cc = INT_MAX, d = INT_MIN;
        for (int w = 0; w < 4; w++) {
            a = min(a, sRed[w][0]); b = max(b, sRed[w][1]);
            cc = min(cc, sRed[w][2]); d = max(d, sRed[w][3]);
        }
        sBox[0] = a; sBox[1] = b; sBox[2] = cc; sBox[3] = d;
    }
    __syncthreads();
    const bool any_interior = sBox[0] <= sBox[1];
    const int Y0 = sBox[0] - kHalf, X0 = sBox[2] - kHalf;
    const int ny = any_interior ? sBox[1] + kHalf + 1 - Y0 + 1 : 0;
    const int nx = any_interior ? sBox[3] + kHalf + 1 - X0 + 1 : 0;
    const bool staged = any_interior && ny <= FT_MAXY && nx <= FT_MAXX;

    if (staged) {
        // per-column phase coefficients, then the deramped window
        for (int i = threadIdx.x; i < nx; i += 256) {
            double a, b, cc;
            phase_coeffs(p, X0 + i, a, b, cc);
            sA[i] = a; sB[i] = b; sC[i] = cc;
        }
        __syncthreads();
        const int n = ny * nx;
        for (int q = threadIdx.x; q < n; q += 256) {
            int yy = q / nx, xx = q - yy * nx;
            int ya = Y0 + yy, xa = X0 + xx;
            double y = (double)(ya - p.first_line);
            float sn, cs;
            sincosf(reduce_2pi(fma(y, fma(y, sC[xx], sB[xx]), sA[xx])), &sn, &cs);
            tops_ci16 s = p.slave[(size_t)(ya - p.sb[2]) * p.s_pixels + (xa - p.sb[0])];
            float re = (float)s.re, im = (float)s.im;
            sD[q] = make_float2(re * cs - im * sn, re * sn + im * cs);
        }
        __syncthreads();
    }

#pragma unroll 1
    for (int k = 0; k < 4; k++) {
        int r = rbase + k;
        if (reg[k] < 0) continue;
        tops_cf32 o = {0.0f, 0.0f};
        if (reg[k] == 1) {
            tops_ci16 s = p.slave[(size_t)(iy[k] - p.sb[2]) * p.s_pixels + (ix[k] - p.sb[0])];
            o.x = (float)s.re;
            o.y = (float)s.im;
        } else if (reg[k] == 2) {
            if (staged) {
                float kl[kTaps], kp[kTaps];
                load_taps(p.kaz, fy[k], kl);
                load_taps(p.krg, fx[k], kp);
                const int oy = iy[k] - kHalf - Y0, ox = ix[k] - kHalf - X0;
                const float2* w = sD + oy * nx + ox;
                float sr = 0.f, si = 0.f;
#pragma unroll 2
                for (int j = 0; j < kTaps; j++) {
                    float tr = 0.f, ti = 0.f;
#pragma unroll
                    for (int i = 0; i < kTaps; i++) {
                        float2 d = w[j * nx + i];
                        tr = fmaf(kp[i], d.x, tr);
                        ti = fmaf(kp[i], d.y, ti);
                    }
                    sr = fmaf(kl[j], tr, sr);
                    si = fmaf(kl[j], ti, si);
                }
                double A = 0.0, B = 0.0, C = 0.0;
#pragma unroll
                for (int i = 0; i < kTaps; i++) {
                    double kk = (double)kp[i];
                    A = fma(kk, sA[ox + i], A);
                    B = fma(kk, sB[ox + i], B);
                    C = fma(kk, sC[ox + i], C);
                }
                double4 m = reinterpret_cast<const double4*>(p.mom)[fy[k]];
                double y0 = (double)(iy[k] - p.first_line);
                double phi = m.x * A + fma(y0, m.x, m.y) * B +
                             (y0 * y0 * m.x + 2.0 * y0 * m.y + m.z) * C;
                o = reramp(sr, si, phi);
            } else {
                o = interior_direct(p, iy[k], fy[k], ix[k], fx[k]);
            }
        }
        p.out[(size_t)r * p.m_pixels + c] = o;
    }
}

int launch_resample_fast(const ResampleParams& p, hipStream_t s)
{
    dim3 grid((p.m_pixels + FT_COLS - 1) / FT_COLS, (p.m_lines + FT_ROWS - 1) / FT_ROWS);
    hipLaunchKernelGGL(resample_fast_kernel, grid, dim3(256), 0, s, p);
    TOPS_HIP(hipGetLastError());
    return TOPS_OK;
}

}  // namespace tops
