// geocoreg.hip -- DEM-assisted geometric coregistration fit (gfx950, FP64).
//
// Semantics of Geocoreg_cuda.cu: per DEM post (row r, col c) at
// (lat0 - r*dlat, lon0 + c*dlon) deg with height h (int, nodata skipped):
//   llh2xyz on WGS84 (87-99), then for master and slave orbits a
//   zero-Doppler Newton solve in azimuth time, at most 11 steps, stop at
//   |dt| < 1e-10 (272-294):  dt = -(v.d) / (a.d - v.v), d = P - S(t)
//   line  = lineOffset + (t - firstLineTime) / azimuthTimeInterval (302)
//   pixel = (|P - S(t)| - slrTimeToFirstPixel * c) / rangeSpacing (304-307)
// then weight = post valid and master (line, pixel) inside the burst box
// (WeighthSet, 548-574) and the least-squares fit of slave line and pixel
// on [1, L, P, L^2, LP, P^2] of the normalised master position (AMatrixSet,
// 576-605; cuBLAS Ddgmm/Dgemm + cuSOLVER potrf/potrs, 819-855).
//
// Here the design matrix is never formed: each thread projects one post
// against both orbits and adds its weighted contribution to the 21 unique
// normal-matrix entries, the 12 right-hand-side entries and a count; the
// workgroup reduces them with wave shuffles and writes one FP64 partial row.
// A second kernel adds the partial rows in a fixed order; the 6x6 Cholesky
// solve runs on the host.
#include "tops_internal.hpp"

namespace tops {

constexpr int GEO_THREADS = 256;

__device__ inline void orbit_pos(const OrbitDev& o, double tn, double s[3])
{
    s[0] = s[1] = s[2] = 0.0;
    for (int i = o.ncoef - 1; i >= 0; --i) {
        s[0] = fma(s[0], tn, o.cx[i]);
        s[1] = fma(s[1], tn, o.cy[i]);
        s[2] = fma(s[2], tn, o.cz[i]);
    }
}

// velocity (/10) and acceleration (/100) of Geocoreg_cuda.cu:149-205, with
// the powers of tn built incrementally instead of pow()
__device__ inline void orbit_vel_acc(const OrbitDev& o, double tn, double v[3], double a[3])
{
    v[0] = o.cx[1]; v[1] = o.cy[1]; v[2] = o.cz[1];
    a[0] = a[1] = a[2] = 0.0;
    double p1 = 1.0;   // tn^(i-2)
    for (int i = 2; i < o.ncoef; ++i) {
        double pv = (double)i * (p1 * tn);             // i tn^(i-1)
        double pa = (double)((i - 1) * i) * p1;        // i(i-1) tn^(i-2)
        v[0] = fma(o.cx[i], pv, v[0]);
        v[1] = fma(o.cy[i], pv, v[1]);
        v[2] = fma(o.cz[i], pv, v[2]);
        a[0] = fma(o.cx[i], pa, a[0]);
        a[1] = fma(o.cy[i], pa, a[1]);
        a[2] = fma(o.cz[i], pa, a[2]);
        p1 *= tn;
    }
    v[0] /= 10.0; v[1] /= 10.0; v[2] /= 10.0;
    a[0] /= 100.0; a[1] /= 100.0; a[2] /= 100.0;
}

__device__ inline void project(const OrbitDev& o, const double P[3], double& line, double& pixel)
{
    double t = o.az_init, s[3], v[3], a[3];
    for (int it = 0; it <= 10; it++) {
        double tn = (t - o.mid_time) / 10.0;
        orbit_pos(o, tn, s);
        orbit_vel_acc(o, tn, v, a);
        double d0 = P[0] - s[0], d1 = P[1] - s[1], d2 = P[2] - s[2];
        double sol = -(v[0] * d0 + v[1] * d1 + v[2] * d2) /
                     ((a[0] * d0 + a[1] * d1 + a[2] * d2) - v[0] * v[0] - v[1] * v[1] - v[2] * v[2]);
        t += sol;
        if (fabs(sol) < 1e-10) break;
    }
    double tn = (t - o.mid_time) / 10.0;
    orbit_pos(o, tn, s);
    line = o.line_offset + (t - o.first_line_time) / o.dt;
    double d0 = P[0] - s[0], d1 = P[1] - s[1], d2 = P[2] - s[2];
    pixel = (sqrt(d0 * d0 + d1 * d1 + d2 * d2) - o.slr_first * kSol) / o.range_spacing;
}

__device__ inline double wave_sum_d(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ void __launch_bounds__(GEO_THREADS) geocoreg_kernel(
    const int32_t* dem, int lines, int pixels, double lat0, double lon0, double dlat,
    double dlon, int nodata, OrbitDev m, OrbitDev s, int xmin, int xmax, int ymin, int ymax,
    double* partials, double* lp)
{
    __shared__ double red[GEO_THREADS / 64][kGeoSums];
    const long n = (long)lines * pixels;
    double acc[kGeoSums];
#pragma unroll
    for (int k = 0; k < kGeoSums; k++) acc[k] = 0.0;

    for (long t = (long)blockIdx.x * GEO_THREADS + threadIdx.x; t < n;
         t += (long)gridDim.x * GEO_THREADS) {
        int row = (int)(t / pixels), col = (int)(t - (long)row * pixels);
        int hgt = dem[t];
        double ml = -99999999.0, mp = -99999999.0, sl = -99999999.0, sp = -99999999.0;
        if ((float)hgt != (float)nodata) {
            double lat = (lat0 - dlat * row) * kPi / 180.0;
            double lon = (lon0 + dlon * col) * kPi / 180.0;
            double sphi, cphi, slam, clam;
            sincos(lat, &sphi, &cphi);
            sincos(lon, &slam, &clam);
            const double e2 = 1.0 - (kEllB * kEllB / kEllA / kEllA);
            double h = (double)(float)hgt;
            double nph = kEllA / sqrt(1.0 - e2 * sphi * sphi) + h;
            double P[3] = {nph * cphi * clam, nph * cphi * slam, (nph - e2 * (nph - h)) * sphi};
            project(m, P, ml, mp);
            project(s, P, sl, sp);
            if (ml >= ymin && ml <= ymax && mp >= xmin && mp <= xmax) {
                double L = norm_warp(ml, ymin, ymax), Pn = norm_warp(mp, xmin, xmax);
                double a[6] = {1.0, L, Pn, L * L, L * Pn, Pn * Pn};
                int k = 0;
#pragma unroll
                for (int i = 0; i < 6; i++)
#pragma unroll
                    for (int j = i; j < 6; j++) acc[k++] += a[i] * a[j];
#pragma unroll
                for (int i = 0; i < 6; i++) {
                    acc[21 + i] += a[i] * sl;
                    acc[27 + i] += a[i] * sp;
                }
                acc[33] += 1.0;
            }
        }
        if (lp) {
            lp[t] = ml;
            lp[n + t] = mp;
            lp[2 * n + t] = sl;
            lp[3 * n + t] = sp;
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kGeoSums; k++) {
        double v = wave_sum_d(acc[k]);
        if (lane == 0) red[wv][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kGeoSums) {
        double v = 0.0;
        for (int w = 0; w < GEO_THREADS / 64; w++) v += red[w][threadIdx.x];
        partials[(size_t)blockIdx.x * kGeoSums + threadIdx.x] = v;
    }
}

// One workgroup per sum.  Thread t adds partials t, t+256, ... in order, then
// a fixed LDS tree adds the 256 thread sums: the result depends only on
// nblocks, never on timing.
__global__ void __launch_bounds__(256) geocoreg_final_kernel(double* partials, int nblocks)
{
    __shared__ double sh[256];
    const int k = blockIdx.x, t = threadIdx.x;
    double v = 0.0;
    for (int b = t; b < nblocks; b += 256) v += partials[(size_t)b * kGeoSums + k];
    sh[t] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    if (t == 0) partials[(size_t)nblocks * kGeoSums + k] = sh[0];
}

int geocoreg_blocks(long posts)
{
    long b = (posts + GEO_THREADS - 1) / GEO_THREADS;
    if (b > 2048) b = 2048;
    if (b < 1) b = 1;
    return (int)b;
}

int launch_geocoreg(const int32_t* dem, int lines, int pixels, double lat0, double lon0,
                    double dlat, double dlon, int nodata, const OrbitDev& m,
                    const OrbitDev& s, int xmin, int xmax, int ymin, int ymax,
                    double* d_partials, int nblocks, double* d_lp, hipStream_t st)
{
    hipLaunchKernelGGL(geocoreg_kernel, dim3(nblocks), dim3(GEO_THREADS), 0, st, dem, lines,
                       pixels, lat0, lon0, dlat, dlon, nodata, m, s, xmin, xmax, ymin, ymax,
                       d_partials, d_lp);
    TOPS_HIP(hipGetLastError());
    hipLaunchKernelGGL(geocoreg_final_kernel, dim3(kGeoSums), dim3(256), 0, st, d_partials,
                       nblocks);
    TOPS_HIP(hipGetLastError());
    return TOPS_OK;
}

}  // namespace tops
