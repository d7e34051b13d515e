// geocoreg.hip -- DEM-assisted geometric coregistration fit (gfx950; FP32
// DEM projection, FP64 orbit, Newton solve and fit).
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
// Precision split: the DEM projection (llh2xyz) runs in FP32 relative to a
// reference point per tile of 4 x 64 posts (its centre post at height 0,
// whose ECEF position Pr each wave forms in FP64).  With dphi = lat - lat_r
// and dlam = lon - lon_r (at most 2 rows and 32 columns of posts), angle addition
// gives the small changes of sin / cos from sin(d) and 1 - cos(d), and the
// prime-vertical radius changes by dN = Nr ((1 - u)^-1/2 - 1) with
// u = e2 d(sin^2 phi) / (1 - e2 sin^2 lat_r) (series in u).  The post's
// offset dP from Pr is a sum of FP32 terms of at most a few hundred metres,
// well under a millimetre in FP32; P = Pr + dP is formed in FP64 for the
// zero-Doppler Newton solve, which stays FP64 with the orbit polynomials and
// the normal-equation sums.
//
// Here the design matrix is never formed: each thread projects one post
// against both orbits and adds its weighted contribution to the 15 moments
// L^i P^j (i + j <= 4) that the 21 unique normal-matrix entries are made of
// (the first is the count of posts used) and to the 12 right-hand-side
// entries; the workgroup reduces them with wave shuffles and writes one FP64
// partial row.
// A second kernel adds the partial rows in a fixed order; the 6x6 Cholesky
// solve runs on the host.
#include "tops_internal.hpp"

namespace tops {

constexpr int GEO_THREADS = 256;
constexpr int GT_ROWS = 4, GT_COLS = 64;   // posts per tile (one FP64 reference each)

// The orbit polynomials run over all 6 coefficient slots, unrolled: the slots
// past ncoef hold zeros, and their terms add exact zeros (fma(0, x, c) = c,
// fma(0, x, v) = v).
constexpr int kOrbitSlots = 6;
static_assert(sizeof(OrbitDev::cx) == kOrbitSlots * sizeof(double), "coefficient slots");

__device__ inline void orbit_pos(const OrbitDev& o, double tn, double s[3])
{
    s[0] = s[1] = s[2] = 0.0;
#pragma unroll
    for (int i = kOrbitSlots - 1; i >= 0; --i) {
        s[0] = fma(s[0], tn, o.cx[i]);
        s[1] = fma(s[1], tn, o.cy[i]);
        s[2] = fma(s[2], tn, o.cz[i]);
    }
}

// velocity x 10 and acceleration x 100: Geocoreg_cuda.cu:149-205 without
// their final / 10 and / 100, which the Newton step folds into one factor;
// the powers of tn built incrementally instead of pow()
__device__ inline void orbit_vel10_acc100(const OrbitDev& o, double tn, double v[3], double a[3])
{
    v[0] = o.cx[1]; v[1] = o.cy[1]; v[2] = o.cz[1];
    a[0] = a[1] = a[2] = 0.0;
    double p1 = 1.0;   // tn^(i-2)
#pragma unroll
    for (int i = 2; i < kOrbitSlots; ++i) {
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
}

// Zero-Doppler Newton solve (Geocoreg_cuda.cu:272-294) from t: at most 11
// steps, stopping at a step below 1e-10 s.  The normalised time
// (t - t_mid) / 10 is formed as a product with 0.1 (within an ulp), and with
// v = V / 10, a = A / 100 the step -(v.d) / (a.d - v.v) is
// -10 (V.d) / (A.d - V.V): one FP64 division per step instead of eight.
__device__ inline double zero_doppler_time(const OrbitDev& o, const double P[3], double t)
{
    double s[3], v[3], a[3];
    for (int it = 0; it <= 10; it++) {
        double tn = (t - o.mid_time) * 0.1;
        orbit_pos(o, tn, s);
        orbit_vel10_acc100(o, tn, v, a);
        double d0 = P[0] - s[0], d1 = P[1] - s[1], d2 = P[2] - s[2];
        double sol = -10.0 * (v[0] * d0 + v[1] * d1 + v[2] * d2) /
                     ((a[0] * d0 + a[1] * d1 + a[2] * d2) - v[0] * v[0] - v[1] * v[1] - v[2] * v[2]);
        t += sol;
        if (fabs(sol) < 1e-10) break;
    }
    return t;
}

// line and pixel of P at its zero-Doppler time t (Geocoreg_cuda.cu:302-307)
__device__ inline void line_pixel(const OrbitDev& o, const double P[3], double t, double& line,
                                  double& pixel)
{
    double s[3];
    double tn = (t - o.mid_time) * 0.1;
    orbit_pos(o, tn, s);
    line = o.line_offset + (t - o.first_line_time) / o.dt;
    double d0 = P[0] - s[0], d1 = P[1] - s[1], d2 = P[2] - s[2];
    pixel = (sqrt(d0 * d0 + d1 * d1 + d2 * d2) - o.slr_first * kSol) / o.range_spacing;
}

// sin x and 1 - cos x of a small angle (|x| <= 1e-2 rad; offsets within a
// tile are a few 1e-4): Taylor terms to x^5 and x^6, so the FP32 results keep
// their relative precision
__device__ inline float sin_small(float x)
{
    const float x2 = x * x;
    return x * (1.0f - x2 * (1.0f / 6.0f) * (1.0f - x2 * (1.0f / 20.0f)));
}
__device__ inline float omc_small(float x)
{
    const float x2 = x * x;
    return 0.5f * x2 * (1.0f - x2 * (1.0f / 12.0f) * (1.0f - x2 * (1.0f / 30.0f)));
}

// FP64 constants of llh2xyz at a tile's reference point (lat_r, lon_r, 0)
// (Geocoreg_cuda.cu:87-99)
struct TileRef {
    double pr[3];   // ECEF of the reference post
    double nv[3];   // the ellipsoid normal there (cos lat cos lon, cos lat sin lon, sin lat)
    float sr, cr, slr, clr, nr, inv_w2;
    // zero-Doppler times of the reference post on the master and slave
    // orbits (solved from the orbits' own starting times): the starting
    // times of the tile's posts, a few posts away, so their solves take
    // fewer steps to the same roots
    double t0[2];
};

__device__ inline TileRef tile_reference(const GeoRef& g, int row, int col)
{
    const double d2r = kPi / 180.0;
    TileRef t;
    double sr, cr, slr, clr;
    sincos((g.lat0 - g.dlat * row) * d2r, &sr, &cr);
    sincos((g.lon0 + g.dlon * col) * d2r, &slr, &clr);
    const double w2 = 1.0 - g.e2 * sr * sr;
    const double nr = kEllA / sqrt(w2);
    t.pr[0] = nr * cr * clr;
    t.pr[1] = nr * cr * slr;
    t.pr[2] = nr * g.one_e2 * sr;
    t.nv[0] = cr * clr;
    t.nv[1] = cr * slr;
    t.nv[2] = sr;
    t.sr = (float)sr; t.cr = (float)cr; t.slr = (float)slr; t.clr = (float)clr;
    t.nr = (float)nr;
    t.inv_w2 = (float)(1.0 / w2);
    return t;
}

// ECEF of the post at (lat_r + dphi, lon_r + dlam, h) less Pr + h n_r (FP32;
// the height along the reference normal is added in FP64 by the caller)
__device__ inline void llh_offset(const TileRef& t, float e2, float one_e2, float dphi,
                                  float dlam, float h, float d[3])
{
    const float sp = sin_small(dphi), sl = sin_small(dlam);
    const float cp1 = omc_small(dphi), cl1 = omc_small(dlam);   // 1 - cos
    const float ds = t.cr * sp - t.sr * cp1;       // sin(lat) - sin(lat_r)
    const float dc = -t.cr * cp1 - t.sr * sp;      // cos(lat) - cos(lat_r)
    const float dsl = t.clr * sl - t.slr * cl1;    // sin(lon) - sin(lon_r)
    const float dcl = -t.clr * cl1 - t.slr * sl;   // cos(lon) - cos(lon_r)
    const float u = e2 * (2.0f * t.sr * ds + ds * ds) * t.inv_w2;
    const float dn = t.nr * (u * (0.5f + u * (0.375f + u * 0.3125f)));
    const float c = t.cr + dc, nh = t.nr + h;
    d[0] = dn * c * (t.clr + dcl) + nh * (t.cr * dcl + dc * t.clr + dc * dcl);
    d[1] = dn * c * (t.slr + dsl) + nh * (t.cr * dsl + dc * t.slr + dc * dsl);
    d[2] = dn * one_e2 * (t.sr + ds) + (t.nr * one_e2 + h) * ds;
}

__device__ inline double wave_sum_d(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// the kernel's arguments in one block, read in place in the kernel-argument
// segment: the two orbits' coefficients become scalar loads where the
// Newton solve uses them instead of SGPRs held for the whole kernel (which
// spilled to VGPR lanes)
struct GeoArgs {
    const int32_t* dem;
    int lines, pixels;
    GeoRef g;
    int nodata;
    OrbitDev m, s;
    int xmin, xmax, ymin, ymax;
    double* partials;
    double* lp;
    const TileRef* refs;   // per tile, from geocoreg_tile_ref_kernel
};

// The FP64 reference point of every tile (its sines and cosines are about a
// fifth of a wave's instructions per tile when every wave forms them)
__global__ void __launch_bounds__(256) geocoreg_tile_ref_kernel(GeoRef g, OrbitDev m,
                                                                OrbitDev s, int lines,
                                                                int pixels, TileRef* refs)
{
    const int tiles_x = (pixels + GT_COLS - 1) / GT_COLS;
    const long ntiles = (long)tiles_x * ((lines + GT_ROWS - 1) / GT_ROWS);
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= ntiles) return;
    const int tr0 = (int)(i / tiles_x) * GT_ROWS, tc0 = (int)(i % tiles_x) * GT_COLS;
    TileRef r = tile_reference(g, tr0 + GT_ROWS / 2, tc0 + GT_COLS / 2);
    r.t0[0] = zero_doppler_time(m, r.pr, m.az_init);
    r.t0[1] = zero_doppler_time(s, r.pr, s.az_init);
    // a solve that did not give a finite time leaves the posts the orbit's
    // own starting time
    if (!isfinite(r.t0[0])) r.t0[0] = m.az_init;
    if (!isfinite(r.t0[1])) r.t0[1] = s.az_init;
    refs[i] = r;
}

// 4 waves per SIMD (127 VGPRs, no spill; unbounded the compiler takes 133
// and 3 waves)
__global__ void __launch_bounds__(GEO_THREADS) __attribute__((amdgpu_waves_per_eu(4)))
geocoreg_kernel(GeoArgs args)
{
    typedef __attribute__((address_space(4))) const GeoArgs CGeoArgs;
    const GeoArgs& A = *(const GeoArgs*)(CGeoArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    (void)args;
    const int32_t* dem = A.dem;
    const int lines = A.lines, pixels = A.pixels, nodata = A.nodata;
    const GeoRef& g = A.g;
    const OrbitDev& m = A.m;
    const OrbitDev& s = A.s;
    const int xmin = A.xmin, xmax = A.xmax, ymin = A.ymin, ymax = A.ymax;
    double* partials = A.partials;
    double* lp = A.lp;
    __shared__ double red[GEO_THREADS / 64][kGeoSums];
    const long n = (long)lines * pixels;
    double acc[kGeoSums];
#pragma unroll
    for (int k = 0; k < kGeoSums; k++) acc[k] = 0.0;

    // tiles of GT_ROWS x GT_COLS posts, one post per thread: wave w takes
    // row w of the tile, lane l its column l
    static_assert(GT_ROWS * GT_COLS == GEO_THREADS && GT_COLS == 64, "one post per thread");
    const int tiles_x = (pixels + GT_COLS - 1) / GT_COLS;
    const long ntiles = (long)tiles_x * ((lines + GT_ROWS - 1) / GT_ROWS);
    const float e2f = (float)g.e2, one_e2f = (float)g.one_e2;
    const float dlatf = (float)(g.dlat * kPi / 180.0), dlonf = (float)(g.dlon * kPi / 180.0);
    for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int tr0 = (int)(tile / tiles_x) * GT_ROWS, tc0 = (int)(tile % tiles_x) * GT_COLS;
        const int row = tr0 + (threadIdx.x >> 6), col = tc0 + (threadIdx.x & 63);
        if (row >= lines || col >= pixels) continue;
        const int rc = tr0 + GT_ROWS / 2, cc = tc0 + GT_COLS / 2;   // reference post
        const long t = (long)row * pixels + col;
        int hgt = dem[t];
        double ml = -99999999.0, mp = -99999999.0, sl = -99999999.0, sp = -99999999.0;
        if ((float)hgt != (float)nodata) {
            const TileRef ref = A.refs[tile];   // tile_reference(g, rc, cc)
            // angles from the reference post: a few posts times the spacing
            const float dphi = -(float)(row - rc) * dlatf;
            const float dlam = (float)(col - cc) * dlonf;
            float d[3];
            llh_offset(ref, e2f, one_e2f, dphi, dlam, (float)hgt, d);
            const double hd = (double)(float)hgt;
            double P[3] = {fma(hd, ref.nv[0], ref.pr[0]) + (double)d[0],
                           fma(hd, ref.nv[1], ref.pr[1]) + (double)d[1],
                           fma(hd, ref.nv[2], ref.pr[2]) + (double)d[2]};
            line_pixel(m, P, zero_doppler_time(m, P, ref.t0[0]), ml, mp);
            line_pixel(s, P, zero_doppler_time(s, P, ref.t0[1]), sl, sp);
            if (ml >= ymin && ml <= ymax && mp >= xmin && mp <= xmax) {
                const double L = norm_warp(ml, ymin, ymax), Pn = norm_warp(mp, xmin, xmax);
                const double L2 = L * L, LP = L * Pn, P2 = Pn * Pn;
                const double a[6] = {1.0, L, Pn, L2, LP, P2};
                // the normal matrix's 21 entries a_i a_j are 15 moments
                // L^i P^j (i + j <= 4): one sum each, the first the count
                const double mom[kGeoMoments] = {1.0,     L,       Pn,      L2,      LP,
                                                 P2,      L2 * L,  L2 * Pn, L * P2,  P2 * Pn,
                                                 L2 * L2, L2 * LP, L2 * P2, LP * P2, P2 * P2};
#pragma unroll
                for (int k = 0; k < kGeoMoments; k++) acc[k] += mom[k];
#pragma unroll
                for (int i = 0; i < 6; i++) {
                    acc[kGeoMoments + i] += a[i] * sl;
                    acc[kGeoMoments + 6 + i] += a[i] * sp;
                }
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

// one workgroup per tile of GT_ROWS x GT_COLS posts, at most 2048 (the
// grid-stride loop takes the rest)
int geocoreg_blocks(int lines, int pixels)
{
    long b = (long)((lines + GT_ROWS - 1) / GT_ROWS) * ((pixels + GT_COLS - 1) / GT_COLS);
    if (b > 2048) b = 2048;
    if (b < 1) b = 1;
    return (int)b;
}

static long geocoreg_tiles(int lines, int pixels)
{
    return (long)((lines + GT_ROWS - 1) / GT_ROWS) * ((pixels + GT_COLS - 1) / GT_COLS);
}

// the partial rows, the final sums and the tiles' reference points
size_t geocoreg_scratch_doubles(int lines, int pixels, int nblocks)
{
    static_assert(sizeof(TileRef) % sizeof(double) == 0, "table in doubles");
    return (size_t)kGeoSums * (nblocks + 1) +
           (size_t)geocoreg_tiles(lines, pixels) * (sizeof(TileRef) / sizeof(double));
}

int launch_geocoreg(const int32_t* dem, int lines, int pixels, double lat0, double lon0,
                    double dlat, double dlon, int nodata, const OrbitDev& m,
                    const OrbitDev& s, int xmin, int xmax, int ymin, int ymax,
                    double* d_partials, int nblocks, double* d_lp, hipStream_t st)
{
    GeoRef g;
    g.lat0 = lat0;
    g.lon0 = lon0;
    g.dlat = dlat;
    g.dlon = dlon;
    g.e2 = 1.0 - (kEllB * kEllB / kEllA / kEllA);   // b_ell_e2, cuConstants.cuh:13
    g.one_e2 = 1.0 - g.e2;
    GeoArgs a;
    a.dem = dem;
    a.lines = lines;
    a.pixels = pixels;
    a.g = g;
    a.nodata = nodata;
    a.m = m;
    a.s = s;
    a.xmin = xmin;
    a.xmax = xmax;
    a.ymin = ymin;
    a.ymax = ymax;
    a.partials = d_partials;
    a.lp = d_lp;
    TileRef* refs = reinterpret_cast<TileRef*>(d_partials + (size_t)kGeoSums * (nblocks + 1));
    a.refs = refs;
    const long ntiles = geocoreg_tiles(lines, pixels);
    hipLaunchKernelGGL(geocoreg_tile_ref_kernel, dim3((unsigned)((ntiles + 255) / 256)), dim3(256),
                       0, st, g, m, s, lines, pixels, refs);
    TOPS_HIP(hipGetLastError());
    hipLaunchKernelGGL(geocoreg_kernel, dim3(nblocks), dim3(GEO_THREADS), 0, st, a);
    TOPS_HIP(hipGetLastError());
    hipLaunchKernelGGL(geocoreg_final_kernel, dim3(kGeoSums), dim3(256), 0, st, d_partials,
                       nblocks);
    TOPS_HIP(hipGetLastError());
    return TOPS_OK;
}

}  // namespace tops
