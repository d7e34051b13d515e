// tops_internal.hpp -- shared definitions of the MI355X TOPS burst library.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <atomic>

#include "tops_abi.h"

namespace tops {

constexpr double kPi = 3.1415926535897931;        // Constants.h:8
constexpr double kTwoPi = 6.283185307179586232;
constexpr double kInvTwoPi = 0.15915494309189533577;
constexpr double kSol = 299792458.0;              // cuConstants.cuh:7
constexpr double kEllA = 6378137.0;               // cuConstants.cuh:5
constexpr double kEllB = 6356752.3142451794975639665996337;  // cuConstants.cuh:6
constexpr int kLutRows = 2048;                    // Interval + 1, Constants.h:9
constexpr int kAzShifts = 5;     // window offsets the shared azimuth pass takes
constexpr int kAzShiftTaps = 16; // taps per shifted azimuth row (the window height)
constexpr int kInterval = 2047;
constexpr int kTaps = 12;                         // Main.cpp:75, 113
constexpr int kHalf = kTaps / 2 - 1;              // c_Npoints2m1
// slave lines a burst geometry may address (0 .. kMaxSlaveLine - 1): the
// resample walker packs (line, LUT row) into one int as line * 2048 + row
constexpr int kMaxSlaveLine = (1 << 20) - 64;

// A value computed once per device from its attributes (occupancy, CU
// count).  Concurrent first calls compute the same value, so relaxed loads
// and stores suffice; devices past the cache take `fallback`.
constexpr int kMaxCachedDevices = 64;
template <class F>
int per_device_cached(std::atomic<int>* cache, int fallback, F&& compute)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxCachedDevices) return fallback;
    int v = cache[dev].load(std::memory_order_relaxed);
    if (!v) {
        v = compute(dev);
        cache[dev].store(v, std::memory_order_relaxed);
    }
    return v;
}
// the fast kernels address the slave raster with 32-bit buffer offsets
constexpr long long kMaxSlaveBytes = (1ll << 31) - 1;

// ---- status (thread-local, read through tops_last_error) ---------------
int set_error(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);
int check_burst_geom(const tops_burst_geom* g, const char* who);

#define TOPS_HIP(call)                                                       \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) return ::tops::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

// ---- resample context ---------------------------------------------------
struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);   // grows with hipMalloc outside any capture
    void release();
};

}  // namespace tops

struct tops_ctx {
    float* d_kaz = nullptr;     // [2048 x 12] azimuth LUT
    float* d_krg = nullptr;     // [2048 x 12] range LUT
    double* d_mom = nullptr;    // [2048 x 4] azimuth LUT moments M0 M1 M2 0
    float* d_kaz_sh = nullptr;  // [5 x 2048 x 16] azimuth rows shifted by 0..4 taps in 16
    int npoints = 12;
    tops::Scratch scratch;      // FP64 baseline intermediates
};

namespace tops {

// 6-term polynomial, cuConstants.cuh:57-66.  Evaluated left to right with
// FMA contraction off: the slave position selects a LUT row by truncation,
// so it must be bit-identical wherever it is computed (positions on an
// exact LUT-row boundary occur for offset fields with short decimal
// coefficients).
// Workgroup i of n runs on XCD i % 8 (the dispatcher deals workgroups to
// the 8 XCDs in turn).  xcd_remap gives each XCD a contiguous range of
// [0, n), so work items numbered next to each other (tiles or strips whose
// halos overlap) run on one XCD and share its L2.  A bijection for any n.
__device__ inline int xcd_remap(int i, int n)
{
    const int x = i & 7, q = n >> 3, r = n & 7;
    return x * q + min(x, r) + (i >> 3);
}

__host__ __device__ inline double polyval6(double x, double y, const double* c)
{
#pragma clang fp contract(off)
    return c[0] + c[1] * x + c[2] * y + c[3] * x * x + c[4] * x * y + c[5] * y * y;
}

// cuConstants.cuh:52-55 (contraction off, as polyval6)
__host__ __device__ inline double norm_warp(double v, double lo, double hi)
{
#pragma clang fp contract(off)
    return (v - (0.5 * (lo + hi))) / (0.25 * (hi - lo));
}

// Per-call resample parameters passed by value to the kernels.
struct ResampleParams {
    int mb[4];          // master box
    int sb[4];          // slave box
    int m_lines, m_pixels, s_lines, s_pixels;
    int first_line;     // first_line_in_burst
    double dt;          // azimuth time interval
    double az_shift;
    double cpm_az[6];
    double cpm_rg[6];
    const tops_ci16* slave;
    const double* kt;
    const double* tref;
    const double* fdc;
    const float* kaz;
    const float* krg;
    const double* mom;
    const float* kaz_sh;    // tops_ctx::d_kaz_sh
    const double* az_shift_dev;   // device azimuth shift used instead of az_shift, or null
    tops_cf32* out;
};

// launchers (resample.hip)
// synchronous copies of large pageable host buffers through pinned staging
// (ordered after the null stream's queued work)
int staged_upload(void* d_dst, const void* h_src, size_t bytes);
int staged_download(void* h_dst, const void* d_src, size_t bytes);

int launch_resample_fp64(tops_ctx* ctx, const ResampleParams& p, hipStream_t s);
int launch_resample_fast(const ResampleParams& p, hipStream_t s, int strip_rows = 0);
// output rows row0 .. row0+rows-1 of the master box only (the rest untouched)
int launch_resample_rows(const ResampleParams& p, int row0, int rows, hipStream_t s,
                         int strip_rows = 0);
int launch_resample_batch(const ResampleParams* d_items, const int4* d_work, int n_work,
                          hipStream_t s);
int resample_strip_count(int m_lines, int m_pixels, int strip_rows, int* nx, int* ny);
int resample_wave_slots();
int resample_rows_for_rounds(const int* lines, const int* pixels, int n, int rounds);

// launchers (coherence.hip)
int launch_coherence(const tops_ci16* m, const tops_cf32* s, int w, int h,
                     int dx, int dy, int win_rg, int win_az, float* coh,
                     int mode, hipStream_t st);
struct CohItem {
    const tops_ci16* m;
    const tops_cf32* s;
    float* coh;
    int w, h;
};
int launch_coherence_batch(const CohItem* d_items, const int4* d_work, int n_work, int dx,
                           int dy, int win_rg, int win_az, int walk_rows, hipStream_t st);
int coherence_tile_count(int w, int h, int* nx, int* ny);
// walking form (3 az x 10 rg window, dx = 4, dy = 1); batched work items are
// {item, column band, row block, coherence_walk_vec(...)}, blocks of
// walk_rows output rows
bool coherence_walk_fits(int dx, int dy, int win_rg, int win_az);
bool coherence_batch_fits(int dx, int dy, int win_rg, int win_az);
int coherence_walk_vec(const void* m, const void* s, const void* coh, int w);
int coherence_walk_count(int w, int h, int rows, int* nx, int* ny);

// launchers (fused.hip): one persistent launch of deramp -> resample ->
// reramp -> coherence over work items {burst, column strip, first row, rows}
// for the 3 az x 10 rg window (dx = 4, dy = 1); coh_of[burst] indexes
// coh_items or is -1 (no coherence); queue: 2 counters, zero before the
// first launch (each launch leaves them zero)
bool fused_window_fits(int dx, int dy, int win_rg, int win_az);
int fused_strip_count(int m_pixels);
int fused_grid();
int launch_fused_batch(const ResampleParams* d_items, const CohItem* d_coh, const int* d_coh_of,
                       const int4* d_work, int n_work, unsigned* d_queue, hipStream_t s);

// launchers (esd.hip): writes 2*n_overlap double sums to d_sums
int launch_esd(const tops_ci16* mf, const tops_ci16* mb, const tops_cf32* sf,
               const tops_cf32* sb, const int* overlap_size, int n_overlap,
               int pixels, int win_az, int win_rg, float thr, double* d_sums,
               double* d_partials, int max_blocks_per_band, hipStream_t st);
int esd_partial_capacity(const int* overlap_size, int n_overlap, int pixels);
// bands at line offsets f_offset[i] (mf, sf) and b_offset[i] (mb, sb)
int launch_esd_bands(const tops_ci16* mf, const tops_ci16* mb, const tops_cf32* sf,
                     const tops_cf32* sb, const int* f_offset, const int* b_offset,
                     const int* overlap_size, int n_overlap, int pixels, int win_az, int win_rg,
                     float thr, double* d_sums, double* d_partials, hipStream_t st);
// device shifts from the band sums (with the |shift| > 100 reset)
int launch_esd_shifts(const double* d_sums, int n, double df, double dt, double* d_shifts,
                      hipStream_t st);

// launchers (geocoreg.hip)
struct OrbitDev {
    double cx[6], cy[6], cz[6];
    int ncoef;
    double mid_time, dt, first_line_time, slr_first, range_spacing, az_init;
    int line_offset;
};
// DEM grid and ellipsoid of the geocoreg kernel (the FP32 projection's
// reference points are formed per tile from these)
struct GeoRef {
    double lat0, lon0, dlat, dlon;   // degrees: post (r, c) at (lat0 - r dlat, lon0 + c dlon)
    double e2, one_e2;               // e^2 and 1 - e^2 of WGS84
};
int launch_geocoreg(const int32_t* dem, int lines, int pixels, double lat0,
                    double lon0, double dlat, double dlon, int nodata,
                    const OrbitDev& m, const OrbitDev& s, int xmin, int xmax,
                    int ymin, int ymax, double* d_partials, int nblocks,
                    double* d_lp, hipStream_t st);
int geocoreg_blocks(int lines, int pixels);
// doubles of device scratch launch_geocoreg takes at d_partials
size_t geocoreg_scratch_doubles(int lines, int pixels, int nblocks);
// geocoreg sums: the moments L^i P^j (i + j <= 4, by degree, then by the
// power of P; the first is the post count), then 6 rhs_az and 6 rhs_rg
constexpr int kGeoMoments = 15;
constexpr int kGeoSums = kGeoMoments + 12;

}  // namespace tops
