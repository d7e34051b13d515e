// stages.cpp -- host stage layer of the TOPS burst path: the metadata
// helpers the reference's Func.cpp runs before the GPU stages (resample
// LUT, orbit fit, deramp parameters, overlap sizes, spectral separation) and
// the three per-subswath stages GeometricCoreg, DerampAndResample and
// ESDAndCoh over in-memory bursts (the reference reads and writes GeoTIFFs
// between them; here the bursts stay in memory and on the device).
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "tops_internal.hpp"

namespace tops {

static double sinc_pi(double x) { return x == 0 ? 1.0 : sin(kPi * x) / (kPi * x); }

static double rect(double x)
{
    if (x < 0.5 && x > -0.5) return 1.0;
    if (x == 0.5 || x == -0.5) return 0.5;
    return 0.0;
}

// Raised-cosine-windowed sinc tap at offset x for oversampling chi.
static double rc_tap(double x, int npoints, double chi)
{
#pragma clang fp contract(off)
    double v = 1.0 - 1.0 / chi;
    return sinc_pi(x) * rect(x / (double)npoints) * cos(v * kPi * x) /
           (1.0 - (2.0 * v * x) * (2.0 * v * x));
}

// Cholesky factorisation + solve of a small SPD system, in place.
static bool spd_solve(double* a, int n, double* const* rhs, int nrhs)
{
    for (int j = 0; j < n; j++) {
        double d = a[j * n + j];
        for (int k = 0; k < j; k++) d -= a[j * n + k] * a[j * n + k];
        if (!(d > 0.0)) return false;
        d = sqrt(d);
        a[j * n + j] = d;
        for (int i = j + 1; i < n; i++) {
            double v = a[i * n + j];
            for (int k = 0; k < j; k++) v -= a[i * n + k] * a[j * n + k];
            a[i * n + j] = v / d;
        }
    }
    for (int r = 0; r < nrhs; r++) {
        double* b = rhs[r];
        for (int i = 0; i < n; i++) {
            double v = b[i];
            for (int k = 0; k < i; k++) v -= a[i * n + k] * b[k];
            b[i] = v / a[i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double v = b[i];
            for (int k = i + 1; k < n; k++) v -= a[k * n + i] * b[k];
            b[i] = v / a[i * n + i];
        }
    }
    return true;
}

}  // namespace tops

using namespace tops;

extern "C" {

// InKernelInit, Func.cpp:2105-2158: row i (fraction i/2047), tap j at
// x = j - (npoints/2 - 1) - i/2047; chi_az = prf/abw, chi_rg = (rsr2x/2)/abw
// (the reference divides the range oversampling by the azimuth bandwidth,
// Func.cpp:2111, and so does this); rows normalised in FP64 and stored as
// float.
int tops_lut_init(int npoints, double prf, double abw, double rsr2x, float* kernel_az,
                  float* kernel_rg)
{
#pragma clang fp contract(off)
    if (npoints <= 0 || npoints > 64 || !kernel_az || !kernel_rg || abw <= 0.0)
        return set_error(TOPS_ERR_ARG, "tops_lut_init: bad arguments");
    const double chi_az = prf / abw, chi_rg = (rsr2x / 2.0) / abw;
    const int half = npoints / 2;
    for (int i = 0; i < kLutRows; i++) {
        double saz = 0.0, srg = 0.0;
        double taz[64], trg[64];
        for (int j = 0; j < npoints && j < 64; j++) {
            double x = 1.0 - half + j - (1.0 / kInterval) * i;
            taz[j] = rc_tap(x, npoints, chi_az);
            trg[j] = rc_tap(x, npoints, chi_rg);
            saz += taz[j];
            srg += trg[j];
        }
        for (int j = 0; j < npoints && j < 64; j++) {
            // stored as float first, then divided by the FP64 row sum
            kernel_az[i * npoints + j] = (float)((double)(float)taz[j] / saz);
            kernel_rg[i * npoints + j] = (float)((double)(float)trg[j] / srg);
        }
    }
    return TOPS_OK;
}

// Polyfit_Orbit, Func.cpp:1050-1126: least-squares polynomial of degree
// min(5, n-2) in t' = (t - t[n/2]) / 10 for x, y, z.  Returns the number of
// coefficients, 0 on failure.
int tops_polyfit_orbit(int n, const double* t, const double* x, const double* y,
                       const double* z, double* coef_x, double* coef_y, double* coef_z)
{
    if (n < 3 || !t || !x || !y || !z || !coef_x || !coef_y || !coef_z) {
        set_error(TOPS_ERR_ARG, "tops_polyfit_orbit: need at least 3 state vectors");
        return 0;
    }
    int deg = n > 6 ? 5 : n - 2;
    int nu = deg + 1;
    double N[36] = {0}, bx[6] = {0}, by[6] = {0}, bz[6] = {0};
    for (int k = 0; k < n; k++) {
        double tn = (t[k] - t[n / 2]) / 10.0, p[6];
        p[0] = 1.0;
        for (int j = 1; j < nu; j++) p[j] = p[j - 1] * tn;
        for (int i = 0; i < nu; i++) {
            for (int j = 0; j < nu; j++) N[i * nu + j] += p[i] * p[j];
            bx[i] += p[i] * x[k];
            by[i] += p[i] * y[k];
            bz[i] += p[i] * z[k];
        }
    }
    double* rhs[3] = {bx, by, bz};
    if (!spd_solve(N, nu, rhs, 3)) {
        set_error(TOPS_ERR_NOT_PD, "tops_polyfit_orbit: normal matrix not positive definite");
        return 0;
    }
    for (int i = 0; i < nu; i++) {
        coef_x[i] = bx[i];
        coef_y[i] = by[i];
        coef_z[i] = bz[i];
    }
    return nu;
}

// Orbit velocity magnitude at absolute time t (getVelocity, Func.cpp:1279-1300).
double tops_orbit_speed(double t, double mid_time, const double* cx, const double* cy,
                        const double* cz, int ncoef)
{
    double tn = (t - mid_time) / 10.0, v[3] = {cx[1], cy[1], cz[1]}, p = 1.0;
    for (int i = 2; i < ncoef; i++) {
        p *= tn;   // tn^(i-1)
        v[0] += cx[i] * i * p;
        v[1] += cy[i] * i * p;
        v[2] += cz[i] * i * p;
    }
    return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / 10.0;
}

// ComputeDerampPara formulas, Func.cpp:1955-2090 (the annotation XML part is
// not here): per burst b and range sample x, with two-way range time
// tau = 2 (slr0 + x dr / c):
//   fdc = dc0 + dc1 (tau - t0) + dc2 (tau - t0)^2, ka likewise from fm,
//   tref = lpb dt / 2 + fdc(fvp)/ka(fvp) - fdc/ka,
//   kt = ka krot / (ka - krot), krot = 2 v steer / lambda.
// dc/fm are [n_bursts][4] = (t0, c0, c1, c2).  Outputs [n_bursts x spb].
int tops_deramp_params(int n_bursts, int spb, double slr_time_first_pixel, double range_spacing,
                       int lines_per_burst, double az_time_interval, int first_valid_pixel,
                       const double* dc, const double* fm, double steering_rate_deg,
                       double wavelength, double orbit_speed, double* kt, double* tref,
                       double* fdc, double* ka)
{
    if (n_bursts <= 0 || spb <= 0 || !dc || !fm || !kt || !tref || !fdc || !ka ||
        first_valid_pixel < 0 || first_valid_pixel >= spb)
        return set_error(TOPS_ERR_ARG, "tops_deramp_params: bad arguments");
    const double krot = 2.0 * orbit_speed * (steering_rate_deg * kPi / 180.0) / wavelength;
    const double half_burst = lines_per_burst * az_time_interval / 2.0;
    for (int b = 0; b < n_bursts; b++) {
        const double* d = dc + 4 * b;
        const double* f = fm + 4 * b;
        double* F = fdc + (size_t)b * spb;
        double* K = ka + (size_t)b * spb;
        for (int x = 0; x < spb; x++) {
            double tau = (slr_time_first_pixel + x * range_spacing / kSol) * 2.0;
            double u = tau - d[0], w = tau - f[0];
            F[x] = d[1] + d[2] * u + d[3] * (u * u);
            K[x] = f[1] + f[2] * w + f[3] * w * w;
        }
        double t0 = half_burst + F[first_valid_pixel] / K[first_valid_pixel];
        for (int x = 0; x < spb; x++) {
            tref[(size_t)b * spb + x] = t0 - F[x] / K[x];
            kt[(size_t)b * spb + x] = K[x] * krot / (K[x] - krot);
        }
    }
    return TOPS_OK;
}

// EstOverlapSize, Func.cpp:2935-2960: n_i = floor((last[i] - first[i+1]) / dt)
// for bursts burst0..burstN-1.  Returns the number of overlaps.
int tops_overlap_sizes(const double* first_line_time, const double* last_line_time, int burst0,
                       int burstN, double az_time_interval, int* sizes, int* total, int* maximum)
{
    int n = 0, tot = 0, mx = -9999;
    for (int i = burst0; i < burstN; i++) {
        int v = (int)floor((last_line_time[i] - first_line_time[i + 1]) / az_time_interval);
        sizes[n++] = v;
        tot += v;
        if (v > mx) mx = v;
    }
    if (total) *total = tot;
    if (maximum) *maximum = mx;
    return n;
}

// GetSpectralSep, Func.cpp:2962-2976: mean of kt * lpb * dt over all bursts
// and range samples.
double tops_spectral_sep(const double* kt, int n_bursts, int spb, int lines_per_burst,
                         double az_time_interval)
{
    double cycle = lines_per_burst * az_time_interval, s = 0.0;
    for (size_t q = 0; q < (size_t)n_bursts * spb; q++) s += kt[q] * cycle;
    return s / spb / n_bursts;
}

}  // extern "C"

// ===========================================================================
// Per-subswath stages over in-memory bursts
// ===========================================================================
namespace {

struct DBuf {
    void* p = nullptr;
    size_t n = 0;
    ~DBuf() { release(); }
    int ensure(size_t bytes)
    {
        if (bytes <= n) return TOPS_OK;
        release();
        TOPS_HIP(hipMalloc(&p, bytes ? bytes : 1));
        n = bytes;
        return TOPS_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <class T> T* as() const { return (T*)p; }
};

int check_swath(const tops_swath* s, const char* who)
{
    if (!s || !s->slc || !s->burst_first_line_time || !s->burst_last_line_time ||
        s->lines_per_burst <= 0 || s->samples_per_burst <= 0 || s->n_bursts <= 0)
        return set_error(TOPS_ERR_ARG, "%s: incomplete swath description", who);
    return TOPS_OK;
}

int check_range(const tops_swath* m, const tops_swath* s, int mb0, int mbN, int sb0, int sbN,
                const char* who)
{
    if (mb0 < 0 || mbN < mb0 || mbN >= m->n_bursts || sb0 < 0 || sbN >= s->n_bursts ||
        sbN - sb0 != mbN - mb0)
        return set_error(TOPS_ERR_ARG, "%s: burst ranges [%d,%d] / [%d,%d] do not match the swaths",
                         who, mb0, mbN, sb0, sbN);
    return TOPS_OK;
}

// RefDem::getData (DataTypes.cpp:240-300): window around the lat/lon box
// extended by `margin` posts, clipped to the DEM; returns its origin.
bool dem_window(const tops_dem* d, double lat_min, double lat_max, double lon_min,
                double lon_max, int margin, int* r0, int* c0, int* lines, int* pixels)
{
    double ul0 = floor((d->lat_max - (lat_max + margin * d->dlat)) / d->dlat);
    double ul1 = floor(((lon_min - margin * d->dlon) - d->lon_min) / d->dlon);
    double lr0 = ceil((d->lat_max - (lat_min - margin * d->dlat)) / d->dlat);
    double lr1 = ceil(((lon_max + margin * d->dlon) - d->lon_min) / d->dlon);
    if (ul0 < 0) ul0 = 0;
    if (ul1 < 0) ul1 = 0;
    if (lr0 > d->lines - 1) lr0 = d->lines - 1;
    if (lr1 > d->pixels - 1) lr1 = d->pixels - 1;
    if (lr0 < ul0 || lr1 < ul1) return false;
    *r0 = (int)ul0;
    *c0 = (int)ul1;
    *lines = (int)(lr0 - ul0) + 1;
    *pixels = (int)(lr1 - ul1) + 1;
    return true;
}

}  // namespace

extern "C" {

// GeometricCoreg, Func.cpp:2215-2297.  For each master burst b of
// [mburst0, mburstN] (slave burst b + sburst0 - mburst0): box x in
// [0, spb-1], y in [b lpb, (b+1) lpb - 1]; DEM window = the burst's lat/lon
// box plus 24 posts; Newton start = mid-burst time.  Differences from the
// reference, which has these index bugs: the slave uses its own burst's
// first-line and mid times (the reference indexes them with the master burst,
// Func.cpp:2235, 2274), the DEM origin passed on is the window's own
// (Func.cpp:2273 passes the burst's lat_max/lon_min), and the whole int32
// window is used (Geocoreg_cuda.cu:788 copies half of it).
// cpm_az/cpm_rg: [(mburstN - mburst0 + 1) x 6].  A burst whose fit fails
// gets zero coefficients; the call then returns TOPS_ERR_NOT_PD after
// processing the other bursts.
int tops_geometric_coreg(const tops_swath* m, const tops_swath* s, const tops_orbit* mo,
                         const tops_orbit* so, const tops_dem* dem, int mburst0, int mburstN,
                         int sburst0, int sburstN, double* cpm_az, double* cpm_rg)
{
    int rc = check_swath(m, "tops_geometric_coreg");
    if (!rc) rc = check_swath(s, "tops_geometric_coreg");
    if (!rc) rc = check_range(m, s, mburst0, mburstN, sburst0, sburstN, "tops_geometric_coreg");
    if (rc) return rc;
    if (!mo || !so || !dem || !dem->data || !cpm_az || !cpm_rg || !m->burst_ll)
        return set_error(TOPS_ERR_ARG, "tops_geometric_coreg: null argument");
    const int gap = sburst0 - mburst0;
    DBuf d_dem;
    int status = TOPS_OK;
    for (int b = mburst0; b <= mburstN; b++) {
        const int sbi = b + gap;
        const double* ll = m->burst_ll + 4 * b;
        int r0, c0, lines, pixels;
        double* ca = cpm_az + 6 * (b - mburst0);
        double* cr = cpm_rg + 6 * (b - mburst0);
        if (!dem_window(dem, ll[0], ll[1], ll[2], ll[3], 24, &r0, &c0, &lines, &pixels)) {
            for (int i = 0; i < 6; i++) ca[i] = cr[i] = 0.0;
            status = set_error(TOPS_ERR_NOT_PD, "burst %d: no DEM coverage", b);
            continue;
        }
        std::vector<int32_t> win((size_t)lines * pixels);
        for (int i = 0; i < lines; i++)
            memcpy(&win[(size_t)i * pixels], dem->data + (size_t)(r0 + i) * dem->pixels + c0,
                   sizeof(int32_t) * pixels);
        if ((rc = d_dem.ensure(win.size() * sizeof(int32_t)))) return rc;
        TOPS_HIP(hipMemcpy(d_dem.p, win.data(), win.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        tops_orbit_geom gm = {mo->coef_x, mo->coef_y, mo->coef_z, mo->ncoef,
                              mo->az_time[mo->num_points / 2], m->az_time_interval,
                              m->burst_first_line_time[b], m->slr_time_first_pixel,
                              m->range_spacing,
                              0.5 * (m->burst_first_line_time[b] + m->burst_last_line_time[b]),
                              b * m->lines_per_burst};
        tops_orbit_geom gs = {so->coef_x, so->coef_y, so->coef_z, so->ncoef,
                              so->az_time[so->num_points / 2], s->az_time_interval,
                              s->burst_first_line_time[sbi], s->slr_time_first_pixel,
                              s->range_spacing,
                              0.5 * (s->burst_first_line_time[sbi] + s->burst_last_line_time[sbi]),
                              sbi * s->lines_per_burst};
        rc = tops_geocoreg(d_dem.as<int32_t>(), lines, pixels, dem->lat_max - r0 * dem->dlat,
                           dem->lon_min + c0 * dem->dlon, dem->dlat, dem->dlon, dem->nodata, &gm,
                           &gs, 0, m->samples_per_burst - 1, b * m->lines_per_burst,
                           (b + 1) * m->lines_per_burst - 1, ca, cr, nullptr, nullptr, nullptr);
        if (rc == TOPS_ERR_NOT_PD) {
            status = rc;
            continue;
        }
        if (rc) return rc;
    }
    return status;
}

// DerampAndResample, Func.cpp:2326-2406: each slave burst b + gap is
// deramped, resampled onto master burst b and reramped; out is the
// burst-stacked [(mburstN - mburst0 + 1) * lpb_m x spb_m] result (the
// reference's ReSlave raster).  The per-range arrays of the slave burst
// itself are used (the reference always passes the burst-0 row,
// Func.cpp:2383-2384).  d_out, when not NULL, receives the same raster on
// the device (dense) and out may then be NULL.
int tops_deramp_and_resample(const tops_swath* m, const tops_swath* s, const double* cpm_az,
                             const double* cpm_rg, int mburst0, int mburstN, int sburst0,
                             int sburstN, const float* kernel_az, const float* kernel_rg,
                             tops_cf32* out, tops_cf32* d_out)
{
    int rc = check_swath(m, "tops_deramp_and_resample");
    if (!rc) rc = check_swath(s, "tops_deramp_and_resample");
    if (!rc) rc = check_range(m, s, mburst0, mburstN, sburst0, sburstN, "tops_deramp_and_resample");
    if (rc) return rc;
    if (!cpm_az || !cpm_rg || !kernel_az || !kernel_rg || (!out && !d_out) ||
        !s->doppler_rate || !s->reference_time || !s->doppler_centroid)
        return set_error(TOPS_ERR_ARG, "tops_deramp_and_resample: null argument");
    tops_ctx* ctx = tops_ctx_create(kernel_az, kernel_rg, kTaps);
    if (!ctx) return tops_last_error();
    const int lpm = m->lines_per_burst, spm = m->samples_per_burst;
    const int lps = s->lines_per_burst, sps = s->samples_per_burst;
    const size_t mbytes = (size_t)lpm * spm * sizeof(tops_cf32);
    DBuf d_s, d_par, d_o;
    rc = d_s.ensure((size_t)lps * sps * sizeof(tops_ci16));
    if (!rc) rc = d_par.ensure(3 * sizeof(double) * sps);
    if (!rc && !d_out) rc = d_o.ensure(mbytes);
    for (int b = mburst0; b <= mburstN && !rc; b++) {
        const int sbi = b + sburst0 - mburst0;
        rc = staged_upload(d_s.p, s->slc + (size_t)sbi * lps * sps,
                           (size_t)lps * sps * sizeof(tops_ci16));
        if (rc) break;
        hipError_t e = hipSuccess;
        double* par = d_par.as<double>();
        if (e == hipSuccess)
            e = hipMemcpy(par, s->doppler_rate + (size_t)sbi * sps, sizeof(double) * sps,
                          hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(par + sps, s->reference_time + (size_t)sbi * sps, sizeof(double) * sps,
                          hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(par + 2 * sps, s->doppler_centroid + (size_t)sbi * sps,
                          sizeof(double) * sps, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            rc = hip_fail(e, "tops_deramp_and_resample upload", __FILE__, __LINE__);
            break;
        }
        tops_burst_geom g;
        int mbox[4] = {0, spm - 1, b * lpm, (b + 1) * lpm - 1};
        int sbox[4] = {0, sps - 1, sbi * lps, (sbi + 1) * lps - 1};
        memcpy(g.master_box, mbox, sizeof(mbox));
        memcpy(g.slave_box, sbox, sizeof(sbox));
        g.first_line_in_burst = sbi * lps;
        g.samples_per_burst = sps;
        g.az_time_interval = s->az_time_interval;
        g.az_shift = 0.0;
        memcpy(g.cpm_az, cpm_az + 6 * (b - mburst0), sizeof(g.cpm_az));
        memcpy(g.cpm_rg, cpm_rg + 6 * (b - mburst0), sizeof(g.cpm_rg));
        tops_cf32* dst = d_out ? d_out + (size_t)(b - mburst0) * lpm * spm : d_o.as<tops_cf32>();
        rc = tops_resample_burst(ctx, &g, d_s.as<tops_ci16>(), par, par + sps, par + 2 * sps, dst,
                                 TOPS_MODE_FAST, nullptr);
        if (!rc && out) rc = staged_download(out + (size_t)(b - mburst0) * lpm * spm, dst, mbytes);
    }
    if (!rc) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = hip_fail(e, "tops_deramp_and_resample", __FILE__, __LINE__);
    }
    tops_ctx_destroy(ctx);
    return rc;
}

// ESDAndCoh, Func.cpp:2496-2763, over the pass-1 result `reslave`
// ([(mburstN - mburst0 + 1) * lpb x spb], as tops_deramp_and_resample
// writes it):
//   1. overlap sizes of the master bursts (EstOverlapSize) and the spectral
//      separation of the master swath (GetSpectralSep);
//   2. forward/backward bands of the master SLC and of reslave, ESD shifts
//      with window esd_win_az x esd_win_rg and threshold (the reference
//      hard-codes 6 x 24 and 0.3, Func.cpp:2583-2585); |shift| > 100 -> 0;
//   3. burst mburst0 keeps its pass-1 result; every later burst b is
//      resampled again with AzimuthShift = shift[b - mburst0 - 1] (not a
//      cumulative shift, as the reference, Func.cpp:2702);
//   4. coherence of every burst against the master, window coh_win_az x
//      coh_win_rg starting (win-1)/2 before the sample.
// Outputs (host): reslave_esd [N lpb x spb], coh [N lpb x spb], shifts
// [mburstN - mburst0] (any may be NULL except shifts when N > 1).
int tops_esd_and_coh(const tops_swath* m, const tops_swath* s, const double* cpm_az,
                     const double* cpm_rg, int mburst0, int mburstN, int sburst0, int sburstN,
                     const float* kernel_az, const float* kernel_rg, const tops_cf32* reslave,
                     int esd_win_az, int esd_win_rg, float threshold, int coh_win_az,
                     int coh_win_rg, tops_cf32* reslave_esd, float* coh, double* shifts)
{
    int rc = check_swath(m, "tops_esd_and_coh");
    if (!rc) rc = check_swath(s, "tops_esd_and_coh");
    if (!rc) rc = check_range(m, s, mburst0, mburstN, sburst0, sburstN, "tops_esd_and_coh");
    if (rc) return rc;
    const int nb = mburstN - mburst0 + 1, nov = nb - 1;
    if (!cpm_az || !cpm_rg || !kernel_az || !kernel_rg || !reslave || (nov > 0 && !shifts) ||
        !m->doppler_rate)
        return set_error(TOPS_ERR_ARG, "tops_esd_and_coh: null argument");
    const int lpb = m->lines_per_burst, spb = m->samples_per_burst;
    const size_t bpix = (size_t)lpb * spb;
    std::vector<double> sh(nov > 0 ? nov : 1, 0.0);
    if (nov > 0) {
        std::vector<int> sizes(nov);
        int total = 0, maximum = 0;
        tops_overlap_sizes(m->burst_first_line_time, m->burst_last_line_time, mburst0, mburstN,
                           m->az_time_interval, sizes.data(), &total, &maximum);
        for (int i = 0; i < nov; i++)
            if (sizes[i] < 0 || sizes[i] > lpb)
                return set_error(TOPS_ERR_ARG, "tops_esd_and_coh: overlap %d has %d lines", i,
                                 sizes[i]);
        double df = tops_spectral_sep(m->doppler_rate, m->n_bursts, spb, lpb, m->az_time_interval);
        // DataTypes.cpp:456-513 band gather: each band is contiguous in its
        // raster, so it goes straight to its place in the device band array
        DBuf a, b2, c, d;
        size_t n = (size_t)total * spb;
        if ((rc = a.ensure(n * 4)) || (rc = b2.ensure(n * 4)) || (rc = c.ensure(n * 8)) ||
            (rc = d.ensure(n * 8)))
            return rc;
        size_t off = 0;
        for (int i = 0; i < nov && !rc; i++) {
            const size_t cnt = (size_t)sizes[i] * spb;
            const size_t fwd = (size_t)(i + 1) * lpb - sizes[i], bwd = (size_t)(i + 1) * lpb;
            const tops_ci16* mr = m->slc + (size_t)mburst0 * lpb * spb;
            if (!(rc = staged_upload(a.as<tops_ci16>() + off, mr + fwd * spb, cnt * 4)) &&
                !(rc = staged_upload(b2.as<tops_ci16>() + off, mr + bwd * spb, cnt * 4)) &&
                !(rc = staged_upload(c.as<tops_cf32>() + off, reslave + fwd * spb, cnt * 8)))
                rc = staged_upload(d.as<tops_cf32>() + off, reslave + bwd * spb, cnt * 8);
            off += cnt;
        }
        if (rc) return rc;
        rc = tops_esd_shifts(a.as<tops_ci16>(), b2.as<tops_ci16>(), c.as<tops_cf32>(),
                             d.as<tops_cf32>(), sizes.data(), nov, spb, esd_win_az, esd_win_rg,
                             threshold, df, m->az_time_interval, sh.data(), nullptr, nullptr);
        if (rc) return rc;
        for (int i = 0; i < nov; i++) {
            if (fabs(sh[i]) > 100) sh[i] = 0;   // Func.cpp:2612-2621
            shifts[i] = sh[i];
        }
    }
    // pass 2: re-resample with the shift, then coherence
    tops_ctx* ctx = tops_ctx_create(kernel_az, kernel_rg, kTaps);
    if (!ctx) return tops_last_error();
    const int lps = s->lines_per_burst, sps = s->samples_per_burst;
    DBuf d_s, d_par, d_res, d_m, d_coh;
    rc = d_s.ensure((size_t)lps * sps * sizeof(tops_ci16));
    if (!rc) rc = d_par.ensure(3 * sizeof(double) * sps);
    if (!rc) rc = d_res.ensure(bpix * sizeof(tops_cf32));
    if (!rc) rc = d_m.ensure(bpix * sizeof(tops_ci16));
    if (!rc) rc = d_coh.ensure(bpix * sizeof(float));
    for (int b = mburst0; b <= mburstN && !rc; b++) {
        const int k = b - mburst0, sbi = b + sburst0 - mburst0;
        hipError_t e = hipSuccess;
        if (k == 0) {
            rc = staged_upload(d_res.p, reslave, bpix * sizeof(tops_cf32));
        } else {
            rc = staged_upload(d_s.p, s->slc + (size_t)sbi * lps * sps,
                               (size_t)lps * sps * sizeof(tops_ci16));
            double* par = d_par.as<double>();
            if (e == hipSuccess)
                e = hipMemcpy(par, s->doppler_rate + (size_t)sbi * sps, sizeof(double) * sps,
                              hipMemcpyHostToDevice);
            if (e == hipSuccess)
                e = hipMemcpy(par + sps, s->reference_time + (size_t)sbi * sps,
                              sizeof(double) * sps, hipMemcpyHostToDevice);
            if (e == hipSuccess)
                e = hipMemcpy(par + 2 * sps, s->doppler_centroid + (size_t)sbi * sps,
                              sizeof(double) * sps, hipMemcpyHostToDevice);
            if (e == hipSuccess && !rc) {
                tops_burst_geom g;
                int mbox[4] = {0, spb - 1, b * lpb, (b + 1) * lpb - 1};
                int sbox[4] = {0, sps - 1, sbi * lps, (sbi + 1) * lps - 1};
                memcpy(g.master_box, mbox, sizeof(mbox));
                memcpy(g.slave_box, sbox, sizeof(sbox));
                // Func.cpp:2701-2705 passes the master's lines per burst and
                // azimuth interval here
                g.first_line_in_burst = sbi * lpb;
                g.samples_per_burst = sps;
                g.az_time_interval = m->az_time_interval;
                g.az_shift = sh[k - 1];
                memcpy(g.cpm_az, cpm_az + 6 * k, sizeof(g.cpm_az));
                memcpy(g.cpm_rg, cpm_rg + 6 * k, sizeof(g.cpm_rg));
                rc = tops_resample_burst(ctx, &g, d_s.as<tops_ci16>(), par, par + sps,
                                         par + 2 * sps, d_res.as<tops_cf32>(), TOPS_MODE_FAST,
                                         nullptr);
            }
        }
        if (e != hipSuccess) rc = hip_fail(e, "tops_esd_and_coh upload", __FILE__, __LINE__);
        if (rc) break;
        if (reslave_esd &&
            (rc = staged_download(reslave_esd + (size_t)k * bpix, d_res.p, bpix * sizeof(tops_cf32))))
            break;
        if (coh) {
            if ((rc = staged_upload(d_m.p, m->slc + (size_t)b * bpix, bpix * sizeof(tops_ci16))))
                break;
            rc = tops_coherence(d_m.as<tops_ci16>(), d_res.as<tops_cf32>(), spb, lpb,
                                (coh_win_rg - 1) / 2, (coh_win_az - 1) / 2, coh_win_rg, coh_win_az,
                                d_coh.as<float>(), TOPS_MODE_FAST, nullptr);
            if (rc) break;
            if ((rc = staged_download(coh + (size_t)k * bpix, d_coh.p, bpix * sizeof(float))))
                break;
        }
    }
    tops_ctx_destroy(ctx);
    return rc;
}

}  // extern "C"

// ReadInput, Func.cpp:7-130: "key=value" (or key<TAB>value) lines with the
// keys of ConfigSet; surrounding blanks are trimmed.  Returns TOPS_ERR_ARG
// when the file cannot be opened.
extern "C" int tops_read_config(const char* path, tops_config* cfg)
{
    if (!path || !cfg) return set_error(TOPS_ERR_ARG, "tops_read_config: null argument");
    FILE* f = fopen(path, "r");
    if (!f) return set_error(TOPS_ERR_ARG, "tops_read_config: cannot open %s", path);
    memset(cfg, 0, sizeof(*cfg));
    char line[2048];
    while (fgets(line, sizeof(line), f)) {
        char* k = line;
        while (*k == ' ' || *k == '\t' || *k == '=') k++;
        char* e = k;
        while (*e && *e != '=' && *e != '\t' && *e != '\n' && *e != '\r') e++;
        char* v = e;
        while (*v == '=' || *v == '\t' || *v == ' ') v++;
        char* ke = e;
        while (ke > k && ke[-1] == ' ') ke--;
        *ke = 0;
        char* ve = v + strlen(v);
        while (ve > v && (ve[-1] == '\n' || ve[-1] == '\r' || ve[-1] == ' ' || ve[-1] == '\t')) ve--;
        *ve = 0;
        struct { const char* key; char* dst; size_t n; } strs[] = {
            {"process_dir", cfg->process_dir, sizeof(cfg->process_dir)},
            {"masterpath", cfg->masterpath, sizeof(cfg->masterpath)},
            {"slavepath", cfg->slavepath, sizeof(cfg->slavepath)},
            {"polarisation", cfg->polarisation, sizeof(cfg->polarisation)},
            {"preciseOrbitMaster", cfg->precise_orbit_master, sizeof(cfg->precise_orbit_master)},
            {"preciseOrbitSlave", cfg->precise_orbit_slave, sizeof(cfg->precise_orbit_slave)},
            {"SpecificDemPath", cfg->specific_dem_path, sizeof(cfg->specific_dem_path)}};
        struct { const char* key; int* dst; } ints[] = {
            {"firstsubswath", &cfg->firstsubswath}, {"lastsubswath", &cfg->lastsubswath},
            {"multi_az", &cfg->multi_az}, {"multi_rg", &cfg->multi_rg},
            {"burst0", &cfg->burst0}, {"burstN", &cfg->burstN}};
        for (auto& sdef : strs)
            if (!strcmp(k, sdef.key)) {
                strncpy(sdef.dst, v, sdef.n - 1);
                sdef.dst[sdef.n - 1] = 0;
            }
        for (auto& idef : ints)
            if (!strcmp(k, idef.key)) *idef.dst = atoi(v);
    }
    fclose(f);
    return TOPS_OK;
}
