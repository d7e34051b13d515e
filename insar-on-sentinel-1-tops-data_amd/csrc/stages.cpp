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
