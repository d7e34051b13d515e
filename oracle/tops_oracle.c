/*
 * tops_oracle.c -- FP64 CPU restatement of the reference TOPS burst path.
 *
 * TEST INFRASTRUCTURE ONLY.  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.  Each function cites the reference file:line it follows.
 */
#include "tops_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

void orc_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
}

int orc_max_threads(void) { return omp_get_max_threads(); }

#define ORC_PI 3.1415926535897931      /* Constants.h:8, CUDART_PI         */
#define ORC_SOL 299792458.0            /* Constants.h:7, cuConstants.cuh:7 */
#define ORC_ELL_A 6378137.0            /* cuConstants.cuh:5                */
#define ORC_ELL_B 6356752.3142451794975639665996337 /* cuConstants.cuh:6 */
#define ORC_INTERVAL 2047              /* Constants.h:9                    */

/* ------------------------------------------------------------------------ */
/* LUT: Func.cpp:2105-2158; sinc/rect Constants.h:15, 27-33                 */
/* ------------------------------------------------------------------------ */
static double orc_sinc(double x) { return x == 0 ? 1 : sin(ORC_PI * x) / (ORC_PI * x); }
static double orc_rect(double x)
{
    double ans = 0.0;
    if (x < 0.5 && x > -0.5) ans = 1;
    else if (x == 0.5 || x == -0.5) ans = 0.5;
    return ans;
}

void orc_lut_init(int npoints, double prf, double abw, double rsr2x,
                  float* kaz, float* krg)
{
    const double chi_az = prf / abw;
    const double chi_rg = (rsr2x / 2.0) / abw;   /* as Func.cpp:2111 (ABW) */
    const double dx = 1.0 / ORC_INTERVAL;
    const int half = npoints / 2;
    const double v_az = 1. - 1. / chi_az;
    const double v_rg = 1. - 1. / chi_rg;
    for (int i = 0; i <= ORC_INTERVAL; i++) {
        double sum_az = 0.0, sum_rg = 0.0;
        for (int j = 0; j < npoints; j++) {
            double x = 1.0 - half + j - dx * i;
            double caz = orc_sinc(x) * orc_rect(x / (double)npoints) *
                         cos(v_az * ORC_PI * x) / (1.0 - (2.0 * v_az * x) * (2.0 * v_az * x));
            double crg = orc_sinc(x) * orc_rect(x / (double)npoints) *
                         cos(v_rg * ORC_PI * x) / (1.0 - (2.0 * v_rg * x) * (2.0 * v_rg * x));
            sum_az += caz;
            sum_rg += crg;
            kaz[i * npoints + j] = (float)caz;
            krg[i * npoints + j] = (float)crg;
        }
        for (int p = 0; p < npoints; p++) {
            kaz[i * npoints + p] = (float)((double)kaz[i * npoints + p] / sum_az);
            krg[i * npoints + p] = (float)((double)krg[i * npoints + p] / sum_rg);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Orbit fit: Func.cpp:940-1126 (matTxmat, choles, solvechol, Polyfit_Orbit) */
/* ------------------------------------------------------------------------ */
static int orc_choles(double* a, int n)          /* Func.cpp:958-981        */
{
    int ok = 1;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            double sum = a[i * n + j];
            for (int k = i - 1; k >= 0; --k) sum -= a[i * n + k] * a[j * n + k];
            if (i == j) {
                if (sum <= 0.) ok = 0;
                a[i * n + i] = sqrt(sum);
            } else {
                a[j * n + i] = sum / a[i * n + i];
            }
        }
    return ok;
}

static void orc_solvechol(const double* a, double* b, int n) /* 983-1010 */
{
    for (int i = 0; i < n; ++i) {
        double sum = b[i];
        for (int j = i - 1; j >= 0; --j) sum -= a[i * n + j] * b[j];
        b[i] = sum / a[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double sum = b[i];
        for (int j = i + 1; j < n; ++j) sum -= a[j * n + i] * b[j];
        b[i] = sum / a[i * n + i];
    }
}

int orc_polyfit_orbit(int n, const double* t, const double* x,
                      const double* y, const double* z,
                      double* cx, double* cy, double* cz)
{
    int degree = (n > 6) ? 5 : n - 2;            /* Func.cpp:1063           */
    int nu = degree + 1;
    double* a = (double*)malloc(sizeof(double) * n * nu);
    double nn[36] = {0}, rx[6] = {0}, ry[6] = {0}, rz[6] = {0};
    for (int i = 0; i < n; i++) {
        double tn = (t[i] - t[n / 2]) / 10.0;    /* Func.cpp:1060           */
        for (int j = 0; j < nu; j++) a[i * nu + j] = pow(tn, j);
    }
    for (int i = 0; i < nu; i++) {               /* matTxmat, 940-956       */
        for (int j = 0; j < nu; j++) {
            double s = 0.0;
            for (int k = 0; k < n; k++) s += a[k * nu + i] * a[k * nu + j];
            nn[i * nu + j] = s;
        }
        double sx = 0, sy = 0, sz = 0;
        for (int k = 0; k < n; k++) {
            sx += a[k * nu + i] * x[k];
            sy += a[k * nu + i] * y[k];
            sz += a[k * nu + i] * z[k];
        }
        rx[i] = sx; ry[i] = sy; rz[i] = sz;
    }
    orc_choles(nn, nu);
    orc_solvechol(nn, rx, nu);
    orc_solvechol(nn, ry, nu);
    orc_solvechol(nn, rz, nu);
    for (int i = 0; i < nu; i++) { cx[i] = rx[i]; cy[i] = ry[i]; cz[i] = rz[i]; }
    free(a);
    return nu;
}

/* Geocoreg_cuda.cu:123-205 */
void orc_orbit_pos(double tn, const double* cx, const double* cy,
                   const double* cz, int n, double out[3])
{
    out[0] = out[1] = out[2] = 0.0;
    for (int i = n - 1; i >= 0; --i) {
        out[0] = out[0] * tn + cx[i];
        out[1] = out[1] * tn + cy[i];
        out[2] = out[2] * tn + cz[i];
    }
}

void orc_orbit_vel(double tn, const double* cx, const double* cy,
                   const double* cz, int n, double out[3])
{
    out[0] = cx[1]; out[1] = cy[1]; out[2] = cz[1];
    for (int i = 2; i < n; ++i) {
        double p = (double)i * pow(tn, i - 1);
        out[0] += cx[i] * p;
        out[1] += cy[i] * p;
        out[2] += cz[i] * p;
    }
    out[0] /= 10.0; out[1] /= 10.0; out[2] /= 10.0;
}

void orc_orbit_acc(double tn, const double* cx, const double* cy,
                   const double* cz, int n, double out[3])
{
    out[0] = out[1] = out[2] = 0.0;
    for (int i = 2; i < n; ++i) {
        double p = (double)((i - 1) * i) * pow(tn, i - 2);
        out[0] += cx[i] * p;
        out[1] += cy[i] * p;
        out[2] += cz[i] * p;
    }
    out[0] /= 100.0; out[1] /= 100.0; out[2] /= 100.0;
}

/* Geocoreg_cuda.cu:87-99; e2 from Geocoreg_cuda.cu:666-668 */
void orc_llh2xyz(double phi, double lam, double h, double out[3])
{
    const double e2 = 1.0 - (ORC_ELL_B * ORC_ELL_B / ORC_ELL_A / ORC_ELL_A);
    double sp = sin(phi), cp = cos(phi), sl = sin(lam), cl = cos(lam);
    double nph = ORC_ELL_A / sqrt(1.0 - e2 * sp * sp) + h;
    out[0] = nph * cp * cl;
    out[1] = nph * cp * sl;
    out[2] = (nph - e2 * (nph - h)) * sp;
}

/* ------------------------------------------------------------------------ */
/* Zero-Doppler Newton per DEM post: Geocoreg_cuda.cu:211-323               */
/* ------------------------------------------------------------------------ */
static double orc_deg2rad(double x) { return x * ORC_PI / 180.0; }

static void orc_project_post(double lat_deg, double lon_deg, double h,
                             const orc_orbit_geom* g, double* line,
                             double* pixel)
{
    double p[3], s[3], v[3], a[3], d[3];
    orc_llh2xyz(orc_deg2rad(lat_deg), orc_deg2rad(lon_deg), h, p);
    double az = g->az_init;
    for (int it = 0; it <= 10; it++) {           /* 272-294                 */
        double tn = (az - g->mid_time) / 10.0;
        orc_orbit_pos(tn, g->cx, g->cy, g->cz, g->ncoef, s);
        orc_orbit_vel(tn, g->cx, g->cy, g->cz, g->ncoef, v);
        orc_orbit_acc(tn, g->cx, g->cy, g->cz, g->ncoef, a);
        d[0] = p[0] - s[0]; d[1] = p[1] - s[1]; d[2] = p[2] - s[2];
        double sol = -(v[0] * d[0] + v[1] * d[1] + v[2] * d[2]) /
                     ((a[0] * d[0] + a[1] * d[1] + a[2] * d[2]) -
                      v[0] * v[0] - v[1] * v[1] - v[2] * v[2]);
        az += sol;
        if (fabs(sol) < 1e-10) break;
    }
    double tn = (az - g->mid_time) / 10.0;
    orc_orbit_pos(tn, g->cx, g->cy, g->cz, g->ncoef, s);
    *line = g->line_offset + (az - g->first_line_time) / g->az_time_interval;
    double r = sqrt((p[0] - s[0]) * (p[0] - s[0]) + (p[1] - s[1]) * (p[1] - s[1]) +
                    (p[2] - s[2]) * (p[2] - s[2]));
    *pixel = (r - g->slr_time_first_pixel * ORC_SOL) / g->range_spacing;
}

void orc_ell2lp(const int32_t* dem, int lines, int pixels, double lat0,
                double lon0, double dlat, double dlon, int nodata,
                const orc_orbit_geom* g, double* line_out, double* pixel_out)
{
    long n = (long)lines * pixels;
#pragma omp parallel for schedule(static)
    for (long t = 0; t < n; t++) {
        int row = (int)(t / pixels), col = (int)(t % pixels);
        float h = (float)dem[t];
        if (h != (float)nodata) {
            orc_project_post(lat0 - dlat * row, lon0 + dlon * col, (double)h, g,
                             &line_out[t], &pixel_out[t]);
        } else {
            line_out[t] = -99999999;
            pixel_out[t] = -99999999;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Geometric coregistration fit: Geocoreg_cuda.cu:548-605 (WeighthSet,
 * AMatrixSet, normalizeWarp cuConstants.cuh:52-55), 812-855 (normal
 * equations + Cholesky).                                                    */
/* ------------------------------------------------------------------------ */
static double orc_norm_warp(double v, double lo, double hi)
{
    return (v - (0.5 * (lo + hi))) / (0.25 * (hi - lo));
}

int orc_geocoreg(const int32_t* dem, int lines, int pixels, double lat0,
                 double lon0, double dlat, double dlon, int nodata,
                 const orc_orbit_geom* m, const orc_orbit_geom* s,
                 int xmin, int xmax, int ymin, int ymax,
                 double cpm_az[6], double cpm_rg[6], int64_t* n_used)
{
    long n = (long)lines * pixels;
    double* ml = (double*)malloc(sizeof(double) * n);
    double* mp = (double*)malloc(sizeof(double) * n);
    double* sl = (double*)malloc(sizeof(double) * n);
    double* sp = (double*)malloc(sizeof(double) * n);
    orc_ell2lp(dem, lines, pixels, lat0, lon0, dlat, dlon, nodata, m, ml, mp);
    orc_ell2lp(dem, lines, pixels, lat0, lon0, dlat, dlon, nodata, s, sl, sp);
    double nn[36] = {0}, ra[6] = {0}, rr[6] = {0};
    int64_t used = 0;
    for (long t = 0; t < n; t++) {
        if (dem[t] == nodata) continue;
        if (!(ml[t] >= ymin && ml[t] <= ymax && mp[t] >= xmin && mp[t] <= xmax)) continue;
        double L = orc_norm_warp(ml[t], ymin, ymax);
        double P = orc_norm_warp(mp[t], xmin, xmax);
        double a[6] = {1.0, L, P, L * L, L * P, P * P};
        for (int i = 0; i < 6; i++) {
            for (int j = 0; j < 6; j++) nn[i * 6 + j] += a[i] * a[j];
            ra[i] += a[i] * sl[t];
            rr[i] += a[i] * sp[t];
        }
        used++;
    }
    free(ml); free(mp); free(sl); free(sp);
    if (n_used) *n_used = used;
    int ok = orc_choles(nn, 6);
    if (!ok || used < 6) {
        for (int i = 0; i < 6; i++) cpm_az[i] = cpm_rg[i] = 0.0;
        return 1;
    }
    orc_solvechol(nn, ra, 6);
    orc_solvechol(nn, rr, 6);
    for (int i = 0; i < 6; i++) { cpm_az[i] = ra[i]; cpm_rg[i] = rr[i]; }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Deramp parameters: Func.cpp:1955-1970 (f_dc), 2020-2031 (k_a),
 * 2037-2054 (t_ref), 2066-2090 (k_t).                                      */
/* ------------------------------------------------------------------------ */
void orc_deramp_params(int n_bursts, int spb, double slr_time_first_pixel,
                       double range_spacing, int lines_per_burst,
                       double az_time_interval, int first_valid_pixel,
                       const orc_poly* dc, const orc_poly* fm,
                       double steering_rate_deg, double wavelength,
                       double orbit_velocity,
                       double* fdc, double* ka, double* tref, double* kt)
{
    for (int b = 0; b < n_bursts; b++)
        for (int x = 0; x < spb; x++) {
            double slrt = (slr_time_first_pixel + x * range_spacing / ORC_SOL) * 2;
            double dt = slrt - dc[b].t0;
            fdc[b * spb + x] = dc[b].c0 + dc[b].c1 * dt + dc[b].c2 * (dt * dt);
            double dt2 = slrt - fm[b].t0;
            ka[b * spb + x] = fm[b].c0 + fm[b].c1 * dt2 + fm[b].c2 * dt2 * dt2;
        }
    double temp1 = lines_per_burst * az_time_interval / 2.0;
    for (int b = 0; b < n_bursts; b++) {
        double temp2 = temp1 + fdc[b * spb + first_valid_pixel] / ka[b * spb + first_valid_pixel];
        for (int x = 0; x < spb; x++)
            tref[b * spb + x] = temp2 - fdc[b * spb + x] / ka[b * spb + x];
    }
    double krot = 2 * orbit_velocity * orc_deg2rad(steering_rate_deg) / wavelength;
    for (int b = 0; b < n_bursts; b++)
        for (int x = 0; x < spb; x++) {
            double r = ka[b * spb + x];
            kt[b * spb + x] = r * krot / (r - krot);
        }
}

/* ------------------------------------------------------------------------ */
/* Deramp + resample + reramp: DerampAndResample.cu:108-183, 348-520        */
/* ------------------------------------------------------------------------ */
static double orc_polyval(double x, double y, const double* c) /* cuConstants.cuh:57-66 */
{
    return c[0] + c[1] * x + c[2] * y + c[3] * x * x + c[4] * x * y + c[5] * y * y;
}

/* DerampDemod_Shared phase, DerampAndResample.cu:151-158 */
static double orc_phase(int y_abs, int x_abs, int first_line, double dt,
                        const double* kt, const double* tref, const double* fdc)
{
    double ta = (y_abs - first_line) * dt;
    double d = ta - tref[x_abs];
    return -ORC_PI * kt[x_abs] * (d * d) - 2.0 * ORC_PI * fdc[x_abs] * ta;
}

static void orc_slave_position(const orc_resample_geom* g, int r, int c,
                               int* iy, int* fy, int* ix, int* fx)
{
    const int* mb = g->master_box;
    double pn = orc_norm_warp((double)(mb[0] + c), mb[0], mb[1]);
    double ln = orc_norm_warp((double)(mb[2] + r), mb[2], mb[3]);
    double ys = orc_polyval(ln, pn, g->cpm_az) + g->az_shift;
    double xs = orc_polyval(ln, pn, g->cpm_rg);
    /* GetIndexes, DerampAndResample.cu:108-112 */
    *iy = (int)trunc(ys);
    *fy = (int)((ys - *iy) * ORC_INTERVAL + 0.5);
    *ix = (int)trunc(xs);
    *fx = (int)((xs - *ix) * ORC_INTERVAL + 0.5);
}

void orc_resample(const int16_t* slave, const orc_resample_geom* g,
                  const float* kaz, const float* krg, int npoints,
                  const double* kt, const double* tref, const double* fdc,
                  int row0, int row1, int ref_float_phase, double* out)
{
    const int* mb = g->master_box;
    const int* sb = g->slave_box;
    const int mp = mb[1] - mb[0] + 1;
    const int sp = sb[1] - sb[0] + 1;
    const int h = npoints / 2 - 1;               /* c_Npoints2m1           */
    if (row1 <= row0) return;

    /* slave rows needed by this block of output rows */
    int lo = sb[3] + 1, hi = sb[2] - 1;
    for (int r = row0; r < row1; r++)
        for (int c = 0; c < mp; c++) {
            int iy, fy, ix, fx;
            orc_slave_position(g, r, c, &iy, &fy, &ix, &fx);
            if (iy < lo) lo = iy;
            if (iy > hi) hi = iy;
        }
    lo -= h; hi += h + 1;                        /* 12-tap support         */
    if (lo < sb[2]) lo = sb[2];
    if (hi > sb[3]) hi = sb[3];
    int nrows = hi - lo + 1;
    double* dre = NULL;
    double* ph = NULL;
    if (nrows > 0) {
        dre = (double*)malloc(sizeof(double) * 2 * (size_t)nrows * sp);
        ph = (double*)malloc(sizeof(double) * (size_t)nrows * sp);
#pragma omp parallel for schedule(static)
        for (int yy = 0; yy < nrows; yy++)
            for (int xx = 0; xx < sp; xx++) {
                int y_abs = lo + yy, x_abs = sb[0] + xx;
                double phi = orc_phase(y_abs, x_abs, g->first_line_in_burst,
                                       g->az_time_interval, kt, tref, fdc);
                const int16_t* s = &slave[2 * ((size_t)(y_abs - sb[2]) * sp + xx)];
                size_t k = (size_t)yy * sp + xx;
                dre[2 * k] = s[0] * cos(phi) - s[1] * sin(phi);
                dre[2 * k + 1] = s[0] * sin(phi) + s[1] * cos(phi);
                ph[k] = ref_float_phase ? (double)(float)phi : phi;
            }
    }

#pragma omp parallel for schedule(dynamic, 1)
    for (int r = row0; r < row1; r++)
        for (int c = 0; c < mp; c++) {
            double* o = &out[2 * ((size_t)(r - row0) * mp + c)];
            int iy, fy, ix, fx;
            orc_slave_position(g, r, c, &iy, &fy, &ix, &fx);
            if (iy > sb[3] || iy < sb[2] || ix > sb[1] || ix < sb[0]) {
                o[0] = 0.0; o[1] = 0.0;
                continue;
            }
            if (iy > sb[3] - h - 2 || iy < sb[2] + h || ix > sb[1] - h - 2 || ix < sb[0] + h) {
                /* edge band: the truncated-index sample with its own phase
                 * removed again (DerampAndResample.cu:390-403) */
                if (iy < lo || iy > hi) { o[0] = o[1] = 0.0; continue; } /* unreachable */
                size_t k = (size_t)(iy - lo) * sp + (ix - sb[0]);
                double p = ph[k];
                double cr = cos(p), ci = -sin(p);
                o[0] = dre[2 * k] * cr - dre[2 * k + 1] * ci;
                o[1] = dre[2 * k] * ci + dre[2 * k + 1] * cr;
                continue;
            }
            const float* kl = &kaz[fy * npoints];
            const float* kp = &krg[fx * npoints];
            double sr = 0.0, si = 0.0, phs = 0.0;
            for (int j = 0; j < npoints; j++) {
                int yy = iy - h + j - lo;
                double tr = 0.0, ti = 0.0, tp = 0.0;
                for (int i = 0; i < npoints; i++) {
                    size_t k = (size_t)yy * sp + (ix - h + i - sb[0]);
                    tr += dre[2 * k] * (double)kp[i];
                    ti += dre[2 * k + 1] * (double)kp[i];
                    tp += ph[k] * (double)kp[i];
                }
                sr += tr * (double)kl[j];
                si += ti * (double)kl[j];
                phs += tp * (double)kl[j];
            }
            double cr = cos(phs), ci = -sin(phs);
            o[0] = sr * cr - si * ci;
            o[1] = sr * ci + si * cr;
        }
    free(dre);
    free(ph);
}

/* ------------------------------------------------------------------------ */
/* Boxcar coherence: EstCoh.cu:101-139 (CohMatSet), 413-456 (Cohbasic)      */
/* ------------------------------------------------------------------------ */
void orc_coherence(const int16_t* master, const double* slave, int w, int h,
                   int dx, int dy, int win_rg, int win_az, double* coh)
{
    size_t n = (size_t)w * h;
    /* per-sample products (CohMatSet) then horizontal window sums */
    double* hs = (double*)malloc(sizeof(double) * 4 * n);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++) {
            double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
            for (int k = 0; k < win_rg; k++) {
                int cc = c - dx + k;
                if (cc < 0 || cc >= w) continue;
                size_t q = (size_t)r * w + cc;
                double mr = master[2 * q], mi = master[2 * q + 1];
                double sr = slave[2 * q], si = slave[2 * q + 1];
                a0 += mr * sr + mi * si;           /* M * conj(S)            */
                a1 += mi * sr - mr * si;
                a2 += sr * sr + si * si;           /* |S|^2                  */
                a3 += mr * mr + mi * mi;           /* |M|^2                  */
            }
            size_t q = (size_t)r * w + c;
            hs[4 * q] = a0; hs[4 * q + 1] = a1; hs[4 * q + 2] = a2; hs[4 * q + 3] = a3;
        }
#pragma omp parallel for schedule(static)
    for (int r = 0; r < h; r++)
        for (int c = 0; c < w; c++) {
            size_t q = (size_t)r * w + c;
            double sr = slave[2 * q], si = slave[2 * q + 1];
            if (sqrt(sr * sr + si * si) < 0.001) { coh[q] = 0.0; continue; }
            double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
            for (int j = 0; j < win_az; j++) {
                int rr = r - dy + j;
                if (rr < 0 || rr >= h) continue;
                size_t p = (size_t)rr * w + c;
                a0 += hs[4 * p]; a1 += hs[4 * p + 1]; a2 += hs[4 * p + 2]; a3 += hs[4 * p + 3];
            }
            double pp = a2 * a3;
            coh[q] = pp > 0.0 ? sqrt(a0 * a0 + a1 * a1) / sqrt(pp) : 0.0;
        }
    free(hs);
}

/* ------------------------------------------------------------------------ */
/* ESD: EstESD.cu:62-91 (DoubleDifference4OneD), 94-128 (CohMatrixSet),
 * 130-161 (MaskOut), 243-259 (Divide2), 297-546 (band coherence, no mask),
 * 1073-1217 (per-overlap flow and shift).                                   */
/* ------------------------------------------------------------------------ */
static void orc_band_coherence(const int16_t* m, const double* s, int n, int w,
                               int win_az, int win_rg, double* coh)
{
    int dh = (win_az - 1) / 2, dw = (win_rg - 1) / 2;
    double* hs = (double*)malloc(sizeof(double) * 4 * (size_t)n * w);
    for (int r = 0; r < n; r++)
        for (int c = 0; c < w; c++) {
            double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
            for (int k = 0; k < win_rg; k++) {
                int cc = c - dw + k;
                if (cc < 0 || cc >= w) continue;
                size_t q = (size_t)r * w + cc;
                double mr = m[2 * q], mi = m[2 * q + 1];
                double sr = s[2 * q], si = s[2 * q + 1];
                a0 += mr * sr + mi * si;
                a1 += mi * sr - mr * si;
                a2 += sr * sr + si * si;
                a3 += mr * mr + mi * mi;
            }
            size_t q = (size_t)r * w + c;
            hs[4 * q] = a0; hs[4 * q + 1] = a1; hs[4 * q + 2] = a2; hs[4 * q + 3] = a3;
        }
    for (int r = 0; r < n; r++)
        for (int c = 0; c < w; c++) {
            double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
            for (int j = 0; j < win_az; j++) {
                int rr = r - dh + j;
                if (rr < 0 || rr >= n) continue;
                size_t p = (size_t)rr * w + c;
                a0 += hs[4 * p]; a1 += hs[4 * p + 1]; a2 += hs[4 * p + 2]; a3 += hs[4 * p + 3];
            }
            double pp = a2 * a3;
            coh[(size_t)r * w + c] = pp > 0.0 ? sqrt(a0 * a0 + a1 * a1) / sqrt(pp) : 0.0;
        }
    free(hs);
}

void orc_esd(const int16_t* m_for, const int16_t* m_back,
             const double* s_for, const double* s_back,
             const int* overlap_size, int n_overlap, int pixels,
             int win_az, int win_rg, double threshold,
             double spectral_sep, double az_time_interval,
             double* shifts, double* sums)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n_overlap; i++) {
        size_t off = 0;
        for (int k = 0; k < i; k++) off += (size_t)overlap_size[k] * pixels;
        int n = overlap_size[i];
        size_t cnt = (size_t)n * pixels;
        double* cf = (double*)malloc(sizeof(double) * (cnt ? cnt : 1));
        double* cb = (double*)malloc(sizeof(double) * (cnt ? cnt : 1));
        orc_band_coherence(m_for + 2 * off, s_for + 2 * off, n, pixels, win_az, win_rg, cf);
        orc_band_coherence(m_back + 2 * off, s_back + 2 * off, n, pixels, win_az, win_rg, cb);
        double sx = 0.0, sy = 0.0;
        for (size_t q = 0; q < cnt; q++) {
            double coh = (cf[q] + cb[q]) / 2.0;
            if (coh < threshold || coh >= 1) continue;
            size_t p = off + q;
            double m1r = m_for[2 * p], m1i = m_for[2 * p + 1];
            double m2r = m_back[2 * p], m2i = m_back[2 * p + 1];
            double s1r = s_for[2 * p], s1i = s_for[2 * p + 1];
            double s2r = s_back[2 * p], s2i = s_back[2 * p + 1];
            double ar = m1r * s1r + m1i * s1i, ai = m1i * s1r - m1r * s1i; /* Mf Sf* */
            double br = m2r * s2r + m2i * s2i, bi = m2i * s2r - m2r * s2i; /* Mb Sb* */
            double dr = ar * br + ai * bi, di = ai * br - ar * bi;         /* a conj(b) */
            double th = atan2(di, dr);
            sx += cos(th);
            sy += sin(th);
        }
        free(cf);
        free(cb);
        if (sums) { sums[2 * i] = sx; sums[2 * i + 1] = sy; }
        shifts[i] = atan2(sy, sx) / (2 * ORC_PI * spectral_sep * az_time_interval);
    }
}

/* Func.cpp:2935-2960 */
void orc_overlap_sizes(const double* first_line_time,
                       const double* last_line_time, int burst0, int burstN,
                       double az_time_interval, int* sizes, int* total,
                       int* maximum)
{
    *total = 0;
    *maximum = -9999;
    for (int i = burst0; i < burstN; i++) {
        int n = (int)floor((last_line_time[i] - first_line_time[i + 1]) / az_time_interval);
        sizes[i - burst0] = n;
        *total += n;
        if (n > *maximum) *maximum = n;
    }
}

/* Func.cpp:2962-2976 */
double orc_spectral_sep(const double* kt, int n_bursts, int spb,
                        int lines_per_burst, double az_time_interval)
{
    double tcycle = lines_per_burst * az_time_interval;
    double sum = 0.0;
    for (int i = 0; i < n_bursts; i++)
        for (int p = 0; p < spb; p++) sum += kt[i * spb + p] * tcycle;
    return sum / spb / n_bursts;
}

/* ------------------------------------------------------------------------ */
/* Deburst: getLineIdxInMetaData (Func.cpp:2766-2819) and                    */
/* Stitch_OneSubswath (Func.cpp:2821-2933)                                   */
/* ------------------------------------------------------------------------ */

/* Func.cpp:2766-2819: subswath line holding line time t; -1 when none. */
int orc_deburst_line(const double* first_line_time, const double* last_line_time,
                     int n_bursts, int lines_per_burst, int num_lines,
                     double az_time_interval, double t, int burst0, int burstN)
{
    int sy0 = -1, sy1 = -1, b0 = -1, b1 = -1, k = 0;
    for (int i = 0; i < n_bursts; i++) {
        if (!(t >= first_line_time[i] && t < last_line_time[i])) continue;
        int sy = i * lines_per_burst +
                 (int)(((t - first_line_time[i]) / az_time_interval) + 0.5);
        if (k == 0) {
            sy0 = sy;
            b0 = i;
        } else {
            sy1 = sy;
            b1 = i;
            break;
        }
        k++;
    }
    if (sy0 != -1 && sy1 != -1) {
        if (b1 == burst0) return sy1;
        if (b1 == burstN + 1) return sy0;
        double mid = (last_line_time[b0] + first_line_time[b1]) / 2.0;
        if (t > mid) return sy1;
    }
    double tl = last_line_time[n_bursts - 1];
    if (t < tl + az_time_interval && t >= tl) sy0 = num_lines - 1;
    return sy0;
}

/* Func.cpp:2825-2843: output rows yMin .. yMax. */
void orc_deburst_extent(double first_valid_time_b0, double last_valid_time_bN,
                        double first_line_time, double az_time_interval,
                        int* y_min, int* height)
{
    int ymin = (int)((first_valid_time_b0 - first_line_time) / az_time_interval);
    int ymax = (int)((last_valid_time_bN - first_line_time) / az_time_interval);
    *y_min = ymin;
    *height = ymax - ymin + 1;
}

/* Func.cpp:2864-2915: out row yy = gapped row (line - burst0 * lpb), or zeros
 * where that row is outside the gapped raster (the failed read leaves the
 * memset line).  Returns -1 (the reference exits) if a line time falls in
 * no burst, else 0.  rows (may be NULL) receives the gapped row of each
 * output row. */
int orc_deburst(const double* first_line_time, const double* last_line_time,
                const double* first_valid_time, const double* last_valid_time,
                int n_bursts, int lines_per_burst, int samples_per_burst,
                int num_lines, double az_time_interval, double swath_first_line_time,
                int burst0, int burstN, const void* gapped, int elem_bytes,
                void* out, int* rows)
{
    int ymin, height;
    orc_deburst_extent(first_valid_time[burst0], last_valid_time[burstN],
                       swath_first_line_time, az_time_interval, &ymin, &height);
    const double tf = first_valid_time[burst0];
    const int gapped_rows = (burstN - burst0 + 1) * lines_per_burst;
    const size_t row_bytes = (size_t)samples_per_burst * elem_bytes;
    for (int yy = 0; yy < height; yy++) {
        double t = tf + yy * az_time_interval;
        int line = orc_deburst_line(first_line_time, last_line_time, n_bursts,
                                    lines_per_burst, num_lines, az_time_interval, t,
                                    burst0, burstN);
        if (line == -1) return -1;
        line -= burst0 * lines_per_burst;
        if (rows) rows[yy] = line;
        if (!out) continue;
        char* dst = (char*)out + (size_t)yy * row_bytes;
        if (line >= 0 && line < gapped_rows)
            memcpy(dst, (const char*)gapped + (size_t)line * row_bytes, row_bytes);
        else
            memset(dst, 0, row_bytes);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Burst matching: Locating / getLatitude (Func.cpp:1133-1246),             */
/* Geodetic2Geographic (1248-1260), orbit polynomial (1262-1345, 1409-1453),*/
/* xyz2aztime (1349-1377), solve33 (1462-1490), lp2xyz (1492-1584),         */
/* ComputerBurstOverlap (1610-1746)                                         */
/* ------------------------------------------------------------------------ */
static double orc_dot3(const double* a, const double* b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* position / velocity / acceleration at absolute time t, polynomial in
 * (t - mid) / 10 with mid = t[num_points / 2] */
static void orc_match_orbit_eval(const orc_match_orbit* o, double t, double* p, double* v,
                                 double* acc)
{
    const double tn = (t - o->t[o->num_points / 2]) / 10.0;
    const double* c[3] = {o->cx, o->cy, o->cz};
    for (int k = 0; k < 3; k++) {
        double s = 0.0;
        for (int i = o->ncoef - 1; i >= 0; i--) s = s * tn + c[k][i];
        if (p) p[k] = s;
        if (v) {
            double d = c[k][1];
            for (int i = 2; i < o->ncoef; i++) d += c[k][i] * ((double)i * pow(tn, i - 1));
            v[k] = d / 10.0;
        }
        if (acc) {
            double d = 0.0;
            for (int i = 2; i < o->ncoef; i++) d += c[k][i] * ((double)((i - 1) * i) * pow(tn, i - 2));
            acc[k] = d / 100.0;
        }
    }
}

double orc_xyz2aztime(const double* xyz, const orc_match_orbit* o, double t_init)
{
    double t = t_init;
    for (int it = 0; it <= 10; it++) {
        double p[3], v[3], a[3], d[3];
        orc_match_orbit_eval(o, t, p, v, a);
        for (int k = 0; k < 3; k++) d[k] = xyz[k] - p[k];
        const double sol = -orc_dot3(v, d) / (orc_dot3(a, d) - orc_dot3(v, v));
        t += sol;
        if (fabs(sol) < 1e-10) break;
    }
    return t;
}

/* Locating + bilinear interpolation of the grid's latitude / longitude
 * (the interpolation weights kept fractional) */
static void orc_grid_ll(const orc_match_swath* s, double az, double rg, double* lat, double* lon)
{
    const int P = s->geo_ppl, G = s->geo_lines;
    int j0 = -1, j1 = -1;
    if (rg < s->geo_rg[0]) {
        j0 = 0;
        j1 = 1;
    } else if (rg > s->geo_rg[P - 1]) {
        j0 = P - 2;
        j1 = P - 1;
    } else {
        for (int j = 0; j < P - 1; j++)
            if (s->geo_rg[j] <= rg && s->geo_rg[j + 1] > rg) {
                j0 = j;
                j1 = j + 1;
                break;
            }
    }
    if (j0 < 0) { j0 = P - 2; j1 = P - 1; }
    const double mx = (rg - s->geo_rg[j0]) / (s->geo_rg[j1] - s->geo_rg[j0]);
    int i0 = -1;
    double my = 0.0;
    for (int i = 0; i < G - 1; i++) {
        const double a0 = (1 - mx) * s->geo_az[i * P + j0] + mx * s->geo_az[i * P + j1];
        const double a1 = (1 - mx) * s->geo_az[(i + 1) * P + j0] + mx * s->geo_az[(i + 1) * P + j1];
        if ((i == 0 && az < a0) || (i == G - 2 && az >= a1) || (a0 <= az && a1 > az)) {
            i0 = i;
            my = (az - a0) / (a1 - a0);
            break;
        }
    }
    if (i0 < 0) i0 = 0;
    const int i1 = i0 + 1;
    const double* g[2] = {s->geo_lat, s->geo_lon};
    double out[2];
    for (int k = 0; k < 2; k++) {
        const double v00 = g[k][i0 * P + j0], v01 = g[k][i0 * P + j1];
        const double v10 = g[k][i1 * P + j0], v11 = g[k][i1 * P + j1];
        out[k] = (1 - my) * ((1 - mx) * v00 + mx * v01) + my * ((1 - mx) * v10 + mx * v11);
    }
    *lat = out[0];
    *lon = out[1];
}

static void orc_llh2xyz_wgs84(double lat, double lon, double h, double* xyz)
{
    const double a = ORC_ELL_A, b = ORC_ELL_B, e2 = 1.0 - (b / a) * (b / a);
    const double N = a / sqrt(1.0 - e2 * (sin(lat) * sin(lat)));
    const double Nph = N + h;
    xyz[0] = Nph * cos(lat) * cos(lon);
    xyz[1] = Nph * cos(lat) * sin(lon);
    xyz[2] = (Nph - e2 * N) * sin(lat);
}

static void orc_solve33(double* x, const double* rhs, const double* A)
{
    const double L10 = A[3] / A[0], L20 = A[6] / A[0];
    const double U11 = A[4] - L10 * A[1];
    const double L21 = (A[7] - A[1] * L20) / U11;
    const double U12 = A[5] - L10 * A[2];
    const double U22 = A[8] - L20 * A[2] - L21 * U12;
    const double b0 = rhs[0], b1 = rhs[1] - b0 * L10, b2 = rhs[2] - b0 * L20 - b1 * L21;
    x[2] = b2 / U22;
    x[1] = (b1 - U12 * x[2]) / U11;
    x[0] = (b0 - A[1] * x[1] - A[2] * x[2]) / A[0];
}

int orc_lp2xyz(double line, double pixel, const orc_match_swath* s, const orc_match_orbit* o,
               int burst, double* xyz)
{
    const double a = ORC_ELL_A, b = ORC_ELL_B;
    const double az = s->burst_first_line_time[burst] + (line - 1.0) / s->prf;
    const double rt = s->slr_first + (pixel - 1.0) / (s->rsr * 2);
    double ps[3], vs[3];
    orc_match_orbit_eval(o, az, ps, vs, NULL);
    int it;
    for (it = 0; it <= 10; it++) {
        double d[3], eq[3], J[9], sol[3];
        for (int k = 0; k < 3; k++) d[k] = xyz[k] - ps[k];
        eq[0] = -orc_dot3(vs, d);
        eq[1] = -(orc_dot3(d, d) - (ORC_SOL * rt) * (ORC_SOL * rt));
        eq[2] = -((xyz[0] * xyz[0] + xyz[1] * xyz[1]) / (a * a) + (xyz[2] / b) * (xyz[2] / b) - 1.0);
        J[0] = vs[0]; J[1] = vs[1]; J[2] = vs[2];
        J[3] = 2 * d[0]; J[4] = 2 * d[1]; J[5] = 2 * d[2];
        J[6] = 2 * xyz[0] / (a * a); J[7] = 2 * xyz[1] / (a * a); J[8] = 2 * xyz[2] / (b * b);
        orc_solve33(sol, eq, J);
        for (int k = 0; k < 3; k++) xyz[k] += sol[k];
        if (fabs(sol[0]) < 1e-6 && fabs(sol[1]) < 1e-6 && fabs(sol[2]) < 1e-6) break;
    }
    return it;
}

/* zero-Doppler time on the other orbit of the centre pixel of burst `b` */
static double orc_centre_zdt(const orc_match_swath* s, const orc_match_orbit* o,
                             const orc_match_orbit* other, int b, double az0)
{
    double lat, lon, xyz[3];
    orc_grid_ll(s, az0, (s->slr_first + s->slr_last) / 2.0, &lat, &lon);
    orc_llh2xyz_wgs84(lat * ORC_PI / 180.0, lon * ORC_PI / 180.0, 0.0, xyz);
    orc_lp2xyz(s->lpb / 2, s->spb / 2, s, o, b, xyz);
    return orc_xyz2aztime(xyz, other, az0);
}

int orc_burst_overlap(const orc_match_swath* m, const orc_match_swath* s,
                      const orc_match_orbit* mo, const orc_match_orbit* so, int in_b0,
                      int in_bn, int* out)
{
    const int B0 = in_b0 - 1, BN = in_bn - 1;
    if (B0 < 0) return -1;
    const double az_m = m->burst_first_line_time[B0] + m->lpb / 2.0 * m->dt;
    const double zdt = orc_centre_zdt(m, mo, so, B0, az_m);
    int m0, mn, s0, sn;
    if (zdt > s->last_line_time) return -2;
    if (zdt < s->last_line_time && zdt > s->first_line_time) {
        m0 = B0;
        int cc = 0;
        while (cc < s->n_bursts && zdt > s->burst_first_line_time[cc]) cc++;
        s0 = cc - 1;
        if (BN - m0 <= s->n_bursts - 1 - s0) {
            mn = BN;
            sn = s0 + (mn - m0);
        } else {
            sn = s->n_bursts - 1;
            mn = m0 + (sn - s0);
        }
    } else {
        const double az_s = s->burst_first_line_time[0] + s->lpb / 2 * s->dt;
        const double zm = orc_centre_zdt(s, so, mo, 0, az_s);
        if (zm < m->first_line_time) return -2;
        if (!(zm < m->last_line_time && zm > m->first_line_time)) return -2;
        s0 = 0;
        int cc = 0;
        while (cc < m->n_bursts && zm > m->burst_first_line_time[cc]) cc++;
        m0 = cc - 1;
        if (BN - m0 <= s->n_bursts - 1) {
            mn = BN;
            sn = s0 + (mn - m0);
        } else {
            sn = s->n_bursts - 1;
            mn = m0 + (sn - s0);
        }
    }
    out[0] = m0;
    out[1] = mn;
    out[2] = s0;
    out[3] = sn;
    return 0;
}
