// abi.cpp -- the C-ABI of libtops_mi355x.so: status handling, the resample
// context, the device-resident entry points and the seven reference entry
// points (host buffers in, host buffers out).
#include <math.h>
#include <mutex>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "tops_internal.hpp"

namespace tops {

static thread_local int g_err = TOPS_OK;
static thread_local char g_msg[512] = "";

int set_error(int code, const char* fmt, ...)
{
    g_err = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    fprintf(stderr, "[tops] %s\n", g_msg);
    return code;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line)
{
    int code = (e == hipErrorOutOfMemory) ? TOPS_ERR_ALLOC
               : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? TOPS_ERR_NODEVICE
                                                                        : TOPS_ERR_HIP;
    return set_error(code, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

int Scratch::ensure(size_t need)
{
    if (need <= bytes) return TOPS_OK;
    release();
    TOPS_HIP(hipMalloc(&ptr, need));
    bytes = need;
    return TOPS_OK;
}

void Scratch::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

// 6x6 (or smaller) Cholesky solve of the symmetric positive-definite normal
// equations; returns false when a pivot is not positive.
static bool chol_solve(double* a, int n, double* b0, double* b1)
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
    double* rhs[2] = {b0, b1};
    for (int r = 0; r < 2; r++) {
        double* b = rhs[r];
        for (int i = 0; i < n; i++) {   // L y = b
            double v = b[i];
            for (int k = 0; k < i; k++) v -= a[i * n + k] * b[k];
            b[i] = v / a[i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {   // L^T x = y
            double v = b[i];
            for (int k = i + 1; k < n; k++) v -= a[k * n + i] * b[k];
            b[i] = v / a[i * n + i];
        }
    }
    return true;
}

static int make_orbit(const tops_orbit_geom* g, OrbitDev& o)
{
    if (!g || !g->coef_x || !g->coef_y || !g->coef_z || g->ncoef < 2 || g->ncoef > 6)
        return set_error(TOPS_ERR_ARG, "orbit: need 2..6 coefficients (got %d)", g ? g->ncoef : -1);
    memset(&o, 0, sizeof(o));
    for (int i = 0; i < g->ncoef; i++) {
        o.cx[i] = g->coef_x[i];
        o.cy[i] = g->coef_y[i];
        o.cz[i] = g->coef_z[i];
    }
    o.ncoef = g->ncoef;
    o.mid_time = g->mid_time;
    o.dt = g->az_time_interval;
    o.first_line_time = g->first_line_time;
    o.slr_first = g->slr_time_first_pixel;
    o.range_spacing = g->range_spacing;
    o.az_init = g->az_init;
    o.line_offset = g->line_offset;
    return TOPS_OK;
}

static int box_ok(const int b[4]) { return b[1] >= b[0] && b[3] >= b[2]; }

// Boxes the resample kernels can take: non-empty, slave columns inside the
// per-range arrays, slave lines in [0, kMaxSlaveLine).  GetIndexes
// (DerampAndResample.cu:108-112) truncates toward zero, so a negative
// interior position would give a negative LUT row; and the walker packs
// (line, LUT row) keys into one int.
int check_burst_geom(const tops_burst_geom* g, const char* who)
{
    if (!box_ok(g->master_box) || !box_ok(g->slave_box))
        return set_error(TOPS_ERR_ARG, "%s: empty master or slave box", who);
    if (g->slave_box[0] < 0 || g->slave_box[1] >= g->samples_per_burst)
        return set_error(TOPS_ERR_ARG, "%s: slave columns [%d,%d] outside the %d per-range samples",
                         who, g->slave_box[0], g->slave_box[1], g->samples_per_burst);
    if (g->slave_box[2] < 0 || g->slave_box[3] >= kMaxSlaveLine)
        return set_error(TOPS_ERR_ARG, "%s: slave lines [%d,%d] outside [0, %d)", who,
                         g->slave_box[2], g->slave_box[3], kMaxSlaveLine);
    if (g->master_box[2] <= -kMaxSlaveLine || g->master_box[3] >= kMaxSlaveLine)
        return set_error(TOPS_ERR_ARG, "%s: master lines [%d,%d] outside +-%d", who,
                         g->master_box[2], g->master_box[3], kMaxSlaveLine);
    if ((long long)(g->slave_box[3] - g->slave_box[2] + 1) *
            (g->slave_box[1] - g->slave_box[0] + 1) * 4 > kMaxSlaveBytes)
        return set_error(TOPS_ERR_ARG, "%s: slave box over 2^31 bytes of CInt16", who);
    return TOPS_OK;
}

}  // namespace tops

using namespace tops;

extern "C" {

int tops_last_error(void) { return g_err; }
const char* tops_last_error_message(void) { return g_msg; }
void tops_clear_error(void)
{
    g_err = TOPS_OK;
    g_msg[0] = 0;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
tops_ctx* tops_ctx_create(const float* kernel_az, const float* kernel_rg, int npoints)
{
    if (!kernel_az || !kernel_rg || npoints != kTaps) {
        set_error(TOPS_ERR_ARG, "tops_ctx_create: only 12-tap LUTs are supported (got %d)", npoints);
        return nullptr;
    }
    tops_ctx* c = new tops_ctx();
    size_t lut = sizeof(float) * kLutRows * kTaps;
    std::vector<double> mom(kLutRows * 4);
    for (int r = 0; r < kLutRows; r++) {
        double m0 = 0, m1 = 0, m2 = 0;
        for (int j = 0; j < kTaps; j++) {
            double k = (double)kernel_az[r * kTaps + j], u = (double)(j - kHalf);
            m0 += k;
            m1 += k * u;
            m2 += k * u * u;
        }
        mom[4 * r] = m0;
        mom[4 * r + 1] = m1;
        mom[4 * r + 2] = m2;
        mom[4 * r + 3] = 0.0;
    }
    // azimuth rows placed at tap offsets 0..4 of a 16-row window (zeros
    // elsewhere), for the fused walker
    std::vector<float> sh((size_t)kAzShifts * kLutRows * kAzShiftTaps, 0.0f);
    for (int o = 0; o < kAzShifts; o++)
        for (int r = 0; r < kLutRows; r++)
            for (int j = 0; j < kTaps; j++)
                sh[((size_t)o * kLutRows + r) * kAzShiftTaps + o + j] = kernel_az[r * kTaps + j];
    hipError_t e = hipMalloc(&c->d_kaz, lut);
    if (e == hipSuccess) e = hipMalloc(&c->d_krg, lut);
    if (e == hipSuccess) e = hipMalloc(&c->d_mom, sizeof(double) * kLutRows * 4);
    if (e == hipSuccess) e = hipMalloc(&c->d_kaz_sh, sizeof(float) * sh.size());
    if (e == hipSuccess)
        e = hipMemcpy(c->d_kaz_sh, sh.data(), sizeof(float) * sh.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_kaz, kernel_az, lut, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->d_krg, kernel_rg, lut, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(c->d_mom, mom.data(), sizeof(double) * kLutRows * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hip_fail(e, "tops_ctx_create", __FILE__, __LINE__);
        tops_ctx_destroy(c);
        return nullptr;
    }
    return c;
}

void tops_ctx_destroy(tops_ctx* c)
{
    if (!c) return;
    if (c->d_kaz) (void)hipFree(c->d_kaz);
    if (c->d_krg) (void)hipFree(c->d_krg);
    if (c->d_mom) (void)hipFree(c->d_mom);
    if (c->d_kaz_sh) (void)hipFree(c->d_kaz_sh);
    c->scratch.release();
    delete c;
}

// ---------------------------------------------------------------------------
// device-resident entry points
// ---------------------------------------------------------------------------
int tops_resample_burst(tops_ctx* ctx, const tops_burst_geom* g, const tops_ci16* d_slave,
                        const double* d_kt, const double* d_tref, const double* d_fdc,
                        tops_cf32* d_out, int mode, void* stream)
{
    if (!ctx || !g || !d_slave || !d_kt || !d_tref || !d_fdc || !d_out)
        return set_error(TOPS_ERR_ARG, "tops_resample_burst: null argument");
    if (int rc = check_burst_geom(g, "tops_resample_burst")) return rc;
    if (mode != TOPS_MODE_FAST && mode != TOPS_MODE_FP64)
        return set_error(TOPS_ERR_ARG, "tops_resample_burst: unknown mode %d", mode);
    ResampleParams p;
    memcpy(p.mb, g->master_box, sizeof(p.mb));
    memcpy(p.sb, g->slave_box, sizeof(p.sb));
    p.m_lines = g->master_box[3] - g->master_box[2] + 1;
    p.m_pixels = g->master_box[1] - g->master_box[0] + 1;
    p.s_lines = g->slave_box[3] - g->slave_box[2] + 1;
    p.s_pixels = g->slave_box[1] - g->slave_box[0] + 1;
    p.first_line = g->first_line_in_burst;
    p.dt = g->az_time_interval;
    p.az_shift = g->az_shift;
    memcpy(p.cpm_az, g->cpm_az, sizeof(p.cpm_az));
    memcpy(p.cpm_rg, g->cpm_rg, sizeof(p.cpm_rg));
    p.slave = d_slave;
    p.kt = d_kt;
    p.tref = d_tref;
    p.fdc = d_fdc;
    p.kaz = ctx->d_kaz;
    p.krg = ctx->d_krg;
    p.mom = ctx->d_mom;
    p.kaz_sh = ctx->d_kaz_sh;
    p.az_shift_dev = nullptr;
    p.out = d_out;
    hipStream_t s = (hipStream_t)stream;
    return mode == TOPS_MODE_FAST ? launch_resample_fast(p, s) : launch_resample_fp64(ctx, p, s);
}

int tops_coherence(const tops_ci16* d_master, const tops_cf32* d_slave, int w, int h, int dx,
                   int dy, int win_rg, int win_az, float* d_coh, int mode, void* stream)
{
    if (!d_master || !d_slave || !d_coh || w <= 0 || h <= 0 || win_rg <= 0 || win_az <= 0)
        return set_error(TOPS_ERR_ARG, "tops_coherence: bad arguments (w=%d h=%d win=%dx%d)", w, h,
                         win_az, win_rg);
    return launch_coherence(d_master, d_slave, w, h, dx, dy, win_rg, win_az, d_coh, mode,
                            (hipStream_t)stream);
}

int tops_esd_shifts(const tops_ci16* d_m_for, const tops_ci16* d_m_back,
                    const tops_cf32* d_s_for, const tops_cf32* d_s_back,
                    const int* overlap_size, int n_overlap, int pixels, int win_az, int win_rg,
                    float threshold, double spectral_sep, double az_time_interval,
                    double* shifts, double* sums, void* stream)
{
    if (n_overlap < 0 || (n_overlap > 0 && (!overlap_size || !shifts || pixels <= 0)))
        return set_error(TOPS_ERR_ARG, "tops_esd_shifts: bad arguments");
    if (n_overlap == 0) return TOPS_OK;
    for (int i = 0; i < n_overlap; i++)
        if (overlap_size[i] < 0) return set_error(TOPS_ERR_ARG, "tops_esd_shifts: negative overlap");
    hipStream_t s = (hipStream_t)stream;
    int cap = esd_partial_capacity(overlap_size, n_overlap, pixels);
    double* d_part = nullptr;
    double* d_sums = nullptr;
    TOPS_HIP(hipMalloc(&d_part, sizeof(double) * 2 * (size_t)(cap > 0 ? cap : 1)));
    hipError_t e = hipMalloc(&d_sums, sizeof(double) * 2 * n_overlap);
    if (e != hipSuccess) {
        (void)hipFree(d_part);
        return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    }
    int rc = launch_esd(d_m_for, d_m_back, d_s_for, d_s_back, overlap_size, n_overlap, pixels,
                        win_az, win_rg, threshold, d_sums, d_part, 0, s);
    std::vector<double> h(2 * n_overlap);
    if (rc == TOPS_OK) {
        e = hipMemcpyAsync(h.data(), d_sums, sizeof(double) * 2 * n_overlap, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = hip_fail(e, "esd readback", __FILE__, __LINE__);
    }
    (void)hipFree(d_part);
    (void)hipFree(d_sums);
    if (rc) return rc;
    for (int i = 0; i < n_overlap; i++) {
        // EstESD.cu:1213-1217
        shifts[i] = atan2(h[2 * i + 1], h[2 * i]) / (2 * kPi * spectral_sep * az_time_interval);
        if (sums) {
            sums[2 * i] = h[2 * i];
            sums[2 * i + 1] = h[2 * i + 1];
        }
    }
    return TOPS_OK;
}

int tops_geocoreg(const int32_t* d_dem, int lines, int pixels, double lat0, double lon0,
                  double dlat, double dlon, int nodata, const tops_orbit_geom* m,
                  const tops_orbit_geom* s, int xmin, int xmax, int ymin, int ymax, double* cpm_az,
                  double* cpm_rg, int64_t* n_used, double* d_lp, void* stream)
{
    if (!d_dem || lines <= 0 || pixels <= 0 || !cpm_az || !cpm_rg || xmax <= xmin || ymax <= ymin)
        return set_error(TOPS_ERR_ARG, "tops_geocoreg: bad arguments");
    OrbitDev om, os;
    int rc = make_orbit(m, om);
    if (rc) return rc;
    rc = make_orbit(s, os);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    int nb = geocoreg_blocks(lines, pixels);
    double* d_part = nullptr;
    TOPS_HIP(hipMalloc(&d_part, sizeof(double) * geocoreg_scratch_doubles(lines, pixels, nb)));
    rc = launch_geocoreg(d_dem, lines, pixels, lat0, lon0, dlat, dlon, nodata, om, os, xmin, xmax,
                         ymin, ymax, d_part, nb, d_lp, st);
    double sum[kGeoSums];
    if (rc == TOPS_OK) {
        hipError_t e = hipMemcpyAsync(sum, d_part + (size_t)nb * kGeoSums, sizeof(sum),
                                      hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = hip_fail(e, "geocoreg readback", __FILE__, __LINE__);
    }
    (void)hipFree(d_part);
    if (rc) return rc;
    // N[i][j] = sum a_i a_j over [1, L, P, L^2, LP, P^2]: the moment of the
    // summed exponents of L and P (degree d = i + j, power of P p: d(d+1)/2 + p)
    static const int eL[6] = {0, 1, 0, 2, 1, 0}, eP[6] = {0, 0, 1, 0, 1, 2};
    double N[36], ba[6], br[6];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            const int d = eL[i] + eL[j] + eP[i] + eP[j], p = eP[i] + eP[j];
            N[i * 6 + j] = sum[d * (d + 1) / 2 + p];
        }
    for (int i = 0; i < 6; i++) {
        ba[i] = sum[kGeoMoments + i];
        br[i] = sum[kGeoMoments + 6 + i];
    }
    const double used = sum[0];   // the moment L^0 P^0
    if (n_used) *n_used = (int64_t)used;
    if (used < 6.0 || !chol_solve(N, 6, ba, br)) {
        for (int i = 0; i < 6; i++) cpm_az[i] = cpm_rg[i] = 0.0;
        return set_error(TOPS_ERR_NOT_PD,
                         "geocoreg: normal matrix not positive definite (%.0f posts used)", used);
    }
    for (int i = 0; i < 6; i++) {
        cpm_az[i] = ba[i];
        cpm_rg[i] = br[i];
    }
    return TOPS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// reference entry points (host buffers)
// ---------------------------------------------------------------------------
namespace {

std::mutex g_ref_mutex;
tops_ctx* g_ref_ctx = nullptr;
std::vector<float> g_ref_lut;

// Context for the LUTs a reference-style call passes; rebuilt only when the
// LUT contents change.
tops_ctx* ref_ctx(const float* kaz, const float* krg, int npoints)
{
    if (npoints != kTaps) {
        set_error(TOPS_ERR_ARG, "Npoints must be 12 (got %d)", npoints);
        return nullptr;
    }
    size_t n = (size_t)kLutRows * kTaps;
    if (g_ref_ctx && g_ref_lut.size() == 2 * n && !memcmp(g_ref_lut.data(), kaz, n * 4) &&
        !memcmp(g_ref_lut.data() + n, krg, n * 4))
        return g_ref_ctx;
    tops_ctx_destroy(g_ref_ctx);
    g_ref_ctx = tops_ctx_create(kaz, krg, npoints);
    g_ref_lut.assign(kaz, kaz + n);
    g_ref_lut.insert(g_ref_lut.end(), krg, krg + n);
    if (!g_ref_ctx) g_ref_lut.clear();
    return g_ref_ctx;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t n)
    {
        TOPS_HIP(hipMalloc(&p, n ? n : 1));
        return TOPS_OK;
    }
    void* release()
    {
        void* q = p;
        p = nullptr;
        return q;
    }
};

tops_cf32* ref_resample(tops_ci16* SlaveArray, double* CpmAz, double* CpmRg, double AzimuthShift,
                        tops_cf32* output, float* KernelAz, float* KernelRg, int sBurstIdx,
                        int MasterBox[4], int SlaveBox[4], int linesPerBurst, int SamplesPerBurst,
                        double azimuthTimeInterval, double* dopplerRate, double* referenceTime,
                        double* dopplerCentroid, int Npoints, bool keep)
{
    std::lock_guard<std::mutex> lock(g_ref_mutex);
    tops_ctx* ctx = ref_ctx(KernelAz, KernelRg, Npoints);
    if (!ctx) return nullptr;
    if (!SlaveArray || !CpmAz || !CpmRg || !output || !MasterBox || !SlaveBox || !dopplerRate ||
        !referenceTime || !dopplerCentroid) {
        set_error(TOPS_ERR_ARG, "DerampDemodResample: null argument");
        return nullptr;
    }
    tops_burst_geom g;
    memcpy(g.master_box, MasterBox, sizeof(g.master_box));
    memcpy(g.slave_box, SlaveBox, sizeof(g.slave_box));
    g.first_line_in_burst = sBurstIdx * linesPerBurst;
    g.samples_per_burst = SamplesPerBurst;
    g.az_time_interval = azimuthTimeInterval;
    g.az_shift = AzimuthShift;
    memcpy(g.cpm_az, CpmAz, sizeof(g.cpm_az));
    memcpy(g.cpm_rg, CpmRg, sizeof(g.cpm_rg));
    if (!box_ok(MasterBox) || !box_ok(SlaveBox)) {
        set_error(TOPS_ERR_ARG, "DerampDemodResample: empty box");
        return nullptr;
    }
    size_t sn = (size_t)(SlaveBox[3] - SlaveBox[2] + 1) * (SlaveBox[1] - SlaveBox[0] + 1);
    size_t mn = (size_t)(MasterBox[3] - MasterBox[2] + 1) * (MasterBox[1] - MasterBox[0] + 1);
    DevBuf ds, dk, dout;
    if (ds.alloc(sn * 4) || dk.alloc(sizeof(double) * 3 * SamplesPerBurst) || dout.alloc(mn * 8))
        return nullptr;
    double* kt = (double*)dk.p;
    hipError_t e = hipMemcpy(ds.p, SlaveArray, sn * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(kt, dopplerRate, sizeof(double) * SamplesPerBurst, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(kt + SamplesPerBurst, referenceTime, sizeof(double) * SamplesPerBurst,
                      hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(kt + 2 * SamplesPerBurst, dopplerCentroid, sizeof(double) * SamplesPerBurst,
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hip_fail(e, "DerampDemodResample upload", __FILE__, __LINE__);
        return nullptr;
    }
    if (tops_resample_burst(ctx, &g, (const tops_ci16*)ds.p, kt, kt + SamplesPerBurst,
                            kt + 2 * SamplesPerBurst, (tops_cf32*)dout.p, TOPS_MODE_FAST, nullptr))
        return nullptr;
    e = hipMemcpy(output, dout.p, mn * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        hip_fail(e, "DerampDemodResample readback", __FILE__, __LINE__);
        return nullptr;
    }
    return keep ? (tops_cf32*)dout.release() : output;
}

int ref_coherence(int Dw, int Dh, int dx, int dy, int cohWinRg, int cohWinAz,
                  tops_ci16* masterArray, tops_cf32* d_slaveArray, float* cohdata)
{
    DevBuf owned;
    owned.p = d_slaveArray;   // consumed, as in EstCoh.cu:708, 900
    if (!masterArray || !d_slaveArray || !cohdata || Dw <= 0 || Dh <= 0)
        return set_error(TOPS_ERR_ARG, "CoherenceEst: bad arguments");
    size_t n = (size_t)Dw * Dh;
    DevBuf dm, dc;
    if (dm.alloc(n * 4) || dc.alloc(n * 4)) return tops_last_error();
    TOPS_HIP(hipMemcpy(dm.p, masterArray, n * 4, hipMemcpyHostToDevice));
    int rc = tops_coherence((const tops_ci16*)dm.p, d_slaveArray, Dw, Dh, dx, dy, cohWinRg, cohWinAz,
                            (float*)dc.p, TOPS_MODE_FAST, nullptr);
    if (rc) return rc;
    TOPS_HIP(hipMemcpy(cohdata, dc.p, n * 4, hipMemcpyDeviceToHost));
    return TOPS_OK;
}

}  // namespace

extern "C" {

void DerampDemodResample(tops_ci16* SlaveArray, double* CpmAz, double* CpmRg, double AzimuthShift,
                         tops_cf32* output, float* KernelAz, float* KernelRg, int sBurstIdx,
                         int slave_pixels, int slave_lines, int MasterBox[4], int SlaveBox[4],
                         int linesPerBurst, int SamplesPerBurst, double azimuthTimeInterval,
                         double* dopplerRate, double* referenceTime, double* dopplerCentroid,
                         int Npoints)
{
    tops_clear_error();
    (void)slave_pixels;
    (void)slave_lines;
    ref_resample(SlaveArray, CpmAz, CpmRg, AzimuthShift, output, KernelAz, KernelRg, sBurstIdx,
                 MasterBox, SlaveBox, linesPerBurst, SamplesPerBurst, azimuthTimeInterval,
                 dopplerRate, referenceTime, dopplerCentroid, Npoints, false);
}

tops_cf32* DerampDemodResample_ESD(tops_ci16* SlaveArray, double* CpmAz, double* CpmRg,
                                   double AzimuthShift, tops_cf32* output, float* KernelAz,
                                   float* KernelRg, int sBurstIdx, int slave_pixels,
                                   int slave_lines, int MasterBox[4], int SlaveBox[4],
                                   int linesPerBurst, int SamplesPerBurst,
                                   double azimuthTimeInterval, double* dopplerRate,
                                   double* referenceTime, double* dopplerCentroid, int Npoints)
{
    tops_clear_error();
    (void)slave_pixels;
    (void)slave_lines;
    return ref_resample(SlaveArray, CpmAz, CpmRg, AzimuthShift, output, KernelAz, KernelRg,
                        sBurstIdx, MasterBox, SlaveBox, linesPerBurst, SamplesPerBurst,
                        azimuthTimeInterval, dopplerRate, referenceTime, dopplerCentroid, Npoints,
                        true);
}

tops_cf32* ResampleFirstBurst(tops_cf32* SlaveArray, int ww, int hh)
{
    tops_clear_error();
    if (!SlaveArray || ww <= 0 || hh <= 0) {
        set_error(TOPS_ERR_ARG, "ResampleFirstBurst: bad arguments");
        return nullptr;
    }
    size_t n = (size_t)ww * hh * 8;
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, n);
    if (e == hipSuccess) e = hipMemcpy(d, SlaveArray, n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hip_fail(e, "ResampleFirstBurst", __FILE__, __LINE__);
        if (d) (void)hipFree(d);
        return nullptr;
    }
    return (tops_cf32*)d;
}

void CoherenceEst(int Dx0, int Dy0, int Dw, int Dh, int dx, int dy, int cohWinRg, int cohWinAz,
                  tops_ci16* masterArray, tops_cf32* d_slaveArray, float* cohdata)
{
    tops_clear_error();
    (void)Dx0;
    (void)Dy0;
    ref_coherence(Dw, Dh, dx, dy, cohWinRg, cohWinAz, masterArray, d_slaveArray, cohdata);
}

void CoherenceEst_ESD(int Dx0, int Dy0, int Dw, int Dh, int dx, int dy, int cohWinRg,
                      int cohWinAz, tops_ci16* masterArray, tops_cf32* d_slaveArray,
                      float* cohdata)
{
    tops_clear_error();
    (void)Dx0;
    (void)Dy0;
    ref_coherence(Dw, Dh, dx, dy, cohWinRg, cohWinAz, masterArray, d_slaveArray, cohdata);
}

void EstESDShifts(tops_ci16* MasterFor, tops_ci16* MasterBack, tops_cf32* SlaveFor,
                  tops_cf32* SlaveBack, int* OverlapSizeArray, double* ShiftArray, int numOverlap,
                  int ArrayLines, int ArrayPixels, int CohWinAz, int CohWinRg, int OverlapMaxLines,
                  float CohThresHold, double spectralSeparation, double azimuthTimeInterval)
{
    tops_clear_error();
    (void)OverlapMaxLines;
    if (numOverlap <= 0) return;
    if (!MasterFor || !MasterBack || !SlaveFor || !SlaveBack || !OverlapSizeArray || !ShiftArray) {
        set_error(TOPS_ERR_ARG, "EstESDShifts: null argument");
        return;
    }
    long total = 0;
    for (int i = 0; i < numOverlap; i++) total += OverlapSizeArray[i];
    if (total > ArrayLines) {
        set_error(TOPS_ERR_ARG, "EstESDShifts: overlaps sum to %ld lines > ArrayLines %d", total,
                  ArrayLines);
        return;
    }
    size_t n = (size_t)total * ArrayPixels;
    DevBuf a, b, c, d;
    if (a.alloc(n * 4) || b.alloc(n * 4) || c.alloc(n * 8) || d.alloc(n * 8)) return;
    hipError_t e = hipMemcpy(a.p, MasterFor, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.p, MasterBack, n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c.p, SlaveFor, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d.p, SlaveBack, n * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hip_fail(e, "EstESDShifts upload", __FILE__, __LINE__);
        return;
    }
    tops_esd_shifts((const tops_ci16*)a.p, (const tops_ci16*)b.p, (const tops_cf32*)c.p,
                    (const tops_cf32*)d.p, OverlapSizeArray, numOverlap, ArrayPixels, CohWinAz,
                    CohWinRg, CohThresHold, spectralSeparation, azimuthTimeInterval, ShiftArray,
                    nullptr, nullptr);
}

void GeoCoreg_warp_cuda(int m_burstIdx, int s_burstIdx, int xmin, int xmax, int ymin, int ymax,
                        double dem_DeltaLat, double dem_DeltaLon, int DemLines, int DemPixels,
                        double* AzCpm, double* RgCpm, double lat0, double lon0, int demnodata,
                        double m_azimuthTimeInterval, double s_azimuthTimeInterval,
                        double wavelength, double m_linesPerBurst, double s_linesPerBurst,
                        double m_burstFirstLineTime, double s_burstFirstLineTime,
                        int m_numOfSamples, int s_numOfSamples, double* m_orbitAzTime,
                        double* s_orbitAzTime, int orb_m_numberofpoints, int orb_m_numbercoefs,
                        int orb_s_numberofpoints, int orb_s_numbercoefs,
                        double m_slrTimeToFirstPixel, double s_slrTimeToFirstPixel,
                        double m_rangeSpacing, double s_rangeSpacing, bool m_nearRangeOnLeft,
                        bool s_nearRangeOnLeft, int* inputDEM, double* m_x_coef,
                        double* m_y_coef, double* m_z_coef, double* s_x_coef, double* s_y_coef,
                        double* s_z_coef, double m_aziInitial, double s_aziInitial)
{
    tops_clear_error();
    (void)wavelength;
    (void)m_numOfSamples;
    (void)s_numOfSamples;
    (void)m_nearRangeOnLeft;
    (void)s_nearRangeOnLeft;
    if (!AzCpm || !RgCpm || !inputDEM || !m_orbitAzTime || !s_orbitAzTime ||
        orb_m_numberofpoints < 1 || orb_s_numberofpoints < 1 || DemLines <= 0 || DemPixels <= 0) {
        set_error(TOPS_ERR_ARG, "GeoCoreg_warp_cuda: bad arguments");
        return;
    }
    tops_orbit_geom m = {m_x_coef, m_y_coef, m_z_coef, orb_m_numbercoefs,
                         m_orbitAzTime[orb_m_numberofpoints / 2], m_azimuthTimeInterval,
                         m_burstFirstLineTime, m_slrTimeToFirstPixel, m_rangeSpacing, m_aziInitial,
                         (int)(m_burstIdx * m_linesPerBurst)};
    tops_orbit_geom s = {s_x_coef, s_y_coef, s_z_coef, orb_s_numbercoefs,
                         s_orbitAzTime[orb_s_numberofpoints / 2], s_azimuthTimeInterval,
                         s_burstFirstLineTime, s_slrTimeToFirstPixel, s_rangeSpacing, s_aziInitial,
                         (int)(s_burstIdx * s_linesPerBurst)};
    size_t n = (size_t)DemLines * DemPixels;
    DevBuf dd;
    if (dd.alloc(n * 4)) return;
    hipError_t e = hipMemcpy(dd.p, inputDEM, n * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hip_fail(e, "GeoCoreg_warp_cuda upload", __FILE__, __LINE__);
        return;
    }
    tops_geocoreg((const int32_t*)dd.p, DemLines, DemPixels, lat0, lon0, dem_DeltaLat, dem_DeltaLon,
                  demnodata, &m, &s, xmin, xmax, ymin, ymax, AzCpm, RgCpm, nullptr, nullptr, nullptr);
}

}  // extern "C"
