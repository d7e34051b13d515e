// subswath.cpp -- ESDAndCoh (Func.cpp:2496-2763) over one subswath held in
// HBM, with no host round trip between its steps:
//   1. pass 1: the DerampAndResample stage (Func.cpp:2326-2406: slave lines
//      and slave azimuth interval), on the rows this flow reads from it:
//      burst 0 whole (its result is final) and, of every later burst, only
//      the lines in an overlap band.  Pass 2 rewrites all other rows, so the
//      outputs equal those of a full pass 1;
//   2. ESD over the overlap bands, read in place from the master subswath
//      and the pass-1 result (the forward band of overlap i is the last n_i
//      lines of burst i, the backward band the first n_i lines of burst
//      i+1; DataTypes.cpp:456-513), shifts formed on the device with the
//      |shift| > 100 reset (Func.cpp:2612-2621);
//   3. pass 2: bursts 1 .. n-1 resampled again with AzimuthShift =
//      shift[b-1] read from device memory, the master's lines per burst and
//      azimuth interval (Func.cpp:2701-2705); burst 0 keeps its pass-1
//      result;
//   4. coherence of every burst against the master.
// Every launch goes to the caller's stream in this order, so the whole flow
// can be captured in one hipGraph.  The workspace (ESD partial sums) comes
// from the caller; nothing is allocated here.
#include "tops_internal.hpp"

#include <string.h>

#include <vector>

using namespace tops;

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// workspace layout: [ESD partials (double2)] [ESD sums (2 x n_overlap double)]
size_t workspace_parts(const tops_subswath_desc* d, size_t* partial_bytes)
{
    const int nov = d->n_bursts - 1;
    size_t pb = 0;
    if (nov > 0) {
        std::vector<int> sizes(d->overlap_lines, d->overlap_lines + nov);
        pb = align_up(sizeof(double) * 2 * (size_t)esd_partial_capacity(sizes.data(), nov,
                                                                         d->m_samples_per_burst));
    }
    if (partial_bytes) *partial_bytes = pb;
    return pb + align_up(sizeof(double) * 2 * (size_t)(nov > 0 ? nov : 1));
}

int check_desc(const tops_subswath_desc* d, const char* who)
{
    if (!d || d->n_bursts < 1 || d->m_lines_per_burst <= 0 || d->m_samples_per_burst <= 0 ||
        d->s_lines_per_burst <= 0 || d->s_samples_per_burst <= 0 ||
        (d->n_bursts > 1 && !d->overlap_lines))
        return set_error(TOPS_ERR_ARG, "%s: bad subswath description", who);
    if (d->n_bursts - 1 > 64)
        return set_error(TOPS_ERR_ARG, "%s: at most 65 bursts per subswath", who);
    if ((long)d->n_bursts * d->m_lines_per_burst >= kMaxSlaveLine ||
        (long)d->n_bursts * d->s_lines_per_burst >= kMaxSlaveLine)
        return set_error(TOPS_ERR_ARG, "%s: more than %d lines in the subswath", who,
                         kMaxSlaveLine - 1);
    for (int i = 0; i + 1 < d->n_bursts; i++)
        if (d->overlap_lines[i] < 0 || d->overlap_lines[i] > d->m_lines_per_burst)
            return set_error(TOPS_ERR_ARG, "%s: overlap %d has %d lines", who, i,
                             d->overlap_lines[i]);
    return TOPS_OK;
}

}  // namespace

extern "C" {

size_t tops_subswath_workspace_size(const tops_subswath_desc* d)
{
    if (check_desc(d, "tops_subswath_workspace_size")) return 0;
    return workspace_parts(d, nullptr);
}

int tops_subswath_esd_coh(tops_ctx* ctx, const tops_subswath_desc* d, const tops_ci16* d_master,
                          const tops_ci16* d_slave, const double* d_kt, const double* d_tref,
                          const double* d_fdc, const double* cpm_az, const double* cpm_rg,
                          tops_cf32* d_reslave, float* d_coh, double* d_shifts,
                          void* d_workspace, size_t workspace_bytes, void* stream)
{
    const char* who = "tops_subswath_esd_coh";
    int rc = check_desc(d, who);
    if (rc) return rc;
    if (!ctx || !d_master || !d_slave || !d_kt || !d_tref || !d_fdc || !cpm_az || !cpm_rg ||
        !d_reslave || (d->n_bursts > 1 && !d_shifts))
        return set_error(TOPS_ERR_ARG, "%s: null argument", who);
    size_t partial_bytes = 0;
    const size_t need = workspace_parts(d, &partial_bytes);
    if (!d_workspace || workspace_bytes < need)
        return set_error(TOPS_ERR_ARG, "%s: workspace of %zu bytes needed (got %zu)", who, need,
                         workspace_bytes);
    if (d->coh_win_az <= 0 || d->coh_win_rg <= 0 || d->esd_win_az <= 0 || d->esd_win_rg <= 0)
        return set_error(TOPS_ERR_ARG, "%s: bad window", who);
    hipStream_t s = (hipStream_t)stream;
    const int nb = d->n_bursts, nov = nb - 1;
    const int lpb = d->m_lines_per_burst, spb = d->m_samples_per_burst;
    const int lps = d->s_lines_per_burst, sps = d->s_samples_per_burst;
    const size_t bpix = (size_t)lpb * spb;

    auto params = [&](int b, bool pass2) {
        ResampleParams p;
        memset(&p, 0, sizeof(p));
        const int mbox[4] = {0, spb - 1, b * lpb, (b + 1) * lpb - 1};
        const int sbox[4] = {0, sps - 1, b * lps, (b + 1) * lps - 1};
        memcpy(p.mb, mbox, sizeof(mbox));
        memcpy(p.sb, sbox, sizeof(sbox));
        p.m_lines = lpb;
        p.m_pixels = spb;
        p.s_lines = lps;
        p.s_pixels = sps;
        p.first_line = pass2 ? b * lpb : b * lps;
        p.dt = pass2 ? d->m_az_time_interval : d->s_az_time_interval;
        p.az_shift = 0.0;
        p.az_shift_dev = pass2 ? d_shifts + (b - 1) : nullptr;
        memcpy(p.cpm_az, cpm_az + 6 * b, sizeof(p.cpm_az));
        memcpy(p.cpm_rg, cpm_rg + 6 * b, sizeof(p.cpm_rg));
        p.slave = d_slave + (size_t)b * lps * sps;
        p.kt = d_kt + (size_t)b * sps;
        p.tref = d_tref + (size_t)b * sps;
        p.fdc = d_fdc + (size_t)b * sps;
        p.kaz = ctx->d_kaz;
        p.krg = ctx->d_krg;
        p.mom = ctx->d_mom;
        p.kaz_sh = ctx->d_kaz_sh;
        p.out = d_reslave + (size_t)b * bpix;
        return p;
    };

    // pass 1: burst 0 whole; burst b > 0 only its backward band (first
    // n_{b-1} lines) and forward band (last n_b lines)
    rc = launch_resample_fast(params(0, false), s, 0);
    for (int b = 1; b < nb && !rc; b++) {
        const int head = d->overlap_lines[b - 1];
        const int tail = b + 1 < nb ? d->overlap_lines[b] : 0;
        const ResampleParams p = params(b, false);
        if (head + tail >= lpb) {
            rc = launch_resample_fast(p, s, 0);
        } else {
            rc = launch_resample_rows(p, 0, head, s);
            if (!rc) rc = launch_resample_rows(p, lpb - tail, tail, s);
        }
    }
    if (rc) return rc;
    if (nov > 0) {
        std::vector<int> f_off(nov), b_off(nov);
        for (int i = 0; i < nov; i++) {
            f_off[i] = (i + 1) * lpb - d->overlap_lines[i];
            b_off[i] = (i + 1) * lpb;
        }
        double* partials = (double*)d_workspace;
        double* sums = (double*)((char*)d_workspace + partial_bytes);
        rc = launch_esd_bands(d_master, d_master, d_reslave, d_reslave, f_off.data(), b_off.data(),
                              d->overlap_lines, nov, spb, d->esd_win_az, d->esd_win_rg,
                              d->threshold, sums, partials, s);
        if (!rc)
            rc = launch_esd_shifts(sums, nov, d->spectral_sep, d->m_az_time_interval, d_shifts, s);
        for (int b = 1; b < nb && !rc; b++) rc = launch_resample_fast(params(b, true), s, 0);
        if (rc) return rc;
    }
    if (d_coh) {
        const int dx = (d->coh_win_rg - 1) / 2, dy = (d->coh_win_az - 1) / 2;
        for (int b = 0; b < nb && !rc; b++)
            rc = launch_coherence(d_master + (size_t)b * bpix, d_reslave + (size_t)b * bpix, spb,
                                  lpb, dx, dy, d->coh_win_rg, d->coh_win_az,
                                  d_coh + (size_t)b * bpix, TOPS_MODE_FAST, s);
    }
    return rc;
}

}  // extern "C"
