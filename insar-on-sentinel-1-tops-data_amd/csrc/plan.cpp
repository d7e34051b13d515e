// plan.cpp -- batched execution of many bursts (all bursts of all subswaths
// of an IW pair, or a stack of slaves against one master): one strip table
// covers every burst, so the whole deramp -> resample -> reramp pass is one
// launch and the coherence pass another; the pair of launches can be
// captured once in a hipGraph and replayed.  The strip height is a tuning
// parameter chosen by timing the candidates on the plan's own bursts.
#include <algorithm>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>

#include "tops_internal.hpp"

using namespace tops;

struct tops_plan {
    tops_ctx* ctx = nullptr;
    std::vector<tops_burst_item> items;
    int coh_win_az = 3, coh_win_rg = 10;
    // schedule: 0 = one launch pair per burst (resample, then coherence) with
    // each burst's automatic heights; > 0 = one batched launch of all bursts'
    // strips of this height, then one batched coherence launch
    int strip_rows = 0;
    std::vector<ResampleParams> h_params;
    std::vector<CohItem> h_coh;
    std::vector<int> coh_of;   // index into h_coh per item, or -1
    ResampleParams* d_params = nullptr;
    int4* d_work_rs = nullptr;
    int n_work_rs = 0;
    CohItem* d_coh = nullptr;
    int4* d_work_coh = nullptr;
    int n_work_coh = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int fill_params(tops_plan* p)
{
    std::vector<ResampleParams> prm(p->items.size());
    for (size_t i = 0; i < p->items.size(); i++) {
        const tops_burst_item& it = p->items[i];
        const tops_burst_geom& g = it.geom;
        ResampleParams& r = prm[i];
        memset(&r, 0, sizeof(r));
        memcpy(r.mb, g.master_box, sizeof(r.mb));
        memcpy(r.sb, g.slave_box, sizeof(r.sb));
        r.m_lines = g.master_box[3] - g.master_box[2] + 1;
        r.m_pixels = g.master_box[1] - g.master_box[0] + 1;
        r.s_lines = g.slave_box[3] - g.slave_box[2] + 1;
        r.s_pixels = g.slave_box[1] - g.slave_box[0] + 1;
        r.first_line = g.first_line_in_burst;
        r.dt = g.az_time_interval;
        r.az_shift = g.az_shift;
        memcpy(r.cpm_az, g.cpm_az, sizeof(r.cpm_az));
        memcpy(r.cpm_rg, g.cpm_rg, sizeof(r.cpm_rg));
        r.slave = it.slave;
        r.kt = it.kt;
        r.tref = it.tref;
        r.fdc = it.fdc;
        r.kaz = p->ctx->d_kaz;
        r.krg = p->ctx->d_krg;
        r.mom = p->ctx->d_mom;
        r.kaz_sh = p->ctx->d_kaz_sh;
        r.az_shift_dev = nullptr;
        r.out = it.out;
    }
    p->h_params = prm;
    TOPS_HIP(hipMalloc(&p->d_params, sizeof(ResampleParams) * prm.size()));
    TOPS_HIP(hipMemcpy(p->d_params, prm.data(), sizeof(ResampleParams) * prm.size(),
                       hipMemcpyHostToDevice));
    std::vector<CohItem> ci;
    std::vector<int4> wc;
    p->coh_of.assign(p->items.size(), -1);
    for (size_t i = 0; i < p->items.size(); i++) {
        const tops_burst_item& it = p->items[i];
        if (!it.master || !it.coh) continue;
        p->coh_of[i] = (int)ci.size();
        CohItem c;
        c.m = it.master;
        c.s = it.out;
        c.coh = it.coh;
        c.w = prm[i].m_pixels;
        c.h = prm[i].m_lines;
        int nx, ny, tag = 0;
        if (coherence_walk_fits((p->coh_win_rg - 1) / 2, (p->coh_win_az - 1) / 2, p->coh_win_rg,
                                p->coh_win_az)) {
            coherence_walk_count(c.w, c.h, &nx, &ny);
            tag = coherence_walk_vec(c.m, c.s, c.coh, c.w);
        } else {
            coherence_tile_count(c.w, c.h, &nx, &ny);
        }
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) wc.push_back(make_int4((int)ci.size(), x, y, tag));
        ci.push_back(c);
    }
    p->h_coh = ci;
    p->n_work_coh = (int)wc.size();
    if (!ci.empty()) {
        TOPS_HIP(hipMalloc(&p->d_coh, sizeof(CohItem) * ci.size()));
        TOPS_HIP(hipMemcpy(p->d_coh, ci.data(), sizeof(CohItem) * ci.size(), hipMemcpyHostToDevice));
        TOPS_HIP(hipMalloc(&p->d_work_coh, sizeof(int4) * wc.size()));
        TOPS_HIP(hipMemcpy(p->d_work_coh, wc.data(), sizeof(int4) * wc.size(), hipMemcpyHostToDevice));
    }
    return TOPS_OK;
}

// strip table for the current strip height; strips of all bursts are
// interleaved row-band by row-band so uneven bursts spread over the GPU
int fill_work(tops_plan* p)
{
    if (p->strip_rows <= 0) return TOPS_OK;   // per-burst launches need no table
    std::vector<int4> w;
    int max_ny = 0;
    std::vector<int> nxs(p->items.size()), nys(p->items.size());
    for (size_t i = 0; i < p->items.size(); i++) {
        const tops_burst_geom& g = p->items[i].geom;
        resample_strip_count(g.master_box[3] - g.master_box[2] + 1,
                             g.master_box[1] - g.master_box[0] + 1, p->strip_rows, &nxs[i], &nys[i]);
        max_ny = std::max(max_ny, nys[i]);
    }
    for (int y = 0; y < max_ny; y++)
        for (size_t i = 0; i < p->items.size(); i++)
            if (y < nys[i])
                for (int x = 0; x < nxs[i]; x++) w.push_back(make_int4((int)i, x, y, p->strip_rows));
    if (p->d_work_rs) (void)hipFree(p->d_work_rs);
    p->d_work_rs = nullptr;
    p->n_work_rs = (int)w.size();
    if (!w.empty()) {
        TOPS_HIP(hipMalloc(&p->d_work_rs, sizeof(int4) * w.size()));
        TOPS_HIP(hipMemcpy(p->d_work_rs, w.data(), sizeof(int4) * w.size(), hipMemcpyHostToDevice));
    }
    return TOPS_OK;
}

int enqueue(tops_plan* p, hipStream_t s)
{
    const int dx = (p->coh_win_rg - 1) / 2, dy = (p->coh_win_az - 1) / 2;
    if (p->strip_rows <= 0) {
        for (size_t i = 0; i < p->h_params.size(); i++) {
            int rc = launch_resample_fast(p->h_params[i], s, 0);
            if (rc) return rc;
            if (p->coh_of[i] >= 0) {
                const CohItem& c = p->h_coh[p->coh_of[i]];
                rc = launch_coherence(c.m, c.s, c.w, c.h, dx, dy, p->coh_win_rg, p->coh_win_az,
                                      c.coh, TOPS_MODE_FAST, s);
                if (rc) return rc;
            }
        }
        return TOPS_OK;
    }
    int rc = launch_resample_batch(p->d_params, p->d_work_rs, p->n_work_rs, s);
    if (rc) return rc;
    return launch_coherence_batch(p->d_coh, p->d_work_coh, p->n_work_coh, dx, dy, p->coh_win_rg,
                                  p->coh_win_az, s);
}

void drop_graph(tops_plan* p)
{
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

}  // namespace

extern "C" {

tops_plan* tops_plan_create(tops_ctx* ctx, const tops_burst_item* items, int n_items,
                            int coh_win_az, int coh_win_rg, int strip_rows)
{
    if (!ctx || !items || n_items <= 0 || coh_win_az <= 0 || coh_win_rg <= 0) {
        set_error(TOPS_ERR_ARG, "tops_plan_create: bad arguments");
        return nullptr;
    }
    for (int i = 0; i < n_items; i++) {
        const tops_burst_item& it = items[i];
        const tops_burst_geom& g = it.geom;
        if (!it.slave || !it.kt || !it.tref || !it.fdc || !it.out ||
            g.master_box[1] < g.master_box[0] || g.master_box[3] < g.master_box[2] ||
            g.slave_box[1] < g.slave_box[0] || g.slave_box[3] < g.slave_box[2] ||
            g.slave_box[0] < 0 || g.slave_box[1] >= g.samples_per_burst) {
            set_error(TOPS_ERR_ARG, "tops_plan_create: burst %d is malformed", i);
            return nullptr;
        }
    }
    tops_plan* p = new tops_plan();
    p->ctx = ctx;
    p->items.assign(items, items + n_items);
    p->coh_win_az = coh_win_az;
    p->coh_win_rg = coh_win_rg;
    p->strip_rows = strip_rows > 0 ? strip_rows : 0;
    if (fill_params(p) || fill_work(p)) {
        tops_plan_destroy(p);
        return nullptr;
    }
    return p;
}

void tops_plan_destroy(tops_plan* p)
{
    if (!p) return;
    drop_graph(p);
    if (p->d_params) (void)hipFree(p->d_params);
    if (p->d_work_rs) (void)hipFree(p->d_work_rs);
    if (p->d_coh) (void)hipFree(p->d_coh);
    if (p->d_work_coh) (void)hipFree(p->d_work_coh);
    delete p;
}

int tops_plan_run(tops_plan* p, void* stream)
{
    if (!p) return set_error(TOPS_ERR_ARG, "tops_plan_run: null plan");
    hipStream_t s = (hipStream_t)stream;
    if (p->exec) {
        TOPS_HIP(hipGraphLaunch(p->exec, s));
        return TOPS_OK;
    }
    return enqueue(p, s);
}

int tops_plan_capture_graph(tops_plan* p, void* stream)
{
    if (!p || !stream)
        return set_error(TOPS_ERR_ARG, "tops_plan_capture_graph: needs a plan and a non-default stream");
    drop_graph(p);
    hipStream_t s = (hipStream_t)stream;
    TOPS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(p, s);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    p->graph = g;
    TOPS_HIP(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
    return TOPS_OK;
}

int tops_plan_strip_rows(const tops_plan* p) { return p ? p->strip_rows : 0; }

// Autotuner: times every candidate schedule (per-burst launches, or one
// batched launch at each candidate strip height) on the plan's bursts,
// resample and coherence together (warm-up + 3 timed runs each, hipEvents
// on `stream`), and keeps the fastest.  With TOPS_TUNE_CACHE set, the choice is read from / appended to
// that text file ("<key> <strip_rows>" lines, key = the burst shapes).
int tops_plan_tune(tops_plan* p, void* stream, int* best_rows)
{
    if (!p) return set_error(TOPS_ERR_ARG, "tops_plan_tune: null plan");
    hipStream_t s = (hipStream_t)stream;
    std::string key = "resample";
    for (auto& it : p->items) {
        char b[64];
        snprintf(b, sizeof(b), "_%dx%d", it.geom.master_box[3] - it.geom.master_box[2] + 1,
                 it.geom.master_box[1] - it.geom.master_box[0] + 1);
        key += b;
        if (key.size() > 400) break;
    }
    const char* cache = getenv("TOPS_TUNE_CACHE");
    if (cache) {
        FILE* f = fopen(cache, "r");
        if (f) {
            char k[512];
            int v;
            while (fscanf(f, "%511s %d", k, &v) == 2)
                if (key == k && v > 0) {
                    fclose(f);
                    drop_graph(p);
                    p->strip_rows = v;
                    if (best_rows) *best_rows = v;
                    return fill_work(p);
                }
            fclose(f);
        }
    }
    // per-burst launches (0), fixed batched heights, and the batched heights
    // that fill 1, 2 and 3 rounds of the GPU's walker wave slots
    std::vector<int> cand = {0, 32, 64, 128, 192, 256, 320, 384, 512};
    {
        std::vector<int> ls, ps;
        for (auto& it : p->items) {
            ls.push_back(it.geom.master_box[3] - it.geom.master_box[2] + 1);
            ps.push_back(it.geom.master_box[1] - it.geom.master_box[0] + 1);
        }
        for (int r = 1; r <= 3; r++) {
            int h = std::max(16, resample_rows_for_rounds(ls.data(), ps.data(), (int)ls.size(), r));
            if (std::find(cand.begin(), cand.end(), h) == cand.end()) cand.push_back(h);
        }
    }
    float best = 1e30f;
    int best_r = p->strip_rows;
    hipEvent_t e0, e1;
    TOPS_HIP(hipEventCreate(&e0));
    TOPS_HIP(hipEventCreate(&e1));
    drop_graph(p);
    int rc = TOPS_OK;
    for (int c : cand) {
        p->strip_rows = c;
        if ((rc = fill_work(p))) break;
        if ((rc = enqueue(p, s))) break;   // warm-up
        float tot = 0.f;
        for (int k = 0; k < 3 && !rc; k++) {
            hipError_t e = hipEventRecord(e0, s);
            if (e == hipSuccess) rc = enqueue(p, s);
            if (e == hipSuccess) e = hipEventRecord(e1, s);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) rc = hip_fail(e, "tops_plan_tune timing", __FILE__, __LINE__);
            tot += ms;
        }
        if (rc) break;
        if (tot < best) {
            best = tot;
            best_r = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    p->strip_rows = best_r;
    if (!rc) rc = fill_work(p);
    if (!rc && cache) {
        FILE* f = fopen(cache, "a");
        if (f) {
            fprintf(f, "%s %d\n", key.c_str(), best_r);
            fclose(f);
        }
    }
    if (best_rows) *best_rows = best_r;
    return rc;
}

}  // extern "C"
