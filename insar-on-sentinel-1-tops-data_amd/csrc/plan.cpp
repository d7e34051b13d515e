// plan.cpp -- batched execution of many bursts (all bursts of all subswaths
// of an IW pair, or a stack of slaves against one master): one strip table
// covers every burst, so the whole deramp -> resample -> reramp pass is one
// launch and the coherence pass another; the pair of launches can be
// captured once in a hipGraph and replayed.  The bursts can be split into
// groups, each group's coherence launch running on a second stream beside
// the next group's resampling.  A third schedule fuses both passes in one
// persistent launch.  The schedule, its strip height or row band, the
// coherence walk's rows per block and the group count are tuning parameters
// chosen by timing the candidates on the plan's own bursts.
#include <algorithm>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>

#include "tops_internal.hpp"

using namespace tops;

constexpr int kDefaultCohRows = 32;   // the coherence walk's rows per block before tuning

struct tops_plan {
    tops_ctx* ctx = nullptr;
    std::vector<tops_burst_item> items;
    int coh_win_az = 3, coh_win_rg = 10;
    // schedule: 0 = one launch pair per burst (resample, then coherence) with
    // each burst's automatic heights; > 0 = one batched launch of all bursts'
    // strips of this height, then one batched coherence launch; < 0 = one
    // fused persistent launch (resample and coherence) over row bands of
    // -strip_rows lines (3 az x 10 rg window only)
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
    // the batched coherence launch serves the window; otherwise the batched
    // schedule forms coherence with one launch per burst
    bool coh_batched = true;
    bool coh_walk = false;     // the window is the walking form's (3 az x 10 rg)
    int coh_rows = kDefaultCohRows;   // output rows per block of the batched coherence walk
    // batched schedule in burst groups: group g (items grp_first[g] ..
    // grp_first[g+1]-1) is resampled by one launch on the caller's stream,
    // and its coherence launch on a second stream waits for that launch
    // only, so it runs beside the next group's resampling.  groups = 1 is
    // the plain two-launch schedule.
    int groups = 1;
    std::vector<int> grp_first;              // G + 1 item indices
    std::vector<int> rs_off, coh_off;        // G + 1 offsets into the two work tables
    std::vector<int> coh_item;               // item index of each coherence item
    hipStream_t s2 = nullptr;                // the coherence stream (default priority)
    std::vector<hipEvent_t> ev_grp;          // per group: its resample launch is done
    hipEvent_t ev_join = nullptr;            // the coherence stream's work is done
    // the tuner's record: every candidate it timed (strip_rows, coh_rows,
    // groups, ms per run of the plan)
    struct TuneEntry { int strip_rows, coh_rows, groups, stage; float ms; };
    std::vector<TuneEntry> tune_log;
    // fused schedule: coherence item of each burst (-1: none), the work
    // table and the queue counters of the persistent launch
    bool fused_ok = false;
    CohItem* d_coh_all = nullptr;
    int* d_coh_of = nullptr;
    int4* d_work_fused = nullptr;
    int n_work_fused = 0;
    unsigned* d_queue = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int fill_coh_work(tops_plan* p);

// groups of consecutive bursts with about equal sample counts
void partition(tops_plan* p)
{
    const int n = (int)p->items.size();
    const int G = std::max(1, std::min(p->groups, n));
    p->groups = G;
    std::vector<double> cum(n + 1, 0.0);
    for (int i = 0; i < n; i++) {
        const tops_burst_geom& g = p->items[i].geom;
        cum[i + 1] = cum[i] + (double)(g.master_box[3] - g.master_box[2] + 1) *
                                  (g.master_box[1] - g.master_box[0] + 1);
    }
    p->grp_first.assign(G + 1, n);
    p->grp_first[0] = 0;
    for (int g = 1; g < G; g++) {
        // the first burst at or past g/G of the samples, with at least one
        // burst in this group and one left for each later group
        const double target = cum[n] * g / G;
        int j = p->grp_first[g - 1] + 1;
        while (j < n - (G - g) && cum[j] < target) j++;
        p->grp_first[g] = j;
    }
}

// the second stream and the events of a grouped schedule
int ensure_pipe(tops_plan* p)
{
    if (p->groups <= 1) return TOPS_OK;
    if (!p->s2) TOPS_HIP(hipStreamCreate(&p->s2));
    if (!p->ev_join) TOPS_HIP(hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming));
    while ((int)p->ev_grp.size() < p->groups) {
        hipEvent_t e;
        TOPS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        p->ev_grp.push_back(e);
    }
    return TOPS_OK;
}

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
    p->coh_of.assign(p->items.size(), -1);
    p->coh_batched = coherence_batch_fits((p->coh_win_rg - 1) / 2, (p->coh_win_az - 1) / 2,
                                          p->coh_win_rg, p->coh_win_az);
    p->coh_walk = coherence_walk_fits((p->coh_win_rg - 1) / 2, (p->coh_win_az - 1) / 2,
                                      p->coh_win_rg, p->coh_win_az);
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
        ci.push_back(c);
    }
    p->h_coh = ci;
    p->coh_item.assign(ci.size(), 0);
    for (size_t i = 0; i < p->items.size(); i++)
        if (p->coh_of[i] >= 0) p->coh_item[p->coh_of[i]] = (int)i;
    partition(p);
    p->fused_ok = fused_window_fits((p->coh_win_rg - 1) / 2, (p->coh_win_az - 1) / 2,
                                    p->coh_win_rg, p->coh_win_az);
    // the fused kernel packs a slave column with its LUT row into one int
    // (Ix * 2048 + fx): slave columns below 2^20, like the lines
    for (const tops_burst_item& it : p->items)
        if (it.geom.slave_box[1] >= kMaxSlaveLine) p->fused_ok = false;
    if (p->fused_ok) {
        if (!ci.empty()) {
            TOPS_HIP(hipMalloc(&p->d_coh_all, sizeof(CohItem) * ci.size()));
            TOPS_HIP(hipMemcpy(p->d_coh_all, ci.data(), sizeof(CohItem) * ci.size(),
                               hipMemcpyHostToDevice));
        }
        TOPS_HIP(hipMalloc(&p->d_coh_of, sizeof(int) * p->coh_of.size()));
        TOPS_HIP(hipMemcpy(p->d_coh_of, p->coh_of.data(), sizeof(int) * p->coh_of.size(),
                           hipMemcpyHostToDevice));
        // the queue is cleared once here; every launch leaves it cleared
        TOPS_HIP(hipMalloc(&p->d_queue, 2 * sizeof(unsigned)));
        TOPS_HIP(hipMemset(p->d_queue, 0, 2 * sizeof(unsigned)));
    }
    if (p->coh_batched && !ci.empty()) {
        TOPS_HIP(hipMalloc(&p->d_coh, sizeof(CohItem) * ci.size()));
        TOPS_HIP(hipMemcpy(p->d_coh, ci.data(), sizeof(CohItem) * ci.size(), hipMemcpyHostToDevice));
    }
    return fill_coh_work(p);
}

// work table of the batched coherence launch: {coherence item, column band,
// row block, vec}, blocks of coh_rows rows for the walking form, 16 x 64
// tiles otherwise
int fill_coh_work(tops_plan* p)
{
    if (p->d_work_coh) (void)hipFree(p->d_work_coh);
    p->d_work_coh = nullptr;
    p->n_work_coh = 0;
    p->coh_off.assign(p->groups + 1, 0);
    if (!p->coh_batched) return TOPS_OK;
    std::vector<int4> wc;
    // group by group (coherence items are in burst order)
    size_t k = 0;
    for (int g = 0; g < p->groups; g++) {
        for (; k < p->h_coh.size() && p->coh_item[k] < p->grp_first[g + 1]; k++) {
            const CohItem& c = p->h_coh[k];
            int nx, ny, tag = 0;
            if (p->coh_walk) {
                coherence_walk_count(c.w, c.h, p->coh_rows, &nx, &ny);
                tag = coherence_walk_vec(c.m, c.s, c.coh, c.w);
            } else {
                coherence_tile_count(c.w, c.h, &nx, &ny);
            }
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) wc.push_back(make_int4((int)k, x, y, tag));
        }
        p->coh_off[g + 1] = (int)wc.size();
    }
    p->n_work_coh = (int)wc.size();
    if (!wc.empty()) {
        TOPS_HIP(hipMalloc(&p->d_work_coh, sizeof(int4) * wc.size()));
        TOPS_HIP(hipMemcpy(p->d_work_coh, wc.data(), sizeof(int4) * wc.size(), hipMemcpyHostToDevice));
    }
    return TOPS_OK;
}

// strip table for the current strip height; within each burst group the
// strips of its bursts are interleaved row-band by row-band so uneven bursts
// spread over the GPU
int fill_work(tops_plan* p)
{
    if (p->strip_rows < 0) {
        // fused: row band by row band, burst by burst, strip by strip
        const int rows = -p->strip_rows;
        std::vector<int4> w;
        int max_lines = 0;
        for (auto& it : p->items)
            max_lines = std::max(max_lines, it.geom.master_box[3] - it.geom.master_box[2] + 1);
        for (int y0 = 0; y0 < max_lines; y0 += rows)
            for (size_t i = 0; i < p->items.size(); i++) {
                const tops_burst_geom& g = p->items[i].geom;
                const int lines = g.master_box[3] - g.master_box[2] + 1;
                if (y0 >= lines) continue;
                const int nx = fused_strip_count(g.master_box[1] - g.master_box[0] + 1);
                for (int x = 0; x < nx; x++)
                    w.push_back(make_int4((int)i, x, y0, std::min(rows, lines - y0)));
            }
        if (p->d_work_fused) (void)hipFree(p->d_work_fused);
        p->d_work_fused = nullptr;
        p->n_work_fused = (int)w.size();
        if (!w.empty()) {
            TOPS_HIP(hipMalloc(&p->d_work_fused, sizeof(int4) * w.size()));
            TOPS_HIP(hipMemcpy(p->d_work_fused, w.data(), sizeof(int4) * w.size(),
                               hipMemcpyHostToDevice));
        }
        return TOPS_OK;
    }
    if (p->strip_rows == 0) return TOPS_OK;   // per-burst launches need no table
    std::vector<int4> w;
    std::vector<int> nxs(p->items.size()), nys(p->items.size());
    for (size_t i = 0; i < p->items.size(); i++) {
        const tops_burst_geom& g = p->items[i].geom;
        resample_strip_count(g.master_box[3] - g.master_box[2] + 1,
                             g.master_box[1] - g.master_box[0] + 1, p->strip_rows, &nxs[i], &nys[i]);
    }
    p->rs_off.assign(p->groups + 1, 0);
    for (int gi = 0; gi < p->groups; gi++) {
        const int i0 = p->grp_first[gi], i1 = p->grp_first[gi + 1];
        int max_ny = 0;
        for (int i = i0; i < i1; i++) max_ny = std::max(max_ny, nys[i]);
        for (int y = 0; y < max_ny; y++)
            for (int i = i0; i < i1; i++)
                if (y < nys[i])
                    for (int x = 0; x < nxs[i]; x++) w.push_back(make_int4(i, x, y, p->strip_rows));
        p->rs_off[gi + 1] = (int)w.size();
    }
    if (p->d_work_rs) (void)hipFree(p->d_work_rs);
    p->d_work_rs = nullptr;
    p->n_work_rs = (int)w.size();
    if (!w.empty()) {
        TOPS_HIP(hipMalloc(&p->d_work_rs, sizeof(int4) * w.size()));
        TOPS_HIP(hipMemcpy(p->d_work_rs, w.data(), sizeof(int4) * w.size(), hipMemcpyHostToDevice));
    }
    return TOPS_OK;
}

// Per-launch timing of a run (tops_plan_time_launches): an event pair
// around every launch, tagged with the launch's kind.
struct Marks {
    std::vector<hipEvent_t> ev;   // two per launch
    std::vector<int> kind;
    size_t n = 0;                 // launches marked in this run
    ~Marks()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    int before(hipStream_t s, int k)
    {
        if (2 * n + 2 > ev.size()) {
            for (int j = 0; j < 2; j++) {
                hipEvent_t e;
                TOPS_HIP(hipEventCreate(&e));
                ev.push_back(e);
            }
            kind.push_back(k);
        }
        kind[n] = k;
        TOPS_HIP(hipEventRecord(ev[2 * n], s));
        return TOPS_OK;
    }
    int after(hipStream_t s)
    {
        TOPS_HIP(hipEventRecord(ev[2 * n + 1], s));
        n++;
        return TOPS_OK;
    }
};

enum { kKindResample = 0, kKindCoherence = 1, kKindFused = 2, kKinds = 3 };

// one launch, with its event pair when the run is timed
template <class F>
int marked(Marks* mk, hipStream_t s, int kind, F&& launch)
{
    int rc = mk ? mk->before(s, kind) : TOPS_OK;
    if (!rc) rc = launch();
    if (!rc && mk) rc = mk->after(s);
    return rc;
}

int enqueue(tops_plan* p, hipStream_t s, Marks* mk = nullptr)
{
    const int dx = (p->coh_win_rg - 1) / 2, dy = (p->coh_win_az - 1) / 2;
    auto coh_one = [&](const CohItem& c) {
        return marked(mk, s, kKindCoherence, [&] {
            return launch_coherence(c.m, c.s, c.w, c.h, dx, dy, p->coh_win_rg, p->coh_win_az,
                                    c.coh, TOPS_MODE_FAST, s);
        });
    };
    int rc = TOPS_OK;
    if (p->strip_rows < 0)
        return marked(mk, s, kKindFused, [&] {
            return launch_fused_batch(p->d_params, p->d_coh_all, p->d_coh_of, p->d_work_fused,
                                      p->n_work_fused, p->d_queue, s);
        });
    if (p->strip_rows == 0) {
        for (size_t i = 0; i < p->h_params.size() && !rc; i++) {
            rc = marked(mk, s, kKindResample,
                        [&] { return launch_resample_fast(p->h_params[i], s, 0); });
            if (!rc && p->coh_of[i] >= 0) rc = coh_one(p->h_coh[p->coh_of[i]]);
        }
        return rc;
    }
    if (p->groups > 1 && p->coh_batched) {
        // group g's coherence (second stream) waits for group g's resample
        // launch only and runs beside the later groups' resampling
        for (int g = 0; g < p->groups && !rc; g++) {
            const int r0 = p->rs_off[g], nr = p->rs_off[g + 1] - r0;
            const int c0 = p->coh_off[g], nc = p->coh_off[g + 1] - c0;
            rc = marked(mk, s, kKindResample, [&] {
                return launch_resample_batch(p->d_params, p->d_work_rs + r0, nr, s);
            });
            if (rc || nc == 0) continue;
            TOPS_HIP(hipEventRecord(p->ev_grp[g], s));
            TOPS_HIP(hipStreamWaitEvent(p->s2, p->ev_grp[g], 0));
            rc = marked(mk, p->s2, kKindCoherence, [&] {
                return launch_coherence_batch(p->d_coh, p->d_work_coh + c0, nc, dx, dy,
                                              p->coh_win_rg, p->coh_win_az, p->coh_rows, p->s2);
            });
        }
        // the caller's stream waits for the coherence stream (also when a
        // launch failed: the fork is always joined)
        TOPS_HIP(hipEventRecord(p->ev_join, p->s2));
        TOPS_HIP(hipStreamWaitEvent(s, p->ev_join, 0));
        return rc;
    }
    rc = marked(mk, s, kKindResample,
                [&] { return launch_resample_batch(p->d_params, p->d_work_rs, p->n_work_rs, s); });
    if (rc) return rc;
    if (p->coh_batched)
        return marked(mk, s, kKindCoherence, [&] {
            return launch_coherence_batch(p->d_coh, p->d_work_coh, p->n_work_coh, dx, dy,
                                          p->coh_win_rg, p->coh_win_az, p->coh_rows, s);
        });
    for (const CohItem& c : p->h_coh)
        if ((rc = coh_one(c))) return rc;
    return TOPS_OK;
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
        if (!it.slave || !it.kt || !it.tref || !it.fdc || !it.out) {
            set_error(TOPS_ERR_ARG, "tops_plan_create: burst %d has a null buffer", i);
            return nullptr;
        }
        if (check_burst_geom(&g, "tops_plan_create")) return nullptr;
    }
    tops_plan* p = new tops_plan();
    p->ctx = ctx;
    p->items.assign(items, items + n_items);
    p->coh_win_az = coh_win_az;
    p->coh_win_rg = coh_win_rg;
    p->strip_rows = strip_rows;
    if (fill_params(p)) {
        tops_plan_destroy(p);
        return nullptr;
    }
    // the fused schedule serves the 3 az x 10 rg window; other windows take
    // the batched schedule at the same height
    if (p->strip_rows < 0 && !p->fused_ok) p->strip_rows = -p->strip_rows;
    if (fill_work(p)) {
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
    if (p->d_coh_all) (void)hipFree(p->d_coh_all);
    if (p->d_coh_of) (void)hipFree(p->d_coh_of);
    if (p->d_work_fused) (void)hipFree(p->d_work_fused);
    if (p->d_queue) (void)hipFree(p->d_queue);
    for (hipEvent_t e : p->ev_grp) (void)hipEventDestroy(e);
    if (p->ev_join) (void)hipEventDestroy(p->ev_join);
    if (p->s2) (void)hipStreamDestroy(p->s2);
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

int tops_plan_set_groups(tops_plan* p, int groups)
{
    if (!p || groups <= 0) return set_error(TOPS_ERR_ARG, "tops_plan_set_groups: bad arguments");
    drop_graph(p);
    p->groups = groups;
    partition(p);
    int rc = ensure_pipe(p);
    if (!rc) rc = fill_work(p);
    if (!rc) rc = fill_coh_work(p);
    return rc;
}

int tops_plan_groups(const tops_plan* p) { return p ? p->groups : 0; }

// Launch durations of the plan's schedule: `reps` eager runs on `stream`,
// an event pair around every launch; ms[k] = the mean per run of the
// summed durations of the launches of kind k (0 resample, 1 coherence,
// 2 fused resample + coherence), and launches[k] their count per run.
int tops_plan_time_launches(tops_plan* p, void* stream, int reps, float* ms, int* launches)
{
    if (!p || reps <= 0 || !ms) return set_error(TOPS_ERR_ARG, "tops_plan_time_launches: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    double tot[kKinds] = {0.0, 0.0, 0.0};
    int cnt[kKinds] = {0, 0, 0};
    Marks mk;
    for (int r = 0; r < reps; r++) {
        mk.n = 0;
        if (int rc = enqueue(p, s, &mk)) return rc;
        TOPS_HIP(hipStreamSynchronize(s));
        for (size_t i = 0; i < mk.n; i++) {
            float t = 0.f;
            TOPS_HIP(hipEventElapsedTime(&t, mk.ev[2 * i], mk.ev[2 * i + 1]));
            tot[mk.kind[i]] += t;
            if (r == 0) cnt[mk.kind[i]]++;
        }
    }
    for (int k = 0; k < kKinds; k++) {
        ms[k] = (float)(tot[k] / reps);
        if (launches) launches[k] = cnt[k];
    }
    return TOPS_OK;
}

// Kernel behind launches of kind `kind` on the plan's current schedule
// (the name rocprofv3 lists), or "" when the schedule has none of that kind.
const char* tops_plan_kernel_name(const tops_plan* p, int kind)
{
    if (!p || kind < 0 || kind >= kKinds) return "";
    const bool walk = coherence_walk_fits((p->coh_win_rg - 1) / 2, (p->coh_win_az - 1) / 2,
                                          p->coh_win_rg, p->coh_win_az);
    if (p->strip_rows < 0) return kind == kKindFused ? "burst_fused_kernel" : "";
    if (kind == kKindFused) return "";
    if (p->strip_rows <= 0 || !p->coh_batched) {
        if (kind == kKindResample) return p->strip_rows <= 0 ? "resample_fast_kernel"
                                                             : "resample_batch_kernel";
        if (!coherence_batch_fits((p->coh_win_rg - 1) / 2, (p->coh_win_az - 1) / 2,
                                  p->coh_win_rg, p->coh_win_az))
            return "coherence_fp64_kernel";
        return walk ? "coherence_walk_kernel" : "coherence_fast_kernel";
    }
    if (kind == kKindResample) return "resample_batch_kernel";
    return walk ? "coherence_walk_batch_kernel" : "coherence_batch_kernel";
}

}  // extern "C"

// Autotuner, in three stages on the plan's own bursts (warm-up + 3 timed
// runs per candidate, hipEvents on `stream`, resample and coherence
// together):
//   1. the schedule: per-burst launches, one batched launch at each
//      candidate strip height, or the fused launch at each candidate row
//      band (3 az x 10 rg window);
//   2. for a batched schedule with the walking coherence (3 az x 10 rg), the
//      coherence walk's rows per block;
//   3. for a batched schedule, the number of burst groups whose coherence
//      runs on the second stream beside the next group's resampling (one
//      group included), timed as replayed graphs.
// It keeps the fastest of each stage and records every candidate timed
// (tops_plan_tune_entry).  With TOPS_TUNE_CACHE set, the choice is read from
// / appended to that text file ("<key> <strip_rows> <coh_rows> <groups>"
// lines; the key names the burst count and coherence window and hashes the
// burst shapes).
namespace {

// mean ms per run of the plan's current schedule over 3 runs after a warm-up
int time_runs(tops_plan* p, hipStream_t s, hipEvent_t e0, hipEvent_t e1, float* out)
{
    int rc = enqueue(p, s);   // warm-up
    if (rc) return rc;
    float tot = 0.f;
    for (int k = 0; k < 3; k++) {
        hipError_t e = hipEventRecord(e0, s);
        if (e == hipSuccess) rc = enqueue(p, s);
        if (rc) return rc;
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) return hip_fail(e, "tops_plan_tune timing", __FILE__, __LINE__);
        tot += ms;
    }
    *out = tot / 3.f;
    return TOPS_OK;
}

// mean ms per replay of a graph of the plan's current schedule, captured on
// s (not the null stream), over 3 replays after a warm-up
int time_graph(tops_plan* p, hipStream_t s, hipEvent_t e0, hipEvent_t e1, float* out)
{
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    TOPS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(p, s);
    hipError_t e = hipStreamEndCapture(s, &g);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "tops_plan_tune capture", __FILE__, __LINE__);
    if (!rc && (e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0)) != hipSuccess)
        rc = hip_fail(e, "tops_plan_tune instantiate", __FILE__, __LINE__);
    float tot = 0.f;
    for (int k = -1; k < 3 && !rc; k++) {   // k = -1: warm-up
        e = hipEventRecord(e0, s);
        if (e == hipSuccess) e = hipGraphLaunch(x, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) rc = hip_fail(e, "tops_plan_tune graph timing", __FILE__, __LINE__);
        if (k >= 0) tot += ms;
    }
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
    if (!rc) *out = tot / 3.f;
    return rc;
}

}  // namespace

extern "C" {

int tops_plan_tune(tops_plan* p, void* stream, int* best_rows)
{
    if (!p) return set_error(TOPS_ERR_ARG, "tops_plan_tune: null plan");
    hipStream_t s = (hipStream_t)stream;
    // key: the burst count, the coherence window and a 64-bit FNV-1a hash of
    // every burst's output shape, in order
    uint64_t hsh = 1469598103934665603ull;
    auto mix = [&](int v) {
        for (int k = 0; k < 4; k++) {
            hsh ^= (uint64_t)((v >> (8 * k)) & 0xff);
            hsh *= 1099511628211ull;
        }
    };
    for (auto& it : p->items) {
        mix(it.geom.master_box[3] - it.geom.master_box[2] + 1);
        mix(it.geom.master_box[1] - it.geom.master_box[0] + 1);
    }
    char kb[96];
    snprintf(kb, sizeof(kb), "plan_n%zu_w%dx%d_%016llx", p->items.size(), p->coh_win_az,
             p->coh_win_rg, (unsigned long long)hsh);
    const std::string key = kb;
    p->tune_log.clear();
    const char* cache = getenv("TOPS_TUNE_CACHE");
    if (cache) {
        FILE* f = fopen(cache, "r");
        if (f) {
            char line[640], k[512];
            int v, cr, ng;
            while (fgets(line, sizeof(line), f)) {
                if (sscanf(line, "%511s %d %d %d", k, &v, &cr, &ng) != 4) continue;
                if (key == k && (v >= 0 || p->fused_ok) && cr > 0 && ng > 0) {
                    fclose(f);
                    drop_graph(p);
                    p->strip_rows = v;
                    p->coh_rows = cr;
                    p->groups = ng;
                    partition(p);
                    if (best_rows) *best_rows = v;
                    int rc = ensure_pipe(p);
                    if (!rc) rc = fill_work(p);
                    return rc ? rc : fill_coh_work(p);
                }
            }
            fclose(f);
        }
    }
    // per-burst launches (0), fixed batched heights, the batched heights
    // that fill 1, 2, 3, 5 and 8 rounds of the GPU's walker wave slots, and
    // the fused launch's row bands (negative)
    std::vector<int> cand = {0, 32, 64, 128, 192, 256, 320, 384, 512};
    if (p->fused_ok)
        for (int r : {64, 128, 192, 256, 384, 512, 768}) cand.push_back(-r);
    {
        std::vector<int> ls, ps;
        for (auto& it : p->items) {
            ls.push_back(it.geom.master_box[3] - it.geom.master_box[2] + 1);
            ps.push_back(it.geom.master_box[1] - it.geom.master_box[0] + 1);
        }
        for (int r : {1, 2, 3, 5, 8}) {
            int h = std::max(16, resample_rows_for_rounds(ls.data(), ps.data(), (int)ls.size(), r));
            if (std::find(cand.begin(), cand.end(), h) == cand.end()) cand.push_back(h);
        }
    }
    p->coh_rows = kDefaultCohRows;
    p->groups = 1;
    partition(p);
    float best = 1e30f;
    int best_r = p->strip_rows;
    hipEvent_t e0, e1;
    TOPS_HIP(hipEventCreate(&e0));
    TOPS_HIP(hipEventCreate(&e1));
    drop_graph(p);
    int rc = TOPS_OK;
    for (int c : cand) {
        p->strip_rows = c;
        float ms = 0.f;
        if ((rc = fill_work(p)) || (rc = time_runs(p, s, e0, e1, &ms))) break;
        p->tune_log.push_back({c, p->coh_rows, p->groups, 1, ms});
        if (ms < best) {
            best = ms;
            best_r = c;
        }
    }
    p->strip_rows = best_r;
    if (!rc) rc = fill_work(p);
    // stage 2: the batched coherence walk's rows per block
    if (!rc && best_r > 0 && p->coh_batched && p->coh_walk) {
        int best_c = p->coh_rows;
        for (int cr : {16, 64, 128, 256}) {
            p->coh_rows = cr;
            float ms = 0.f;
            if ((rc = fill_coh_work(p)) || (rc = time_runs(p, s, e0, e1, &ms))) break;
            p->tune_log.push_back({best_r, cr, p->groups, 2, ms});
            if (ms < best) {
                best = ms;
                best_c = cr;
            }
        }
        p->coh_rows = best_c;
        if (!rc) rc = fill_coh_work(p);
    }
    // stage 3: burst groups pipelined over two streams (batched schedule),
    // one group included, timed as replayed graphs when the caller's stream
    // can capture one (the overlap of the two streams is not the same in a
    // graph as in eager launches: 9 groups of the IW pair measured 11.7 ms
    // eager and 12.5 ms replayed, so graphs decide, and the search stops at
    // 6 groups)
    if (!rc && best_r > 0 && p->coh_batched && p->items.size() > 1) {
        int best_g = 1;
        float best3 = 1e30f;
        std::vector<int> gs;
        for (int g : {1, 2, 3, 4, 6})
            if (g <= (int)p->items.size() && std::find(gs.begin(), gs.end(), g) == gs.end())
                gs.push_back(g);
        for (int g : gs) {
            p->groups = g;
            partition(p);
            float ms = 0.f;
            if ((rc = ensure_pipe(p)) || (rc = fill_work(p)) || (rc = fill_coh_work(p)) ||
                (rc = s ? time_graph(p, s, e0, e1, &ms) : time_runs(p, s, e0, e1, &ms)))
                break;
            p->tune_log.push_back({best_r, p->coh_rows, g, 3, ms});
            if (ms < best3) {
                best3 = ms;
                best_g = g;
            }
        }
        p->groups = best_g;
        partition(p);
        if (!rc) rc = ensure_pipe(p);
        if (!rc) rc = fill_work(p);
        if (!rc) rc = fill_coh_work(p);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!rc && cache) {
        FILE* f = fopen(cache, "a");
        if (f) {
            fprintf(f, "%s %d %d %d\n", key.c_str(), p->strip_rows, p->coh_rows, p->groups);
            fclose(f);
        }
    }
    if (best_rows) *best_rows = p->strip_rows;
    return rc;
}

int tops_plan_coh_rows(const tops_plan* p) { return p ? p->coh_rows : 0; }

int tops_plan_tune_count(const tops_plan* p) { return p ? (int)p->tune_log.size() : 0; }

int tops_plan_tune_entry(const tops_plan* p, int i, int* strip_rows, int* coh_rows, int* groups,
                         int* stage, float* ms)
{
    if (!p || i < 0 || i >= (int)p->tune_log.size())
        return set_error(TOPS_ERR_ARG, "tops_plan_tune_entry: no entry %d", i);
    if (strip_rows) *strip_rows = p->tune_log[i].strip_rows;
    if (coh_rows) *coh_rows = p->tune_log[i].coh_rows;
    if (groups) *groups = p->tune_log[i].groups;
    if (stage) *stage = p->tune_log[i].stage;
    if (ms) *ms = p->tune_log[i].ms;
    return TOPS_OK;
}

}  // extern "C"
