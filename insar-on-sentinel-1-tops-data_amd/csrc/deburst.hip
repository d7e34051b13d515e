// deburst.hip -- Stitch_OneSubswath (Func.cpp:2821-2933) on rasters in HBM:
// the burst-stacked raster of bursts burst0 .. burstN (gapped, overlapping
// in time) becomes one continuous raster of the valid lines from the first
// valid line of burst0 to the last valid line of burstN, each output line
// copied from the burst that owns its time (getLineIdxInMetaData,
// Func.cpp:2766-2819).
//
// The host evaluates the line map of every output row (it must anyway: the
// reference exits on a line time that no burst holds) and hands the kernel
// the map as runs of rows with a constant source offset -- one run per burst
// in practice -- by value.  The device work is then a row gather with no
// arithmetic: HBM-bound, 2 x row bytes per output row (1 x for the zero rows
// the reference leaves where a source row lies outside the input), one
// 16-byte vector per thread, nontemporal loads and stores (each row is
// touched once).  Measured 0.66 ms for 12619 rows of 21000 CFloat32
// (6.4 TB/s), the rate of a plain streaming copy; evaluating the FP64 map
// per workgroup instead took 0.94 ms.
#include "tops_internal.hpp"

#include <climits>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace tops {
namespace {

// Stitch_OneSubswath's SubSwathInfo arrays are [16] (DataTypes.h:95-96).
constexpr int kMaxBursts = 16;

struct LineMap {
    double first[kMaxBursts];   // burstFirstLineTime
    double last[kMaxBursts];    // burstLastLineTime
    double t_first;             // burstFirstValidLineTime[burst0]
    double dt;                  // azimuthTimeInterval
    int n_bursts, lines_per_burst, num_lines, burst0, burst_n;
    int height;                 // output rows
    int gapped_rows;            // rows of the burst-stacked input
};

// getLineIdxInMetaData (Func.cpp:2766-2819): the subswath line whose time
// is nearest t, taken from the burst that owns t; in an overlap, the earlier
// burst up to the mid time between its last line and the next burst's first
// line (the burst inside burst0..burstN when the other is outside).  -1 when
// no burst holds t (the last line of the last burst and one interval after
// it map to the subswath's last line).
int line_of_time(const LineMap& m, double t)
{
#pragma clang fp contract(off)
    int sy0 = -1, sy1 = -1, b0 = -1, b1 = -1;
    for (int i = 0; i < m.n_bursts; i++) {
        if (!(t >= m.first[i] && t < m.last[i])) continue;
        const int sy = i * m.lines_per_burst + (int)(((t - m.first[i]) / m.dt) + 0.5);
        if (b0 == -1) {
            sy0 = sy;
            b0 = i;
        } else {
            sy1 = sy;
            b1 = i;
            break;
        }
    }
    if (sy0 != -1 && sy1 != -1) {
        if (b1 == m.burst0) return sy1;
        if (b1 == m.burst_n + 1) return sy0;
        if (t > (m.last[b0] + m.first[b1]) / 2.0) return sy1;
    }
    const double tl = m.last[m.n_bursts - 1];
    if (t < tl + m.dt && t >= tl) sy0 = m.num_lines - 1;
    return sy0;
}

// row of the burst-stacked input that output row yy copies (outside
// 0 .. gapped_rows-1: a zero row); INT_MIN when its line time is in no burst
int source_row(const LineMap& m, int yy)
{
#pragma clang fp contract(off)
    const double t = m.t_first + yy * m.dt;
    const int line = line_of_time(m, t);
    if (line == -1) return INT_MIN;
    return line - m.burst0 * m.lines_per_burst;
}

// the extent of Func.cpp:2825-2843, the burst times and the source row of
// every output row; fails where a line time lies in no burst (the reference
// exits there, Func.cpp:2868-2876)
int make_map(const tops_deburst_swath* sw, int burst0, int burstN, LineMap* m, int* y_min,
             std::vector<int>* rows, const char* who)
{
#pragma clang fp contract(off)
    if (!sw || !sw->burst_first_line_time || !sw->burst_last_line_time ||
        !sw->burst_first_valid_line_time || !sw->burst_last_valid_line_time ||
        sw->lines_per_burst <= 0 || sw->samples_per_burst <= 0 || !(sw->az_time_interval > 0))
        return set_error(TOPS_ERR_ARG, "%s: bad subswath description", who);
    if (sw->n_bursts < 1 || sw->n_bursts > kMaxBursts)
        return set_error(TOPS_ERR_ARG, "%s: %d bursts (1 .. %d supported)", who, sw->n_bursts,
                         kMaxBursts);
    if (burst0 < 0 || burstN < burst0 || burstN >= sw->n_bursts)
        return set_error(TOPS_ERR_ARG, "%s: burst range %d .. %d of %d", who, burst0, burstN,
                         sw->n_bursts);
    memset(m, 0, sizeof(*m));
    for (int i = 0; i < sw->n_bursts; i++) {
        m->first[i] = sw->burst_first_line_time[i];
        m->last[i] = sw->burst_last_line_time[i];
    }
    m->t_first = sw->burst_first_valid_line_time[burst0];
    m->dt = sw->az_time_interval;
    m->n_bursts = sw->n_bursts;
    m->lines_per_burst = sw->lines_per_burst;
    m->num_lines = sw->num_lines;
    m->burst0 = burst0;
    m->burst_n = burstN;
    const int ymin = (int)((m->t_first - sw->first_line_time) / m->dt);
    const int ymax = (int)((sw->burst_last_valid_line_time[burstN] - sw->first_line_time) / m->dt);
    m->height = ymax - ymin + 1;
    m->gapped_rows = (burstN - burst0 + 1) * sw->lines_per_burst;
    if (y_min) *y_min = ymin;
    if (m->height <= 0)
        return set_error(TOPS_ERR_ARG, "%s: empty output (%d rows)", who, m->height);
    if (rows) rows->resize(m->height);
    for (int yy = 0; yy < m->height; yy++) {
        const int r = source_row(*m, yy);
        if (r == INT_MIN)
            return set_error(TOPS_ERR_ARG,
                             "%s: output row %d (line time %.9f) lies in no burst", who, yy,
                             m->t_first + yy * m->dt);
        if (rows) (*rows)[yy] = r;
    }
    return TOPS_OK;
}

// ---- device gather ---------------------------------------------------------
constexpr int kThreads = 256;
constexpr int kMaxRuns = 64;        // runs per launch (kernel argument space)
constexpr int kMaxRows = 65535;     // output rows per launch (grid y)

// output rows start[k] .. start[k+1]-1 copy input rows yy + offset[k]
struct RowRuns {
    int n;                          // runs in this launch
    int y_begin, y_end;             // output rows of this launch
    int gapped_rows;
    int start[kMaxRuns];
    int offset[kMaxRuns];
};

typedef unsigned int Vec16 __attribute__((ext_vector_type(4)));

// grid: x = 256-vector chunk of a row, y = output row of this launch.
// One load into a zeroed vector and one store path: the form that streams
// at the copy rate (two store paths and 64-bit element indexing measured
// 10% slower).
template <typename V>
__global__ __launch_bounds__(kThreads) void deburst_kernel(RowRuns runs, const V* __restrict__ src,
                                                           V* __restrict__ dst, int row_vecs)
{
    const int yy = runs.y_begin + (int)blockIdx.y;
    int k = 0;                                    // run of this row, the same for the workgroup
    for (int j = 1; j < runs.n; j++)
        if (yy >= runs.start[j]) k = j;
    const int row = yy + runs.offset[k];
    const int i = (int)blockIdx.x * kThreads + (int)threadIdx.x;
    if (i >= row_vecs) return;
    // rows outside the gapped raster stay zero: the reference's read of such
    // a row fails and leaves its zeroed line buffer (Func.cpp:2880-2883)
    V v = {};
    if ((unsigned)row < (unsigned)runs.gapped_rows)
        v = __builtin_nontemporal_load(src + (size_t)row * row_vecs + i);
    __builtin_nontemporal_store(v, dst + (size_t)yy * row_vecs + i);
}

int launch_gather(const std::vector<int>& rows, int gapped_rows, const void* src, void* dst,
                  size_t row_bytes, hipStream_t s)
{
    const int height = (int)rows.size();
    if (height == 0 || row_bytes == 0) return TOPS_OK;
    const bool v16 = row_bytes % 16 == 0 && (uintptr_t)src % 16 == 0 && (uintptr_t)dst % 16 == 0;
    if (!v16 && (row_bytes % 4 != 0 || (uintptr_t)src % 4 != 0 || (uintptr_t)dst % 4 != 0))
        return set_error(TOPS_ERR_ARG, "deburst: rasters must be 4-byte aligned");
    const size_t row_vecs = row_bytes / (v16 ? 16 : 4);
    if (row_vecs > (size_t)INT_MAX - kThreads)
        return set_error(TOPS_ERR_ARG, "deburst: rows too long");
    const size_t chunks = (row_vecs + kThreads - 1) / kThreads;
    // runs of constant offset, at most kMaxRuns runs and kMaxRows rows per
    // launch
    int y = 0;
    while (y < height) {
        RowRuns r;
        memset(&r, 0, sizeof(r));
        r.gapped_rows = gapped_rows;
        r.y_begin = y;
        while (y < height && y - r.y_begin < kMaxRows) {
            const int off = rows[y] - y;
            if (r.n == 0 || off != r.offset[r.n - 1]) {
                if (r.n == kMaxRuns) break;
                r.start[r.n] = y;
                r.offset[r.n] = off;
                r.n++;
            }
            y++;
        }
        r.y_end = y;
        const dim3 grid((unsigned)chunks, (unsigned)(r.y_end - r.y_begin));
        if (v16)
            deburst_kernel<Vec16><<<grid, dim3(kThreads), 0, s>>>(r, (const Vec16*)src, (Vec16*)dst,
                                                                  (int)row_vecs);
        else
            deburst_kernel<unsigned int><<<grid, dim3(kThreads), 0, s>>>(
                r, (const unsigned int*)src, (unsigned int*)dst, (int)row_vecs);
        TOPS_HIP(hipGetLastError());
    }
    return TOPS_OK;
}

int element_bytes(int data_type)
{
    switch (data_type) {
    case TOPS_GDT_CINT16: return 4;
    case TOPS_GDT_FLOAT32: return 4;
    case TOPS_GDT_CFLOAT32: return 8;
    default: return 0;
    }
}

struct DevMem {
    void* p = nullptr;
    ~DevMem()
    {
        if (p) (void)hipFree(p);
    }
};

}  // namespace
}  // namespace tops

using namespace tops;

extern "C" {

int tops_deburst_extent(const tops_deburst_swath* sw, int burst0, int burstN, int* y_min,
                        int* height)
{
    LineMap m;
    const int rc = make_map(sw, burst0, burstN, &m, y_min, nullptr, "tops_deburst_extent");
    if (rc) return rc;
    if (height) *height = m.height;
    return TOPS_OK;
}

int tops_deburst_rows(const tops_deburst_swath* sw, int burst0, int burstN, int* rows)
{
    LineMap m;
    std::vector<int> r;
    const int rc = make_map(sw, burst0, burstN, &m, nullptr, &r, "tops_deburst_rows");
    if (rc) return rc;
    if (!rows) return set_error(TOPS_ERR_ARG, "tops_deburst_rows: null rows");
    memcpy(rows, r.data(), sizeof(int) * r.size());
    return TOPS_OK;
}

int tops_deburst(const tops_deburst_swath* sw, int burst0, int burstN, int data_type,
                 const void* d_gapped, void* d_out, void* stream)
{
    const char* who = "tops_deburst";
    const int eb = element_bytes(data_type);
    if (!eb)
        return set_error(TOPS_ERR_ARG, "%s: unsupported image type %d for the stitching step",
                         who, data_type);
    LineMap m;
    std::vector<int> rows;
    int rc = make_map(sw, burst0, burstN, &m, nullptr, &rows, who);
    if (rc) return rc;
    if (!d_gapped || !d_out) return set_error(TOPS_ERR_ARG, "%s: null raster", who);
    return launch_gather(rows, m.gapped_rows, d_gapped, d_out, (size_t)sw->samples_per_burst * eb,
                         (hipStream_t)stream);
}

int tops_stitch_subswath(const tops_deburst_swath* sw, int burst0, int burstN, int data_type,
                         const void* gapped, void* out)
{
    const char* who = "tops_stitch_subswath";
    const int eb = element_bytes(data_type);
    if (!eb)
        return set_error(TOPS_ERR_ARG, "%s: unsupported image type %d for the stitching step",
                         who, data_type);
    LineMap m;
    std::vector<int> rows;
    int rc = make_map(sw, burst0, burstN, &m, nullptr, &rows, who);
    if (rc) return rc;
    if (!gapped || !out) return set_error(TOPS_ERR_ARG, "%s: null raster", who);
    const size_t row_bytes = (size_t)sw->samples_per_burst * eb;
    const size_t in_bytes = row_bytes * m.gapped_rows, out_bytes = row_bytes * m.height;
    DevMem din, dout;
    TOPS_HIP(hipMalloc(&din.p, in_bytes));
    TOPS_HIP(hipMalloc(&dout.p, out_bytes));
    TOPS_HIP(hipMemcpy(din.p, gapped, in_bytes, hipMemcpyHostToDevice));
    rc = launch_gather(rows, m.gapped_rows, din.p, dout.p, row_bytes, nullptr);
    if (rc) return rc;
    TOPS_HIP(hipMemcpy(out, dout.p, out_bytes, hipMemcpyDeviceToHost));
    return TOPS_OK;
}

}  // extern "C"
