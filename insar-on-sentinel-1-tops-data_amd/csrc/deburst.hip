// deburst.hip -- Stitch_OneSubswath (Func.cpp:2821-2933) on rasters in HBM:
// the burst-stacked raster of bursts burst0 .. burstN (gapped, overlapping
// in time) becomes one continuous raster of the valid lines from the first
// valid line of burst0 to the last valid line of burstN, each output line
// copied from the burst that owns its time (getLineIdxInMetaData,
// Func.cpp:2766-2819).
//
// The work is a row gather: HBM-bound, 2 x row bytes per output row (1 x for
// the zero rows the reference leaves where a source row lies outside the
// input).  One launch covers the whole output; each workgroup moves one
// 16 KiB chunk of one output row and evaluates that row's source line
// itself from the burst times it receives by value, so the launch needs no
// table in device memory and can be captured in a hipGraph.
#include "deburst.hpp"

#include <climits>
#include <stdint.h>
#include <string.h>

namespace tops {
namespace {

constexpr int kDbThreads = 256;
constexpr int kDbUnroll = 4;        // vectors in flight per thread

typedef unsigned int DbVec16 __attribute__((ext_vector_type(4)));

template <typename V>
__global__ __launch_bounds__(kDbThreads) void deburst_kernel(DeburstMap m, const V* src, V* dst,
                                                             long row_vecs, int chunks)
{
    const int yy = (int)(blockIdx.x / (unsigned)chunks);
    const int chunk = (int)(blockIdx.x - (unsigned)yy * (unsigned)chunks);
    const int row = deburst_source_row(m, yy);          // the same for the whole workgroup
    V* out = dst + (size_t)yy * row_vecs;
    const long v0 = (long)chunk * (kDbThreads * kDbUnroll) + threadIdx.x;
    if (row >= 0 && row < m.gapped_rows) {
        const V* in = src + (size_t)row * row_vecs;
        V v[kDbUnroll];
#pragma unroll
        for (int k = 0; k < kDbUnroll; k++) {
            const long i = v0 + (long)k * kDbThreads;
            if (i < row_vecs) v[k] = in[i];
        }
#pragma unroll
        for (int k = 0; k < kDbUnroll; k++) {
            const long i = v0 + (long)k * kDbThreads;
            if (i < row_vecs) out[i] = v[k];
        }
    } else {
        // the reference's read of a row outside the gapped raster fails and
        // leaves the zeroed line buffer (Func.cpp:2880-2883)
        const V z = {};
#pragma unroll
        for (int k = 0; k < kDbUnroll; k++) {
            const long i = v0 + (long)k * kDbThreads;
            if (i < row_vecs) out[i] = z;
        }
    }
}

}  // namespace

int launch_deburst(const DeburstMap& m, const void* src, void* dst, size_t row_bytes,
                   hipStream_t s)
{
    if (m.height <= 0 || row_bytes == 0) return TOPS_OK;
    const bool v16 = row_bytes % 16 == 0 && (uintptr_t)src % 16 == 0 && (uintptr_t)dst % 16 == 0;
    if (!v16 && (row_bytes % 4 != 0 || (uintptr_t)src % 4 != 0 || (uintptr_t)dst % 4 != 0))
        return set_error(TOPS_ERR_ARG, "deburst: rasters must be 4-byte aligned");
    const size_t vec = v16 ? 16 : 4;
    const size_t row_vecs = row_bytes / vec;
    const size_t per = (size_t)kDbThreads * kDbUnroll;
    const size_t chunks = (row_vecs + per - 1) / per;
    const size_t blocks = chunks * (size_t)m.height;
    if (blocks > (size_t)INT_MAX) return set_error(TOPS_ERR_ARG, "deburst: output too large");
    if (v16)
        deburst_kernel<DbVec16><<<dim3((unsigned)blocks), dim3(kDbThreads), 0, s>>>(
            m, (const DbVec16*)src, (DbVec16*)dst, (long)row_vecs, (int)chunks);
    else
        deburst_kernel<unsigned int><<<dim3((unsigned)blocks), dim3(kDbThreads), 0, s>>>(
            m, (const unsigned int*)src, (unsigned int*)dst, (long)row_vecs, (int)chunks);
    TOPS_HIP(hipGetLastError());
    return TOPS_OK;
}

namespace {

int element_bytes(int data_type)
{
    switch (data_type) {
    case TOPS_GDT_CINT16: return 4;
    case TOPS_GDT_FLOAT32: return 4;
    case TOPS_GDT_CFLOAT32: return 8;
    default: return 0;
    }
}

// the extent of Func.cpp:2825-2843 and the burst times; every output row
// checked to have a source line (the reference exits where one has none,
// Func.cpp:2868-2876)
int make_map(const tops_deburst_swath* sw, int burst0, int burstN, DeburstMap* m, int* y_min,
             const char* who)
{
#pragma clang fp contract(off)
    if (!sw || !sw->burst_first_line_time || !sw->burst_last_line_time ||
        !sw->burst_first_valid_line_time || !sw->burst_last_valid_line_time ||
        sw->lines_per_burst <= 0 || sw->samples_per_burst <= 0 || !(sw->az_time_interval > 0))
        return set_error(TOPS_ERR_ARG, "%s: bad subswath description", who);
    if (sw->n_bursts < 1 || sw->n_bursts > kDeburstMaxBursts)
        return set_error(TOPS_ERR_ARG, "%s: %d bursts (1 .. %d supported)", who, sw->n_bursts,
                         kDeburstMaxBursts);
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
    for (int yy = 0; yy < m->height; yy++)
        if (deburst_source_row(*m, yy) == INT_MIN)
            return set_error(TOPS_ERR_ARG,
                             "%s: output row %d (line time %.9f) lies in no burst", who, yy,
                             m->t_first + yy * m->dt);
    return TOPS_OK;
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
    DeburstMap m;
    const int rc = make_map(sw, burst0, burstN, &m, y_min, "tops_deburst_extent");
    if (rc) return rc;
    if (height) *height = m.height;
    return TOPS_OK;
}

int tops_deburst_rows(const tops_deburst_swath* sw, int burst0, int burstN, int* rows)
{
    DeburstMap m;
    const int rc = make_map(sw, burst0, burstN, &m, nullptr, "tops_deburst_rows");
    if (rc) return rc;
    if (!rows) return set_error(TOPS_ERR_ARG, "tops_deburst_rows: null rows");
    for (int yy = 0; yy < m.height; yy++) rows[yy] = deburst_source_row(m, yy);
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
    DeburstMap m;
    int rc = make_map(sw, burst0, burstN, &m, nullptr, who);
    if (rc) return rc;
    if (!d_gapped || !d_out) return set_error(TOPS_ERR_ARG, "%s: null raster", who);
    return launch_deburst(m, d_gapped, d_out, (size_t)sw->samples_per_burst * eb,
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
    DeburstMap m;
    int rc = make_map(sw, burst0, burstN, &m, nullptr, who);
    if (rc) return rc;
    if (!gapped || !out) return set_error(TOPS_ERR_ARG, "%s: null raster", who);
    const size_t row_bytes = (size_t)sw->samples_per_burst * eb;
    const size_t in_bytes = row_bytes * m.gapped_rows, out_bytes = row_bytes * m.height;
    DevMem din, dout;
    TOPS_HIP(hipMalloc(&din.p, in_bytes));
    TOPS_HIP(hipMalloc(&dout.p, out_bytes));
    TOPS_HIP(hipMemcpy(din.p, gapped, in_bytes, hipMemcpyHostToDevice));
    rc = launch_deburst(m, din.p, dout.p, row_bytes, nullptr);
    if (rc) return rc;
    TOPS_HIP(hipMemcpy(out, dout.p, out_bytes, hipMemcpyDeviceToHost));
    return TOPS_OK;
}

}  // extern "C"
