// deburst.hpp -- the line map of the deburst (Stitch_OneSubswath,
// Func.cpp:2821-2933), shared by the host validation and the device copy so
// both evaluate the same FP64 expressions.
#pragma once

#include "tops_internal.hpp"

namespace tops {

// Stitch_OneSubswath's SubSwathInfo arrays are [16] (DataTypes.h:95-96);
// the kernel receives the burst times by value.
constexpr int kDeburstMaxBursts = 16;

struct DeburstMap {
    double first[kDeburstMaxBursts];   // burstFirstLineTime
    double last[kDeburstMaxBursts];    // burstLastLineTime
    double t_first;                    // burstFirstValidLineTime[burst0]
    double dt;                         // azimuthTimeInterval
    int n_bursts, lines_per_burst, num_lines, burst0, burst_n;
    int height;                        // output rows
    int gapped_rows;                   // rows of the burst-stacked input
};

// getLineIdxInMetaData (Func.cpp:2766-2819): the subswath line whose time
// is nearest t, taken from the burst that owns t; in an overlap, the earlier
// burst up to the mid time between its last line and the next burst's first
// line (the burst inside burst0..burstN when the other is outside).  -1 when
// no burst holds t (the last line of the last burst and one interval after
// it map to the subswath's last line).
__host__ __device__ inline int deburst_line(const DeburstMap& m, double t)
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

// row of the burst-stacked input copied to output row yy, before the range
// check (outside 0 .. gapped_rows-1 the output row is zero); INT_MIN when
// the line time falls in no burst
__host__ __device__ inline int deburst_source_row(const DeburstMap& m, int yy)
{
#pragma clang fp contract(off)
    const double t = m.t_first + yy * m.dt;
    const int line = deburst_line(m, t);
    if (line == -1) return (int)0x80000000;
    return line - m.burst0 * m.lines_per_burst;
}

int launch_deburst(const DeburstMap& m, const void* src, void* dst, size_t row_bytes,
                   hipStream_t s);

}  // namespace tops
