// burst_match.cpp -- ComputerBurstOverlap (Func.cpp:1610-1746): which master
// and slave bursts image the same ground, for pairs whose burst grids are
// offset (BurstGap != 0).  The centre pixel of the first requested master
// burst is put on the ground with the range-Doppler + ellipsoid Newton of
// lp2xyz (Func.cpp:1492-1584; start value from the annotation's
// geolocation grid, Locating / getLatitude, 1133-1246), its zero-Doppler time
// on the slave orbit (xyz2aztime, 1349-1377) picks the slave burst; when the
// slave starts later, the roles swap.  Host code: one call per pair.
//
// Where the reference's code cannot have meant what it does, this follows
// the evident intent (none of it changes a converged position):
//   * the grid interpolation weights stay fractional (getLatitude truncates
//     them to int, 1148-1149) -- they only seed the Newton solve;
//   * the start slant-range time is the swath's mid time in both branches
//     (the first branch averages the first pixel with itself, 1638-1639);
//   * the orbit polynomial's centre time stays a double (xyz2aztime takes
//     it as int, 1351);
//   * the burst search stops at the last burst, and a zero-Doppler time past
//     the other acquisition's end is "no overlap" (both fall off the
//     reference's arrays / function end).
#include "tops_internal.hpp"

#include <math.h>
#include <string.h>

namespace tops {
namespace {

double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// getxyz / getxyzdot / getAcceleration (Func.cpp:1262-1345, 1409-1453):
// the fitted polynomial in (t - t[num_points / 2]) / 10
void orbit_state(const tops_orbit& o, double t, double* p, double* v, double* a)
{
    const double tn = (t - o.az_time[o.num_points / 2]) / 10.0;
    const double* c[3] = {o.coef_x, o.coef_y, o.coef_z};
    for (int k = 0; k < 3; k++) {
        if (p) {
            double s = 0.0;
            for (int i = o.ncoef - 1; i >= 0; --i) {
                s *= tn;
                s += c[k][i];
            }
            p[k] = s;
        }
        if (v) {
            double s = c[k][1];
            for (int i = 2; i < o.ncoef; ++i) s += c[k][i] * (double(i) * pow(tn, i - 1));
            v[k] = s / 10.0;
        }
        if (a) {
            double s = 0.0;
            for (int i = 2; i < o.ncoef; ++i) s += c[k][i] * (double((i - 1) * i) * pow(tn, i - 2));
            a[k] = s / 100.0;
        }
    }
}

// Locating + bilinear interpolation over the geolocation grid
void grid_lat_lon(const tops_annotation& s, double az, double rg, double* lat, double* lon)
{
    const int P = s.geo_points_per_line, G = s.geo_lines;
    const double* srt = s.geo_slant_range_time;
    int j0 = P - 2, j1 = P - 1;
    if (rg < srt[0]) {
        j0 = 0;
        j1 = 1;
    } else if (!(rg > srt[P - 1])) {
        for (int j = 0; j < P - 1; j++)
            if (srt[j] <= rg && srt[j + 1] > rg) {
                j0 = j;
                j1 = j + 1;
                break;
            }
    }
    const double mx = (rg - srt[j0]) / (srt[j1] - srt[j0]);
    int i0 = 0;
    double my = 0.0;
    for (int i = 0; i < G - 1; i++) {
        const double a0 = (1 - mx) * s.geo_az_time[i * P + j0] + mx * s.geo_az_time[i * P + j1];
        const double a1 =
            (1 - mx) * s.geo_az_time[(i + 1) * P + j0] + mx * s.geo_az_time[(i + 1) * P + j1];
        if ((i == 0 && az < a0) || (i == G - 2 && az >= a1) || (a0 <= az && a1 > az)) {
            i0 = i;
            my = (az - a0) / (a1 - a0);
            break;
        }
    }
    const int i1 = i0 + 1;
    auto interp = [&](const double* g) {
        return (1 - my) * ((1 - mx) * g[i0 * P + j0] + mx * g[i0 * P + j1]) +
               my * ((1 - mx) * g[i1 * P + j0] + mx * g[i1 * P + j1]);
    };
    *lat = interp(s.geo_latitude);
    *lon = interp(s.geo_longitude);
}

// Geodetic2Geographic (Func.cpp:1248-1260), WGS84
void llh_to_xyz(double lat, double lon, double h, double* xyz)
{
    const double e2 = 1.0 - (kEllB / kEllA) * (kEllB / kEllA);
    const double N = kEllA / sqrt(1.0 - e2 * (sin(lat) * sin(lat)));
    const double nph = N + h;
    xyz[0] = nph * cos(lat) * cos(lon);
    xyz[1] = nph * cos(lat) * sin(lon);
    xyz[2] = (nph - e2 * N) * sin(lat);
}

// solve33 (Func.cpp:1462-1490): LU of the 3 x 3 system without pivoting
void solve3(double* x, const double* rhs, const double* A)
{
    const double l10 = A[3] / A[0], l20 = A[6] / A[0];
    const double u11 = A[4] - l10 * A[1];
    const double l21 = (A[7] - A[1] * l20) / u11;
    const double u12 = A[5] - l10 * A[2];
    const double u22 = A[8] - l20 * A[2] - l21 * u12;
    const double b0 = rhs[0], b1 = rhs[1] - b0 * l10, b2 = rhs[2] - b0 * l20 - b1 * l21;
    x[2] = b2 / u22;
    x[1] = (b1 - u12 * x[2]) / u11;
    x[0] = (b0 - A[1] * x[1] - A[2] * x[2]) / A[0];
}

bool swath_ok(const tops_annotation* a)
{
    return a && a->n_bursts > 0 && a->lines_per_burst > 0 && a->samples_per_burst > 0 &&
           a->burst_first_line_time && a->geo_az_time && a->geo_slant_range_time &&
           a->geo_latitude && a->geo_longitude && a->geo_lines >= 2 &&
           a->geo_points_per_line >= 2 && a->prf > 0 && a->range_sampling_rate > 0;
}

bool orbit_ok(const tops_orbit* o)
{
    return o && o->num_points > 0 && o->az_time && o->ncoef >= 2 && o->coef_x && o->coef_y &&
           o->coef_z;
}

}  // namespace
}  // namespace tops

using namespace tops;

extern "C" {

double tops_xyz2aztime(const double* xyz, const tops_orbit* o, double t_init)
{
    if (!xyz || !orbit_ok(o)) {
        set_error(TOPS_ERR_ARG, "tops_xyz2aztime: bad arguments");
        return t_init;
    }
    double t = t_init;
    for (int it = 0; it <= 10; ++it) {
        double p[3], v[3], a[3], d[3];
        orbit_state(*o, t, p, v, a);
        for (int k = 0; k < 3; k++) d[k] = xyz[k] - p[k];
        const double sol = -dot3(v, d) / (dot3(a, d) - dot3(v, v));
        t += sol;
        if (fabs(sol) < 1e-10) break;
    }
    return t;
}

long tops_lp2xyz(double line, double pixel, const tops_annotation* a, const tops_orbit* o,
                 int burst, double* xyz, int max_iter, double criter)
{
    if (!swath_ok(a) || !orbit_ok(o) || !xyz || burst < 0 || burst >= a->n_bursts) {
        set_error(TOPS_ERR_ARG, "tops_lp2xyz: bad arguments");
        return -1;
    }
    const double az = a->burst_first_line_time[burst] + (line - 1.0) / a->prf;
    const double rt = a->slr_time_first_pixel + (pixel - 1.0) / (a->range_sampling_rate * 2);
    double ps[3], vs[3];
    orbit_state(*o, az, ps, vs, nullptr);
    long it;
    for (it = 0; it <= max_iter; ++it) {
        double d[3], eq[3], J[9], sol[3];
        for (int k = 0; k < 3; k++) d[k] = xyz[k] - ps[k];
        eq[0] = -dot3(vs, d);
        eq[1] = -1 * (dot3(d, d) - (kSol * rt) * (kSol * rt));
        eq[2] = -1 * ((xyz[0] * xyz[0] + xyz[1] * xyz[1]) / (kEllA * kEllA) +
                      (xyz[2] / kEllB) * (xyz[2] / kEllB) - 1.0);
        J[0] = vs[0];
        J[1] = vs[1];
        J[2] = vs[2];
        J[3] = 2 * d[0];
        J[4] = 2 * d[1];
        J[5] = 2 * d[2];
        J[6] = (2 * xyz[0]) / (kEllA * kEllA);
        J[7] = (2 * xyz[1]) / (kEllA * kEllA);
        J[8] = (2 * xyz[2]) / (kEllB * kEllB);
        solve3(sol, eq, J);
        for (int k = 0; k < 3; k++) xyz[k] += sol[k];
        if (fabs(sol[0]) < criter && fabs(sol[1]) < criter && fabs(sol[2]) < criter) break;
    }
    return it;
}

int tops_burst_overlap(const tops_annotation* m, const tops_annotation* s, const tops_orbit* mo,
                       const tops_orbit* so, int burst0, int burstN, int* m_burst0,
                       int* m_burstN, int* s_burst0, int* s_burstN)
{
    const char* who = "tops_burst_overlap";
    if (!swath_ok(m) || !swath_ok(s) || !orbit_ok(mo) || !orbit_ok(so) || !m_burst0 ||
        !m_burstN || !s_burst0 || !s_burstN)
        return set_error(TOPS_ERR_ARG, "%s: bad arguments", who);
    const int b0 = burst0 - 1, bn = burstN - 1;
    if (b0 < 0 || b0 >= m->n_bursts)
        return set_error(TOPS_ERR_ARG, "%s: started number of burst %d is not valid", who, burst0);
    // ground point of a swath's burst centre, then its zero-Doppler time on
    // the other orbit
    auto centre_zdt = [](const tops_annotation* a, const tops_orbit* own, const tops_orbit* other,
                         int b, double az0) {
        double lat, lon, xyz[3];
        grid_lat_lon(*a, az0, (a->slr_time_first_pixel + a->slr_time_last_pixel) / 2.0, &lat,
                     &lon);
        llh_to_xyz(lat * kPi / 180.0, lon * kPi / 180.0, 0.0, xyz);
        tops_lp2xyz(a->lines_per_burst / 2, a->samples_per_burst / 2, a, own, b, xyz, 10, 1e-6);
        return tops_xyz2aztime(xyz, other, az0);
    };
    const double az_m = m->burst_first_line_time[b0] + m->lines_per_burst / 2.0 * m->az_time_interval;
    const double zs = centre_zdt(m, mo, so, b0, az_m);
    int m0, mn, s0, sn;
    if (zs > s->last_line_time)
        return set_error(TOPS_ERR_ARG, "%s: no overlap (the slave ends before master burst %d)",
                         who, burst0);
    if (zs < s->last_line_time && zs > s->first_line_time) {
        // the slave covers the master's start burst
        m0 = b0;
        int cc = 0;
        while (cc < s->n_bursts && zs > s->burst_first_line_time[cc]) cc++;
        s0 = cc - 1;
        if (bn - m0 <= s->n_bursts - 1 - s0) {
            mn = bn;
            sn = s0 + (mn - m0);
        } else {
            sn = s->n_bursts - 1;
            mn = m0 + (sn - s0);
        }
    } else {
        // the slave starts later: locate its first burst in the master
        const double az_s =
            s->burst_first_line_time[0] + s->lines_per_burst / 2 * s->az_time_interval;
        const double zm = centre_zdt(s, so, mo, 0, az_s);
        if (!(zm < m->last_line_time && zm > m->first_line_time))
            return set_error(TOPS_ERR_ARG, "%s: no overlap of the image pair", who);
        s0 = 0;
        int cc = 0;
        while (cc < m->n_bursts && zm > m->burst_first_line_time[cc]) cc++;
        m0 = cc - 1;
        if (bn - m0 <= s->n_bursts - 1) {
            mn = bn;
            sn = s0 + (mn - m0);
        } else {
            sn = s->n_bursts - 1;
            mn = m0 + (sn - s0);
        }
    }
    *m_burst0 = m0;
    *m_burstN = mn;
    *s_burst0 = s0;
    *s_burstN = sn;
    return TOPS_OK;
}

}  // extern "C"
