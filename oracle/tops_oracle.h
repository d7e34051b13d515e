/*
 * tops_oracle.h -- FP64 CPU restatement of the TOPS burst path of
 * a787854/InSAR-on-Sentinel-1-TOPS-data.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (the HIP library in the package) never links or calls it.
 *
 * Every function restates one reference function in double precision and
 * cites the reference file:line it follows.  Deliberate differences from the
 * reference (its undefined behaviour and float storage) are listed next to
 * the function that has them.
 *
 * Pinning: each function is checked by tests/test_oracle_*.py against closed
 * forms, invariants and brute force (see each test's docstring).
 */
#ifndef TOPS_ORACLE_H
#define TOPS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* --- resample look-up table: Func.cpp:2105-2158, Constants.h:15-33 ------- */
/* Fills kaz/krg [2048 x npoints] as the reference stores them (float).      */
/* Threads of the oracle's parallel loops (used only to time the oracle at a
 * chosen core count); the current maximum. */
void orc_set_threads(int n);
int orc_max_threads(void);

void orc_lut_init(int npoints, double prf, double abw, double rsr2x,
                  float* kaz, float* krg);

/* --- orbit: Func.cpp:940-1126 (Polyfit_Orbit + Cholesky helpers) --------- */
/* Returns the number of coefficients (degree+1) and fills cx/cy/cz[6].      */
int orc_polyfit_orbit(int n, const double* t, const double* x,
                      const double* y, const double* z,
                      double* cx, double* cy, double* cz);
/* Geocoreg_cuda.cu:123-205 (getPos/getVelo/getAcc); tn = (t-tmid)/10.       */
void orc_orbit_pos(double tn, const double* cx, const double* cy,
                   const double* cz, int ncoef, double out[3]);
void orc_orbit_vel(double tn, const double* cx, const double* cy,
                   const double* cz, int ncoef, double out[3]);
void orc_orbit_acc(double tn, const double* cx, const double* cy,
                   const double* cz, int ncoef, double out[3]);

/* --- geodesy: Geocoreg_cuda.cu:87-99 (llh2xyz), WGS84 cuConstants.cuh:5-6 - */
void orc_llh2xyz(double lat_rad, double lon_rad, double h, double out[3]);

/* --- zero-Doppler projection of DEM posts: Geocoreg_cuda.cu:211-323 ------- */
typedef struct {
    const double* cx; const double* cy; const double* cz; int ncoef;
    double mid_time;             /* orbitAzTime[N/2]                        */
    double az_time_interval;     /* s                                       */
    double first_line_time;      /* s, of the burst                         */
    double slr_time_first_pixel; /* one-way, s                              */
    double range_spacing;        /* m                                       */
    double az_init;              /* Newton start, s                         */
    int    line_offset;          /* burstIdx * linesPerBurst                */
} orc_orbit_geom;

/* line/pixel [lines*pixels]; nodata posts get -99999999.                   */
void orc_ell2lp(const int32_t* dem, int lines, int pixels, double lat0,
                double lon0, double dlat, double dlon, int nodata,
                const orc_orbit_geom* g, double* line_out, double* pixel_out);

/* --- geometric coregistration: Geocoreg_cuda.cu:548-605, 812-855 --------- */
/* Returns 0 on success, 1 when the 6x6 normal matrix is not positive
 * definite (the reference prints and leaves Cpm unset; here Cpm is zeroed). */
int orc_geocoreg(const int32_t* dem, int lines, int pixels, double lat0,
                 double lon0, double dlat, double dlon, int nodata,
                 const orc_orbit_geom* m, const orc_orbit_geom* s,
                 int xmin, int xmax, int ymin, int ymax,
                 double cpm_az[6], double cpm_rg[6], int64_t* n_used);

/* --- deramp parameters: Func.cpp:1868-2099 (formulas; XML out of scope) -- */
typedef struct { double t0, c0, c1, c2; } orc_poly;
void orc_deramp_params(int n_bursts, int spb, double slr_time_first_pixel,
                       double range_spacing, int lines_per_burst,
                       double az_time_interval, int first_valid_pixel,
                       const orc_poly* dc, const orc_poly* fm,
                       double steering_rate_deg, double wavelength,
                       double orbit_velocity,
                       double* fdc, double* ka, double* tref, double* kt);

/* --- deramp + 12x12 resample + reramp --------------------------------------
 * DerampAndResample.cu:108-183 (GetIndexes, DerampDemod_Shared) and
 * 348-520 (Interpolation_12p); boxes as in Func.cpp:2356-2371.
 * out: interleaved complex double for output rows [row0, row1) of the
 * master box, full width.  The deramp phase is kept in double (the reference
 * stores it as float, DerampAndResample.cu:163-164: see ref_float_phase).
 * Columns are normalised as absolute col = MasterBox[0] + c.
 * ref_float_phase != 0 rounds the per-sample phase to float as the
 * reference does (used only to quantify that rounding).                    */
typedef struct {
    int master_box[4];   /* x0 x1 y0 y1, inclusive, absolute                */
    int slave_box[4];
    int first_line_in_burst;
    double az_time_interval;
    double az_shift;
    double cpm_az[6], cpm_rg[6];
} orc_resample_geom;

void orc_resample(const int16_t* slave, const orc_resample_geom* g,
                  const float* kaz, const float* krg, int npoints,
                  const double* kt, const double* tref, const double* fdc,
                  int row0, int row1, int ref_float_phase, double* out);

/* --- boxcar coherence: EstCoh.cu:101-139 (CohMatSet, zero padding) and
 * 413-456 (Cohbasic).  master CInt16 [h*w], slave complex double [h*w].
 * coh = 0 where |slave| < 0.001.                                           */
void orc_coherence(const int16_t* master, const double* slave, int w, int h,
                   int dx, int dy, int win_rg, int win_az, double* coh);

/* --- ESD: EstESD.cu:62-161, 243-546, 914-1217 -----------------------------
 * Band arrays are the concatenated overlap bands (Func.cpp:2564-2577).
 * shifts[i] = atan2(sum_i)/(2 pi df dt), in lines.                         */
void orc_esd(const int16_t* m_for, const int16_t* m_back,
             const double* s_for, const double* s_back,
             const int* overlap_size, int n_overlap, int pixels,
             int win_az, int win_rg, double threshold,
             double spectral_sep, double az_time_interval,
             double* shifts, double* sums /* 2*n_overlap, may be NULL */);

/* Func.cpp:2935-2960 and 2962-2976. */
void orc_overlap_sizes(const double* first_line_time,
                       const double* last_line_time, int burst0, int burstN,
                       double az_time_interval, int* sizes, int* total,
                       int* maximum);
double orc_spectral_sep(const double* kt, int n_bursts, int spb,
                        int lines_per_burst, double az_time_interval);

/* --- deburst: getLineIdxInMetaData (Func.cpp:2766-2819), the extent and
 * row loop of Stitch_OneSubswath (Func.cpp:2821-2933).                     */
int orc_deburst_line(const double* first_line_time, const double* last_line_time,
                     int n_bursts, int lines_per_burst, int num_lines,
                     double az_time_interval, double t, int burst0, int burstN);
void orc_deburst_extent(double first_valid_time_b0, double last_valid_time_bN,
                        double first_line_time, double az_time_interval,
                        int* y_min, int* height);
int orc_deburst(const double* first_line_time, const double* last_line_time,
                const double* first_valid_time, const double* last_valid_time,
                int n_bursts, int lines_per_burst, int samples_per_burst,
                int num_lines, double az_time_interval, double swath_first_line_time,
                int burst0, int burstN, const void* gapped, int elem_bytes,
                void* out, int* rows);

/* --- burst matching: Func.cpp:1133-1260, 1349-1377, 1462-1584, 1610-1746 */
typedef struct {
    int n_bursts, lpb, spb;
    double dt, prf, rsr, slr_first, slr_last, first_line_time, last_line_time;
    const double* burst_first_line_time;
    int geo_lines, geo_ppl;
    const double *geo_az, *geo_rg, *geo_lat, *geo_lon;
} orc_match_swath;
typedef struct {
    int num_points;
    const double* t;
    int ncoef;
    const double *cx, *cy, *cz;
} orc_match_orbit;
double orc_xyz2aztime(const double* xyz, const orc_match_orbit* o, double t_init);
int orc_lp2xyz(double line, double pixel, const orc_match_swath* s, const orc_match_orbit* o,
               int burst, double* xyz);
/* out = (m_burst0, m_burstN, s_burst0, s_burstN), 0-based; returns 0, or
 * -1 (bad start burst) / -2 (no overlap) */
int orc_burst_overlap(const orc_match_swath* m, const orc_match_swath* s,
                      const orc_match_orbit* mo, const orc_match_orbit* so, int in_b0,
                      int in_bn, int* out);

#ifdef __cplusplus
}
#endif
#endif
