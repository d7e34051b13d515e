/*
 * tops_abi.h -- C-ABI of the MI355X TOPS burst library (libtops_mi355x.so).
 *
 * Part 1 keeps the seven extern "C" entry points of the reference
 * a787854/InSAR-on-Sentinel-1-TOPS-data with identical names, parameter
 * order and meaning, so its Func.cpp can link against this library
 * unchanged.  std::complex<short> is laid out as tops_ci16 and cuComplex /
 * std::complex<float> as tops_cf32.  Differences in behaviour:
 *   - no exit() and no device reset inside the library; errors set a status
 *     readable with tops_last_error() and are printed once to stderr (each
 *     Part 1 entry point clears the status on entry, so a check after the
 *     call reports that call);
 *   - per-call parameter blocks instead of module globals (reentrant):
 *     calls may run on several streams and from several host threads at
 *     once, each thread with its own status; a context (tops_ctx) may be
 *     shared, while a plan (tops_plan) and the fixed scratch of
 *     tops_esd_shifts_noalloc serve one call at a time;
 *   - device buffers returned by DerampDemodResample_ESD/ResampleFirstBurst
 *     are dense [rows x cols] (pitch = cols) and are released by the
 *     CoherenceEst* call that consumes them, as in the reference.
 *
 * Part 2 is the device-resident API: inputs and outputs already in HBM,
 * asynchronous on a caller stream (hipStream_t passed as void*).
 *
 * Part 3 is the host stage layer: the reference's GeometricCoreg,
 * DerampAndResample and ESDAndCoh over in-memory bursts.
 */
#ifndef TOPS_ABI_H
#define TOPS_ABI_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { int16_t re, im; } tops_ci16;  /* CInt16, 4 B */
typedef struct { float x, y; } tops_cf32;      /* CF32, 8 B   */

/* ---- status ------------------------------------------------------------ */
enum {
    TOPS_OK = 0,
    TOPS_ERR_ARG = 1,        /* invalid argument (sizes, boxes, npoints)   */
    TOPS_ERR_HIP = 2,        /* HIP runtime error                          */
    TOPS_ERR_NOT_PD = 3,     /* geocoreg normal matrix not positive definite */
    TOPS_ERR_NODEVICE = 4,   /* no GPU visible                             */
    TOPS_ERR_ALLOC = 5       /* device allocation failed                   */
};
int tops_last_error(void);
const char* tops_last_error_message(void);
void tops_clear_error(void);

/* ======================================================================== */
/* Part 1: reference entry points                                           */
/* ======================================================================== */

/* Replaces Geocoreg_cuda.cu:22-68 (definition 610-913).  Writes the six
 * azimuth and six range coefficients of the slave position polynomial in
 * normalised master coordinates.  inputDEM is [DemLines x DemPixels] int32;
 * post (r, c) lies at (lat0 - r*dem_DeltaLat, lon0 + c*dem_DeltaLon) deg.
 * The full int32 DEM is used (the reference copies half of it,
 * Geocoreg_cuda.cu:788).  On a singular fit AzCpm/RgCpm are zeroed and the
 * status is TOPS_ERR_NOT_PD. */
void GeoCoreg_warp_cuda(int m_burstIdx, int s_burstIdx, int xmin, int xmax,
                        int ymin, int ymax, double dem_DeltaLat,
                        double dem_DeltaLon, int DemLines, int DemPixels,
                        double* AzCpm, double* RgCpm, double lat0, double lon0,
                        int demnodata, double m_azimuthTimeInterval,
                        double s_azimuthTimeInterval, double wavelength,
                        double m_linesPerBurst, double s_linesPerBurst,
                        double m_burstFirstLineTime, double s_burstFirstLineTime,
                        int m_numOfSamples, int s_numOfSamples,
                        double* m_orbitAzTime, double* s_orbitAzTime,
                        int orb_m_numberofpoints, int orb_m_numbercoefs,
                        int orb_s_numberofpoints, int orb_s_numbercoefs,
                        double m_slrTimeToFirstPixel, double s_slrTimeToFirstPixel,
                        double m_rangeSpacing, double s_rangeSpacing,
                        bool m_nearRangeOnLeft, bool s_nearRangeOnLeft,
                        int* inputDEM, double* m_x_coef, double* m_y_coef,
                        double* m_z_coef, double* s_x_coef, double* s_y_coef,
                        double* s_z_coef, double m_aziInitial, double s_aziInitial);

/* Replaces DerampAndResample.cu:27-47 (definition 1112-1426): deramp the
 * slave burst SlaveArray [SlaveBox rows x cols], resample it onto the master
 * box with the 12-tap LUTs KernelAz/KernelRg [2048 x Npoints], reramp, and
 * write output [MasterBox rows x cols].  dopplerRate/referenceTime/
 * dopplerCentroid are indexed by absolute slave column.  Npoints must be 12
 * (the reference's only working value, DerampAndResample.cu:1295-1298). */
void DerampDemodResample(tops_ci16* SlaveArray, double* CpmAz, double* CpmRg,
                         double AzimuthShift, tops_cf32* output,
                         float* KernelAz, float* KernelRg, int sBurstIdx,
                         int slave_pixels, int slave_lines, int MasterBox[4],
                         int SlaveBox[4], int linesPerBurst,
                         int SamplesPerBurst, double azimuthTimeInterval,
                         double* dopplerRate, double* referenceTime,
                         double* dopplerCentroid, int Npoints);

/* Replaces DerampAndResample.cu:49-69 (definition 1430-1749): as above, and
 * also returns the resampled burst as a device buffer owned by the caller
 * until passed to CoherenceEst/CoherenceEst_ESD.  NULL on error. */
tops_cf32* DerampDemodResample_ESD(tops_ci16* SlaveArray, double* CpmAz,
                                   double* CpmRg, double AzimuthShift,
                                   tops_cf32* output, float* KernelAz,
                                   float* KernelRg, int sBurstIdx,
                                   int slave_pixels, int slave_lines,
                                   int MasterBox[4], int SlaveBox[4],
                                   int linesPerBurst, int SamplesPerBurst,
                                   double azimuthTimeInterval,
                                   double* dopplerRate, double* referenceTime,
                                   double* dopplerCentroid, int Npoints);

/* Replaces DerampAndResample.cu:71-75 (definition 1751-1769): upload an
 * already resampled burst [hh x ww] to a device buffer.  NULL on error. */
tops_cf32* ResampleFirstBurst(tops_cf32* SlaveArray, int ww, int hh);

/* Replaces EstESD.cu:28-44 (definition 914-1299).  Band arrays are the
 * numOverlap overlap bands concatenated (ArrayLines = sum of
 * OverlapSizeArray, ArrayPixels columns).  ShiftArray[i] is in lines. */
void EstESDShifts(tops_ci16* MasterFor, tops_ci16* MasterBack,
                  tops_cf32* SlaveFor, tops_cf32* SlaveBack,
                  int* OverlapSizeArray, double* ShiftArray, int numOverlap,
                  int ArrayLines, int ArrayPixels, int CohWinAz, int CohWinRg,
                  int OverlapMaxLines, float CohThresHold,
                  double spectralSeparation, double azimuthTimeInterval);

/* Replaces EstCoh.cu:919-951 (CoherenceEst) and 462-736 (CoherenceEst_ESD):
 * boxcar coherence of masterArray [Dh x Dw] against the device buffer
 * d_slaveArray [Dh x Dw] (consumed and freed), window cohWinAz x cohWinRg
 * starting dy rows / dx columns before the output sample, zero outside the
 * burst; 0 where |slave| < 0.001.  cohdata [Dh x Dw] on the host. */
void CoherenceEst(int Dx0, int Dy0, int Dw, int Dh, int dx, int dy,
                  int cohWinRg, int cohWinAz, tops_ci16* masterArray,
                  tops_cf32* d_slaveArray, float* cohdata);
void CoherenceEst_ESD(int Dx0, int Dy0, int Dw, int Dh, int dx, int dy,
                      int cohWinRg, int cohWinAz, tops_ci16* masterArray,
                      tops_cf32* d_slaveArray, float* cohdata);

/* ======================================================================== */
/* Part 2: device-resident API (no counterpart in the reference; it is the
 * form the reference's per-call host round trips, DerampAndResample.cu:
 * 1150-1156, 1373, 1421, are replaced by)                                   */
/* ======================================================================== */

/* Arithmetic modes. */
enum {
    TOPS_MODE_FAST = 0,   /* fused FP32 MACs, FP64 phase reduced mod 2*pi    */
    TOPS_MODE_FP64 = 1    /* plain all-FP64 per-stage kernels (baseline)     */
};

/* Resample context: LUTs on the device plus their FP64 tap moments. */
typedef struct tops_ctx tops_ctx;
tops_ctx* tops_ctx_create(const float* kernel_az, const float* kernel_rg,
                          int npoints);
void tops_ctx_destroy(tops_ctx* ctx);

/* Geometry of one burst: what DerampDemodResample receives besides data. */
typedef struct {
    int master_box[4];        /* x0 x1 y0 y1 inclusive, absolute master     */
    int slave_box[4];         /* x0 x1 y0 y1 inclusive, absolute slave      */
    int first_line_in_burst;  /* sBurstIdx * linesPerBurst                  */
    int samples_per_burst;    /* length of the per-range arrays             */
    double az_time_interval;  /* s                                          */
    double az_shift;          /* lines, added to the azimuth polynomial     */
    double cpm_az[6];
    double cpm_rg[6];
} tops_burst_geom;

/* Deramp -> resample -> reramp of one burst.  d_slave [sLines x sPixels],
 * d_out [mLines x mPixels], dense.  d_kt/d_tref/d_fdc [samples_per_burst]. */
int tops_resample_burst(tops_ctx* ctx, const tops_burst_geom* geom,
                        const tops_ci16* d_slave, const double* d_kt,
                        const double* d_tref, const double* d_fdc,
                        tops_cf32* d_out, int mode, void* stream);

/* Boxcar coherence (Cohbasic semantics, zero padding, |S|<0.001 mask).
 * d_master [h x w] CInt16, d_slave [h x w] CF32, d_coh [h x w] float. */
int tops_coherence(const tops_ci16* d_master, const tops_cf32* d_slave,
                   int w, int h, int dx, int dy, int win_rg, int win_az,
                   float* d_coh, int mode, void* stream);

/* ESD over concatenated device bands; shifts written to host array
 * (synchronises the stream).  Same band layout as EstESDShifts. */
int tops_esd_shifts(const tops_ci16* d_m_for, const tops_ci16* d_m_back,
                    const tops_cf32* d_s_for, const tops_cf32* d_s_back,
                    const int* overlap_size, int n_overlap, int pixels,
                    int win_az, int win_rg, float threshold,
                    double spectral_sep, double az_time_interval,
                    double* shifts, double* sums, void* stream);

/* tops_esd_shifts without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate (code that runs beside HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as tops_esd_shifts; the tile partials
 * and band sums live in a fixed per-GPU device array of 4 MiB (at most
 * 262144 tiles of 16 lines x 64 samples, i.e. n_overlap x ceil(pixels/64) x
 * ceil(max overlap/16), and at most 64 overlaps), held by each GPU the
 * library runs on until the process ends.  Larger calls return TOPS_ERR_ARG.
 * Calls that overlap on one GPU share that array and race. */
int tops_esd_shifts_noalloc(const tops_ci16* d_m_for, const tops_ci16* d_m_back,
                            const tops_cf32* d_s_for, const tops_cf32* d_s_back,
                            const int* overlap_size, int n_overlap, int pixels,
                            int win_az, int win_rg, float threshold,
                            double spectral_sep, double az_time_interval,
                            double* shifts, double* sums, void* stream);

/* ---- batched bursts (no counterpart in the reference, which runs one
 * burst per call with a device reset in between) ---------------------- */
typedef struct {
    tops_burst_geom geom;
    const tops_ci16* slave;      /* device [sLines x sPixels]                */
    const double* kt;            /* device [samples_per_burst] (+ tref, fdc) */
    const double* tref;
    const double* fdc;
    tops_cf32* out;              /* device [mLines x mPixels]                */
    const tops_ci16* master;     /* device [mLines x mPixels]; NULL: no coherence */
    float* coh;                  /* device [mLines x mPixels]                */
} tops_burst_item;

typedef struct tops_plan tops_plan;
/* A plan resamples every burst and forms its coherence (window coh_win_az
 * x coh_win_rg starting (win-1)/2 before the sample) on one stream, on one
 * of three schedules: strip_rows == 0, one resample and one coherence
 * launch per burst with heights sized to each burst; strip_rows > 0, one
 * batched resample launch over every burst's strips of that height, then
 * one batched coherence launch; strip_rows < 0, one persistent launch that
 * resamples and forms the coherence of row bands of -strip_rows lines (the
 * 3 az x 10 rg window; other windows take the batched schedule at height
 * -strip_rows).  No counterpart in the reference, which runs one burst per
 * call. */
tops_plan* tops_plan_create(tops_ctx* ctx, const tops_burst_item* items, int n_items,
                            int coh_win_az, int coh_win_rg, int strip_rows);
int tops_plan_run(tops_plan* plan, void* stream);
/* Capture the plan's launches in a hipGraph on a non-default stream; later
 * tops_plan_run calls replay it. */
int tops_plan_capture_graph(tops_plan* plan, void* stream);
/* Time the candidate schedules on this plan (per-burst, batched at several
 * strip heights, fused at several row bands) and keep the fastest; then,
 * for a batched schedule, time the batched coherence's rows per block (3 az
 * x 10 rg window) and the burst group count (as replayed graphs), keeping
 * the fastest of each.
 * *best_strip_rows gets the chosen strip_rows value.  Cached in the file
 * named by TOPS_TUNE_CACHE when set. */
int tops_plan_tune(tops_plan* plan, void* stream, int* best_strip_rows);
int tops_plan_strip_rows(const tops_plan* plan);
/* Rows per block of the plan's batched coherence walk (default 32). */
int tops_plan_coh_rows(const tops_plan* plan);
/* Burst groups of the batched schedule (default 1): the bursts split into
 * that many runs of consecutive bursts with about equal sample counts; each
 * group is resampled by one launch on the caller's stream and its coherence
 * launch, on a second stream of the plan, waits for that launch only, so it
 * runs beside the next group's resampling.  The plan's run (and its graph)
 * ends with the caller's stream waiting for the second stream. */
int tops_plan_set_groups(tops_plan* plan, int groups);
int tops_plan_groups(const tops_plan* plan);
/* The candidates the last tops_plan_tune timed, in order: entry i gives
 * its strip_rows, coherence rows per block, burst groups, tuning stage (1
 * schedule, 2 coherence rows, 3 burst groups) and mean ms per run (eager
 * runs in stages 1 and 2; replays of a captured graph in stage 3 unless the
 * stream was the null stream). */
int tops_plan_tune_count(const tops_plan* plan);
int tops_plan_tune_entry(const tops_plan* plan, int i, int* strip_rows, int* coh_rows,
                         int* groups, int* stage, float* ms);
/* Durations of the plan's launches: `reps` eager runs of its schedule on
 * `stream` with an event pair around every launch.  ms[k] (k = 0 resample,
 * 1 coherence, 2 fused resample + coherence; 3 entries) gets the mean per
 * run of the summed durations of that kind's launches, launches[k]
 * (optional, 3 entries) their count per run. */
int tops_plan_time_launches(tops_plan* plan, void* stream, int reps, float* ms,
                            int* launches);
/* Kernel symbol behind the launches of kind `kind` (as above) on the
 * plan's current schedule, "" when it has none of that kind. */
const char* tops_plan_kernel_name(const tops_plan* plan, int kind);
void tops_plan_destroy(tops_plan* plan);

/* ESDAndCoh (Func.cpp:2496-2763) over one subswath held in HBM, with no
 * host round trip between its steps: pass-1 resample (burst 0 whole, of
 * later bursts only the lines in an overlap band: pass 2 rewrites the
 * rest, so the outputs equal a full pass 1's), ESD over the overlap bands
 * read in place, shifts formed on the device
 * (|shift| > 100 -> 0), pass-2 resample of bursts 1..n-1 with
 * AzimuthShift = shift[b-1] (master lines per burst and azimuth interval,
 * Func.cpp:2701-2705), coherence of every burst.  All launches go to
 * `stream` in that order (capturable in one hipGraph); nothing is
 * allocated.  Rasters are whole subswaths, burst after burst. */
typedef struct tops_subswath_desc {
    int n_bursts;
    int m_lines_per_burst, m_samples_per_burst;
    int s_lines_per_burst, s_samples_per_burst;
    double m_az_time_interval, s_az_time_interval;
    double spectral_sep;            /* GetSpectralSep of the master swath   */
    const int* overlap_lines;       /* host [n_bursts-1], EstOverlapSize    */
    int esd_win_az, esd_win_rg;     /* reference: 6 x 24                    */
    float threshold;                /* reference: 0.3                       */
    int coh_win_az, coh_win_rg;
} tops_subswath_desc;
/* Bytes of device workspace tops_subswath_esd_coh needs (0: bad desc). */
size_t tops_subswath_workspace_size(const tops_subswath_desc* desc);
int tops_subswath_esd_coh(tops_ctx* ctx, const tops_subswath_desc* desc,
                          const tops_ci16* d_master,  /* [n*mLpb x mSpb]        */
                          const tops_ci16* d_slave,   /* [n*sLpb x sSpb]        */
                          const double* d_kt,         /* [n x sSpb] (+ tref, fdc) */
                          const double* d_tref, const double* d_fdc,
                          const double* cpm_az,       /* host [n x 6]           */
                          const double* cpm_rg,
                          tops_cf32* d_reslave,       /* out [n*mLpb x mSpb]    */
                          float* d_coh,               /* out [n*mLpb x mSpb] or NULL */
                          double* d_shifts,           /* out device [n-1]       */
                          void* d_workspace, size_t workspace_bytes, void* stream);

/* Geocoreg orbit description for one acquisition (device-side copy is made). */
typedef struct {
    const double* coef_x; const double* coef_y; const double* coef_z;
    int ncoef;
    double mid_time;             /* orbitAzTime[N/2]                        */
    double az_time_interval;
    double first_line_time;      /* burst first line time                  */
    double slr_time_first_pixel; /* one-way                                 */
    double range_spacing;
    double az_init;              /* Newton start                            */
    int line_offset;             /* burstIdx * linesPerBurst                */
} tops_orbit_geom;

/* DEM-assisted coregistration fit on a device DEM [lines x pixels] int32.
 * Writes cpm_az/cpm_rg (host), and optionally the number of posts used.
 * d_lp (optional, device, 4*lines*pixels doubles) receives master line,
 * master pixel, slave line, slave pixel planes. */
int tops_geocoreg(const int32_t* d_dem, int lines, int pixels, double lat0,
                  double lon0, double dlat, double dlon, int nodata,
                  const tops_orbit_geom* m, const tops_orbit_geom* s,
                  int xmin, int xmax, int ymin, int ymax, double* cpm_az,
                  double* cpm_rg, int64_t* n_used, double* d_lp, void* stream);

/* Deburst (Stitch_OneSubswath, Func.cpp:2821-2933): the SubSwathInfo
 * fields it reads (DataTypes.h:51-140).  Times in seconds; the valid-line
 * times are burstFirstLineTime + firstValidLine * dt and likewise for the
 * last valid line (Func.cpp:401-403); num_lines is the subswath's
 * numberOfLines.  At most 16 bursts (the reference's arrays are [16]). */
typedef struct tops_deburst_swath {
    int lines_per_burst, samples_per_burst, n_bursts, num_lines;
    double az_time_interval, first_line_time;
    const double* burst_first_line_time;        /* host [n_bursts] */
    const double* burst_last_line_time;         /* host [n_bursts] */
    const double* burst_first_valid_line_time;  /* host [n_bursts] */
    const double* burst_last_valid_line_time;   /* host [n_bursts] */
} tops_deburst_swath;

/* data_type codes of Stitch_OneSubswath (GDAL's GDT_* values). */
#define TOPS_GDT_FLOAT32 6
#define TOPS_GDT_CINT16 8
#define TOPS_GDT_CFLOAT32 10

/* Output extent of the deburst of bursts burst0 .. burstN: first subswath
 * line y_min and row count (yMin, yMax of Func.cpp:2837-2838).  Fails
 * (TOPS_ERR_ARG) where an output line time lies in no burst: the reference
 * exits there (Func.cpp:2868-2876). */
int tops_deburst_extent(const tops_deburst_swath* sw, int burst0, int burstN, int* y_min,
                        int* height);

/* Row of the burst-stacked input each output row copies
 * (getLineIdxInMetaData, Func.cpp:2766-2819, minus burst0 * lines per
 * burst); rows outside 0 .. (burstN-burst0+1)*lpb-1 give zero lines.
 * rows: host [height]. */
int tops_deburst_rows(const tops_deburst_swath* sw, int burst0, int burstN, int* rows);

/* Stitch_OneSubswath on rasters in HBM: d_gapped [(burstN-burst0+1)*lpb x
 * spb] of data_type elements -> d_out [height x spb].  One launch on
 * `stream`; nothing is allocated (capturable). */
int tops_deburst(const tops_deburst_swath* sw, int burst0, int burstN, int data_type,
                 const void* d_gapped, void* d_out, void* stream);

/* ======================================================================== */
/* Part 3: host stage layer                                                 */
/* ======================================================================== */

/* Replaces InKernelInit, Func.cpp:2105-2158: fills kernel_az/kernel_rg
 * [2048 x npoints] (chi_az = prf/abw, chi_rg = (rsr2x/2)/abw as the
 * reference, Func.cpp:2110-2111). */
int tops_lut_init(int npoints, double prf, double abw, double rsr2x, float* kernel_az,
                  float* kernel_rg);

/* Replaces Polyfit_Orbit, Func.cpp:1050-1126: returns the number of
 * coefficients (min(5, n-2) + 1), 0 on failure. */
int tops_polyfit_orbit(int n, const double* t, const double* x, const double* y,
                       const double* z, double* coef_x, double* coef_y, double* coef_z);

/* |velocity| at absolute time t (getVelocity, Func.cpp:1279-1300). */
double tops_orbit_speed(double t, double mid_time, const double* cx, const double* cy,
                        const double* cz, int ncoef);

/* Replaces the formulas of ComputeDerampPara, Func.cpp:1868-2099 (the
 * annotation XML reading is not part of this library): dc/fm are
 * [n_bursts][4] = (t0, c0, c1, c2); outputs [n_bursts x spb]. */
int tops_deramp_params(int n_bursts, int spb, double slr_time_first_pixel,
                       double range_spacing, int lines_per_burst, double az_time_interval,
                       int first_valid_pixel, const double* dc, const double* fm,
                       double steering_rate_deg, double wavelength, double orbit_speed,
                       double* kt, double* tref, double* fdc, double* ka);

/* Replaces EstOverlapSize, Func.cpp:2935-2960; returns the overlap count. */
int tops_overlap_sizes(const double* first_line_time, const double* last_line_time,
                       int burst0, int burstN, double az_time_interval, int* sizes,
                       int* total, int* maximum);

/* Replaces GetSpectralSep, Func.cpp:2962-2976. */
double tops_spectral_sep(const double* kt, int n_bursts, int spb, int lines_per_burst,
                         double az_time_interval);

/* SubSwathInfo subset (DataTypes.h:51-140) the stages read.  Arrays are
 * host memory; the per-range arrays are [n_bursts x samples_per_burst] as
 * ComputeDerampPara leaves them; slc is the burst-stacked raster
 * [n_bursts * lines_per_burst x samples_per_burst] of the SLC. */
typedef struct {
    int lines_per_burst, samples_per_burst, n_bursts;
    double az_time_interval, range_spacing, slr_time_first_pixel, wavelength;
    const double* burst_first_line_time;   /* [n_bursts] */
    const double* burst_last_line_time;    /* [n_bursts] */
    const double* burst_ll;                /* [n_bursts][4] lat_min lat_max lon_min lon_max */
    const double* doppler_rate;            /* kt */
    const double* reference_time;          /* t_ref */
    const double* doppler_centroid;        /* f_dc */
    const tops_ci16* slc;
} tops_swath;

/* S1PreciseOrbit after Polyfit_Orbit (DataTypes.h:162-181). */
typedef struct {
    int num_points;
    const double* az_time;                 /* [num_points] */
    int ncoef;
    const double* coef_x;
    const double* coef_y;
    const double* coef_z;
} tops_orbit;

/* In-memory DEM (what RefDem reads from a GeoTIFF, DataTypes.cpp:201-307):
 * post (r, c) at (lat_max - r dlat, lon_min + c dlon) degrees. */
typedef struct {
    const int32_t* data;
    int lines, pixels;
    double lat_max, lon_min, dlat, dlon;
    int nodata;
} tops_dem;

/* Replaces GeometricCoreg, Func.cpp:2215-2297.  cpm_az/cpm_rg receive
 * [(mburstN - mburst0 + 1) x 6] coefficients. */
int tops_geometric_coreg(const tops_swath* m, const tops_swath* s, const tops_orbit* mo,
                         const tops_orbit* so, const tops_dem* dem, int mburst0, int mburstN,
                         int sburst0, int sburstN, double* cpm_az, double* cpm_rg);

/* Replaces DerampAndResample, Func.cpp:2326-2406: out (host) and/or d_out
 * (device) receive the burst-stacked resampled slave
 * [(mburstN - mburst0 + 1) * lpb x spb]. */
int tops_deramp_and_resample(const tops_swath* m, const tops_swath* s, const double* cpm_az,
                             const double* cpm_rg, int mburst0, int mburstN, int sburst0,
                             int sburstN, const float* kernel_az, const float* kernel_rg,
                             tops_cf32* out, tops_cf32* d_out);

/* Replaces ESDAndCoh, Func.cpp:2496-2763 (reslave = the pass-1 result). */
int tops_esd_and_coh(const tops_swath* m, const tops_swath* s, const double* cpm_az,
                     const double* cpm_rg, int mburst0, int mburstN, int sburst0, int sburstN,
                     const float* kernel_az, const float* kernel_rg, const tops_cf32* reslave,
                     int esd_win_az, int esd_win_rg, float threshold, int coh_win_az,
                     int coh_win_rg, tops_cf32* reslave_esd, float* coh, double* shifts);

/* config.txt (Func.cpp:7-130 ReadInput, ConfigSet DataTypes.h:24-40):
 * "key=value" lines; unknown keys are ignored. */
typedef struct {
    char process_dir[512];
    char masterpath[512];
    char slavepath[512];
    char polarisation[32];
    char precise_orbit_master[512];
    char precise_orbit_slave[512];
    char specific_dem_path[512];
    int firstsubswath, lastsubswath, multi_az, multi_rg, burst0, burstN;
} tops_config;
int tops_read_config(const char* path, tops_config* cfg);

/* Replaces Stitch_OneSubswath, Func.cpp:2821-2933, with host buffers for
 * its two TIFF files: gapped [(burstN-burst0+1)*lpb x spb] -> out
 * [height x spb] (tops_deburst_extent), through the device kernel. */
int tops_stitch_subswath(const tops_deburst_swath* sw, int burst0, int burstN, int data_type,
                         const void* gapped, void* out);

/* ======================================================================== */
/* Part 4: ingest without GDAL or tinyxml2                                  */
/* ======================================================================== */

/* Replaces utc2seconds, Func.cpp:629-635: seconds of the day of an
 * "YYYY-MM-DDTHH:MM:SS.ffffff" time (the date is ignored, as there). */
double tops_utc2seconds(const char* utc);

/* One subswath's annotation XML: the fields S1_Initialize reads
 * (Func.cpp:205-556) and the per-burst Doppler centroid and FM-rate
 * polynomials of ComputeDerampPara (Func.cpp:1868-2031, computeDC
 * 1828-1866).  Times are seconds of the day; slant-range times one way.
 * Arrays are allocated by tops_annotation_read and released by
 * tops_annotation_free. */
typedef struct tops_annotation {
    char polarisation[8];
    int num_lines, num_samples;
    double first_line_time, last_line_time;
    double az_time_interval, prf;
    double range_pixel_spacing, azimuth_pixel_spacing;
    double radar_frequency, wavelength, range_sampling_rate;
    double azimuth_steering_rate, heading, ascending_node_time;
    double azimuth_bandwidth, range_bandwidth;
    double slr_time_first_pixel, slr_time_last_pixel;
    int n_bursts, lines_per_burst, samples_per_burst;
    double* burst_first_line_time;        /* [n_bursts] */
    double* burst_last_line_time;         /* [n_bursts] */
    double* burst_first_valid_line_time;  /* [n_bursts] */
    double* burst_last_valid_line_time;   /* [n_bursts] */
    int* first_valid_line;                /* [n_bursts] */
    int* last_valid_line;                 /* [n_bursts] */
    int* first_valid_sample;              /* [n_bursts * lines_per_burst] */
    int* last_valid_sample;               /* [n_bursts * lines_per_burst] */
    int first_valid_pixel, last_valid_pixel;
    double slr_time_first_valid_pixel, slr_time_last_valid_pixel;
    int geo_lines, geo_points_per_line;   /* geolocation grid */
    double* geo_az_time;                  /* [geo_lines * geo_points_per_line] */
    double* geo_slant_range_time;
    double* geo_latitude;
    double* geo_longitude;
    double* geo_incidence_angle;
    int near_range_on_left;
    double* burst_ll;                     /* [n_bursts][4] lat_min lat_max lon_min lon_max */
    double* dc;                           /* [n_bursts][4] t0 c0 c1 c2 (tops_deramp_params) */
    double* fm;                           /* [n_bursts][4] t0 c0 c1 c2 */
    char first_line_utc[32];              /* productFirstLineUtcTime text (the   */
    char last_line_utc[32];               /* orbit window's times, Main.cpp:39)  */
} tops_annotation;
int tops_annotation_read(const char* xml_path, tops_annotation* a);
void tops_annotation_free(tops_annotation* a);

/* Precise orbit (EOF): the state vectors S1_OrbitInitialize takes
 * (Func.cpp:704-819): from 4 vectors before the one in the 10 s slot of
 * first_utc, tt/10 + 10 vectors (tt = whole seconds from first_utc to
 * last_utc).  Feed t, x, y, z to tops_polyfit_orbit. */
typedef struct tops_eof_orbit {
    int n;
    double* t;                            /* seconds of the day */
    double *x, *y, *z, *vx, *vy, *vz;
    double time_interval;
} tops_eof_orbit;
int tops_eof_read(const char* eof_path, const char* first_utc, const char* last_utc,
                  tops_eof_orbit* o);
void tops_eof_free(tops_eof_orbit* o);

/* Stripped, uncompressed, single-band TIFF / BigTIFF (TiffRead / TiffWrite,
 * DataTypes.cpp:417-620) of TOPS_GDT_* types CInt16, Float32, CFloat32 and
 * 7 (Float64).  Windows are (x0, y0, lines, pixels) as GDALRasterIO takes
 * them; a window outside the raster is an error and leaves the buffer
 * untouched. */
typedef struct tops_tiff tops_tiff;
#define TOPS_GDT_FLOAT64 7
tops_tiff* tops_tiff_open(const char* path);
int tops_tiff_info(const tops_tiff* t, int* pixels, int* lines, int* data_type);
int tops_tiff_read(tops_tiff* t, int x0, int y0, int lines, int pixels, int data_type,
                   void* buffer);
/* The overlap-band read of DataTypes.cpp:456-513: for overlap i, lines
 * y0+(i+1)*lpb-n_i .. (forward) and y0+(i+1)*lpb .. (backward), n_i lines
 * each, concatenated. */
int tops_tiff_read_overlaps(tops_tiff* t, void* data_for, void* data_back, int n_overlap,
                            const int* overlap_size, int lines_per_burst, int x0, int y0,
                            int width, int data_type);
/* The window straight into device memory: file reads into pinned staging
 * buffers overlapped with their copies to d_buffer on `stream`; returns
 * after the last copy has completed. */
int tops_tiff_read_device(tops_tiff* t, int x0, int y0, int lines, int pixels, int data_type,
                          void* d_buffer, void* stream);
tops_tiff* tops_tiff_create(const char* path, int data_type, int pixels, int lines);
int tops_tiff_write(tops_tiff* t, int x0, int y0, int lines, int pixels, int data_type,
                    const void* buffer);
int tops_tiff_close(tops_tiff* t);

/* Burst matching for pairs whose burst grids are offset.  Replaces
 * xyz2aztime (Func.cpp:1349-1377): zero-Doppler time of xyz on a fitted
 * orbit, Newton from t_init. */
double tops_xyz2aztime(const double* xyz, const tops_orbit* orbit, double t_init);
/* Replaces lp2xyz (Func.cpp:1492-1584): ground position (WGS84 ellipsoid)
 * of 1-based (line, pixel) of burst `burst`, Newton from the position in
 * xyz; returns the iteration count (-1 on bad arguments). */
long tops_lp2xyz(double line, double pixel, const tops_annotation* a, const tops_orbit* orbit,
                 int burst, double* xyz, int max_iter, double criter);
/* Replaces ComputerBurstOverlap (Func.cpp:1610-1746): burst0 / burstN are
 * the 1-based master bursts asked for (config.txt); outputs are the 0-based
 * matching master and slave burst ranges.  TOPS_ERR_ARG when the pair does
 * not overlap. */
int tops_burst_overlap(const tops_annotation* master, const tops_annotation* slave,
                       const tops_orbit* master_orbit, const tops_orbit* slave_orbit,
                       int burst0, int burstN, int* m_burst0, int* m_burstN, int* s_burst0,
                       int* s_burstN);

#ifdef __cplusplus
}
#endif
#endif
