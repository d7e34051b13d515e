/*
 * tops_abi.h -- C-ABI of the MI355X TOPS burst library (libtops_mi355x.so).
 *
 * Part 1 keeps the seven extern "C" entry points of the reference
 * a787854/InSAR-on-Sentinel-1-TOPS-data with identical names, parameter
 * order and meaning, so its Func.cpp can link against this library
 * unchanged.  std::complex<short> is laid out as tops_ci16 and cuComplex /
 * std::complex<float> as tops_cf32.  Differences in behaviour:
 *   - no exit() and no device reset inside the library; errors set a status
 *     readable with tops_last_error() and are printed once to stderr;
 *   - per-call parameter blocks instead of module globals (reentrant);
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

#ifdef __cplusplus
}
#endif
#endif
