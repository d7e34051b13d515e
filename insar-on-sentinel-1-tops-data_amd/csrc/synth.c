/*
 * synth.c -- seeded synthetic Sentinel-1 TOPS burst pairs and DEM tiles.
 *
 * The scene is a band-limited complex Gaussian field Z(T, X) on the
 * absolute azimuth-line / range-pixel grid of one subswath.  A burst's raw
 * sample is Z at its scene position times exp(-j*phi), phi being the TOPS
 * deramp phase of DerampAndResample.cu:151-158, so that deramping with the
 * same parameters returns the band-limited scene (what the 12-tap resampler
 * of the reference assumes).  Bursts of one subswath read the same scene at
 * their own absolute lines, so burst overlaps carry the TOPS spectral
 * separation that the reference's ESD (EstESD.cu) measures.
 *
 * A slave sample at (r_s, x_s) is the master raw signal evaluated at the
 * continuous master position (rho, xi) that the offset field maps onto
 * (r_s - delta, x_s): Z is interpolated with an 8x8 Kaiser-windowed sinc and
 * phi is evaluated analytically.  delta is an injected azimuth
 * misregistration (lines) that ESD must recover.
 *
 * Everything is a pure function of (seed, stream, absolute indices), so any
 * burst or DEM window is reproducible on its own.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define TSYN_PI 3.1415926535897931
#define TSYN_HALF 8            /* low-pass half length (17 taps)           */
#define TSYN_ITAPS 8           /* interpolation taps per axis              */

static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline uint64_t key4(uint64_t seed, uint64_t stream, int64_t a, int64_t b)
{
    uint64_t h = mix64(seed ^ mix64(stream + 0x1234567ULL));
    h = mix64(h ^ (uint64_t)a);
    return mix64(h ^ ((uint64_t)b * 0x9E3779B97F4A7C15ULL));
}

/* two independent N(0,1) from one counter (Box-Muller) */
static inline void gauss2(uint64_t k, double* g0, double* g1)
{
    uint64_t a = mix64(k), b = mix64(k ^ 0xD1B54A32D192ED03ULL);
    double u1 = ((a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    double u2 = ((b >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    double r = sqrt(-2.0 * log(u1));
    *g0 = r * cos(2.0 * TSYN_PI * u2);
    *g1 = r * sin(2.0 * TSYN_PI * u2);
}

static double bessel_i0(double x)
{
    double s = 1.0, t = 1.0;
    for (int k = 1; k < 40; k++) {
        t *= (x / (2.0 * k)) * (x / (2.0 * k));
        s += t;
    }
    return s;
}

static double kaiser(double x, double half, double beta)
{
    double r = x / half;
    if (r <= -1.0 || r >= 1.0) return 0.0;
    return bessel_i0(beta * sqrt(1.0 - r * r)) / bessel_i0(beta);
}

static double sincd(double x) { return x == 0.0 ? 1.0 : sin(TSYN_PI * x) / (TSYN_PI * x); }

/* low-pass FIR of band fraction fc (of Nyquist), unit output variance for
 * unit-variance white input */
static void lowpass(double fc, double* h)
{
    double e = 0.0;
    for (int k = -TSYN_HALF; k <= TSYN_HALF; k++) {
        double v = fc * sincd(fc * k) * kaiser((double)k, TSYN_HALF + 1.0, 5.0);
        h[k + TSYN_HALF] = v;
        e += v * v;
    }
    e = 1.0 / sqrt(e);
    for (int k = 0; k <= 2 * TSYN_HALF; k++) h[k] *= e;
}

/* 8-tap interpolation weights for fractional position f in [0,1): taps at
 * offsets -3..4.  Kaiser-windowed sinc, normalised to unit DC gain,
 * tabulated at 4096 steps and linearly interpolated (table error < 1e-7). */
#define TSYN_WSTEPS 4096
static double g_wtab[(TSYN_WSTEPS + 1) * TSYN_ITAPS];
static int g_wtab_ready = 0;

static void interp_table_init(void)
{
    if (g_wtab_ready) return;
    for (int k = 0; k <= TSYN_WSTEPS; k++) {
        double f = (double)k / TSYN_WSTEPS, s = 0.0;
        double* w = &g_wtab[k * TSYN_ITAPS];
        for (int i = 0; i < TSYN_ITAPS; i++) {
            double x = (i - 3) - f;
            w[i] = sincd(x) * kaiser(x, TSYN_ITAPS / 2 + 0.5, 4.0);
            s += w[i];
        }
        for (int i = 0; i < TSYN_ITAPS; i++) w[i] /= s;
    }
    g_wtab_ready = 1;
}

static inline void interp_weights(double f, double* w)
{
    double p = f * TSYN_WSTEPS;
    int k = (int)p;
    if (k >= TSYN_WSTEPS) k = TSYN_WSTEPS - 1;
    double t = p - k;
    const double* a = &g_wtab[k * TSYN_ITAPS];
    const double* b = a + TSYN_ITAPS;
    for (int i = 0; i < TSYN_ITAPS; i++) w[i] = a[i] * (1.0 - t) + b[i] * t;
}

/* Scene block: Z on rows [t0, t0+rows) and cols [c0, c0+cols) of the
 * absolute scene grid, interleaved complex float. */
static float* scene_block(uint64_t seed, uint64_t stream, double az_band,
                          double rg_band, int64_t t0, int rows, int64_t c0,
                          int cols)
{
    double haz[2 * TSYN_HALF + 1], hrg[2 * TSYN_HALF + 1];
    lowpass(az_band, haz);
    lowpass(rg_band, hrg);
    int wcols = cols + 2 * TSYN_HALF;
    int wrows = rows + 2 * TSYN_HALF;
    /* white noise, then range filter: tmp[wrows][cols] */
    float* tmp = (float*)malloc(sizeof(float) * 2 * (size_t)wrows * cols);
#pragma omp parallel
    {
        double* wr = (double*)malloc(sizeof(double) * 2 * wcols);
#pragma omp for schedule(static)
        for (int r = 0; r < wrows; r++) {
            int64_t ra = t0 - TSYN_HALF + r;
            for (int c = 0; c < wcols; c++) {
                int64_t ca = c0 - TSYN_HALF + c;
                gauss2(key4(seed, stream, ra, ca), &wr[2 * c], &wr[2 * c + 1]);
            }
            for (int c = 0; c < cols; c++) {
                double sr = 0.0, si = 0.0;
                for (int k = 0; k <= 2 * TSYN_HALF; k++) {
                    sr += hrg[k] * wr[2 * (c + k)];
                    si += hrg[k] * wr[2 * (c + k) + 1];
                }
                tmp[2 * ((size_t)r * cols + c)] = (float)sr;
                tmp[2 * ((size_t)r * cols + c) + 1] = (float)si;
            }
        }
        free(wr);
    }
    float* z = (float*)malloc(sizeof(float) * 2 * (size_t)rows * cols);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; r++)
        for (int c = 0; c < cols; c++) {
            double sr = 0.0, si = 0.0;
            for (int k = 0; k <= 2 * TSYN_HALF; k++) {
                size_t q = (size_t)(r + k) * cols + c;
                sr += haz[k] * tmp[2 * q];
                si += haz[k] * tmp[2 * q + 1];
            }
            z[2 * ((size_t)r * cols + c)] = (float)sr;
            z[2 * ((size_t)r * cols + c) + 1] = (float)si;
        }
    free(tmp);
    return z;
}

static inline int16_t clamp16(double v)
{
    double r = nearbyint(v);
    if (r > 32767.0) r = 32767.0;
    if (r < -32768.0) r = -32768.0;
    return (int16_t)r;
}

/* linear interpolation of a per-column array at continuous column xi */
static inline double col_interp(const double* a, int n, double xi)
{
    if (xi <= 0.0) return a[0];
    if (xi >= n - 1) return a[n - 1];
    int i = (int)xi;
    double f = xi - i;
    return a[i] * (1.0 - f) + a[i + 1] * f;
}

/* TOPS deramp phase at burst-relative continuous line rho (DerampAndResample.cu:151-158) */
static inline double tops_phase(double rho, double xi, int spb, double dt,
                                const double* kt, const double* tref,
                                const double* fdc)
{
    double ta = rho * dt;
    double d = ta - col_interp(tref, spb, xi);
    return -TSYN_PI * col_interp(kt, spb, xi) * d * d -
           2.0 * TSYN_PI * col_interp(fdc, spb, xi) * ta;
}

/* offset polynomial in u = rho/(L-1), v = xi/(P-1) */
static inline double off_eval(const double* c, double u, double v)
{
    return c[0] + c[1] * u + c[2] * v + c[3] * u * u + c[4] * u * v + c[5] * v * v;
}

/*
 * Generate one burst [lines x pixels] of CInt16.
 *   scene_line0   : absolute scene line of the burst's first line
 *   az_off/rg_off : NULL for a master burst, else the offset field (lines,
 *                   pixels) from master to slave position, 6 coefficients in
 *                   (u, v) each
 *   delta         : extra azimuth misregistration of the slave (lines)
 *   kt/tref/fdc   : per-column ramp of the raw signal [pixels]
 *   phase_noise   : std (rad) of i.i.d. phase noise (slave only)
 */
void tsyn_burst(uint64_t seed, uint64_t stream, double az_band, double rg_band,
                double sigma, int64_t scene_line0, int lines, int pixels,
                const double* az_off, const double* rg_off, double delta,
                double dt, const double* kt, const double* tref,
                const double* fdc, double phase_noise, uint64_t noise_stream,
                int16_t* out)
{
    const int margin = 12;
    int64_t t0 = scene_line0 - margin;
    int rows = lines + 2 * margin;
    int64_t c0 = -margin;
    int cols = pixels + 2 * margin;
    interp_table_init();
    float* z = scene_block(seed, stream, az_band, rg_band, t0, rows, c0, cols);
    const double lden = lines > 1 ? (double)(lines - 1) : 1.0;
    const double pden = pixels > 1 ? (double)(pixels - 1) : 1.0;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < lines; r++)
        for (int x = 0; x < pixels; x++) {
            double rho = r, xi = x;
            double zr, zi;
            if (az_off) {
                /* invert the smooth offset field by fixed-point iteration */
                for (int it = 0; it < 6; it++) {
                    double u = rho / lden, v = xi / pden;
                    rho = r - delta - off_eval(az_off, u, v);
                    xi = x - off_eval(rg_off, u, v);
                }
                double fr = floor(rho), fc = floor(xi);
                double wy[TSYN_ITAPS], wx[TSYN_ITAPS];
                interp_weights(rho - fr, wy);
                interp_weights(xi - fc, wx);
                int ry = (int)fr - 3 + margin, cx = (int)fc - 3 + margin;
                zr = zi = 0.0;
                for (int j = 0; j < TSYN_ITAPS; j++) {
                    int yy = ry + j;
                    if (yy < 0 || yy >= rows) continue;
                    double ar = 0.0, ai = 0.0;
                    for (int i = 0; i < TSYN_ITAPS; i++) {
                        int xx = cx + i;
                        if (xx < 0 || xx >= cols) continue;
                        size_t q = (size_t)yy * cols + xx;
                        ar += wx[i] * z[2 * q];
                        ai += wx[i] * z[2 * q + 1];
                    }
                    zr += wy[j] * ar;
                    zi += wy[j] * ai;
                }
            } else {
                size_t q = (size_t)(r + margin) * cols + (x + margin);
                zr = z[2 * q];
                zi = z[2 * q + 1];
            }
            double ph = -tops_phase(rho, xi, pixels, dt, kt, tref, fdc);
            if (phase_noise > 0.0) {
                double g0, g1;
                gauss2(key4(seed, noise_stream, scene_line0 + r, x), &g0, &g1);
                ph += phase_noise * g0;
            }
            double c = cos(ph), s = sin(ph);
            size_t o = (size_t)r * pixels + x;
            out[2 * o] = clamp16(sigma * (zr * c - zi * s));
            out[2 * o + 1] = clamp16(sigma * (zr * s + zi * c));
        }
    free(z);
}

/* ---- fractal DEM ------------------------------------------------------- */
static inline double lattice(uint64_t seed, int64_t i, int64_t j, int oct)
{
    return (double)(mix64(key4(seed, 777 + oct, i, j)) >> 11) * (1.0 / 9007199254740992.0);
}

static inline double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

static double value_noise(uint64_t seed, double y, double x, int oct)
{
    double fy = floor(y), fx = floor(x);
    int64_t iy = (int64_t)fy, ix = (int64_t)fx;
    double ty = smooth(y - fy), tx = smooth(x - fx);
    double a = lattice(seed, iy, ix, oct), b = lattice(seed, iy, ix + 1, oct);
    double c = lattice(seed, iy + 1, ix, oct), d = lattice(seed, iy + 1, ix + 1, oct);
    return (a * (1 - tx) + b * tx) * (1 - ty) + (c * (1 - tx) + d * tx) * ty;
}

/*
 * DEM window: row i, col j of the output is the global post
 * (row0 + i, col0 + j) of a grid with the given spacing (arc-seconds);
 * heights in metres in [0, hmax], rounded to int.  Posts for which
 * nodata_every > 0 and (global row + col) % nodata_every == 0 get nodata.
 */
void tsyn_dem(uint64_t seed, int64_t row0, int64_t col0, int lines,
              int pixels, double spacing_arcsec, double hmax,
              int nodata_every, int nodata, int32_t* out)
{
    /* feature scale in posts: about 0.02 deg for the coarsest octave */
    double base = 72.0 / spacing_arcsec * 3.0;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < lines; i++)
        for (int j = 0; j < pixels; j++) {
            double y = (double)(row0 + i), x = (double)(col0 + j);
            double h = 0.0, amp = 1.0, norm = 0.0, sc = base;
            for (int o = 0; o < 6; o++) {
                h += amp * value_noise(seed, y / sc, x / sc, o);
                norm += amp;
                amp *= 0.5;
                sc *= 0.5;
            }
            int32_t v = (int32_t)nearbyint(hmax * h / norm);
            if (nodata_every > 0 && ((row0 + i + col0 + j) % nodata_every) == 0) v = nodata;
            out[(size_t)i * pixels + j] = v;
        }
}
