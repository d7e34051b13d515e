// walker.hpp -- device pieces shared by the column-walking resample kernels
// (resample.hip: one wave per strip; fused.hip: the persistent fused
// resample + coherence kernel): slave position and region rules
// (DerampAndResample.cu:108-112, 348-403), the deramp phase coefficients
// (DerampAndResample.cu:117-183), the per-lane column walker with its FP64
// phasor recurrence, the 12-tap azimuth sums, the reramp and the direct
// per-output path.
#pragma once

#include "tops_internal.hpp"

namespace tops {

// AzimuthShift: a device value (the device-resident ESD flow) or the
// parameter-block value
__device__ inline double az_shift_of(const ResampleParams& p)
{
    return p.az_shift_dev ? *p.az_shift_dev : p.az_shift;
}

__device__ inline void slave_pos(const ResampleParams& p, int r, int c, int& iy,
                                 int& fy, int& ix, int& fx)
{
#pragma clang fp contract(off)
    double pn = norm_warp((double)(p.mb[0] + c), (double)p.mb[0], (double)p.mb[1]);
    double ln = norm_warp((double)(p.mb[2] + r), (double)p.mb[2], (double)p.mb[3]);
    double ys = polyval6(ln, pn, p.cpm_az) + az_shift_of(p);
    double xs = polyval6(ln, pn, p.cpm_rg);
    // keep the int conversion defined for wild polynomials; anything this
    // far out lands in the zero region anyway
    ys = fmin(fmax(ys, -1.0e9), 1.0e9);
    xs = fmin(fmax(xs, -1.0e9), 1.0e9);
    iy = (int)trunc(ys);
    fy = (int)((ys - iy) * kInterval + 0.5);
    ix = (int)trunc(xs);
    fx = (int)((xs - ix) * kInterval + 0.5);
}

// 0: outside, 1: edge band, 2: interior (DerampAndResample.cu:383-403)
__device__ inline int region_of(const ResampleParams& p, int iy, int ix)
{
    if (iy > p.sb[3] || iy < p.sb[2] || ix > p.sb[1] || ix < p.sb[0]) return 0;
    if (iy > p.sb[3] - kHalf - 2 || iy < p.sb[2] + kHalf || ix > p.sb[1] - kHalf - 2 ||
        ix < p.sb[0] + kHalf)
        return 1;
    return 2;
}

// range of c0 + c1 x + c2 y + c3 x^2 + c4 x y + c5 y^2 over the box
// [x0, x1] x [y0, y1] (x0 <= x1, y0 <= y1), by interval arithmetic on each
// term; lo > hi (an empty range) when an input is not finite
__device__ inline void poly6_range(const double* c, double x0, double x1, double y0, double y1,
                                   double& lo, double& hi)
{
    auto lin = [](double k, double a, double b, double& l, double& h) {
        l = fmin(k * a, k * b);
        h = fmax(k * a, k * b);
    };
    auto sq = [](double a, double b, double& l, double& h) {   // range of t^2, t in [a, b]
        h = fmax(a * a, b * b);
        l = (a <= 0.0 && b >= 0.0) ? 0.0 : fmin(a * a, b * b);
    };
    double l, h, tl, th;
    lo = hi = c[0];
    lin(c[1], x0, x1, l, h); lo += l; hi += h;
    lin(c[2], y0, y1, l, h); lo += l; hi += h;
    sq(x0, x1, tl, th); lin(c[3], tl, th, l, h); lo += l; hi += h;
    const double p0 = x0 * y0, p1 = x0 * y1, p2 = x1 * y0, p3 = x1 * y1;
    tl = fmin(fmin(p0, p1), fmin(p2, p3));
    th = fmax(fmax(p0, p1), fmax(p2, p3));
    lin(c[4], tl, th, l, h); lo += l; hi += h;
    sq(y0, y1, tl, th); lin(c[5], tl, th, l, h); lo += l; hi += h;
    if (!(lo <= hi) || !isfinite(lo) || !isfinite(hi)) {
        lo = 1.0;
        hi = 0.0;
    }
}

// Whether every output of rows r0 .. r1-1, columns c0 .. c1 of the master
// box is interior (region 2): the ranges of both position polynomials over
// the block lie, with a margin, inside the slave box's interior band.  Then
// the walker need not test the region per output.
__device__ inline bool block_interior(const ResampleParams& p, int r0, int r1, int c0, int c1)
{
#pragma clang fp contract(off)
    const double l0 = norm_warp((double)(p.mb[2] + r0), (double)p.mb[2], (double)p.mb[3]);
    const double l1 = norm_warp((double)(p.mb[2] + r1 - 1), (double)p.mb[2], (double)p.mb[3]);
    const double q0 = norm_warp((double)(p.mb[0] + c0), (double)p.mb[0], (double)p.mb[1]);
    const double q1 = norm_warp((double)(p.mb[0] + c1), (double)p.mb[0], (double)p.mb[1]);
    double ylo, yhi, xlo, xhi;
    poly6_range(p.cpm_az, fmin(l0, l1), fmax(l0, l1), fmin(q0, q1), fmax(q0, q1), ylo, yhi);
    poly6_range(p.cpm_rg, fmin(l0, l1), fmax(l0, l1), fmin(q0, q1), fmax(q0, q1), xlo, xhi);
    if (!(ylo <= yhi) || !(xlo <= xhi)) return false;
    const double azs = az_shift_of(p);
    ylo += azs;
    yhi += azs;
    // interior: kHalf <= trunc(pos) - box start <= box end - kHalf - 2, so
    // pos in [start + kHalf, end - kHalf - 1); a margin of 1e-3 px plus 1e-9
    // of the magnitude covers the rounding of both evaluations
    const double my = 1e-3 + 1e-9 * (fabs(ylo) + fabs(yhi));
    const double mx = 1e-3 + 1e-9 * (fabs(xlo) + fabs(xhi));
    return ylo >= p.sb[2] + kHalf + my && yhi < p.sb[3] - kHalf - 1 - my &&
           xlo >= p.sb[0] + kHalf + mx && xhi < p.sb[1] - kHalf - 1 - mx;
}

__device__ inline double deramp_phase_direct(const ResampleParams& p, int ya, int xa)
{
    double ta = (ya - p.first_line) * p.dt;
    double d = ta - p.tref[xa];
    return -kPi * p.kt[xa] * (d * d) - 2.0 * kPi * p.fdc[xa] * ta;
}

__device__ inline void phase_coeffs(const ResampleParams& p, int xa, double& a,
                                    double& b, double& c)
{
    double kt = p.kt[xa], tr = p.tref[xa], fd = p.fdc[xa];
    a = -kPi * kt * (tr * tr);
    b = 2.0 * kPi * p.dt * (kt * tr - fd);
    c = -kPi * kt * (p.dt * p.dt);
}

__device__ inline float reduce_2pi(double ph)
{
    double k = rint(ph * kInvTwoPi);
    return (float)fma(-k, kTwoPi, ph);
}

// output columns per strip (64-lane wave): 52 + 11 tap columns + 1 column of
// slack for the slave column drifting along the strip.  52 outputs are 416
// bytes of CF32: with strips starting on 4-column boundaries and rows of a
// multiple of 4 samples, every strip row fills whole 32-byte sectors.
constexpr int WK_OUT = 52;
constexpr int WK_WIN = 16;     // register window rows
constexpr int WK_KEYS = 4;     // azimuth keys per row served by the shared pass
constexpr int kWalkWavesPerSimd = 4;   // occupancy the walker is compiled for

__device__ inline void load_taps(const float* lut, int row, float k[kTaps])
{
    const float4* q = reinterpret_cast<const float4*>(lut + row * kTaps);
    float4 a = q[0], b = q[1], c = q[2];
    k[0] = a.x; k[1] = a.y; k[2] = a.z; k[3] = a.w;
    k[4] = b.x; k[5] = b.y; k[6] = b.z; k[7] = b.w;
    k[8] = c.x; k[9] = c.y; k[10] = c.z; k[11] = c.w;
}

// out = (sr + j si) * exp(-j phi).  The phase goes to revolutions and is
// reduced to [-1/2, 1/2] in FP64; the hardware sine and cosine take
// revolutions directly (absolute error about 1.3e-7 on that interval).
__device__ inline tops_cf32 reramp(float sr, float si, double phi)
{
    double t = phi * kInvTwoPi;
    float tf = (float)(t - rint(t));
    float sn = __builtin_amdgcn_sinf(tf), cs = __builtin_amdgcn_cosf(tf);
    tops_cf32 o;
    o.x = sr * cs + si * sn;
    o.y = si * cs - sr * sn;
    return o;
}

// Interior output computed straight from HBM (outputs the shared pass does
// not serve).  Inline with rolled loops and one tap loaded per step, so it
// stays small in registers; it takes plain values, not the parameter block.
__device__ __forceinline__ tops_cf32 interior_direct(const tops_ci16* slave, const double* kt,
                                                     const double* tref, const double* fdc,
                                                     const double* mom, const float* kaz,
                                                     const float* krg, int s_pixels, int sx0,
                                                     int sy0, int first_line, double dt, int iy,
                                                     int fy, int ix, int fx)
{
    float sr = 0.f, si = 0.f;
    double A = 0.0, B = 0.0, C = 0.0;
#pragma unroll 1
    for (int j = 0; j < kTaps; j++) {
        const int ya = iy - kHalf + j;
        const double y = (double)(ya - first_line);
        float tr = 0.f, ti = 0.f;
#pragma unroll 1
        for (int i = 0; i < kTaps; i++) {
            const int xa = ix - kHalf + i;
            const float kpi = krg[fx * kTaps + i];
            double k = kt[xa], t = tref[xa], f = fdc[xa];
            double a = -kPi * k * (t * t);
            double b = 2.0 * kPi * dt * (k * t - f);
            double c = -kPi * k * (dt * dt);
            float sn, cs;
            sincosf(reduce_2pi(fma(y, fma(y, c, b), a)), &sn, &cs);
            tops_ci16 v = slave[(size_t)(ya - sy0) * s_pixels + (xa - sx0)];
            float dr = (float)v.re * cs - (float)v.im * sn;
            float di = (float)v.re * sn + (float)v.im * cs;
            tr = fmaf(kpi, dr, tr);
            ti = fmaf(kpi, di, ti);
            if (j == 0) {
                A = fma((double)kpi, a, A);
                B = fma((double)kpi, b, B);
                C = fma((double)kpi, c, C);
            }
        }
        const float klj = kaz[fy * kTaps + j];
        sr = fmaf(klj, tr, sr);
        si = fmaf(klj, ti, si);
    }
    double4 m = reinterpret_cast<const double4*>(mom)[fy];
    double y0 = (double)(iy - first_line);
    double phi = m.x * A + fma(y0, m.x, m.y) * B + (y0 * y0 * m.x + 2.0 * y0 * m.y + m.z) * C;
    return reramp(sr, si, phi);
}

// Column state of one lane: deramp recurrence along its slave column, with
// the raw samples of the next four lines already requested from HBM.
// word 3 of a raw (untyped, unswizzled) buffer resource on gfx9 parts
constexpr int kRawBufferWord3 = 0x00020000;

struct ColumnWalker {
    int xcol;            // absolute slave column
    bool in_box;         // column inside the slave box
    unsigned xoff;       // byte offset in a slave row of the column clamped into the box
    double pr, pi;       // exp(j phi(ynext))
    double ir, ii;       // exp(j (b + c (2 ynext + 1)))
    double qr, qi;       // exp(j 2c)
    int ynext;           // absolute slave line deramped next
    int pend[4];         // raw CInt16 of lines ynext .. ynext+3

    // raw CInt16 of line y (inside the box) of this column: a buffer load
    // with the line's byte offset as the scalar offset and the column's as
    // the lane offset, so it takes no vector address arithmetic.  The
    // launchers keep the slave raster under 2^31 bytes.
    __device__ int load_line(const ResampleParams& p, int y) const
    {
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<tops_ci16*>(p.slave), 0, p.s_lines * p.s_pixels * 4, kRawBufferWord3);
        return __builtin_amdgcn_raw_buffer_load_b32(r, (int)xoff, (y - p.sb[2]) * p.s_pixels * 4, 0);
    }

    __device__ void init(const ResampleParams& p, int col)
    {
        xcol = col;
        in_box = col >= p.sb[0] && col <= p.sb[1];
        xoff = 4u * (unsigned)(min(max(col, p.sb[0]), p.sb[1]) - p.sb[0]);
    }

    // raw CInt16 (packed re | im << 16) of line y of this column, 0 for a
    // line outside the box (y is wave-uniform: a scalar branch).  A column
    // outside the box loads its clamped neighbour; deramp() drops that
    // value when it consumes it, so the load stays in flight until then.
    __device__ int fetch(const ResampleParams& p, int y) const
    {
        if (y < p.sb[2] || y > p.sb[3]) return 0;
        return load_line(p, y);
    }

    // phi(y) = a + b y + c y^2 (y burst-relative line), anchored at y_abs
    __device__ void anchor(const ResampleParams& p, int y_abs, double a, double b, double c)
    {
        ynext = y_abs;
        double y = (double)(y_abs - p.first_line);
        double s, cc;
        double ph = fma(y, fma(y, c, b), a);
        sincos(fma(-rint(ph * kInvTwoPi), kTwoPi, ph), &s, &cc);
        pr = cc; pi = s;
        double d = fma(c, 2.0 * y + 1.0, b);
        sincos(fma(-rint(d * kInvTwoPi), kTwoPi, d), &s, &cc);
        ir = cc; ii = s;
    }

    // the raw value as floats (0 for a column outside the box)
    __device__ float2 unpack(int raw) const
    {
        raw = in_box ? raw : 0;
        return make_float2((float)(short)(raw & 0xffff), (float)(short)(raw >> 16));
    }

    // deramped sample of line ynext from its raw value, then advance
    __device__ float2 deramp(int raw) { return deramp(unpack(raw)); }

    __device__ float2 deramp(float2 v)
    {
        const float re = v.x, im = v.y;
        const float cr = (float)pr, ci = (float)pi;
        float2 d = make_float2(re * cr - im * ci, re * ci + im * cr);
        double nr = pr * ir - pi * ii, ni = pr * ii + pi * ir;
        pr = nr; pi = ni;
        double jr = ir * qr - ii * qi, ji = ir * qi + ii * qr;
        ir = jr; ii = ji;
        ynext++;
        return d;
    }

    // request lines ynext + skip .. ynext + skip + 3 (one scalar test when
    // all four lie inside the box): skip = 0 after the pending lines were
    // deramped, 4 while they still wait
    __device__ void prefetch(const ResampleParams& p, int skip = 0)
    {
        const int y0 = ynext + skip;
        if (y0 >= p.sb[2] && y0 + 3 <= p.sb[3]) {
#pragma unroll
            for (int k = 0; k < 4; k++) pend[k] = load_line(p, y0 + k);
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) pend[k] = fetch(p, y0 + k);
        }
    }
};

// A window row: one deramped sample as a 64-bit value (real part in the low
// half).  Kept as one 64-bit register pair, the window's slide by 4 rows is
// 12 64-bit moves instead of 24 32-bit ones.
typedef unsigned long long WinRow;
__device__ __forceinline__ WinRow win_pack(float2 v)
{
    return (WinRow)__float_as_uint(v.x) | ((WinRow)__float_as_uint(v.y) << 32);
}
__device__ __forceinline__ float win_re(WinRow w) { return __uint_as_float((unsigned)w); }
__device__ __forceinline__ float win_im(WinRow w) { return __uint_as_float((unsigned)(w >> 32)); }

// Slide a window by 4 rows: the 4 oldest rows out, the 4 pending rows
// deramped in.  The raw values are unpacked and only then are the next 4
// lines requested (the asm statement orders the two): the loads then land
// in the registers the raw values held, and stay in flight until the next
// slide.  (Requested earlier, the compiler copies the landing registers at
// once, which waits for the loads.)
template <typename Win>
__device__ __forceinline__ void slide4(Win (&win)[WK_WIN], ColumnWalker& w,
                                       const ResampleParams& p)
{
    float2 v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = w.unpack(w.pend[k]);
    asm volatile("" : "+v"(v[0].x), "+v"(v[0].y), "+v"(v[1].x), "+v"(v[1].y), "+v"(v[2].x),
                 "+v"(v[2].y), "+v"(v[3].x), "+v"(v[3].y)::"memory");
    w.prefetch(p, 4);   // the 4 lines after the ones unpacked
#pragma unroll
    for (int k = 0; k < WK_WIN - 4; k++) win[k] = win[k + 4];
#pragma unroll
    for (int k = 0; k < 4; k++) win[WK_WIN - 4 + k] = win_pack(w.deramp(v[k]));
}

// read-only tables addressed with wave-uniform indices
typedef __attribute__((address_space(4))) const float CFloat;
typedef __attribute__((address_space(4))) const double CDouble;

// 12-tap azimuth sum of one column over window rows oq .. oq+11: the key's
// taps pre-shifted by oq into a 16-tap row of the context's table (zeros
// elsewhere), so one unrolled 16-term sum serves every window offset.  (A
// 12-term sum per offset behind a scalar branch measured 6% slower on the
// IW pair.)
static_assert(kAzShiftTaps == WK_WIN, "a shifted tap row spans the window");
__device__ __forceinline__ float2 az_taps_shifted(const WinRow (&win)[WK_WIN], const float* kaz_sh,
                                                  int oq, int fyq)
{
    CFloat* taps = (CFloat*)(kaz_sh + ((size_t)oq * kLutRows + fyq) * kAzShiftTaps);
    float2 u = make_float2(0.f, 0.f);
#pragma unroll
    for (int t = 0; t < kAzShiftTaps; t++) {
        const float k = taps[t];
        u.x = fmaf(k, win_re(win[t]), u.x);
        u.y = fmaf(k, win_im(win[t]), u.y);
    }
    return u;
}

}  // namespace tops
