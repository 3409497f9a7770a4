/*
 * CPU oracle: a plain-C restatement of the AutoRIFT.jl dense ZNCC tracker.
 *
 * TEST INFRASTRUCTURE ONLY -- the checker, never the product. Only the test suite, the smoke
 * check and the CPU-baseline leg of the benchmark load it. Each function cites the reference
 * file:line it follows (paths relative to the AutoRIFT.jl checkout); operation order follows the
 * reference so that results can be compared bit for bit. Build without FP contraction.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------ */
/* element access                                                                             */
/* ------------------------------------------------------------------------------------------ */

static inline double ldv(const void* p, int dt, int64_t idx) {
    switch (dt) {
        case ORC_U8: return (double)((const uint8_t*)p)[idx];
        case ORC_F32: return (double)((const float*)p)[idx];
        case ORC_F64: return ((const double*)p)[idx];
        case ORC_I16: return (double)((const int16_t*)p)[idx];
    }
    return NAN;
}

/* Float32(x) of an element: exact for u8/i16/f32, rounds for f64 (Julia's Float32(::Float64)). */
static inline float ldf(const void* p, int dt, int64_t idx) {
    switch (dt) {
        case ORC_U8: return (float)((const uint8_t*)p)[idx];
        case ORC_F32: return ((const float*)p)[idx];
        case ORC_F64: return (float)((const double*)p)[idx];
        case ORC_I16: return (float)((const int16_t*)p)[idx];
    }
    return NAN;
}

static inline size_t esize(int dt) {
    switch (dt) {
        case ORC_U8: return 1;
        case ORC_F32: return 4;
        case ORC_F64: return 8;
        case ORC_I16: return 2;
        case ORC_C64: return 8;
    }
    return 0;
}

static inline const void* elem_ptr(const void* p, int dt, int64_t idx) {
    return (const char*)p + (size_t)idx * esize(dt);
}

/* Julia max(x, 0.0) / min semantics: NaN propagates. */
static inline double jmax0(double x) { return isnan(x) ? x : (x < 0.0 ? 0.0 : x); }
static inline double jmaxd(double a, double b) {
    if (isnan(a) || isnan(b)) return NAN;
    return a < b ? b : a;
}

/* ------------------------------------------------------------------------------------------ */
/* next_fft_size -- src/correlate.jl:373-403                                                  */
/* ------------------------------------------------------------------------------------------ */

int64_t orc_next_fft_size(int64_t n) {
    if (n <= 1) return 1;
    int64_t best = 0;
    double best_cost = INFINITY;
    int64_t hi = (int64_t)ceil(1.5 * (double)n);
    for (int64_t m = n; m <= hi; m++) {
        int64_t r = m, twos = 0;
        while ((r & 1) == 0) { r /= 2; twos++; }
        const int64_t ps[3] = {3, 5, 7};
        for (int k = 0; k < 3; k++)
            while (r % ps[k] == 0) r /= ps[k];
        if (r != 1) continue;
        double l2 = log2((double)m);
        double frac2 = (double)twos / l2;
        double cost = (double)(m * m) * l2 * (1.0 - 0.35 * frac2);
        if (cost < best_cost) { best = m; best_cost = cost; }
    }
    return best;
}

/* ------------------------------------------------------------------------------------------ */
/* prepare_chip! -- src/correlate.jl:423-453                                                   */
/* ------------------------------------------------------------------------------------------ */

static int prepare_chip(const void* chip, int64_t ld, int dt, int64_t ch, int64_t cw, float* dst,
                        double* cnorm) {
    double s = 0.0;
    for (int64_t j = 0; j < cw; j++)
        for (int64_t i = 0; i < ch; i++) s += ldv(chip, dt, i + j * ld);
    double mean = s / (double)(ch * cw);
    double sq = 0.0;
    for (int64_t j = 0; j < cw; j++)
        for (int64_t i = 0; i < ch; i++) {
            double d = ldv(chip, dt, i + j * ld) - mean;
            dst[i + j * ch] = (float)d;
            sq += d * d;
        }
    int ok = sq > DBL_EPSILON * (double)ch * (double)cw;
    *cnorm = sqrt(sq);
    return ok;
}

/* ------------------------------------------------------------------------------------------ */
/* integral_both! -- src/integral.jl:141-170; boxsum -- src/integral.jl:244-247               */
/* ------------------------------------------------------------------------------------------ */

void orc_integral_both(const void* a, int64_t ld, int dt, int64_t m, int64_t n, double* S,
                       double* S2) {
    int64_t ls = m + 1;
    for (int64_t j = 0; j <= n; j++) { S[j * ls] = 0.0; S2[j * ls] = 0.0; }
    for (int64_t i = 0; i <= m; i++) { S[i] = 0.0; S2[i] = 0.0; }
    for (int64_t j = 1; j <= n; j++) {
        double run = 0.0, runsq = 0.0;
        for (int64_t i = 1; i <= m; i++) {
            double v = ldv(a, dt, (i - 1) + (j - 1) * ld);
            run += v;
            runsq += v * v;
            S[i + j * ls] = run + S[i + (j - 1) * ls];
            S2[i + j * ls] = runsq + S2[i + (j - 1) * ls];
        }
    }
}

/* 1-based (i, j) into a table with leading dimension ls. */
static inline double boxsum(const double* S, int64_t ls, int64_t i, int64_t j, int64_t h,
                            int64_t w) {
#define SX(r, c) S[((r) - 1) + ((c) - 1) * ls]
    return SX(i + h, j + w) - SX(i, j + w) - SX(i + h, j) + SX(i, j);
#undef SX
}

/* ------------------------------------------------------------------------------------------ */
/* numerators: direct (src/correlate.jl:799-814) and FFT (src/correlate.jl:816-848)           */
/* ------------------------------------------------------------------------------------------ */

static void numerators_direct(double* out, const void* search, int64_t lds, int dt,
                              const float* cd, int64_t nr, int64_t nc, int64_t ch, int64_t cw) {
    for (int64_t j = 0; j < nc; j++)
        for (int64_t i = 0; i < nr; i++) {
            double acc = 0.0;
            for (int64_t jj = 0; jj < cw; jj++)
                for (int64_t ii = 0; ii < ch; ii++)
                    acc += (double)cd[ii + jj * ch] * ldv(search, dt, (i + ii) + (j + jj) * lds);
            out[i + j * nr] = acc;
        }
}

/* Mixed-radix (2,3,5,7) complex FFT in single precision, standing in for FFTW's Float32
 * transforms. Not FFTW's arithmetic: FFT-mode numerators agree with the exact numerator to
 * single-precision rounding, as FFTW's do (the reference's own tests pin this mode only to 1e-4). */
typedef struct { float re, im; } cf;

typedef struct {
    int64_t n;
    int nf;
    int64_t fac[32];
    cf* tw; /* n twiddles exp(-2 pi i k / n) */
} fft_plan;

static void plan_init(fft_plan* p, int64_t n) {
    p->n = n;
    p->nf = 0;
    int64_t r = n;
    const int64_t ps[4] = {4, 2, 3, 5};
    while (r % 4 == 0) { p->fac[p->nf++] = 4; r /= 4; }
    while (r % 2 == 0) { p->fac[p->nf++] = 2; r /= 2; }
    for (int k = 2; k < 4; k++)
        while (r % ps[k] == 0) { p->fac[p->nf++] = ps[k]; r /= ps[k]; }
    while (r % 7 == 0) { p->fac[p->nf++] = 7; r /= 7; }
    p->tw = (cf*)malloc(sizeof(cf) * (size_t)n);
    for (int64_t k = 0; k < n; k++) {
        double a = -2.0 * M_PI * (double)k / (double)n;
        p->tw[k].re = (float)cos(a);
        p->tw[k].im = (float)sin(a);
    }
}

static void plan_free(fft_plan* p) { free(p->tw); }

/* Recursive decimation in time: x (stride xs) -> y (contiguous), length n, sign -1 forward. */
static void fft_rec(const fft_plan* p, int fi, const cf* x, int64_t xs, cf* y, int64_t n,
                    int inverse) {
    if (n == 1) { y[0] = x[0]; return; }
    int64_t rad = p->fac[fi];
    int64_t m = n / rad;
    for (int64_t q = 0; q < rad; q++) fft_rec(p, fi + 1, x + q * xs, xs * rad, y + q * m, m, inverse);
    int64_t tstep = p->n / n;
    cf tmp[8];
    for (int64_t k = 0; k < m; k++) {
        cf in[8];
        for (int64_t q = 0; q < rad; q++) {
            cf v = y[q * m + k];
            int64_t ti = (q * k * tstep) % p->n;
            cf w = p->tw[ti];
            if (inverse) w.im = -w.im;
            in[q].re = v.re * w.re - v.im * w.im;
            in[q].im = v.re * w.im + v.im * w.re;
        }
        for (int64_t o = 0; o < rad; o++) {
            float sr = 0.f, si = 0.f;
            for (int64_t q = 0; q < rad; q++) {
                int64_t ti = ((q * o) % rad) * (p->n / rad);
                cf w = p->tw[ti];
                if (inverse) w.im = -w.im;
                sr += in[q].re * w.re - in[q].im * w.im;
                si += in[q].re * w.im + in[q].im * w.re;
            }
            tmp[o].re = sr;
            tmp[o].im = si;
        }
        for (int64_t o = 0; o < rad; o++) y[o * m + k] = tmp[o];
    }
}

/* In-place 2-D transform of an F x F column-major complex array. */
static void fft2(const fft_plan* p, cf* a, cf* line_in, cf* line_out, int inverse) {
    int64_t F = p->n;
    for (int64_t j = 0; j < F; j++) { /* columns are contiguous */
        memcpy(line_in, a + j * F, sizeof(cf) * (size_t)F);
        fft_rec(p, 0, line_in, 1, a + j * F, F, inverse);
    }
    for (int64_t i = 0; i < F; i++) {
        fft_rec(p, 0, a + i, F, line_out, F, inverse);
        for (int64_t j = 0; j < F; j++) a[i + j * F] = line_out[j];
    }
}

typedef struct {
    int64_t F;
    fft_plan plan;
    cf* z;
    cf* line_in;
    cf* line_out;
} fft_ws;

static void fftws_ensure(fft_ws* w, int64_t F) {
    if (w->F == F) return;
    if (w->F) { plan_free(&w->plan); free(w->z); free(w->line_in); free(w->line_out); }
    w->F = F;
    plan_init(&w->plan, F);
    w->z = (cf*)malloc(sizeof(cf) * (size_t)(F * F));
    w->line_in = (cf*)malloc(sizeof(cf) * (size_t)F);
    w->line_out = (cf*)malloc(sizeof(cf) * (size_t)F);
}

static void fftws_free(fft_ws* w) {
    if (w->F) { plan_free(&w->plan); free(w->z); free(w->line_in); free(w->line_out); }
    w->F = 0;
}

/* Transforms the zero-padded window and the reversed chip separately (real input, zero imaginary
 * part), multiplies the spectra and inverts. The valid block is read at offset (ch, cw) as in
 * src/correlate.jl:844-846. Square transform of side F = next_fft_size(max extent). */
static void numerators_fft(double* out, fft_ws* w, const void* search, int64_t lds, int dt,
                           const float* cd, int64_t nr, int64_t nc, int64_t ch, int64_t cw) {
    int64_t sh = ch + nr - 1, sw = cw + nc - 1;
    int64_t fy = orc_next_fft_size(sh), fx = orc_next_fft_size(sw);
    int64_t F = fy > fx ? fy : fx;
    fftws_ensure(w, F);
    cf* a = w->z;
    cf* b = (cf*)calloc((size_t)(F * F), sizeof(cf));
    memset(a, 0, sizeof(cf) * (size_t)(F * F));
    for (int64_t j = 0; j < sw; j++)
        for (int64_t i = 0; i < sh; i++) a[i + j * F].re = (float)ldv(search, dt, i + j * lds);
    for (int64_t j = 0; j < cw; j++)
        for (int64_t i = 0; i < ch; i++) b[i + j * F].re = cd[(ch - 1 - i) + (cw - 1 - j) * ch];
    fft2(&w->plan, a, w->line_in, w->line_out, 0);
    fft2(&w->plan, b, w->line_in, w->line_out, 0);
    for (int64_t k = 0; k < F * F; k++) {
        cf u = a[k], v = b[k];
        a[k].re = u.re * v.re - u.im * v.im;
        a[k].im = u.re * v.im + u.im * v.re;
    }
    fft2(&w->plan, a, w->line_in, w->line_out, 1);
    float scale = 1.0f / (float)(F * F);
    for (int64_t j = 0; j < nc; j++)
        for (int64_t i = 0; i < nr; i++)
            out[i + j * nr] = (double)(a[(i + ch - 1) + (j + cw - 1) * F].re * scale);
    free(b);
}

/* ------------------------------------------------------------------------------------------ */
/* correlate! -- src/correlate.jl:528-557; ZNCC surface 575-602; NCC surface 604-630          */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    float* chip;
    double* S;
    double* S2;
    double* num;
    fft_ws fft;
} corr_ws;

static void cws_init(corr_ws* w, int64_t maxch, int64_t maxcw, int64_t maxry, int64_t maxrx) {
    int64_t sh = maxch + 2 * maxry, sw = maxcw + 2 * maxrx;
    /* room for a complex chip (two floats per sample) and a complex sum table (S holds the real
     * parts, S + (sh + 1)(sw + 1) the imaginary ones) */
    w->chip = (float*)malloc(sizeof(float) * (size_t)(2 * maxch * maxcw + 1));
    w->S = (double*)malloc(sizeof(double) * (size_t)(2 * (sh + 1) * (sw + 1)));
    w->S2 = (double*)malloc(sizeof(double) * (size_t)((sh + 1) * (sw + 1)));
    w->num = (double*)malloc(sizeof(double) * (size_t)(8 * maxry * maxrx + 1));
    w->fft.F = 0;
}

static void cws_free(corr_ws* w) {
    free(w->chip); free(w->S); free(w->S2); free(w->num);
    fftws_free(&w->fft);
}

/* ------------------------------------------------------------------------------------------ */
/* Coherence over ComplexF32 images                                                            */
/* ------------------------------------------------------------------------------------------ */

/* prepare_chip! for complex input -- src/correlate.jl:480-502: ComplexF64 mean in column-major
 * order, the chip stored as ComplexF32(z - mean), cnorm = sqrt(sum |z - mean|^2) in Float64
 * (abs2 = re*re + im*im, two roundings and a third for the sum). */
static int prepare_cchip(const float* chip, int64_t ld, int64_t ch, int64_t cw, float* dst,
                         double* cnorm) {
    double sr = 0.0, si = 0.0;
    for (int64_t j = 0; j < cw; j++)
        for (int64_t i = 0; i < ch; i++) {
            sr += (double)chip[2 * (i + j * ld)];
            si += (double)chip[2 * (i + j * ld) + 1];
        }
    const double n = (double)(ch * cw);
    const double mr = sr / n, mi = si / n;
    double sq = 0.0;
    for (int64_t j = 0; j < cw; j++)
        for (int64_t i = 0; i < ch; i++) {
            const double dr = (double)chip[2 * (i + j * ld)] - mr;
            const double di = (double)chip[2 * (i + j * ld) + 1] - mi;
            dst[2 * (i + j * ch)] = (float)dr;
            dst[2 * (i + j * ch) + 1] = (float)di;
            sq += dr * dr + di * di;
        }
    *cnorm = sqrt(sq);
    return sq > DBL_EPSILON * n;
}

/* integral_both! for complex input -- src/integral.jl:172-200: Sr / Si the complex running sum,
 * S2 the sum of |w|^2 (abs2 of a ComplexF32 value is exact product terms, one rounding). */
static void cintegral_both(const float* a, int64_t ld, int64_t m, int64_t n, double* Sr,
                           double* Si, double* S2) {
    int64_t ls = m + 1;
    for (int64_t j = 0; j <= n; j++) { Sr[j * ls] = 0.0; Si[j * ls] = 0.0; S2[j * ls] = 0.0; }
    for (int64_t i = 0; i <= m; i++) { Sr[i] = 0.0; Si[i] = 0.0; S2[i] = 0.0; }
    for (int64_t j = 1; j <= n; j++) {
        double rr = 0.0, ri = 0.0, runsq = 0.0;
        for (int64_t i = 1; i <= m; i++) {
            const double vr = (double)a[2 * ((i - 1) + (j - 1) * ld)];
            const double vi = (double)a[2 * ((i - 1) + (j - 1) * ld) + 1];
            rr += vr;
            ri += vi;
            runsq += vr * vr + vi * vi;
            Sr[i + j * ls] = rr + Sr[i + (j - 1) * ls];
            Si[i + j * ls] = ri + Si[i + (j - 1) * ls];
            S2[i + j * ls] = runsq + S2[i + (j - 1) * ls];
        }
    }
}

/* _correlate_surface!(::Coherence) -- src/correlate.jl:659-683, with the direct complex
 * numerator of _cnumerators_direct! (src/correlate.jl:712-730): per shift a ComplexF64
 * accumulation of conj(T') W in column-major order (conj(c) w = (cr wr + ci wi) + i (cr wi -
 * ci wr); the Float32 products are exact in Float64, so each component rounds once before it is
 * added), narrowed to ComplexF32. |num| is Julia's Float32 hypot, which squares and sums in
 * Float64 and rounds the square root to Float32. wvar = max(s2 - |s1|^2 / n, 0), den = cnorm
 * sqrt(wvar), surface = den > 0 ? Float32(|num| / den) : 0. */
static int ccorrelate_ws(corr_ws* w, const float* search, int64_t lds, const float* chip,
                         int64_t ldc, int64_t ch, int64_t cw, int64_t rx, int64_t ry,
                         float* surface) {
    int64_t nr = 2 * ry, nc = 2 * rx;
    double cnorm;
    if (!prepare_cchip(chip, ldc, ch, cw, w->chip, &cnorm)) {
        for (int64_t k = 0; k < nr * nc; k++) surface[k] = 0.0f;
        return 0;
    }
    int64_t sh = ch + nr - 1, sw = cw + nc - 1, ls = sh + 1;
    double* Sr = w->S;
    double* Si = w->S + (size_t)((sh + 1) * (sw + 1));
    cintegral_both(search, lds, sh, sw, Sr, Si, w->S2);
    const double n = (double)(ch * cw);
    for (int64_t j = 1; j <= nc; j++)
        for (int64_t i = 1; i <= nr; i++) {
            double ar = 0.0, ai = 0.0;
            for (int64_t jj = 0; jj < cw; jj++)
                for (int64_t ii = 0; ii < ch; ii++) {
                    const double cr = (double)w->chip[2 * (ii + jj * ch)];
                    const double ci = (double)w->chip[2 * (ii + jj * ch) + 1];
                    const int64_t q = 2 * ((i - 1 + ii) + (j - 1 + jj) * lds);
                    const double wr = (double)search[q], wi = (double)search[q + 1];
                    ar += cr * wr + ci * wi;
                    ai += cr * wi - ci * wr;
                }
            const float nr32 = (float)ar, ni32 = (float)ai;
            const float mag =
                (float)sqrt((double)nr32 * (double)nr32 + (double)ni32 * (double)ni32);
            const double s1r = boxsum(Sr, ls, i, j, ch, cw);
            const double s1i = boxsum(Si, ls, i, j, ch, cw);
            const double s2 = boxsum(w->S2, ls, i, j, ch, cw);
            const double a2 = s1r * s1r + s1i * s1i;
            const double wvar = jmax0(s2 - a2 / n);
            const double den = cnorm * sqrt(wvar);
            surface[(i - 1) + (j - 1) * nr] = den > 0 ? (float)((double)mag / den) : 0.0f;
        }
    return 1;
}

/* ramp_phase -- src/preprocess.jl:233-245: the phase of sum conj(z_i) z_(i+1) over valid
 * neighbour pairs down columns (phase_y) and along rows (phase_x), ComplexF64 sums in
 * column-major order. */
static void ramp_sums(const float* z, const uint8_t* m, int64_t nr, int64_t nc, int axis_y,
                      double* re, double* im) {
    double ar = 0.0, ai = 0.0;
    const int64_t di = axis_y ? 1 : 0, dj = axis_y ? 0 : 1;
    for (int64_t j = 0; j < nc - dj; j++)
        for (int64_t i = 0; i < nr - di; i++) {
            const int64_t k = i + j * nr, l = (i + di) + (j + dj) * nr;
            if (!(m[k] && m[l])) continue;
            const double zr = (double)z[2 * k], zi = (double)z[2 * k + 1];
            const double ur = (double)z[2 * l], ui = (double)z[2 * l + 1];
            ar += zr * ur + zi * ui;
            ai += zr * ui - zi * ur;
        }
    *re = ar;
    *im = ai;
}

void orc_ramp_phase(const float* z, const uint8_t* mask, int64_t nr, int64_t nc, double* phase_x,
                    double* phase_y) {
    double r, i;
    ramp_sums(z, mask, nr, nc, 0, &r, &i);
    *phase_x = atan2(i, r);
    ramp_sums(z, mask, nr, nc, 1, &r, &i);
    *phase_y = atan2(i, r);
}

/* deramp -- src/preprocess.jl:184-231: out = ComplexF32(z * cis(-((i-1) phase_y + (j-1)
 * phase_x))) with the phases of the axes asked for (the other is zero). */
void orc_deramp(const float* z, const uint8_t* mask, int64_t nr, int64_t nc, int axis,
                float* out) {
    double py = 0.0, px = 0.0, r, im;
    if (axis != 1) { ramp_sums(z, mask, nr, nc, 1, &r, &im); py = atan2(im, r); }
    if (axis != 2) { ramp_sums(z, mask, nr, nc, 0, &r, &im); px = atan2(im, r); }
    for (int64_t j = 0; j < nc; j++)
        for (int64_t i = 0; i < nr; i++) {
            const int64_t k = i + j * nr;
            const double th = -((double)i * py + (double)j * px);
            const double c = cos(th), s = sin(th);
            const double zr = (double)z[2 * k], zi = (double)z[2 * k + 1];
            out[2 * k] = (float)(zr * c - zi * s);
            out[2 * k + 1] = (float)(zr * s + zi * c);
        }
}

static int correlate_ws(corr_ws* w, const void* search, int64_t lds, const void* chip,
                        int64_t ldc, int dt, int cdt, int64_t ch, int64_t cw, int64_t rx,
                        int64_t ry, int measure, int num_mode, float* surface) {
    if (dt == ORC_C64)
        return ccorrelate_ws(w, (const float*)search, lds, (const float*)chip, ldc, ch, cw, rx,
                             ry, surface);
    int64_t nr = 2 * ry, nc = 2 * rx;
    double cnorm;
    int ok = prepare_chip(chip, ldc, cdt, ch, cw, w->chip, &cnorm);
    if (!ok) {
        for (int64_t k = 0; k < nr * nc; k++) surface[k] = 0.0f;
        return 0;
    }
    int64_t sh = ch + nr - 1, sw = cw + nc - 1, ls = sh + 1;
    orc_integral_both(search, lds, dt, sh, sw, w->S, w->S2);
    /* DIRECT_THRESHOLD (src/correlate.jl:785) chooses the direct path for tiny work. */
    if (num_mode == ORC_NUM_FFT && nr * nc * ch * cw > 20000)
        numerators_fft(w->num, &w->fft, search, lds, dt, w->chip, nr, nc, ch, cw);
    else
        numerators_direct(w->num, search, lds, dt, w->chip, nr, nc, ch, cw);
    double n = (double)(ch * cw);
    for (int64_t j = 1; j <= nc; j++)
        for (int64_t i = 1; i <= nr; i++) {
            double den;
            if (measure == ORC_NCC) {
                double wsq = boxsum(w->S2, ls, i, j, ch, cw);
                den = cnorm * sqrt(jmax0(wsq));
            } else {
                double s1 = boxsum(w->S, ls, i, j, ch, cw);
                double s2 = boxsum(w->S2, ls, i, j, ch, cw);
                double wvar = jmax0(s2 - s1 * s1 / n);
                den = cnorm * sqrt(wvar);
            }
            double num = w->num[(i - 1) + (j - 1) * nr];
            surface[(i - 1) + (j - 1) * nr] = den > 0 ? (float)(num / den) : 0.0f;
        }
    return 1;
}

int orc_correlate(const void* search, int64_t ld_search, const void* chip, int64_t ld_chip,
                  int dtype, int64_t ch, int64_t cw, int64_t rx, int64_t ry, int measure,
                  int num_mode, float* surface) {
    corr_ws w;
    cws_init(&w, ch, cw, ry, rx);
    int ok = correlate_ws(&w, search, ld_search, chip, ld_chip, dtype, dtype, ch, cw, rx, ry,
                          measure, num_mode, surface);
    cws_free(&w);
    return ok;
}

/* ------------------------------------------------------------------------------------------ */
/* peak_index -- src/peak.jl:32-61; peak_offset -- src/peak.jl:102-106                         */
/* ------------------------------------------------------------------------------------------ */

void orc_peak_index(const float* s, int64_t nr, int64_t nc, int64_t* row, int64_t* col) {
    int64_t bi = 1, bj = 1;
    float best = -INFINITY;
    for (int64_t j = 1; j <= nc; j++)
        for (int64_t i = 1; i <= nr; i++) {
            float v = s[(i - 1) + (j - 1) * nr];
            if (v > best || (v == best && (i < bi || (i == bi && j < bj)))) {
                best = v;
                bi = i;
                bj = j;
            }
        }
    *row = bi;
    *col = bj;
}

/* ------------------------------------------------------------------------------------------ */
/* pyrup! -- src/peak.jl:196-294; taps 297-337; reflect101_0 374-382                           */
/* ------------------------------------------------------------------------------------------ */

static const float W_C = 6.0f / 8;
static const float W_S = 1.0f / 8;
static const float W_H = 4.0f / 8;

static int64_t reflect101_0(int64_t p, int64_t n) {
    if (n == 1) return 0;
    int64_t m = n - 1;
    while (p < 0 || p > m) {
        if (p < 0) p = -p;
        if (p > m) p = 2 * m - p;
    }
    return p;
}

/* (minus, centre, plus, ntaps), 1-based output position k on an axis of length n. */
static void pyrup_taps(int64_t k, int64_t n, int64_t* m, int64_t* c, int64_t* q, int* nt) {
    int64_t p = k - 1;
    if ((p & 1) == 0) {
        *c = p / 2 + 1;
        *m = reflect101_0(p - 2, 2 * n) / 2 + 1;
        *q = reflect101_0(p + 2, 2 * n) / 2 + 1;
        *nt = 3;
    } else {
        *c = reflect101_0(p - 1, 2 * n) / 2 + 1;
        *q = reflect101_0(p + 1, 2 * n) / 2 + 1;
        *m = *c;
        *nt = 2;
    }
}

static void pyrup_scratch(float* dst, const float* src, int64_t sh, int64_t sw, float* tmp) {
    int64_t th = 2 * sh;
    /* vertical: tmp (2sh x sw) */
    for (int64_t j = 1; j <= sw; j++)
        for (int64_t i = 1; i <= th; i++) {
            int64_t im, i0, ip;
            int nt;
            pyrup_taps(i, sh, &im, &i0, &ip, &nt);
            const float* col = src + (j - 1) * sh;
            float v = nt == 3 ? W_C * col[i0 - 1] + W_S * (col[im - 1] + col[ip - 1])
                              : W_H * (col[i0 - 1] + col[ip - 1]);
            tmp[(i - 1) + (j - 1) * th] = v;
        }
    /* horizontal: dst (2sh x 2sw) */
    for (int64_t j = 1; j <= 2 * sw; j++) {
        int64_t jm, j0, jp;
        int nt;
        pyrup_taps(j, sw, &jm, &j0, &jp, &nt);
        for (int64_t i = 1; i <= th; i++) {
            float v = nt == 3 ? W_C * tmp[(i - 1) + (j0 - 1) * th] +
                                    W_S * (tmp[(i - 1) + (jm - 1) * th] + tmp[(i - 1) + (jp - 1) * th])
                              : W_H * (tmp[(i - 1) + (j0 - 1) * th] + tmp[(i - 1) + (jp - 1) * th]);
            dst[(i - 1) + (j - 1) * th] = v;
        }
    }
}

void orc_pyrup(float* dst, const float* src, int64_t sh, int64_t sw) {
    float* tmp = (float*)malloc(sizeof(float) * (size_t)(2 * sh * sw));
    pyrup_scratch(dst, src, sh, sw, tmp);
    free(tmp);
}

/* ------------------------------------------------------------------------------------------ */
/* subpixel_peak -- src/peak.jl:456-522                                                        */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    float* a;
    float* b;
    float* tmp;
    int64_t up;
} refine_ws;

static void rws_init(refine_ws* r, int64_t up) {
    int64_t n = 5 * up;
    r->up = up;
    r->a = (float*)malloc(sizeof(float) * (size_t)(n * n));
    r->b = (float*)malloc(sizeof(float) * (size_t)(n * n));
    r->tmp = (float*)malloc(sizeof(float) * (size_t)(n * n));
}

static void rws_free(refine_ws* r) { free(r->a); free(r->b); free(r->tmp); }

static void peak_offset(const float* s, int64_t nr, int64_t nc, int64_t rx, int64_t ry, double* dx,
                        double* dy, float* val) {
    int64_t i, j;
    orc_peak_index(s, nr, nc, &i, &j);
    *dx = (double)(j - rx - 1);
    *dy = (double)(i - ry - 1);
    *val = s[(i - 1) + (j - 1) * nr];
}

static void subpixel_ws(refine_ws* rw, const float* s, int64_t nr, int64_t nc, int64_t rx,
                        int64_t ry, int64_t up, double* dx, double* dy, float* val) {
    if (up == 1) { peak_offset(s, nr, nc, rx, ry, dx, dy, val); return; }
    int64_t pi, pj;
    orc_peak_index(s, nr, nc, &pi, &pj);
    const int64_t p = 5, half = 2;
    int64_t hi_r = nr - p + 1 > 1 ? nr - p + 1 : 1;
    int64_t hi_c = nc - p + 1 > 1 ? nc - p + 1 : 1;
    int64_t i0 = pi - half < 1 ? 1 : (pi - half > hi_r ? hi_r : pi - half);
    int64_t j0 = pj - half < 1 ? 1 : (pj - half > hi_c ? hi_c : pj - half);
    if (nr < p || nc < p) { peak_offset(s, nr, nc, rx, ry, dx, dy, val); return; }
    float patch[25];
    for (int64_t j = 0; j < p; j++)
        for (int64_t i = 0; i < p; i++) patch[i + j * p] = s[(i0 - 1 + i) + (j0 - 1 + j) * nr];
    int64_t n = p, factor = 1;
    const float* cur = patch;
    int use_a = 1;
    while (factor < up) {
        float* dst = use_a ? rw->a : rw->b;
        pyrup_scratch(dst, cur, n, n, rw->tmp);
        n *= 2;
        factor *= 2;
        cur = dst;
        use_a = !use_a;
    }
    int64_t si, sj;
    orc_peak_index(cur, n, n, &si, &sj);
    *val = cur[(si - 1) + (sj - 1) * n];
    double row = (double)(si - 1) / (double)factor + (double)(i0 - 1);
    double col = (double)(sj - 1) / (double)factor + (double)(j0 - 1);
    *dx = col - (double)rx;
    *dy = row - (double)ry;
}

void orc_subpixel_peak(const float* surface, int64_t nr, int64_t nc, int64_t rx, int64_t ry,
                       int64_t up, double* dx, double* dy, float* val) {
    refine_ws rw;
    rws_init(&rw, up < 1 ? 1 : up);
    subpixel_ws(&rw, surface, nr, nc, rx, ry, up, dx, dy, val);
    rws_free(&rw);
}

/* ------------------------------------------------------------------------------------------ */
/* track! -- src/track.jl:103-305; geometry src/points.jl:116-117, 358-425                     */
/* ------------------------------------------------------------------------------------------ */

static inline int64_t jfloor(double v) { return (int64_t)floor(v); }

static void* zeropad(const void* a, int dt, int64_t nr, int64_t nc, int64_t px, int64_t py) {
    size_t es = esize(dt);
    int64_t NR = nr + 2 * py, NC = nc + 2 * px;
    char* out = (char*)calloc((size_t)(NR * NC), es);
    for (int64_t j = 0; j < nc; j++)
        memcpy(out + ((py) + (j + px) * NR) * es, (const char*)a + (j * nr) * es, (size_t)nr * es);
    return out;
}

/* ------------------------------------------------------------------------------------------ */
/* _rotate_chip -- src/track.jl:382-409: the chip rotated about its centre by `deg` degrees,   */
/* bilinear in Float64 from the chip's samples, Float32 out, zero where a source corner falls  */
/* outside the chip. deg2rad is deg * (pi / 180) (Julia's deg2rad for Float64); sin and cos    */
/* are the C library's.                                                                        */
/* ------------------------------------------------------------------------------------------ */
void orc_rotate_chip(const void* chip, int64_t ldc, int dt, int64_t ch, int64_t cw, double deg,
                     float* dst) {
    const double rad = deg * (M_PI / 180.0);
    const double s = sin(rad), c = cos(rad);
    const double yc = (double)(ch + 1) / 2.0, xc = (double)(cw + 1) / 2.0;
    for (int64_t j = 1; j <= cw; j++)
        for (int64_t i = 1; i <= ch; i++) {
            const double dyy = (double)i - yc, dxx = (double)j - xc;
            const double sy = yc + c * dyy + s * dxx;
            const double sx = xc - s * dyy + c * dxx;
            const int64_t i0 = jfloor(sy), j0 = jfloor(sx);
            float* o = dst + (i - 1) + (j - 1) * ch;
            if (i0 < 1 || j0 < 1 || i0 + 1 > ch || j0 + 1 > cw) {
                *o = 0.0f;
                continue;
            }
            const double fy = sy - (double)i0, fx = sx - (double)j0;
            const double v00 = ldv(chip, dt, (i0 - 1) + (j0 - 1) * ldc);
            const double v10 = ldv(chip, dt, i0 + (j0 - 1) * ldc);
            const double v01 = ldv(chip, dt, (i0 - 1) + j0 * ldc);
            const double v11 = ldv(chip, dt, i0 + j0 * ldc);
            *o = (float)((1 - fy) * ((1 - fx) * v00 + fx * v01) + fy * ((1 - fx) * v10 + fx * v11));
        }
}

int orc_track(const void* ref, const void* sec, const uint8_t* ok_mask, int dt, int64_t ny,
              int64_t nx, const orc_points* pts, int64_t n, int64_t up, int measure,
              int num_mode, int nthreads, int nrot, const double* angles, float* dx, float* dy,
              float* corr, uint8_t* searched) {
    for (int64_t i = 0; i < n; i++) {
        dx[i] = NAN; dy[i] = NAN; corr[i] = NAN; searched[i] = 0;
    }
    /* Coherence needs complex images (src/correlate.jl:739-748), the real measures real ones */
    if ((dt == ORC_C64) != (measure == ORC_COHERENCE)) return ORC_ERR_MEASURE;
    if (dt == ORC_C64 && nrot > 0) return ORC_ERR_ROTATION;
    /* _pass_geometry -- src/track.jl:176-197 */
    int64_t cx = 0, cy = 0, mrx = 0, mry = 0, px = 0, py = 0;
    int fits = 1;
    for (int64_t i = 0; i < n; i++) {
        if (!(pts->radius_x[i] > 0 && pts->radius_y[i] > 0)) continue;
        int64_t csx = pts->chip_size_x[i], csy = pts->chip_size_y[i];
        int64_t prx = pts->radius_x[i], pry = pts->radius_y[i];
        if (csx > cx) cx = csx;
        if (csy > cy) cy = csy;
        if (prx > mrx) mrx = prx;
        if (pry > mry) mry = pry;
        int64_t qx = csx / 2 + prx + (int64_t)ceil(fabs(pts->dx_prior[i]));
        int64_t qy = csy / 2 + pry + (int64_t)ceil(fabs(pts->dy_prior[i]));
        if (qx > px) px = qx;
        if (qy > py) py = qy;
        /* inbounds -- src/points.jl:421-425, on the unshifted coordinates */
        int64_t hx = csx / 2, hy = csy / 2;
        int64_t fx = jfloor(pts->x[i]), fy = jfloor(pts->y[i]);
        int64_t r0 = fy - hy - pry, r1 = fy + hy + pry - 2;
        int64_t c0 = fx - hx - prx, c1 = fx + hx + prx - 2;
        fits = fits && (r0 >= 1 && r1 <= ny && c0 >= 1 && c1 <= nx);
    }
    if (cx == 0) return 0;
    int64_t padx = 0, pady = 0;
    const void *R = ref, *S = sec;
    const uint8_t* M = ok_mask;
    void *Rp = NULL, *Sp = NULL;
    uint8_t* Mp = NULL;
    if (!fits) {
        padx = px + 2;
        pady = py + 2;
        Rp = zeropad(ref, dt, ny, nx, padx, pady);
        Sp = zeropad(sec, dt, ny, nx, padx, pady);
        Mp = (uint8_t*)zeropad(ok_mask, ORC_U8, ny, nx, padx, pady);
        R = Rp; S = Sp; M = Mp;
    }
    int64_t NY = ny + 2 * pady, NX = nx + 2 * padx;
    double shx = (double)padx + 0.5, shy = (double)pady + 0.5;

#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
    {
        corr_ws w;
        cws_init(&w, cy, cx, mry, mrx);
        refine_ws rw;
        rws_init(&rw, up < 1 ? 1 : up);
        float* surf = (float*)malloc(sizeof(float) * (size_t)(4 * mrx * mry + 1));
        float* hold = (float*)malloc(sizeof(float) * (size_t)(4 * mrx * mry + 1));
        float* rot = (float*)malloc(sizeof(float) * (size_t)(cx * cy + 1));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 8)
#endif
        for (int64_t i = 0; i < n; i++) {
            int64_t prx = pts->radius_x[i], pry = pts->radius_y[i];
            if (!(prx > 0 && pry > 0)) continue;
            double X = pts->x[i] + shx, Y = pts->y[i] + shy;
            int64_t hx = pts->chip_size_x[i] / 2, hy = pts->chip_size_y[i] / 2;
            int64_t ccx = jfloor(X - pts->dx_prior[i]), ccy = jfloor(Y - pts->dy_prior[i]);
            int64_t cr0 = ccy - hy, cr1 = ccy + hy - 1, cc0 = ccx - hx, cc1 = ccx + hx - 1;
            int64_t wcx = jfloor(X), wcy = jfloor(Y);
            int64_t wr0 = wcy - hy - pry, wr1 = wcy + hy + pry - 2;
            int64_t wc0 = wcx - hx - prx, wc1 = wcx + hx + prx - 2;
            /* checkbounds: an empty range is always in bounds */
            int chip_in = (cr1 < cr0 || cc1 < cc0) ||
                          (cr0 >= 1 && cr1 <= NY && cc0 >= 1 && cc1 <= NX);
            int win_in = (wr1 < wr0 || wc1 < wc0) ||
                         (wr0 >= 1 && wr1 <= NY && wc0 >= 1 && wc1 <= NX);
            if (!chip_in || !win_in) continue;
            /* _any_valid -- src/track.jl:416-421 */
            int any = 0;
            for (int64_t c = cc0; c <= cc1 && !any; c++)
                for (int64_t r = cr0; r <= cr1; r++)
                    if (M[(r - 1) + (c - 1) * NY]) { any = 1; break; }
            if (!any) continue;
            searched[i] = 1;
            int64_t ch = 2 * hy, cw = 2 * hx;
            const void* chip = elem_ptr(S, dt, (cr0 - 1) + (cc0 - 1) * NY);
            const void* win = elem_ptr(R, dt, (wr0 - 1) + (wc0 - 1) * NY);
            int ok;
            if (nrot <= 0) {
                ok = correlate_ws(&w, win, NY, chip, NY, dt, dt, ch, cw, prx, pry, measure,
                                  num_mode, surf);
            } else {
                /* _correlate_rotations! -- src/track.jl:316-351: the surface whose maximum is
                 * highest (strictly: an earlier angle wins a tie); every angle degenerate is a
                 * degenerate chip */
                float best = -INFINITY;
                ok = 0;
                const int64_t ns = 4 * prx * pry;
                for (int k = 0; k < nrot; k++) {
                    const void* cp = chip;
                    int cdt = dt;
                    int64_t cld = NY;
                    if (angles[k] != 0.0) {
                        orc_rotate_chip(chip, NY, dt, ch, cw, angles[k], rot);
                        cp = rot;
                        cdt = ORC_F32;
                        cld = ch;
                    }
                    if (!correlate_ws(&w, win, NY, cp, cld, dt, cdt, ch, cw, prx, pry, measure,
                                      num_mode, surf))
                        continue;
                    float pk = surf[0];
                    for (int64_t q = 1; q < ns; q++)
                        if (surf[q] > pk || surf[q] != surf[q]) pk = surf[q];
                    if (pk > best) {
                        best = pk;
                        memcpy(hold, surf, sizeof(float) * (size_t)ns);
                        ok = 1;
                    }
                }
                if (ok) memcpy(surf, hold, sizeof(float) * (size_t)ns);
            }
            if (!ok) continue;
            double pdx, pdy;
            float val;
            if (up <= 1)
                peak_offset(surf, 2 * pry, 2 * prx, prx, pry, &pdx, &pdy, &val);
            else
                subpixel_ws(&rw, surf, 2 * pry, 2 * prx, prx, pry, up, &pdx, &pdy, &val);
            dx[i] = (float)(pdx + pts->dx_prior[i]);
            dy[i] = (float)(pdy + pts->dy_prior[i]);
            corr[i] = val;
        }
        free(surf);
        free(hold);
        free(rot);
        rws_free(&rw);
        cws_free(&w);
    }
    free(Rp); free(Sp); free(Mp);
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* window reductions -- src/window.jl                                                          */
/* ------------------------------------------------------------------------------------------ */

static inline void window_margins(int64_t wx, int64_t wy, int64_t* lx, int64_t* ly, int64_t* rx,
                                  int64_t* ry) {
    *lx = wx / 2;
    *ly = wy / 2;
    *rx = wx - 1 - *lx;
    *ry = wy - 1 - *ly;
}

/* _windowmean_masked! -- src/window.jl:289-366 (generic element type) */
static void windowmean_masked(const void* A, int dt, int64_t nr, int64_t nc, int64_t wx,
                              int64_t wy, float* out) {
    int64_t lx, ly, rx, ry;
    window_margins(wx, wy, &lx, &ly, &rx, &ry);
    int64_t nb = nr < 16 ? nr : 16;
    double* sums = (double*)malloc(sizeof(double) * (size_t)(nb * nc));
    int32_t* counts = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nb * nc));
    double* rowbuf = (double*)malloc(sizeof(double) * (size_t)nc);
    int32_t* cntbuf = (int32_t*)malloc(sizeof(int32_t) * (size_t)nc);
#define AV(i, j) ldv(A, dt, ((i) - 1) + ((j) - 1) * nr)
    int64_t i0 = 1;
    while (i0 <= nr) {
        int64_t i1 = i0 + nb - 1 < nr ? i0 + nb - 1 : nr;
        for (int64_t j = 1; j <= nc; j++) {
            double s = 0.0;
            int32_t c = 0;
            int64_t a0 = i0 - ly > 1 ? i0 - ly : 1, a1 = i0 + ry < nr ? i0 + ry : nr;
            for (int64_t i = a0; i <= a1; i++) {
                double v = AV(i, j);
                if (!isnan(v)) { s += v; c += 1; }
            }
            for (int64_t i = i0; i <= i1; i++) {
                sums[(i - i0) + (j - 1) * nb] = s;
                counts[(i - i0) + (j - 1) * nb] = c;
                int64_t oi = i - ly;
                if (1 <= oi && oi <= nr) {
                    double v = AV(oi, j);
                    if (!isnan(v)) { s -= v; c -= 1; }
                }
                int64_t ii = i + ry + 1;
                if (ii <= nr) {
                    double v = AV(ii, j);
                    if (!isnan(v)) { s += v; c += 1; }
                }
            }
        }
        for (int64_t i = i0; i <= i1; i++) {
            for (int64_t j = 1; j <= nc; j++) {
                rowbuf[j - 1] = sums[(i - i0) + (j - 1) * nb];
                cntbuf[j - 1] = counts[(i - i0) + (j - 1) * nb];
            }
            double s = 0.0;
            int32_t c = 0;
            int64_t b1 = rx + 1 < nc ? rx + 1 : nc;
            for (int64_t j = 1; j <= b1; j++) { s += rowbuf[j - 1]; c += cntbuf[j - 1]; }
            for (int64_t j = 1; j <= nc; j++) {
                out[(i - 1) + (j - 1) * nr] = c > 0 ? (float)(s / (double)c) : NAN;
                int64_t oj = j - lx;
                if (1 <= oj && oj <= nc) { s -= rowbuf[oj - 1]; c -= cntbuf[oj - 1]; }
                int64_t jj = j + rx + 1;
                if (jj <= nc) { s += rowbuf[jj - 1]; c += cntbuf[jj - 1]; }
            }
        }
        i0 = i1 + 1;
    }
#undef AV
    free(sums); free(counts); free(rowbuf); free(cntbuf);
}

/* _windowmean_dense! -- src/window.jl:388-439 */
static void windowmean_dense(const void* A, int dt, int64_t nr, int64_t nc, int64_t wx,
                             int64_t wy, float* out) {
    int64_t lx, ly, rx, ry;
    window_margins(wx, wy, &lx, &ly, &rx, &ry);
    int64_t nb = nr < 16 ? nr : 16;
    double* sums = (double*)malloc(sizeof(double) * (size_t)(nb * nc));
    double* rowbuf = (double*)malloc(sizeof(double) * (size_t)nc);
#define AV(i, j) ldv(A, dt, ((i) - 1) + ((j) - 1) * nr)
    int64_t i0 = 1;
    while (i0 <= nr) {
        int64_t i1 = i0 + nb - 1 < nr ? i0 + nb - 1 : nr;
        for (int64_t j = 1; j <= nc; j++) {
            double s = 0.0;
            int64_t a0 = i0 - ly > 1 ? i0 - ly : 1, a1 = i0 + ry < nr ? i0 + ry : nr;
            for (int64_t i = a0; i <= a1; i++) s += AV(i, j);
            for (int64_t i = i0; i <= i1; i++) {
                sums[(i - i0) + (j - 1) * nb] = s;
                int64_t oi = i - ly;
                if (1 <= oi && oi <= nr) s -= AV(oi, j);
                int64_t ii = i + ry + 1;
                if (ii <= nr) s += AV(ii, j);
            }
        }
        for (int64_t i = i0; i <= i1; i++) {
            int64_t ci = (i + ry < nr ? i + ry : nr) - (i - ly > 1 ? i - ly : 1) + 1;
            for (int64_t j = 1; j <= nc; j++) rowbuf[j - 1] = sums[(i - i0) + (j - 1) * nb];
            double s = 0.0;
            int64_t b1 = rx + 1 < nc ? rx + 1 : nc;
            for (int64_t j = 1; j <= b1; j++) s += rowbuf[j - 1];
            for (int64_t j = 1; j <= nc; j++) {
                int64_t cj = (j + rx < nc ? j + rx : nc) - (j - lx > 1 ? j - lx : 1) + 1;
                out[(i - 1) + (j - 1) * nr] = (float)(s / (double)(ci * cj));
                int64_t oj = j - lx;
                if (1 <= oj && oj <= nc) s -= rowbuf[oj - 1];
                int64_t jj = j + rx + 1;
                if (jj <= nc) s += rowbuf[jj - 1];
            }
        }
        i0 = i1 + 1;
    }
#undef AV
    free(sums); free(rowbuf);
}

void orc_windowmean(const float* a, int64_t nr, int64_t nc, int64_t wx, int64_t wy, int hasnan,
                    float* out) {
    if (wx > nc) wx = nc;
    if (wy > nr) wy = nr;
    if (hasnan < 0) {
        hasnan = 0;
        for (int64_t k = 0; k < nr * nc; k++)
            if (isnan(a[k])) { hasnan = 1; break; }
    }
    if (hasnan) windowmean_masked(a, ORC_F32, nr, nc, wx, wy, out);
    else windowmean_dense(a, ORC_F32, nr, nc, wx, wy, out);
}

/* _deque_pass! -- src/window.jl:150-194 (1-based indices into vals) */
static void deque_pass(float* buf, const float* vals, int64_t n, int64_t w, int64_t* dq,
                       int is_max) {
    int64_t left = w / 2, right = w - 1 - left;
    int64_t head = 1, tail = 0, admitted = 0;
    for (int64_t i = 1; i <= n; i++) {
        int64_t target = i + right < n ? i + right : n;
        while (admitted < target) {
            admitted += 1;
            float v = vals[admitted - 1];
            if (isnan(v)) continue;
            while (tail >= head) {
                float t = vals[dq[tail - 1] - 1];
                int better = is_max ? (t > v) : (t < v);
                if (better) break;
                tail -= 1;
            }
            tail += 1;
            dq[tail - 1] = admitted;
        }
        int64_t oldest = i - left;
        while (tail >= head && dq[head - 1] < oldest) head += 1;
        buf[i - 1] = tail >= head ? vals[dq[head - 1] - 1] : NAN;
    }
}

/* _window_extremum! -- src/window.jl:103-148 */
void orc_window_extremum(const float* A, int64_t nr, int64_t nc, int64_t wx, int64_t wy,
                         int is_max, float* out) {
    if (wx > nc) wx = nc;
    if (wy > nr) wy = nr;
    int64_t n = nr > nc ? nr : nc;
    float* buf = (float*)malloc(sizeof(float) * (size_t)n);
    float* vals = (float*)malloc(sizeof(float) * (size_t)n);
    int64_t* dq = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    float* tmp = (float*)malloc(sizeof(float) * (size_t)(nr * nc));
    for (int64_t j = 0; j < nc; j++) {
        for (int64_t i = 0; i < nr; i++) vals[i] = A[i + j * nr];
        deque_pass(buf, vals, nr, wy, dq, is_max);
        for (int64_t i = 0; i < nr; i++) tmp[i + j * nr] = buf[i];
    }
    for (int64_t i = 0; i < nr; i++) {
        for (int64_t j = 0; j < nc; j++) vals[j] = tmp[i + j * nr];
        deque_pass(buf, vals, nc, wx, dq, is_max);
        for (int64_t j = 0; j < nc; j++) out[i + j * nr] = buf[j];
    }
    free(buf); free(vals); free(dq); free(tmp);
}

/* ------------------------------------------------------------------------------------------ */
/* pre-filters -- src/preprocess.jl:275-288 (_filtered), 304-312 (highpass), 333-351 (wallis), */
/* 371-399 (_masked_boxmean), 416-438 (_masked_boxstd), 447-460 (_erode_mask)                 */
/* ------------------------------------------------------------------------------------------ */

/* _masked_boxmean -- src/preprocess.jl:371-399: the box mean of the valid pixels. With every
 * pixel marked valid the image itself is averaged (windowmean! decides the NaN-aware path from
 * the data); otherwise a Float32 copy with NaN where the mask is false. */
static void masked_boxmean(const void* img, const uint8_t* mask, int dt, int64_t nr, int64_t nc,
                           int64_t wx, int64_t wy, int allmask, float* out) {
    int64_t N = nr * nc;
    if (allmask) {
        int gaps = 0;
        for (int64_t k = 0; k < N; k++)
            if (isnan(ldv(img, dt, k))) { gaps = 1; break; }
        if (gaps) windowmean_masked(img, dt, nr, nc, wx, wy, out);
        else windowmean_dense(img, dt, nr, nc, wx, wy, out);
    } else {
        float* masked = (float*)malloc(sizeof(float) * (size_t)N);
        for (int64_t k = 0; k < N; k++) masked[k] = mask[k] ? ldf(img, dt, k) : NAN;
        windowmean_masked(masked, ORC_F32, nr, nc, wx, wy, out);
        free(masked);
    }
}

/* _filtered -- src/preprocess.jl:275-288, with _erode_mask 447-460: the mask eroded by the
 * filter width, then non-finite outputs become 0 and invalid. */
static void filtered_finish(float* out, const uint8_t* mask, int allmask, int64_t nr, int64_t nc,
                            int64_t w, uint8_t* vmask) {
    int64_t N = nr * nc;
    if (allmask) {
        memcpy(vmask, mask, (size_t)N);
    } else {
        float* f = (float*)malloc(sizeof(float) * (size_t)N);
        float* lo = (float*)malloc(sizeof(float) * (size_t)N);
        for (int64_t k = 0; k < N; k++) f[k] = mask[k] ? 1.0f : 0.0f;
        orc_window_extremum(f, nr, nc, w, w, 0, lo);
        for (int64_t k = 0; k < N; k++) vmask[k] = lo[k] > 0.5f;
        free(f); free(lo);
    }
    for (int64_t k = 0; k < N; k++)
        if (!isfinite(out[k])) { out[k] = 0.0f; vmask[k] = 0; }
}

static int all_set(const uint8_t* mask, int64_t N) {
    for (int64_t k = 0; k < N; k++)
        if (!mask[k]) return 0;
    return 1;
}

/* preprocess(img, mask, ::Highpass) -- src/preprocess.jl:155-156, highpass 304-312 */
void orc_highpass_filtered(const void* img, const uint8_t* mask, int dt, int64_t nr, int64_t nc,
                           int64_t w, float* out, uint8_t* vmask) {
    int64_t N = nr * nc;
    int64_t wx = w < nc ? w : nc, wy = w < nr ? w : nr;
    int allmask = all_set(mask, N);
    masked_boxmean(img, mask, dt, nr, nc, wx, wy, allmask, out);
    for (int64_t k = 0; k < N; k++)
        out[k] = (mask[k] && isfinite(out[k])) ? ldf(img, dt, k) - out[k] : NAN;
    filtered_finish(out, mask, allmask, nr, nc, w, vmask);
}

/* preprocess(img, mask, ::Wallis) -- src/preprocess.jl:158-159; wallis 333-351; _masked_boxstd
 * 416-438: the population standard deviation about the measured local mean, from the box mean
 * (NaN-aware when any squared deviation is NaN) of Float32 squared deviations; out =
 * (Float32(x) - mean) / max(sd, Float32(min_std)), NaN where the mask is false, the mean or sd is
 * not finite, or the divisor is 0. Julia's max propagates NaN, and max(x, 0f0) of a negative x
 * (a running sum's rounding) is 0. */
void orc_wallis_filtered(const void* img, const uint8_t* mask, int dt, int64_t nr, int64_t nc,
                         int64_t w, double min_std, float* out, uint8_t* vmask) {
    int64_t N = nr * nc;
    int64_t wx = w < nc ? w : nc, wy = w < nr ? w : nr;
    int allmask = all_set(mask, N);
    float* mean = (float*)malloc(sizeof(float) * (size_t)N);
    float* dev2 = (float*)malloc(sizeof(float) * (size_t)N);
    float* msd = (float*)malloc(sizeof(float) * (size_t)N);
    masked_boxmean(img, mask, dt, nr, nc, wx, wy, allmask, mean);
    int gaps = 0;
    for (int64_t k = 0; k < N; k++) {
        if (mask[k] && isfinite(mean[k])) {
            float d = ldf(img, dt, k) - mean[k];
            float v = d * d;
            dev2[k] = v;
            gaps |= isnan(v);
        } else {
            dev2[k] = NAN;
            gaps = 1;
        }
    }
    if (gaps) windowmean_masked(dev2, ORC_F32, nr, nc, wx, wy, msd);
    else windowmean_dense(dev2, ORC_F32, nr, nc, wx, wy, msd);
    const float floor_sd = (float)min_std;
    for (int64_t k = 0; k < N; k++) {
        float m = msd[k];
        float sd = isnan(m) ? NAN : sqrtf(m < 0.0f ? 0.0f : m);
        if (!mask[k] || !isfinite(mean[k]) || !isfinite(sd)) { out[k] = NAN; continue; }
        float sdiv = sd < floor_sd ? floor_sd : sd;
        out[k] = sdiv > 0 ? (ldf(img, dt, k) - mean[k]) / sdiv : NAN;
    }
    free(mean); free(dev2); free(msd);
    filtered_finish(out, mask, allmask, nr, nc, w, vmask);
}

/* The statistics of _to_uint8 -- src/preprocess.jl:547-570: valid count, mean and Bessel standard
 * deviation, each a left-to-right Float64 sum over the valid finite values v (given as Float64). */
void orc_quant_stats(const double* v, const uint8_t* mask, int64_t N, double* out) {
    int64_t n = 0;
    double s = 0.0;
    for (int64_t k = 0; k < N; k++)
        if (mask[k] && isfinite(v[k])) { n += 1; s += v[k]; }
    out[0] = (double)n;
    out[1] = out[2] = 0.0;
    if (n == 0) return;
    double mean = s / (double)n;
    double ss = 0.0;
    for (int64_t k = 0; k < N; k++)
        if (mask[k] && isfinite(v[k])) { double d = v[k] - mean; ss += d * d; }
    out[1] = mean;
    out[2] = n > 1 ? sqrt(ss / (double)(n - 1)) : 0.0;
}

/* _to_uint8 -- src/preprocess.jl:546-592 */
void orc_to_uint8(const float* img, const uint8_t* mask, int64_t nr, int64_t nc, uint8_t* out) {
    int64_t N = nr * nc;
    double* v = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
    for (int64_t k = 0; k < N; k++) v[k] = (double)img[k];
    double st[3];
    orc_quant_stats(v, mask, N, st);
    memset(out, 0, (size_t)N);
    if (st[0] == 0) { free(v); return; }
    double mean = st[1], sd = st[2];
    if (sd == 0) {
        for (int64_t k = 0; k < N; k++) out[k] = mask[k] ? 0x80 : 0x00;
        free(v);
        return;
    }
    double lo = mean - 3 * sd;
    double scale = 255.0 / (6 * sd);
    for (int64_t k = 0; k < N; k++) {
        if (!mask[k] || !isfinite(v[k])) { out[k] = 0; continue; }
        double t = (v[k] - lo) * scale;
        t = t < 0.0 ? 0.0 : (t > 255.0 ? 255.0 : t);
        out[k] = (uint8_t)nearbyint(t);
    }
    free(v);
}

/* quantize(::NoQuantize) for floating images -- src/preprocess.jl:497-507 */
void orc_noquantize(const float* img, int64_t nr, int64_t nc, float* out) {
    for (int64_t k = 0; k < nr * nc; k++) out[k] = isfinite(img[k]) ? img[k] : 0.0f;
}

/* ------------------------------------------------------------------------------------------ */
/* grid reductions: count_agreeing src/window.jl:594-619; windowmedmad 501, 513-582           */
/* ------------------------------------------------------------------------------------------ */

void orc_count_agreeing(const float* A, int64_t nr, int64_t nc, int64_t w, float t, float* out) {
    int64_t wx = w < nc ? w : nc, wy = w < nr ? w : nr;
    int64_t lx, ly, rx, ry;
    window_margins(wx, wy, &lx, &ly, &rx, &ry);
    for (int64_t j = 1; j <= nc; j++)
        for (int64_t i = 1; i <= nr; i++) {
            float centre = A[(i - 1) + (j - 1) * nr];
            if (isnan(centre)) { out[(i - 1) + (j - 1) * nr] = 0.0f; continue; }
            int64_t n = 0;
            for (int64_t jj = (j - lx > 1 ? j - lx : 1); jj <= (j + rx < nc ? j + rx : nc); jj++)
                for (int64_t ii = (i - ly > 1 ? i - ly : 1); ii <= (i + ry < nr ? i + ry : nr); ii++) {
                    float v = A[(ii - 1) + (jj - 1) * nr];
                    if (fabsf(v - centre) < t) n += 1;
                }
            out[(i - 1) + (j - 1) * nr] = (float)n;
        }
}

static void insertion_sort(float* buf, int64_t n) {
    for (int64_t k = 1; k < n; k++) {
        float v = buf[k];
        int64_t m = k - 1;
        while (m >= 0 && buf[m] > v) { buf[m + 1] = buf[m]; m -= 1; }
        buf[m + 1] = v;
    }
}

static float sorted_median(const float* buf, int64_t n) {
    return (n & 1) ? buf[(n + 1) / 2 - 1] : (buf[n / 2 - 1] + buf[n / 2]) / 2;
}

void orc_windowmedmad(const float* A, int64_t nr, int64_t nc, int64_t w, float* med, float* mad) {
    int64_t wx = w < nc ? w : nc, wy = w < nr ? w : nr;
    int64_t lx, ly, rx, ry;
    window_margins(wx, wy, &lx, &ly, &rx, &ry);
    float* buf = (float*)malloc(sizeof(float) * (size_t)(wx * wy));
    for (int64_t j = 1; j <= nc; j++)
        for (int64_t i = 1; i <= nr; i++) {
            int64_t n = 0;
            for (int64_t jj = (j - lx > 1 ? j - lx : 1); jj <= (j + rx < nc ? j + rx : nc); jj++)
                for (int64_t ii = (i - ly > 1 ? i - ly : 1); ii <= (i + ry < nr ? i + ry : nr); ii++) {
                    float v = A[(ii - 1) + (jj - 1) * nr];
                    if (isnan(v)) continue;
                    buf[n++] = v;
                }
            int64_t o = (i - 1) + (j - 1) * nr;
            if (n == 0) { med[o] = NAN; mad[o] = NAN; continue; }
            insertion_sort(buf, n);
            float m = sorted_median(buf, n);
            for (int64_t k = 0; k < n; k++) buf[k] = fabsf(buf[k] - m);
            insertion_sort(buf, n);
            med[o] = m;
            mad[o] = sorted_median(buf, n);
        }
    free(buf);
}

/* _reject(..., ::GardnerFilter) -- src/outliers.jl:90-160; _mask_to_nan! 165-173 */
void orc_reject_gardner(const float* dx, const float* dy, const int64_t* radius_x,
                        const int64_t* radius_y, const uint8_t* valid, int64_t nr, int64_t nc,
                        int64_t up, const orc_gardner* f, uint8_t* keep) {
    int64_t N = nr * nc;
    float* nx = (float*)malloc(sizeof(float) * (size_t)N);
    float* ny = (float*)malloc(sizeof(float) * (size_t)N);
    float* ax = (float*)malloc(sizeof(float) * (size_t)N);
    float* ay = (float*)malloc(sizeof(float) * (size_t)N);
    float* medx = (float*)malloc(sizeof(float) * (size_t)N);
    float* madx = (float*)malloc(sizeof(float) * (size_t)N);
    float* medy = (float*)malloc(sizeof(float) * (size_t)N);
    float* mady = (float*)malloc(sizeof(float) * (size_t)N);
    for (int64_t k = 0; k < N; k++) {
        nx[k] = radius_x[k] > 0 ? dx[k] / (float)radius_x[k] : NAN;
        ny[k] = radius_y[k] > 0 ? dy[k] / (float)radius_y[k] : NAN;
        keep[k] = valid[k] ? 1 : 0;
    }
    float needed = (float)(f->min_agree_fraction * (double)(f->window * f->window));
    float tol = (float)f->agree_tolerance;
    for (int64_t it = 0; it < f->iterations; it++) {
        for (int64_t k = 0; k < N; k++)
            if (!keep[k]) { nx[k] = NAN; ny[k] = NAN; }
        orc_count_agreeing(nx, nr, nc, f->window, tol, ax);
        orc_count_agreeing(ny, nr, nc, f->window, tol, ay);
        for (int64_t k = 0; k < N; k++) keep[k] = ax[k] >= needed && ay[k] >= needed;
    }
    int64_t n2 = f->iterations - 1 > 1 ? f->iterations - 1 : 1;
    double scale = f->mad_scale;
    for (int64_t it = 0; it < n2; it++) {
        for (int64_t k = 0; k < N; k++)
            if (!keep[k]) { nx[k] = NAN; ny[k] = NAN; }
        orc_windowmedmad(nx, nr, nc, f->window, medx, madx);
        orc_windowmedmad(ny, nr, nc, f->window, medy, mady);
        for (int64_t k = 0; k < N; k++) {
            if (!keep[k]) continue;
            int64_t rx = radius_x[k], ry = radius_y[k];
            double floor_x = rx > 0 ? 2.0 / (double)(up * rx) : INFINITY;
            double floor_y = ry > 0 ? 2.0 / (double)(up * ry) : INFINITY;
            double tolx = jmaxd(scale * (double)madx[k], floor_x);
            double toly = jmaxd(scale * (double)mady[k], floor_y);
            keep[k] = (double)fabsf(nx[k] - medx[k]) <= tolx &&
                      (double)fabsf(ny[k] - medy[k]) <= toly;
        }
    }
    free(nx); free(ny); free(ax); free(ay); free(medx); free(madx); free(medy); free(mady);
}

/* ------------------------------------------------------------------------------------------ */
/* dilate_within / EDT -- src/resample.jl:211-325; Nearest resample 56-71                     */
/* ------------------------------------------------------------------------------------------ */

static void lower_envelope(double* dd, const double* f, int64_t* v, double* z, int64_t n) {
    int allinf = 1;
    for (int64_t j = 0; j < n; j++)
        if (!isinf(f[j])) { allinf = 0; break; }
    if (allinf) {
        for (int64_t j = 0; j < n; j++) dd[j] = INFINITY;
        return;
    }
    /* 1-based arrays: v[1..n], z[1..n+1] stored at index-1 */
    int64_t k = 1;
    v[0] = 1;
    z[0] = -INFINITY;
    z[1] = INFINITY;
    for (int64_t q = 2; q <= n; q++) {
        if (isinf(f[q - 1])) continue;
        double s = 0.0;
        for (;;) {
            int64_t p = v[k - 1];
            s = ((f[q - 1] + (double)(q * q)) - (f[p - 1] + (double)(p * p))) / (double)(2 * q - 2 * p);
            if (s > z[k - 1]) break;
            k -= 1;
            if (k == 0) break;
        }
        if (k == 0) {
            k = 1;
            v[0] = q;
            z[0] = -INFINITY;
            z[1] = INFINITY;
        } else {
            k += 1;
            v[k - 1] = q;
            z[k - 1] = s;
            z[k] = INFINITY;
        }
    }
    k = 1;
    for (int64_t q = 1; q <= n; q++) {
        while (z[k] < (double)q) k += 1;
        int64_t p = v[k - 1];
        dd[q - 1] = (double)((q - p) * (q - p)) + f[p - 1];
    }
}

void orc_sqdist_transform(const uint8_t* mask, int64_t nr, int64_t nc, double* d2) {
    for (int64_t j = 0; j < nc; j++) {
        double run = INFINITY;
        for (int64_t i = 0; i < nr; i++) {
            run = mask[i + j * nr] ? 0.0 : run + 1.0;
            d2[i + j * nr] = run;
        }
        run = INFINITY;
        for (int64_t i = nr - 1; i >= 0; i--) {
            run = mask[i + j * nr] ? 0.0 : run + 1.0;
            double cur = d2[i + j * nr];
            d2[i + j * nr] = cur < run ? cur : run;
        }
        for (int64_t i = 0; i < nr; i++) {
            double v = d2[i + j * nr];
            d2[i + j * nr] = isinf(v) ? INFINITY : v * v;
        }
    }
    double* f = (double*)malloc(sizeof(double) * (size_t)nc);
    double* dd = (double*)malloc(sizeof(double) * (size_t)nc);
    int64_t* v = (int64_t*)malloc(sizeof(int64_t) * (size_t)nc);
    double* z = (double*)malloc(sizeof(double) * (size_t)(nc + 1));
    for (int64_t i = 0; i < nr; i++) {
        for (int64_t j = 0; j < nc; j++) f[j] = d2[i + j * nr];
        lower_envelope(dd, f, v, z, nc);
        for (int64_t j = 0; j < nc; j++) d2[i + j * nr] = dd[j];
    }
    free(f); free(dd); free(v); free(z);
}

void orc_dilate_within(const uint8_t* mask, int64_t nr, int64_t nc, double radius, uint8_t* out) {
    double* d2 = (double*)malloc(sizeof(double) * (size_t)(nr * nc));
    orc_sqdist_transform(mask, nr, nc, d2);
    double r2 = radius * radius;
    for (int64_t k = 0; k < nr * nc; k++) out[k] = d2[k] <= r2;
    free(d2);
}

void orc_resample_nearest(const float* A, int64_t sr, int64_t sc, int64_t dr, int64_t dc,
                          float* out) {
    double ys = (double)sr / (double)dr, xs = (double)sc / (double)dc;
    for (int64_t j = 1; j <= dc; j++) {
        int64_t sj = jfloor(((double)j - 0.5) * xs) + 1;
        sj = sj < 1 ? 1 : (sj > sc ? sc : sj);
        for (int64_t i = 1; i <= dr; i++) {
            int64_t si = jfloor(((double)i - 0.5) * ys) + 1;
            si = si < 1 ? 1 : (si > sr ? sr : si);
            out[(i - 1) + (j - 1) * dr] = A[(si - 1) + (sj - 1) * sr];
        }
    }
}

/* _fill_holes! -- src/multichip.jl:347-390. Returns the count of filled (0-based) indices. */
int64_t orc_fill_holes(float* dx, float* dy, int64_t nr, int64_t nc, int64_t w, int64_t* filled) {
    int64_t lo = w / 2, hi = w - 1 - lo;
    int64_t needed = 2 * w * w / 3;
    float* bx = (float*)malloc(sizeof(float) * (size_t)(w * w));
    float* by = (float*)malloc(sizeof(float) * (size_t)(w * w));
    int64_t* pidx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nr * nc));
    float* pmx = (float*)malloc(sizeof(float) * (size_t)(nr * nc));
    float* pmy = (float*)malloc(sizeof(float) * (size_t)(nr * nc));
    int64_t nfilled = 0;
    for (int pass = 0; pass < 3; pass++) {
        int64_t np = 0;
        for (int64_t j = 1; j <= nc; j++)
            for (int64_t i = 1; i <= nr; i++) {
                if (!isnan(dx[(i - 1) + (j - 1) * nr])) continue;
                int64_t n = 0;
                for (int64_t jj = (j - lo > 1 ? j - lo : 1); jj <= (j + hi < nc ? j + hi : nc); jj++)
                    for (int64_t ii = (i - lo > 1 ? i - lo : 1); ii <= (i + hi < nr ? i + hi : nr); ii++) {
                        float v = dx[(ii - 1) + (jj - 1) * nr];
                        if (isnan(v)) continue;
                        bx[n] = v;
                        by[n] = dy[(ii - 1) + (jj - 1) * nr];
                        n++;
                    }
                if (n < needed) continue;
                insertion_sort(bx, n);
                insertion_sort(by, n);
                pidx[np] = (i - 1) + (j - 1) * nr;
                pmx[np] = sorted_median(bx, n);
                pmy[np] = sorted_median(by, n);
                np++;
            }
        if (np == 0) break;
        for (int64_t k = 0; k < np; k++) {
            dx[pidx[k]] = pmx[k];
            dy[pidx[k]] = pmy[k];
            filled[nfilled++] = pidx[k];
        }
    }
    free(bx); free(by); free(pidx); free(pmx); free(pmy);
    return nfilled;
}

/* ------------------------------------------------------------------------------------------ */
/* correlate_multichip -- src/multichip.jl:106-422; chip_sizes src/types.jl:842-850           */
/* ------------------------------------------------------------------------------------------ */

static int64_t chip_size_y_of(const orc_params* p, int64_t csx) {
    /* 2 * round(Int, csx * aspect / 2), round-half-even -- src/types.jl:914-915 */
    return 2 * (int64_t)nearbyint((double)csx * p->chip_aspect / 2.0);
}

/* sanitize! -- src/points.jl:332-346 */
static int64_t sanitize(int64_t* rx, int64_t* ry, int64_t n, int64_t min_radius) {
    int64_t k = 0;
    for (int64_t i = 0; i < n; i++) {
        if (rx[i] <= 0 || ry[i] <= 0) { rx[i] = 0; ry[i] = 0; }
        else {
            if (rx[i] < min_radius) rx[i] = min_radius;
            if (ry[i] < min_radius) ry[i] = min_radius;
            k++;
        }
    }
    return k;
}

static int64_t nsearchable(const int64_t* rx, const int64_t* ry, int64_t n) {
    int64_t k = 0;
    for (int64_t i = 0; i < n; i++) k += (rx[i] > 0 && ry[i] > 0);
    return k;
}

static void keep_from_filter(const orc_params* p, const orc_gardner* g, const float* dx,
                             const float* dy, const int64_t* rx, const int64_t* ry,
                             const uint8_t* valid, int64_t nr, int64_t nc, uint8_t* keep) {
    if (p->outliers_on) orc_reject_gardner(dx, dy, rx, ry, valid, nr, nc, p->upsampling, g, keep);
    else memcpy(keep, valid, (size_t)(nr * nc));
}

/* _cell_max_radius! -- src/multichip.jl:292-306: the largest radius over each coarse cell, rows
 * (i - stride/2):(i + stride - 1 - stride/2) clipped to the grid (left-biased, like the sliding
 * windows). out is (nr / stride) x (nc / stride), column-major. */
void orc_cell_max_radius(const int64_t* radius, int64_t nr, int64_t nc, int64_t stride,
                         int64_t* out) {
    int64_t nrc = nr / stride, ncc = nc / stride;
    int64_t lo = stride / 2, hi = stride - 1 - lo;
    for (int64_t jo = 1; jo <= ncc; jo++)
        for (int64_t io = 1; io <= nrc; io++) {
            int64_t i = io * stride, j = jo * stride;
            int64_t m = 0;
            for (int64_t jj = (j - lo > 1 ? j - lo : 1); jj <= (j + hi < nc ? j + hi : nc); jj++)
                for (int64_t ii = (i - lo > 1 ? i - lo : 1); ii <= (i + hi < nr ? i + hi : nr); ii++) {
                    int64_t v = radius[(ii - 1) + (jj - 1) * nr];
                    if (v > m) m = v;
                }
            out[(io - 1) + (jo - 1) * nrc] = m;
        }
}

/* The coarse pass's coherence gate -- src/multichip.jl:265-279: over the points with a positive
 * x radius that were measured, the kept fraction must not fall below min_fraction; no such point
 * at all also fails. Returns 1 when the level goes on. */
int orc_coarse_fraction_ok(const uint8_t* keep, const uint8_t* measured, const int64_t* radius_x,
                           int64_t n, double min_fraction) {
    int64_t denom = 0, numer = 0;
    for (int64_t k = 0; k < n; k++) {
        if (!(radius_x[k] > 0)) continue;
        if (!measured[k]) continue;
        denom++;
        if (keep[k]) numer++;
    }
    if (denom == 0) return 0;
    return !((double)numer / (double)denom < min_fraction);
}

/* _coarse_pass -- src/multichip.jl:224-288. Returns 1 with mask filled, 0 for "nothing". */
int orc_coarse_pass(const void* ref, const void* sec, const uint8_t* ok, int dt, int64_t ny,
                    int64_t nx, const orc_points* lv, int64_t nr, int64_t nc, int64_t csx,
                    int64_t csy, const orc_params* p, int num_mode, int nthreads,
                    uint8_t* mask) {
    int64_t stride = p->coarse_stride;
    int64_t nrc = nr / stride, ncc = nc / stride;
    orc_gardner filt = p->gardner;
    filt.iterations = filt.iterations - 1 > 1 ? filt.iterations - 1 : 1;
    int64_t w = p->outliers_on ? filt.window : 0;
    if (nrc < w || ncc < w) { memset(mask, 1, (size_t)(nr * nc)); return 1; }
    int64_t M = nrc * ncc;
    double* x = malloc(sizeof(double) * M); double* y = malloc(sizeof(double) * M);
    double* pdx = malloc(sizeof(double) * M); double* pdy = malloc(sizeof(double) * M);
    int64_t* rx = malloc(sizeof(int64_t) * M); int64_t* ry = malloc(sizeof(int64_t) * M);
    int64_t* cx = malloc(sizeof(int64_t) * M); int64_t* cy = malloc(sizeof(int64_t) * M);
    for (int64_t jo = 1; jo <= ncc; jo++)
        for (int64_t io = 1; io <= nrc; io++) {
            int64_t i = io * stride, j = jo * stride;
            int64_t src = (i - 1) + (j - 1) * nr, dst = (io - 1) + (jo - 1) * nrc;
            x[dst] = lv->x[src]; y[dst] = lv->y[src];
            pdx[dst] = lv->dx_prior[src]; pdy[dst] = lv->dy_prior[src];
            cx[dst] = csx; cy[dst] = csy;
        }
    orc_cell_max_radius(lv->radius_x, nr, nc, stride, rx);
    orc_cell_max_radius(lv->radius_y, nr, nc, stride, ry);
    int result = 0;
    float *cdx = NULL, *cdy = NULL, *ccorr = NULL, *gm = NULL;
    uint8_t *srch = NULL, *meas = NULL, *keep = NULL, *grown = NULL;
    if (nsearchable(rx, ry, M) == 0) goto done;
    cdx = malloc(sizeof(float) * M); cdy = malloc(sizeof(float) * M);
    ccorr = malloc(sizeof(float) * M); srch = malloc(M);
    orc_points cp = {x, y, rx, ry, pdx, pdy, cx, cy};
    orc_track(ref, sec, ok, dt, ny, nx, &cp, M, 1, p->measure, num_mode, nthreads, p->nrot,
              p->angles, cdx, cdy,
              ccorr, srch);
    meas = malloc(M); keep = malloc(M);
    int anym = 0;
    for (int64_t k = 0; k < M; k++) { meas[k] = !isnan(cdx[k]); anym |= meas[k]; }
    if (!anym) goto done;
    keep_from_filter(p, &filt, cdx, cdy, rx, ry, meas, nrc, ncc, keep);
    for (int64_t k = 0; k < M; k++) if (!meas[k]) keep[k] = 0;
    if (!orc_coarse_fraction_ok(keep, meas, rx, M, p->min_coarse_valid_fraction)) goto done;
    grown = malloc(M);
    orc_dilate_within(keep, nrc, ncc, (double)p->coarse_buffer, grown);
    float* gf = malloc(sizeof(float) * M);
    for (int64_t k = 0; k < M; k++) gf[k] = grown[k] ? 1.0f : 0.0f;
    gm = malloc(sizeof(float) * nr * nc);
    orc_resample_nearest(gf, nrc, ncc, nr, nc, gm);
    free(gf);
    for (int64_t k = 0; k < nr * nc; k++) mask[k] = gm[k] > 0.5f;
    result = 1;
done:
    free(x); free(y); free(pdx); free(pdy); free(rx); free(ry); free(cx); free(cy);
    free(cdx); free(cdy); free(ccorr); free(srch); free(meas); free(keep); free(grown); free(gm);
    return result;
}

/* _merge_level! -- src/multichip.jl:401-422: the level's value goes wherever no finer level owns
 * the point (chip_size 0) and the level measured or filled it; interpolated marks the filled
 * indices. wasfilled is n bytes of scratch. */
void orc_merge_level(float* rdx, float* rdy, float* rcorr, uint16_t* rcs, uint8_t* rinterp,
                     const float* fdx, const float* fdy, const float* fcorr,
                     const int64_t* filled, int64_t nfilled, int64_t n, int64_t chip_size,
                     uint8_t* wasfilled) {
    memset(wasfilled, 0, (size_t)n);
    for (int64_t k = 0; k < nfilled; k++) wasfilled[filled[k]] = 1;
    for (int64_t k = 0; k < n; k++) {
        if (rcs[k] != 0) continue;
        if (isnan(fdx[k])) continue;
        rdx[k] = fdx[k]; rdy[k] = fdy[k]; rcorr[k] = fcorr[k];
        rcs[k] = (uint16_t)chip_size;
        rinterp[k] = wasfilled[k];
    }
}

int orc_multichip(const void* ref, const void* sec, const uint8_t* ok_mask, int dt, int64_t ny,
                  int64_t nx, const orc_points* grid, int64_t nr, int64_t nc,
                  const orc_params* p, int num_mode, int nthreads, float* rdx, float* rdy,
                  float* rcorr, uint16_t* rcs, uint8_t* rinterp) {
    int64_t N = nr * nc;
    for (int64_t k = 0; k < N; k++) {
        rdx[k] = NAN; rdy[k] = NAN; rcorr[k] = NAN; rcs[k] = 0; rinterp[k] = 0;
    }
    int64_t* rx = malloc(sizeof(int64_t) * N); int64_t* ry = malloc(sizeof(int64_t) * N);
    int64_t* cx = malloc(sizeof(int64_t) * N); int64_t* cy = malloc(sizeof(int64_t) * N);
    uint8_t* cmask = malloc(N); uint8_t* meas = malloc(N); uint8_t* keep = malloc(N);
    uint8_t* srch = malloc(N); uint8_t* wasfilled = malloc(N);
    float* fdx = malloc(sizeof(float) * N); float* fdy = malloc(sizeof(float) * N);
    float* fcorr = malloc(sizeof(float) * N);
    int64_t* filled = malloc(sizeof(int64_t) * N);
    const orc_params* p0 = p;
    int level = 0;
    for (int64_t cs = p0->chip_size_base; cs <= p0->chip_size_max; cs *= 2) {
        if (cs < p0->chip_size_min) continue;
        /* the level's measure (measure_at): the tuple's entry, the last one repeating */
        orc_params pl = *p0;
        if (p0->nmeas > 0) pl.measure = p0->measures[level < p0->nmeas ? level : p0->nmeas - 1];
        level++;
        const orc_params* p = &pl;
        int anyw = 0;
        for (int64_t k = 0; k < N; k++) anyw |= (rcs[k] == 0);
        if (!anyw) break;
        int64_t csx = cs, csy = chip_size_y_of(p, cs);
        for (int64_t k = 0; k < N; k++) {
            int wanted = rcs[k] == 0;
            rx[k] = wanted ? grid->radius_x[k] : 0;
            ry[k] = wanted ? grid->radius_y[k] : 0;
            cx[k] = csx; cy[k] = csy;
        }
        if (sanitize(rx, ry, N, p->min_search_radius) == 0) continue;
        orc_points lv = {grid->x, grid->y, rx, ry, grid->dx_prior, grid->dy_prior, cx, cy};
        if (!orc_coarse_pass(ref, sec, ok_mask, dt, ny, nx, &lv, nr, nc, csx, csy, p, num_mode,
                         nthreads, cmask))
            continue;
        for (int64_t k = 0; k < N; k++)
            if (!cmask[k]) { rx[k] = 0; ry[k] = 0; }
        if (nsearchable(rx, ry, N) == 0) continue;
        orc_track(ref, sec, ok_mask, dt, ny, nx, &lv, N, p->upsampling, p->measure, num_mode,
                  nthreads, p->nrot, p->angles, fdx, fdy, fcorr, srch);
        /* _reject_and_fill! -- src/multichip.jl:318-330 */
        for (int64_t k = 0; k < N; k++) meas[k] = !isnan(fdx[k]);
        keep_from_filter(p, &p->gardner, fdx, fdy, rx, ry, meas, nr, nc, keep);
        for (int64_t k = 0; k < N; k++)
            if (!keep[k]) { fdx[k] = NAN; fdy[k] = NAN; fcorr[k] = NAN; }
        int64_t nf = orc_fill_holes(fdx, fdy, nr, nc, p->fill_window, filled);
        orc_merge_level(rdx, rdy, rcorr, rcs, rinterp, fdx, fdy, fcorr, filled, nf, N, cs,
                        wasfilled);
    }
    free(rx); free(ry); free(cx); free(cy); free(cmask); free(meas); free(keep); free(srch);
    free(wasfilled); free(fdx); free(fdy); free(fcorr); free(filled);
    return 0;
}
