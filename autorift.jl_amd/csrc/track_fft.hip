// FFT tracker for Float32 prepared images (precision = :fft): the reference's numerator by
// transform (_numerators_fft!, src/correlate.jl:816-848) in single precision, one workgroup per
// point with the whole transform in LDS.
//
// Per point: the window goes in the real part and the reversed mean-removed chip
// Float32(x - mean) in the imaginary part of one zero-padded fy x fx complex array (fy, fx =
// next_fft_size of the pass's window extents, src/correlate.jl:237-238); one forward 2-D
// transform gives both spectra (A = (Z[k] + conj Z[-k]) / 2, B = (Z[k] - conj Z[-k]) / 2i), their
// product P = A B is Hermitian, and the real part of its inverse, times 1 / (fy fx), is the
// correlation; the valid block starts at (ch - 1, cw - 1) as in the reference's read. The 1-D
// transforms are mixed-radix Stockham passes (radices 4, 2, 3, 5, 7) over every column, then
// every row; each wave takes a few lines at a time through all their passes, reading each pass's
// inputs into registers before writing its outputs in place. Twiddles are Float64 sincospi
// values rounded once, held in LDS. The normalisation is the reference's Float64 formula on the pass's Float64 box images
// (src/correlate.jl:575-630), the argmax follows peak_index (src/peak.jl:32-61), and the 5x5
// patch handed to the subpixel refinement is read from the same surface, as the reference does.
// Results agree with the reference's FFTW numerators to single-precision rounding, not bit for
// bit (FFTW's butterfly order is its own).
#include <hip/hip_runtime.h>

#include <cmath>

#include "ar_internal.h"
#include "layout.h"
#include "refine.h"
#include "track_common.h"

namespace ar {

__device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

// floor(a / d) for 0 <= a < 256 and 1 <= d <= 128 as (a m) >> 16, m = ceil(2^16 / d): with
// e = m d - 2^16 < d, a e < 2^16 keeps the product's excess over a / d below 1 / d, so the floor
// is exact.
struct SmallDiv {
    int d, m;
};
__device__ inline SmallDiv small_div(int d) { return SmallDiv{d, (65536 + d - 1) / d}; }
__device__ inline int sdiv(int a, SmallDiv v) { return (a * v.m) >> 16; }
__device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Radices of n, largest powers of 4 first, as 4-bit fields of the result (stage s in bits
// 4s .. 4s + 3; a zero field ends the list); 0 when n has a prime factor above 7.
__device__ __host__ inline uint64_t fft_radices(int n) {
    uint64_t r = 0;
    int k = 0;
    while (n % 4 == 0) { r |= (uint64_t)4 << (4 * k++); n /= 4; }
    while (n % 2 == 0) { r |= (uint64_t)2 << (4 * k++); n /= 2; }
    for (int p = 3; p <= 7; p += 2)
        while (n % p == 0) { r |= (uint64_t)p << (4 * k++); n /= p; }
    return n == 1 ? r : 0;
}

// cos and sin of 2 pi k / R for the odd radices (k < R)
template <int R> struct Roots;
template <> struct Roots<3> {
    static constexpr float c[3] = {1.0f, -0.5f, -0.5f};
    static constexpr float s[3] = {0.0f, 0.86602540378443864676f, -0.86602540378443864676f};
};
template <> struct Roots<5> {
    static constexpr float c[5] = {1.0f, 0.30901699437494742410f, -0.80901699437494742410f,
                                   -0.80901699437494742410f, 0.30901699437494742410f};
    static constexpr float s[5] = {0.0f, 0.95105651629515357212f, 0.58778525229247312917f,
                                   -0.58778525229247312917f, -0.95105651629515357212f};
};
template <> struct Roots<7> {
    static constexpr float c[7] = {1.0f, 0.62348980185873353053f, -0.22252093395631440429f,
                                   -0.90096886790241912624f, -0.90096886790241912624f,
                                   -0.22252093395631440429f, 0.62348980185873353053f};
    static constexpr float s[7] = {0.0f, 0.78183148246802980871f, 0.97492791218182360702f,
                                   0.43388373911755812048f, -0.43388373911755812048f,
                                   -0.97492791218182360702f, -0.78183148246802980871f};
};

// DFT of odd size R by its symmetric form: with s_r = v_r + v_{R-r}, d_r = v_r - v_{R-r},
// out[q] = v_0 + sum_r (s_r cos - i d_r sin)(2 pi r q / R) and out[R - q] with +i (forward).
template <int R, bool INV>
__device__ inline void dft_odd(const float2 (&v)[R], float2 (&o)[R]) {
    constexpr int H = (R - 1) / 2;
    float2 sm[H + 1], df[H + 1];
    float2 o0 = v[0];
#pragma unroll
    for (int r = 1; r <= H; r++) {
        sm[r] = cadd(v[r], v[R - r]);
        df[r] = csub(v[r], v[R - r]);
        o0 = cadd(o0, sm[r]);
    }
    o[0] = o0;
#pragma unroll
    for (int q = 1; q <= H; q++) {
        float2 av = v[0], bv = make_float2(0.0f, 0.0f);
#pragma unroll
        for (int r = 1; r <= H; r++) {
            const float c = Roots<R>::c[(r * q) % R], sn = Roots<R>::s[(r * q) % R];
            av.x = fmaf(sm[r].x, c, av.x);
            av.y = fmaf(sm[r].y, c, av.y);
            bv.x = fmaf(df[r].x, sn, bv.x);
            bv.y = fmaf(df[r].y, sn, bv.y);
        }
        // forward: out[q] = a - i b, out[R - q] = a + i b; inverse the other way round
        const float2 mib = make_float2(bv.y, -bv.x), pib = make_float2(-bv.y, bv.x);
        o[q] = cadd(av, INV ? pib : mib);
        o[R - q] = cadd(av, INV ? mib : pib);
    }
}

// DFT of size R (forward: exp(-2 pi i r q / R); INV: the conjugate kernel)
template <int R, bool INV>
__device__ inline void dft(const float2 (&v)[R], float2 (&o)[R]) {
    if constexpr (R == 2) {
        o[0] = cadd(v[0], v[1]);
        o[1] = csub(v[0], v[1]);
    } else if constexpr (R == 4) {
        const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        const float2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
        // forward: (-i) b1; inverse: (+i) b1
        const float2 jb = INV ? make_float2(-b1.y, b1.x) : make_float2(b1.y, -b1.x);
        o[0] = cadd(a0, b0);
        o[2] = csub(a0, b0);
        o[1] = cadd(a1, jb);
        o[3] = csub(a1, jb);
    } else {
        dft_odd<R, INV>(v, o);
    }
}

// Threads of the FFT tracker's workgroup: eight waves (the transform's LDS admits two
// workgroups per CU; more waves hide more of its latency).
#define FFT_WG AR_FFT_WG
#define FFT_NW (FFT_WG / 64)

// The workgroup's items of one pass: item = line + nl * butterfly, thread t takes items
// t, t + FFT_WG, ...; consecutive lanes take consecutive lines, so every access of a stage
// moves along the line stride (1, or the odd column stride) and no two lanes of a half-wave meet
// on an LDS bank.
struct PassWalk {
    int line, bf, dl, db;
};
__device__ inline PassWalk pass_walk(int nl) {
    PassWalk w;
    w.bf = threadIdx.x / nl;
    w.line = threadIdx.x - w.bf * nl;
    w.db = FFT_WG / nl;
    w.dl = FFT_WG - w.db * nl;
    return w;
}

// One in-place radix-R stage over lines l0 .. l0 + nl - 1 (element t of line l at
// Z[l ls + t es]) of length n: blocks of L = R m elements, butterfly (blk, j) on the elements
// blk L + j + r m. Forward (decimation in frequency): DFT_R, then output q times
// exp(-2 pi i j q / L). Inverse (decimation in time, the adjoint): inputs times the conjugate
// factors, then the conjugate DFT_R. T holds exp(-2 pi i k / n), k < n.
template <int R, bool INV>
__device__ inline void fft_stage(float2* Z, const float2* T, int n, int L, int m, int l0, int nl,
                                 int es, int ls, PassWalk w) {
    const int nbf = n / R, tw = n / L, st = m * es;
    const SmallDiv dm = small_div(m);
    for (; w.bf < nbf;) {
        const int blk = sdiv(w.bf, dm), j = w.bf - blk * m;
        const int base = (l0 + w.line) * ls + (blk * L + j) * es;
        float2 v[R], o[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = Z[base + r * st];
        if (INV && m > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) v[r] = cmul(v[r], cconj(T[j * r * tw]));
        }
        dft<R, INV>(v, o);
        if (!INV && m > 1) {
#pragma unroll
            for (int q = 1; q < R; q++) o[q] = cmul(o[q], T[j * q * tw]);
        }
#pragma unroll
        for (int q = 0; q < R; q++) Z[base + q * st] = o[q];
        w.line += w.dl;
        w.bf += w.db;
        if (w.line >= nl) {
            w.line -= nl;
            w.bf++;
        }
    }
}

// The 1-D transform of length n (at most 128, radices rad) of lines l0 .. l0 + nl - 1 along
// one axis. Forward leaves frequency f at its mixed-radix digit-reversed position (f = q1 + R1
// (q2 + R2 (...)) at q1 m1 + q2 m2 + ..., m_s = n / (R1 ... R_s)); inverse takes that order back
// to natural order. A workgroup barrier ends every stage.
template <bool INV>
__device__ __attribute__((always_inline)) inline void fft_pass(float2* Z, const float2* T, int n,
                                                               uint64_t rad, int l0, int nl, int es,
                                                               int ls) {
    const PassWalk w = pass_walk(nl);
    int ns = 0;
    for (uint64_t rr = rad; rr != 0; rr >>= 4) ns++;
    int m = INV ? 1 : n;
    for (int k = 0; k < ns; k++) {
        const int R = (int)((rad >> (4 * (INV ? ns - 1 - k : k))) & 15);
        int L;
        if (INV) {
            L = m * R;
        } else {
            L = m;
            m = L / R;
        }
        switch (R) {
            case 4: fft_stage<4, INV>(Z, T, n, L, m, l0, nl, es, ls, w); break;
            case 2: fft_stage<2, INV>(Z, T, n, L, m, l0, nl, es, ls, w); break;
            case 3: fft_stage<3, INV>(Z, T, n, L, m, l0, nl, es, ls, w); break;
            case 5: fft_stage<5, INV>(Z, T, n, L, m, l0, nl, es, ls, w); break;
            default: fft_stage<7, INV>(Z, T, n, L, m, l0, nl, es, ls, w); break;
        }
        if (INV) m = L;
        __syncthreads();
    }
}

// Position of -f for the frequency f held at digit-reversed position p (see fft_pass).
__device__ inline int fft_mirror_pos(int p, int n, uint64_t rad) {
    int m = n, P = 1, f = 0, rem = p;
    for (uint64_t rr = rad; rr != 0; rr >>= 4) {
        const int R = (int)(rr & 15);
        m /= R;
        const int q = rem / m;
        rem -= q * m;
        f += q * P;
        P *= R;
    }
    int mf = f == 0 ? 0 : n - f, pos = 0;
    m = n;
    for (uint64_t rr = rad; rr != 0; rr >>= 4) {
        const int R = (int)(rr & 15);
        m /= R;
        pos += (mf % R) * m;
        mf /= R;
    }
    return pos;
}

// LDS of the FFT tracker: the complex array (columns of ld = fy | 1 elements: an odd column
// stride keeps the column passes' lanes on distinct banks), the two twiddle tables, the two mirror
// tables (position of -f per position of f), and one Float32 plane for the chip.
struct FftLayout {
    size_t off_z, off_ty, off_tx, off_my, off_mx, off_s, total;
};
__host__ __device__ inline FftLayout fft_layout(int fy, int fx, int max_ch, int max_cw) {
    FftLayout L;
    L.off_z = 0;
    L.off_ty = (size_t)(fy | 1) * fx * 8;
    L.off_tx = L.off_ty + (size_t)fy * 8;
    L.off_my = L.off_tx + (size_t)fx * 8;
    L.off_mx = L.off_my + 128;
    L.off_s = L.off_mx + 128;
    L.total = L.off_s + 4 * (size_t)max_ch * max_cw;
    return L;
}

// Block reductions over the FFT_NW waves (every thread gets the result).
__device__ inline double fft_block_sum(double v, double* red) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = red[0];
    for (int k = 1; k < FFT_NW; k++) t += red[k];
    return t;
}
__device__ inline Best fft_block_best(Best b, Best* red) {
    b = wave_best(b);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
    __syncthreads();
    Best r = red[0];
    for (int k = 1; k < FFT_NW; k++) r = best_merge(r, red[k]);
    return r;
}
// _any_valid (src/track.jl:416-421) over the chip rectangle, block-uniform.
__device__ inline bool fft_any_valid(const TrackArgs& a, const Geom& g) {
    if (a.ok == nullptr) {
        const int64_t r0 = g.cr0 - a.pady - 1, c0 = g.cc0 - a.padx - 1;
        return r0 + g.ch - 1 >= 0 && r0 < a.ny && c0 + g.cw - 1 >= 0 && c0 < a.nx;
    }
    int found = 0;
    for (int k = threadIdx.x; k < g.ch * g.cw; k += FFT_WG) {
        const int v = k / g.ch, u = k - v * g.ch;
        const int64_t q = pix(a, g.cr0 + u, g.cc0 + v);
        if (q >= 0 && a.ok[q]) found = 1;
    }
    return __syncthreads_or(found) != 0;
}

__global__ __launch_bounds__(FFT_WG) __attribute__((amdgpu_waves_per_eu(4))) void track_fft_kernel(
    TrackArgs a0) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red_d[FFT_NW];
    __shared__ Best red[FFT_NW];
    TrackArgs a = a0;
    const int tid = threadIdx.x;
    const int fy = a.fft_fy, fx = a.fft_fx;
    const uint64_t rady = fft_radices(fy), radx = fft_radices(fx);
    const FftLayout L = fft_layout(fy, fx, a.max_ch, a.max_cw);
    float2* Z = (float2*)(smem + L.off_z);
    float2* Ty = (float2*)(smem + L.off_ty);
    float2* Tx = (float2*)(smem + L.off_tx);
    unsigned char* mir_y = smem + L.off_my;
    unsigned char* mir_x = smem + L.off_mx;
    float* plane = (float*)(smem + L.off_s);
    // twiddles exp(-2 pi i m / n): Float64 sincospi, rounded once; mirror positions
    for (int m = tid; m < fy + fx; m += FFT_WG) {
        const bool y = m < fy;
        const int n = y ? fy : fx, mm = y ? m : m - fy;
        double sv, cv;
        sincospi(2.0 * (double)mm / (double)n, &sv, &cv);
        (y ? Ty : Tx)[mm] = make_float2((float)cv, (float)-sv);
        (y ? mir_y : mir_x)[mm] = (unsigned char)fft_mirror_pos(mm, n, y ? rady : radx);
    }
    __syncthreads();
    const int ld = fy | 1;
    const float inv_nz = 1.0f / (float)(fy * fx);

    const PointRange pr = point_range(a.n);
    for (int64_t p = pr.first; p < pr.end; p += pr.step) {
        // The argument block is read again from the kernarg segment at each phase and the lane
        // index is hidden from the optimiser per point: values the loop does not need across
        // phases are formed where they are used instead of being held in (spilled) registers.
        a = reload_args();
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int lane = tid & 63, wave = tid >> 6;
        const float* ref = (const float*)a.ref;
        const float* sec = (const float*)a.sec;
        const Geom g = point_geom(a, p);
        if (!g.status || !fft_any_valid(a, g)) { write_skip(a, p, false); continue; }
        const int ch = g.ch, cw = g.cw, NR = 2 * g.ry, NC = 2 * g.rx;
        const int winh = ch + NR - 1, winw = cw + NC - 1;

        // ---- chip: Float32 samples, Float64 mean and energy (block sums in a fixed order).
        // Extents are at most 128 (two rows per lane); four columns per wave at a time keep
        // eight loads in flight.
        const int nn = ch * cw;
        for (int v0 = wave; v0 < cw; v0 += 4 * FFT_NW) {
            float c[4][2];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int v = v0 + (e >> 1) * FFT_NW, u = lane + 64 * (e & 1);
                c[e >> 1][e & 1] = v < cw && u < ch ? mp_pix(a, sec, g.cr0 + u, g.cc0 + v) : 0.0f;
            }
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int v = v0 + (e >> 1) * FFT_NW, u = lane + 64 * (e & 1);
                if (v < cw && u < ch) plane[u + v * ch] = c[e >> 1][e & 1];
            }
        }
        __syncthreads();
        double s = 0.0;
        for (int k = tid; k < nn; k += FFT_WG) s += (double)plane[k];
        const double mean = fft_block_sum(s, red_d) / (double)nn;
        double q = 0.0;
        for (int k = tid; k < nn; k += FFT_WG) {
            const double d = (double)plane[k] - mean;
            q += d * d;
        }
        const double sq = fft_block_sum(q, red_d);
        if (!(sq > 2.220446049250313e-16 * (double)ch * (double)cw)) {   // degenerate chip
            write_skip(a, p, true);
            __syncthreads();
            continue;
        }
        const double cnorm = sqrt(sq);

        // ---- Z = window + i * reversed Float32(chip - mean), zero-padded to fy x fx
        {
            const int64_t orow = g.wr0 - a.pady - 1, ocol = g.wc0 - a.padx - 1;
            const bool inside = orow >= 0 && ocol >= 0 && orow + winh <= a.ny &&
                                ocol + winw <= a.nx;
            const float* wb = ref + (inside ? orow + ocol * a.ny : 0);
            for (int j0 = wave; j0 < fx; j0 += 4 * FFT_NW) {
                float re[4][2];
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int j = j0 + (e >> 1) * FFT_NW, i = lane + 64 * (e & 1);
                    float x = 0.0f;
                    if (i < winh && j < winw)
                        x = inside ? wb[i + (int64_t)j * a.ny] : mp_pix(a, ref, g.wr0 + i, g.wc0 + j);
                    re[e >> 1][e & 1] = x;
                }
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int j = j0 + (e >> 1) * FFT_NW, i = lane + 64 * (e & 1);
                    if (j >= fx || i >= fy) continue;
                    float im = 0.0f;
                    if (i < ch && j < cw)
                        im = (float)((double)plane[(ch - 1 - i) + (cw - 1 - j) * ch] - mean);
                    Z[i + j * ld] = make_float2(re[e >> 1][e & 1], im);
                }
            }
        }
        __syncthreads();

        a = reload_args();
        // ---- forward transform: columns (length fy; columns from winw on are zero), then rows
        fft_pass<false>(Z, Ty, fy, rady, 0, winw, 1, ld);
        fft_pass<false>(Z, Tx, fx, radx, 0, fy, ld, 1);

        // ---- the two spectra and their product, in place: A = (Z[k] + conj Z[-k]) / 2 and
        // B = (Z[k] - conj Z[-k]) / 2i are Hermitian, so P[-k] = conj P[k]; each thread owns the
        // pairs {k, -k} whose first position is the smaller (positions are digit-reversed)
        for (int j = wave; j < fx; j += FFT_NW)
            for (int i = lane; i < fy; i += 64) {
                const int mi = mir_y[i], mj = mir_x[j];
                if (mi + mj * fy < i + j * fy) continue;
                const int k = i + j * ld, mk = mi + mj * ld;
                const float2 z = Z[k], zm = cconj(Z[mk]);
                const float2 A = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y + zm.y));
                const float2 d = csub(z, zm);
                const float2 Bq = make_float2(0.5f * d.y, -0.5f * d.x);   // d / 2i
                const float2 P = cmul(A, Bq);
                Z[k] = P;
                if (mk != k) Z[mk] = cconj(P);
            }
        __syncthreads();

        // ---- inverse transform (natural order again); only rows ch - 1 .. ch + NR - 2 of the
        // row pass are read. The correlation is the real part over fy fx.
        fft_pass<true>(Z, Ty, fy, rady, 0, fx, 1, ld);
        fft_pass<true>(Z, Tx, fx, radx, ch - 1, NR, ld, 1);

        // ---- score per shift and the argmax (peak_index rule): the FFT numerator over the
        // window deviation from the pass's Float32 window variances (formed in Float64 by the box
        // pass); the reported peak and 5x5 patch take the reference's Float64 normalisation
        // (src/correlate.jl:575-630) below
        a = reload_args();
        const double n = (double)(ch * cw);
        const int64_t bb = (g.wr0 - 1) + (g.wc0 - 1) * a.box_ld;
        const float inv_cn = (float)(1.0 / cnorm);
        Best b = best_empty();
        for (int j0 = wave; j0 < NC; j0 += 4 * FFT_NW) {
            float den[4][2];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int j = j0 + (e >> 1) * FFT_NW, i = lane + 64 * (e & 1);
                den[e >> 1][e & 1] = j < NC && i < NR ? a.fwv[bb + i + (int64_t)j * a.box_ld] : 0.0f;
            }
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int j = j0 + (e >> 1) * FFT_NW, i = lane + 64 * (e & 1);
                if (j >= NC || i >= NR) continue;
                const float d = den[e >> 1][e & 1];
                const float num = Z[(i + ch - 1) + (j + cw - 1) * ld].x * inv_nz;
                const float v = d > 0.0f ? num * inv_cn * __builtin_amdgcn_rsqf(d) : 0.0f;
                const int idx = i * NC + j;
                if (better(v, idx, b)) { b.v = v; b.idx = idx; }
            }
        }
        b = fft_block_best(b, red);
        if (b.idx == 0x7fffffff) b.idx = 0;
        const int pi = b.idx / NC + 1, pj = b.idx % NC + 1;
        int i0 = 1, j0 = 1;
        const bool refine = a.up > 1 && patch_origin(pi, pj, NR, NC, i0, j0);
        // exact values: the peak (every thread) and the patch (threads 0..24, row-major)
        auto exact = [&](int i, int j) {   // 0-based shift
            const double num = (double)(Z[(i + ch - 1) + (j + cw - 1) * ld].x * inv_nz);
            const int64_t qb = bb + i + (int64_t)j * a.box_ld;
            return mp_normalise(num, a.fbox1[qb], a.fbox2[qb], n, cnorm, a.measure);
        };
        const float pval = exact(pi - 1, pj - 1);
        float pv = 0.0f;
        if (refine && tid < 25) pv = exact(i0 - 1 + tid / 5, j0 - 1 + tid % 5);
        a = reload_args();
        finish_point(a, p, g, pi, pj, pval, i0, j0, refine, pv);
        __syncthreads();
    }
}


}  // namespace ar

using namespace ar;

// next_fft_size (src/correlate.jl:357-403): the 7-smooth m in [n, 1.5 n] of least
// m^2 log2(m) (1 - 0.35 (twos / log2 m)).
int64_t ar_next_fft_size(int64_t n) {
    if (n <= 1) return 1;
    int64_t best = 0;
    double best_cost = INFINITY;
    const int64_t hi = (int64_t)std::ceil(1.5 * (double)n);
    for (int64_t m = n; m <= hi; m++) {
        int64_t r = m;
        int twos = 0;
        while (r % 2 == 0) { r /= 2; twos++; }
        for (int64_t p : {3, 5, 7})
            while (r % p == 0) r /= p;
        if (r != 1) continue;
        const double lg = std::log2((double)m);
        const double cost = (double)m * (double)m * lg * (1.0 - 0.35 * ((double)twos / lg));
        if (cost < best_cost) {
            best = m;
            best_cost = cost;
        }
    }
    return best;
}

ar_status ar_fft_setup(ar_context* ctx, TrackArgs& a, void (**kern)(TrackArgs), size_t* lds) {
    const int64_t fy = ar_next_fft_size(a.max_ch + 2 * a.max_ry - 1);
    const int64_t fx = ar_next_fft_size(a.max_cw + 2 * a.max_rx - 1);
    if (fy > 128 || fx > 128 || !fft_radices((int)fy) || !fft_radices((int)fx))
        return AR_ERR_UNSUPPORTED;
    const FftLayout L = fft_layout((int)fy, (int)fx, a.max_ch, a.max_cw);
    if (L.total > AR_LDS_LIMIT) return AR_ERR_UNSUPPORTED;
    a.fft_fy = (int32_t)fy;
    a.fft_fx = (int32_t)fx;
    if (ar_status st = ar_f64_box_images(ctx, a)) return st;
    *kern = track_fft_kernel;
    *lds = L.total;
    return AR_OK;
}
