// The per-point ZNCC tracker: one 256-thread workgroup correlates one point at a time and the
// grid is persistent (static stride over points, so the assignment is deterministic).
//
// Two numerator paths, both exact with respect to the CPU restatement of the reference:
//
//  * u8 (the default pipeline: high-pass then UInt8 quantisation) with a power-of-two chip
//    area: the mean-removed chip equals (n*x - sum)/n exactly, so the numerator is an integer
//    identity.  Sum(W*T) runs on the int8 matrix cores (v_mfma_i32_16x16x64_i8, operands centred
//    by -128) as a banded-Toeplitz GEMM: rows are column shifts j, columns are row shifts i and
//    the reduction runs over (chip column v, window row r).  Box sums of W and W^2 are integer
//    sliding sums.  Every shift is first scored in FP32 with a proven error bound; only the
//    shifts that can still be the maximum, and the 5x5 refinement patch, are evaluated in the
//    reference's FP64 normalisation (src/correlate.jl:591-600).
//
//  * generic (float images, or chip areas that are not a power of two): the reference's FP64
//    arithmetic in its own order -- sequential chip mean and energy (src/correlate.jl:423-453),
//    the summed-area recurrence of src/integral.jl:141-170 and the direct numerator of
//    src/correlate.jl:799-814 -- with per-block scratch in global memory.
#include "layout.h"
#include "refine.h"
#include "track_common.h"

namespace ar {

typedef int v4i __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------------------------
// shared tail: peak, refinement, outputs
// ---------------------------------------------------------------------------------------------

__device__ inline void write_skip(const TrackArgs& a, int64_t p, bool searched) {
    if (threadIdx.x == 0) {
        a.dx[p] = __int_as_float(0x7fc00000);
        a.dy[p] = __int_as_float(0x7fc00000);
        a.corr[p] = __int_as_float(0x7fc00000);
        a.searched[p] = searched ? 1 : 0;
    }
}

// Offsets of src/peak.jl:102-106 and src/peak.jl:515-521, plus the prior (src/track.jl:293-295).
__device__ inline void write_result(const TrackArgs& a, int64_t p, const Geom& g, double ddx,
                                    double ddy, float val) {
    if (threadIdx.x == 0) {
        a.dx[p] = (float)(ddx + g.pdx);
        a.dy[p] = (float)(ddy + g.pdy);
        a.corr[p] = val;
        a.searched[p] = 1;
    }
}

// After the integer peak (pi, pj) (1-based) and its exact value are known, and with the 5x5
// patch (exact values, row-major) in `reg[0..24]`, finish the point.
__device__ inline void finish_point(const TrackArgs& a, int64_t p, const Geom& g, int pi, int pj,
                                    float pval, int i0, int j0, bool refine, float* reg,
                                    Best* red) {
    if (!refine) {
        write_result(a, p, g, (double)(pj - g.rx - 1), (double)(pi - g.ry - 1), pval);
        return;
    }
    CascadeOut c = cascade(reg, a.up, red);
    double factor = (double)a.up;
    double row = (double)c.si / factor + (double)(i0 - 1);
    double col = (double)c.sj / factor + (double)(j0 - 1);
    write_result(a, p, g, col - (double)g.rx, row - (double)g.ry, c.val);
}

// Patch origin of src/peak.jl:473-478 (1-based); returns false when the surface is smaller than
// the patch, in which case the integer peak is reported.
__device__ inline bool patch_origin(int pi, int pj, int nr, int nc, int& i0, int& j0) {
    int hr = nr - 5 + 1 > 1 ? nr - 5 + 1 : 1;
    int hc = nc - 5 + 1 > 1 ? nc - 5 + 1 : 1;
    i0 = pi - 2 < 1 ? 1 : (pi - 2 > hr ? hr : pi - 2);
    j0 = pj - 2 < 1 ? 1 : (pj - 2 > hc ? hc : pj - 2);
    return !(nr < 5 || nc < 5);
}

// ---------------------------------------------------------------------------------------------
// u8 path
// ---------------------------------------------------------------------------------------------

// Upper bound on |approximate - exact| for the FP32 screening score (see u8 epilogue).
#define AR_SCREEN_EPS 4e-6f

struct U8Point {
    int64_t n;       // chip area (power of two)
    int64_t sumT;    // sum of chip samples
    double cnorm;    // sqrt(sum (x - mean)^2), src/correlate.jl:440-452
    float inv_cnorm_f;
};

__device__ inline float exact_u8(const U8Point& P, int32_t acc, int32_t s1, int32_t s2,
                                 int measure) {
    int64_t n = P.n;
    int64_t swt = (int64_t)acc + 128 * (P.sumT + (int64_t)s1) - 16384 * n;  // sum W*T
    int64_t nn = n * swt - P.sumT * (int64_t)s1;                              // n * numerator
    double invn = 1.0 / (double)n;                                            // exact: n = 2^k
    double num = (double)nn * invn;
    double den;
    if (measure == 1) {
        den = P.cnorm * sqrt((double)s2);
    } else {
        int64_t wint = n * (int64_t)s2 - (int64_t)s1 * (int64_t)s1;          // n * variance
        double wvar = (double)wint * invn;
        den = P.cnorm * sqrt(wvar);
    }
    return den > 0 ? (float)(num / den) : 0.0f;
}

__device__ inline float approx_u8(const U8Point& P, int32_t acc, int32_t s1, int32_t s2,
                                  int measure) {
    int64_t n = P.n;
    int64_t swt = (int64_t)acc + 128 * (P.sumT + (int64_t)s1) - 16384 * n;
    int64_t nn = n * swt - P.sumT * (int64_t)s1;
    int64_t wint = measure == 1 ? n * (int64_t)s2 : n * (int64_t)s2 - (int64_t)s1 * (int64_t)s1;
    if (wint == 0) return 0.0f;
    // (nn / n) / (cnorm * sqrt(wint / n)) = nn / (cnorm * sqrt(wint * n))
    float d = (float)wint * (float)n;
    return (float)nn * P.inv_cnorm_f * __builtin_amdgcn_rsqf(d);
}

__global__ __launch_bounds__(AR_WG) void track_u8_kernel(TrackArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const U8Layout L = u8_layout(a.max_ch, a.max_cw, a.max_rx, a.max_ry, a.up);
    float* reg = (float*)smem;   // cascade region (aliases everything below once used)
    __shared__ Best red[AR_NWAVES];
    __shared__ int64_t red64[AR_NWAVES];
    __shared__ int sc[8];
    int8_t* Wl = (int8_t*)(smem + L.off_w);
    int8_t* Tp = (int8_t*)(smem + L.off_t);
    int32_t* S1 = (int32_t*)(smem + L.off_s1);
    int32_t* S2 = (int32_t*)(smem + L.off_s2);
    int32_t* colS = (int32_t*)(smem + L.off_tmp);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint8_t* ref = (const uint8_t*)a.ref;
    const uint8_t* sec = (const uint8_t*)a.sec;

    for (int64_t p = blockIdx.x; p < a.n; p += gridDim.x) {
        Geom g = point_geom(a, p);
        if (!g.status) { write_skip(a, p, false); continue; }
        if (!any_valid(a, g)) { write_skip(a, p, false); continue; }
        const int ch = g.ch, cw = g.cw, NR = 2 * g.ry, NC = 2 * g.rx;
        const int winh = ch + NR - 1, winw = cw + NC - 1;
        const int nti = (NR + 15) / 16, ntj = (NC + 15) / 16;
        const int nrb = (ch + 15 + 15) / 16;          // band rows r - i in [0, ch), i < 16
        const int rows_used = 16 * (nti + nrb);
        const int cols_used = 16 * ntj + cw;

        // ---- stage window (int8, value ^ 0x80; virtual zero padding -> 0x80) and chip
        for (int t = tid; t < rows_used * cols_used; t += AR_WG) {
            int c = t / rows_used, r = t - c * rows_used;
            int8_t v = 0;
            if (r < winh && c < winw) {
                int64_t k = pix(a, g.wr0 + r, g.wc0 + c);
                v = (int8_t)((k >= 0 ? ref[k] : 0) ^ 0x80);
            }
            Wl[c * L.ldw + r] = v;
        }
        int64_t st = 0;
        for (int t = tid; t < cw * L.cpad; t += AR_WG) {
            int v = t / L.cpad, x = t - v * L.cpad, u = x - 16;
            int8_t b = 0;
            if (u >= 0 && u < ch) {
                int64_t k = pix(a, g.cr0 + u, g.cc0 + v);
                uint8_t raw = k >= 0 ? sec[k] : 0;
                st += raw;
                b = (int8_t)(raw ^ 0x80);
            }
            Tp[t] = b;
        }
        U8Point P;
        P.n = (int64_t)ch * cw;
        P.sumT = block_sum<int64_t>(st, red64);
        __syncthreads();
        // ---- chip energy, exact: sum (n x - sumT)^2 over n^2
        int64_t e = 0;
        for (int t = tid; t < ch * cw; t += AR_WG) {
            int v = t / ch, u = t - v * ch;
            int64_t x = (int64_t)(uint8_t)(Tp[v * L.cpad + 16 + u] ^ 0x80);
            int64_t k = P.n * x - P.sumT;
            e += k * k;
        }
        int64_t sumk2 = block_sum<int64_t>(e, red64);
        double sq = (double)sumk2 / ((double)P.n * (double)P.n);
        bool ok = sq > 2.220446049250313e-16 * (double)ch * (double)cw;
        if (!ok) { write_skip(a, p, true); __syncthreads(); continue; }
        P.cnorm = sqrt(sq);
        P.inv_cnorm_f = (float)(1.0 / P.cnorm);

        // ---- integer box sums S1 (W) and S2 (W^2) over every shift, sliding
        for (int pass = 0; pass < 2; pass++) {
            int32_t* S = pass == 0 ? S1 : S2;
            for (int c = tid; c < winw; c += AR_WG) {
                const int8_t* col = Wl + c * L.ldw;
                int32_t s = 0;
                for (int u = 0; u < ch; u++) {
                    int32_t w = (uint8_t)(col[u] ^ 0x80);
                    s += pass == 0 ? w : w * w;
                }
                for (int r0 = 0; r0 < NR; r0++) {
                    colS[c * NR + r0] = s;
                    int32_t wo = (uint8_t)(col[r0] ^ 0x80), wi = (uint8_t)(col[r0 + ch] ^ 0x80);
                    s += pass == 0 ? wi - wo : wi * wi - wo * wo;
                }
            }
            __syncthreads();
            for (int i = tid; i < NR; i += AR_WG) {
                int32_t s = 0;
                for (int v = 0; v < cw; v++) s += colS[v * NR + i];
                for (int j = 0; j < NC; j++) {
                    S[j * NR + i] = s;
                    if (j + 1 < NC) s += colS[(j + cw) * NR + i] - colS[j * NR + i];
                }
            }
            __syncthreads();
        }

        // ---- numerator on int8 MFMA; epilogue scores every shift in FP32
        int32_t* accL = colS;                          // int32 [NR*NC] (colS is dead now)
        float* apx = (float*)(colS + NR * NC);         // float [NR*NC]
        const int ntiles = nti * ntj;
        const int nq = cw * nrb;                       // 16-wide k-blocks
        const int nm = (nq + 3) / 4;                   // MFMAs per tile
        const int g16 = lane >> 4, l16 = lane & 15;
        for (int tb = wave; tb < ntiles; tb += 4 * 8) {
            v4i acc[8];
            int nt_here = 0;
            for (int k = 0; k < 8; k++) {
                acc[k] = (v4i){0, 0, 0, 0};
                if (tb + 4 * k < ntiles) nt_here = k + 1;
            }
            for (int m = 0; m < nm; m++) {
                int q = 4 * m + g16;
                v4i bf = (v4i){0, 0, 0, 0};
                int v = 0, rb = 0;
                if (q < nq) {
                    v = q / nrb;
                    rb = q - v * nrb;
                    int off = v * L.cpad + 16 + 16 * rb - l16;
                    int base = off & ~7, sh = (off & 7) * 8;
                    const uint64_t* qp = (const uint64_t*)(Tp + base);
                    uint64_t q0 = qp[0], q1 = qp[1], q2 = qp[2];
                    uint64_t lo = sh ? (q0 >> sh) | (q1 << (64 - sh)) : q0;
                    uint64_t hi = sh ? (q1 >> sh) | (q2 << (64 - sh)) : q1;
                    bf = (v4i){(int)(uint32_t)lo, (int)(uint32_t)(lo >> 32), (int)(uint32_t)hi,
                               (int)(uint32_t)(hi >> 32)};
                }
                for (int k = 0; k < 8; k++) {
                    if (k >= nt_here) break;
                    int t = tb + 4 * k;
                    int ti = t % nti, tj = t / nti;
                    v4i af = (v4i){0, 0, 0, 0};
                    if (q < nq) {
                        const int8_t* ap = Wl + (16 * tj + l16 + v) * L.ldw + 16 * ti + 16 * rb;
                        af = *(const v4i*)ap;
                    }
                    acc[k] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af, bf, acc[k], 0, 0, 0);
                }
            }
            for (int k = 0; k < nt_here; k++) {
                int t = tb + 4 * k;
                int ti = t % nti, tj = t / nti;
                int i = 16 * ti + l16;
                for (int r = 0; r < 4; r++) {
                    int j = 16 * tj + 4 * g16 + r;
                    if (i < NR && j < NC) {
                        int o = j * NR + i;
                        int32_t cv = acc[k][r];
                        accL[o] = cv;
                        apx[o] = approx_u8(P, cv, S1[o], S2[o], a.measure);
                    }
                }
            }
        }
        __syncthreads();

        // ---- integer peak: screen, then exact values for the contenders
        Best b = best_empty();
        for (int o = tid; o < NR * NC; o += AR_WG) {
            int j = o / NR, i = o - j * NR;
            float v = apx[o];
            int idx = i * NC + j;
            if (better(v, idx, b)) { b.v = v; b.idx = idx; }
        }
        b = block_best(b, red);
        float thr = b.v - 2.0f * AR_SCREEN_EPS;
        Best e2 = best_empty();
        for (int o = tid; o < NR * NC; o += AR_WG) {
            if (!(apx[o] >= thr)) continue;
            int j = o / NR, i = o - j * NR;
            float v = exact_u8(P, accL[o], S1[o], S2[o], a.measure);
            int idx = i * NC + j;
            if (better(v, idx, e2)) { e2.v = v; e2.idx = idx; }
        }
        e2 = block_best(e2, red);
        int pi = e2.idx / NC + 1, pj = e2.idx % NC + 1;
        float pval = e2.v;
        int i0 = 1, j0 = 1;
        bool refine = a.up > 1 && patch_origin(pi, pj, NR, NC, i0, j0);
        float pv = 0.0f;
        if (refine && tid < 25) {
            int ii = i0 - 1 + tid / 5, jj = j0 - 1 + tid % 5;
            int o = jj * NR + ii;
            pv = exact_u8(P, accL[o], S1[o], S2[o], a.measure);
        }
        __syncthreads();   // everything above is dead from here: the cascade may alias it
        if (refine && tid < 25) reg[tid] = pv;
        __syncthreads();
        finish_point(a, p, g, pi, pj, pval, i0, j0, refine, reg, red);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------------
// generic path (FP64, reference order)
// ---------------------------------------------------------------------------------------------

__device__ inline float load_pix_f(const TrackArgs& a, const void* img, int is_u8, int64_t r,
                                   int64_t c) {
    int64_t k = pix(a, r, c);
    if (k < 0) return 0.0f;
    return is_u8 ? (float)((const uint8_t*)img)[k] : ((const float*)img)[k];
}

template <int IS_U8>
__global__ __launch_bounds__(AR_WG) void track_generic_kernel(TrackArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* reg = (float*)smem;
    __shared__ Best red[AR_NWAVES];
    __shared__ double shd[4];
    const int tid = threadIdx.x;
    const GenScratch G = gen_scratch(a.max_ch, a.max_cw, a.max_rx, a.max_ry);
    unsigned char* base = a.scratch + (size_t)blockIdx.x * a.scratch_per_block;
    float* chip = (float*)(base + G.off_chip);
    float* win = (float*)(base + G.off_win);
    double* S = (double*)(base + G.off_s);
    double* S2 = (double*)(base + G.off_s2);
    float* surf = (float*)(base + G.off_surf);

    for (int64_t p = blockIdx.x; p < a.n; p += gridDim.x) {
        Geom g = point_geom(a, p);
        if (!g.status) { write_skip(a, p, false); continue; }
        if (!any_valid(a, g)) { write_skip(a, p, false); continue; }
        const int ch = g.ch, cw = g.cw, NR = 2 * g.ry, NC = 2 * g.rx;
        const int winh = ch + NR - 1, winw = cw + NC - 1, ls = winh + 1;
        for (int t = tid; t < winh * winw; t += AR_WG) {
            int c = t / winh, r = t - c * winh;
            win[t] = load_pix_f(a, a.ref, IS_U8, g.wr0 + r, g.wc0 + c);
        }
        for (int t = tid; t < ch * cw; t += AR_WG) {
            int v = t / ch, u = t - v * ch;
            chip[t] = load_pix_f(a, a.sec, IS_U8, g.cr0 + u, g.cc0 + v);
        }
        __syncthreads();
        // prepare_chip!: sequential Float64 mean and energy in column-major order
        if (tid == 0) {
            double s = 0.0;
            for (int k = 0; k < ch * cw; k++) s += (double)chip[k];
            double mean = s / (double)(ch * cw);
            double sq = 0.0;
            for (int k = 0; k < ch * cw; k++) {
                double d = (double)chip[k] - mean;
                sq += d * d;
            }
            shd[0] = mean;
            shd[1] = sq;
        }
        __syncthreads();
        double mean = shd[0], sq = shd[1];
        bool ok = sq > 2.220446049250313e-16 * (double)ch * (double)cw;
        __syncthreads();
        if (!ok) { write_skip(a, p, true); continue; }
        double cnorm = sqrt(sq);
        for (int k = tid; k < ch * cw; k += AR_WG) chip[k] = (float)((double)chip[k] - mean);
        // integral_both!: column running sums, then the row recurrence, in the reference order
        for (int j = tid; j <= winw; j += AR_WG) { S[j * ls] = 0.0; S2[j * ls] = 0.0; }
        for (int i = tid; i <= winh; i += AR_WG) { S[i] = 0.0; S2[i] = 0.0; }
        for (int j = tid; j < winw; j += AR_WG) {
            double run = 0.0, runsq = 0.0;
            for (int i = 0; i < winh; i++) {
                double v = (double)win[i + j * winh];
                run += v;
                runsq += v * v;
                S[(i + 1) + (j + 1) * ls] = run;
                S2[(i + 1) + (j + 1) * ls] = runsq;
            }
        }
        __syncthreads();
        for (int i = tid + 1; i <= winh; i += AR_WG) {
            for (int j = 1; j <= winw; j++) {
                S[i + j * ls] = S[i + j * ls] + S[i + (j - 1) * ls];
                S2[i + j * ls] = S2[i + j * ls] + S2[i + (j - 1) * ls];
            }
        }
        __syncthreads();
        // direct numerator (src/correlate.jl:799-814) and the normalisation, per shift
        double n = (double)(ch * cw);
        for (int o = tid; o < NR * NC; o += AR_WG) {
            int j = o / NR, i = o - j * NR;
            double acc = 0.0;
            for (int jj = 0; jj < cw; jj++) {
                const float* wc = win + (j + jj) * winh + i;
                const float* cc = chip + jj * ch;
                for (int ii = 0; ii < ch; ii++) acc += (double)cc[ii] * (double)wc[ii];
            }
            int I = i + 1, J = j + 1;
#define BOX(T) (T[(I + ch - 1) + (J + cw - 1) * ls] - T[(I - 1) + (J + cw - 1) * ls] - \
                T[(I + ch - 1) + (J - 1) * ls] + T[(I - 1) + (J - 1) * ls])
            double den;
            if (a.measure == 1) {
                double wsq = BOX(S2);
                den = cnorm * sqrt(wsq < 0.0 ? 0.0 : wsq);
            } else {
                double s1 = BOX(S), s2 = BOX(S2);
                double wv = s2 - s1 * s1 / n;
                den = cnorm * sqrt(wv < 0.0 ? 0.0 : wv);
            }
#undef BOX
            surf[o] = den > 0 ? (float)(acc / den) : 0.0f;
        }
        __syncthreads();
        Best b = best_empty();
        for (int o = tid; o < NR * NC; o += AR_WG) {
            int j = o / NR, i = o - j * NR;
            float v = surf[o];
            int idx = i * NC + j;
            if (better(v, idx, b)) { b.v = v; b.idx = idx; }
        }
        b = block_best(b, red);
        if (b.idx == 0x7fffffff) b.idx = 0;   // all NaN: (1, 1)
        int pi = b.idx / NC + 1, pj = b.idx % NC + 1;
        float pval = surf[(pj - 1) * NR + (pi - 1)];
        int i0 = 1, j0 = 1;
        bool refine = a.up > 1 && patch_origin(pi, pj, NR, NC, i0, j0);
        if (refine && tid < 25) reg[tid] = surf[(j0 - 1 + tid % 5) * NR + (i0 - 1 + tid / 5)];
        __syncthreads();
        finish_point(a, p, g, pi, pj, pval, i0, j0, refine, reg, red);
        __syncthreads();
    }
}

template __global__ void track_generic_kernel<0>(TrackArgs);
template __global__ void track_generic_kernel<1>(TrackArgs);

}  // namespace ar
