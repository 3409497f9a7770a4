// Integer peak selection and the 64x pyrUp subpixel cascade, run by one 256-thread workgroup per
// point. Follows src/peak.jl:32-106 (argmax rule, offsets) and src/peak.jl:196-522 (pyrUp steps
// and subpixel_peak) element for element: every upsampled sample is computed by the reference's
// formula in FP32 from the same inputs, so the result is bit-identical whatever the order in which
// samples are produced.
//
// Cost control without approximation: levels up to 80x80 are materialised in LDS; the last one or
// two levels are produced in 16-row strips, strongest first, and a strip is skipped only when a
// proven upper bound on every sample it could hold is strictly below the best sample already
// found. Each pyrUp pass forms convex combinations (weights 3/4, 1/8, 1/8 or 1/2, 1/2), so a
// pass raises the maximum by at most two units in the last place; the bound adds a margin of
// 1e-6 relative, far above that.
#pragma once
#include "track_common.h"

namespace ar {

// Floats needed by the cascade region for upsampling `up` (power of two, 1..64).
__host__ __device__ inline int cascade_floats(int up) {
    // header: patch(32) rowmax(96) ub(64) order/scalars(64)
    int hdr = 256;
    if (up <= 1) return hdr;
    // materialised levels: Y (final materialised level), X (previous), T (vertical tmp)
    int L = 0;
    while ((1 << L) < up) L++;
    int LM = L < 4 ? L : 4;
    int nM = 5 << LM;
    int nP = 5 << (LM > 0 ? LM - 1 : 0);
    int Y = nM * nM;
    int X = nP * nP;
    int T = nM * nP;
    int stream = 0;
    if (L == 5) stream = 16 * 80;                         // tmp5 strip
    if (L == 6) stream = 12 * 160 + 12 * 80 + 16 * 160;  // level-5 rows, tmp5 rows, tmp6 strip
    int body = Y + (X + T > stream ? X + T : stream);
    return hdr + body;
}

struct CascadeOut {
    int si, sj;      // 0-based sample on the final level
    float val;
};

// Pair form of the taps: source sample k of an axis of n samples feeds outputs 2k and 2k+1,
//   out[2k]   = tap3(src[km], src[k], src[kq]),   out[2k+1] = tap2(src[k], src[kq]),
// with km = k == 0 ? 1 : k - 1 and kq = min(k + 1, n - 1) -- exactly taps(2k) and taps(2k+1).
// Three loads give two outputs.
__device__ inline void pair_src(int k, int n, int& km, int& kq) {
    km = k == 0 ? 1 : k - 1;
    kq = k + 1 < n - 1 ? k + 1 : n - 1;
}

// The passes below give each thread U tasks at a time (t, t + 256, ...), issuing all loads
// before any arithmetic so that several LDS reads are in flight per thread.
#define AR_U 4

// Vertical pass: rows 2k, 2k+1 for k in [k0, k1) of an axis of n source rows, W columns.
// src holds source rows from s_off (pitch W); dst holds output rows from d_off (pitch W).
template <int W>
__device__ inline void vert_pairs(const float* src, int n, int s_off, float* dst, int d_off,
                                  int k0, int k1) {
    const int total = (k1 - k0) * W;
    for (int t0 = threadIdx.x; t0 < total; t0 += AR_U * AR_WG) {
        float a[AR_U], b[AR_U], q[AR_U];
        int o[AR_U];
#pragma unroll
        for (int u = 0; u < AR_U; u++) {
            int t = t0 + u * AR_WG;
            t = t < total ? t : total - 1;
            int kk = t / W, c = t - kk * W, k = k0 + kk;
            int km, kq;
            pair_src(k, n, km, kq);
            a[u] = src[(km - s_off) * W + c];
            b[u] = src[(k - s_off) * W + c];
            q[u] = src[(kq - s_off) * W + c];
            o[u] = (2 * k - d_off) * W + c;
        }
#pragma unroll
        for (int u = 0; u < AR_U; u++) {
            if (t0 + u * AR_WG < total) {
                dst[o[u]] = tap3(a[u], b[u], q[u]);
                dst[o[u] + W] = tap2(b[u], q[u]);
            }
        }
    }
}

// Horizontal pass: `rows` rows of WS source columns -> 2 WS output columns.
template <int WS>
__device__ inline void horiz_pairs(const float* src, float* dst, int rows) {
    const int total = rows * WS;
    for (int t0 = threadIdx.x; t0 < total; t0 += AR_U * AR_WG) {
        float a[AR_U], b[AR_U], q[AR_U];
        int o[AR_U];
#pragma unroll
        for (int u = 0; u < AR_U; u++) {
            int t = t0 + u * AR_WG;
            t = t < total ? t : total - 1;
            int r = t / WS, k = t - r * WS;
            int km, kq;
            pair_src(k, WS, km, kq);
            const float* row = src + r * WS;
            a[u] = row[km];
            b[u] = row[k];
            q[u] = row[kq];
            o[u] = r * 2 * WS + 2 * k;
        }
#pragma unroll
        for (int u = 0; u < AR_U; u++) {
            if (t0 + u * AR_WG < total) {
                dst[o[u]] = tap3(a[u], b[u], q[u]);
                dst[o[u] + 1] = tap2(b[u], q[u]);
            }
        }
    }
}

// Horizontal pass of the final level straight into the argmax: row r of the strip is output row
// row0 + r; sample (row, col) has row-major index row * 2WS + col.
template <int WS>
__device__ inline Best horiz_pairs_argmax(const float* src, int rows, int row0) {
    Best best = best_empty();
    const int total = rows * WS;
    for (int t0 = threadIdx.x; t0 < total; t0 += AR_U * AR_WG) {
        float a[AR_U], b[AR_U], q[AR_U];
        int idx[AR_U];
#pragma unroll
        for (int u = 0; u < AR_U; u++) {
            int t = t0 + u * AR_WG;
            t = t < total ? t : total - 1;
            int r = t / WS, k = t - r * WS;
            int km, kq;
            pair_src(k, WS, km, kq);
            const float* row = src + r * WS;
            a[u] = row[km];
            b[u] = row[k];
            q[u] = row[kq];
            idx[u] = (row0 + r) * 2 * WS + 2 * k;
        }
#pragma unroll
        for (int u = 0; u < AR_U; u++) {
            if (t0 + u * AR_WG < total) {
                float e = tap3(a[u], b[u], q[u]), o = tap2(b[u], q[u]);
                if (better(e, idx[u], best)) { best.v = e; best.idx = idx[u]; }
                if (better(o, idx[u] + 1, best)) { best.v = o; best.idx = idx[u] + 1; }
            }
        }
    }
    return best;
}

// One full pyrUp step in LDS: src (N x N, row-major) -> dst (2N x 2N), via tmp (2N x N).
template <int N>
__device__ inline void pyrup_full(const float* src, float* tmp, float* dst) {
    vert_pairs<N>(src, N, 0, tmp, 0, 0, N);
    __syncthreads();
    horiz_pairs<N>(tmp, dst, 2 * N);
    __syncthreads();
}

__device__ inline float ub_margin(float m) {
    if (!(m > -INFINITY)) return m;   // no finite feeder: nothing can beat anything
    return m + fabsf(m) * 1e-6f + 1e-30f;
}

// Source-row range [lo, hi] (0-based) feeding output rows [a, b) of a pass over n source rows.
__device__ inline void feed_rows(int a, int b, int n, int& lo, int& hi) {
    lo = 0x7fffffff;
    hi = -1;
    for (int i = a; i < b; i++) {
        int m, c, q, nt;
        taps(i, n, m, c, q, nt);
        int l = m < c ? m : c;
        l = l < q ? l : q;
        int h = m > c ? m : c;
        h = h > q ? h : q;
        lo = l < lo ? l : lo;
        hi = h > hi ? h : hi;
    }
}

// The cascade. `patch` (5x5, row-major) is in LDS; `reg` is the cascade region
// (cascade_floats(up) floats, 16-byte aligned). Returns the final-level argmax under the
// peak_index rule; identical for every thread.
__device__ inline CascadeOut cascade(float* reg, int up, Best* red) {
    float* patch = reg;          // 32
    float* rowmax = reg + 32;    // 96
    float* ubs = reg + 128;      // 64
    int* sc = (int*)(reg + 192); // 64 ints of scalars
    float* body = reg + 256;

    int L = 0;
    while ((1 << L) < up) L++;
    int LM = L < 4 ? L : 4;
    int nM = 5 << LM;
    int nP = 5 << (LM > 0 ? LM - 1 : 0);
    float* Y = body;                       // nM x nM
    float* X = Y + nM * nM;                // nP x nP; after level LM, the streaming buffers
    float* T = X + nP * nP;                // vertical-pass tmp (2n x n) while materialising

    // Levels 1..LM, ping-ponging so that level LM lands in Y.
    const float* cur = patch;
    if (LM >= 1) { float* d = ((LM - 1) & 1) ? X : Y; pyrup_full<5>(cur, T, d); cur = d; }
    if (LM >= 2) { float* d = ((LM - 2) & 1) ? X : Y; pyrup_full<10>(cur, T, d); cur = d; }
    if (LM >= 3) { float* d = ((LM - 3) & 1) ? X : Y; pyrup_full<20>(cur, T, d); cur = d; }
    if (LM >= 4) { float* d = ((LM - 4) & 1) ? X : Y; pyrup_full<40>(cur, T, d); cur = d; }
    int n = 5 << LM;

    CascadeOut out;
    if (L <= 4) {
        // final level materialised: plain argmax
        Best b = best_empty();
        int nf = n;
        for (int t = threadIdx.x; t < nf * nf; t += AR_WG) {
            float v = cur[t];
            if (better(v, t, b)) { b.v = v; b.idx = t; }
        }
        b = block_best(b, red);
        if (b.idx == 0x7fffffff) b.idx = 0;
        out.si = b.idx / nf;
        out.sj = b.idx - out.si * nf;
        out.val = cur[b.idx];
        __syncthreads();
        return out;
    }

    // Row maxima of level 4 (80 x 80, in Y), ignoring NaN.
    const int n4 = 80;
    for (int r = threadIdx.x; r < n4; r += AR_WG) {
        float m = -INFINITY;
        for (int c = 0; c < n4; c++) {
            float v = Y[r * n4 + c];
            m = v > m ? v : m;
        }
        rowmax[r] = m;
    }
    __syncthreads();

    const int nfin = 5 * up;          // 160 or 320
    const int NS = nfin / 16;         // strips
    // Upper bound per strip from the level-4 rows that feed it.
    for (int s = threadIdx.x; s < NS; s += AR_WG) {
        int lo, hi;
        if (L == 5) {
            feed_rows(16 * s, 16 * s + 16, n4, lo, hi);
        } else {
            int lo5, hi5;
            feed_rows(16 * s, 16 * s + 16, 160, lo5, hi5);
            feed_rows(lo5, hi5 + 1, n4, lo, hi);
        }
        float m = -INFINITY;
        for (int r = lo; r <= hi; r++) m = rowmax[r] > m ? rowmax[r] : m;
        ubs[s] = ub_margin(m);
    }
    __syncthreads();

    Best best = best_empty();
    float* lv5 = X;                  // level-5 rows (L == 6)
    float* t5 = X + 12 * 160;        // tmp5 rows
    float* t6 = t5 + 12 * 80;        // tmp6 strip
    if (L == 5) t5 = X;

    for (int iter = 0; iter < NS; iter++) {
        // pick the unprocessed strip with the largest bound (lowest index on ties)
        if (threadIdx.x == 0) {
            int bs = -1;
            float bu = 0;
            for (int s = 0; s < NS; s++) {
                float u = ubs[s];
                if (u != u) continue;   // processed (NaN marker)
                if (bs < 0 || u > bu) { bs = s; bu = u; }
            }
            int go = bs >= 0 && !(bu < best.v);
            sc[0] = go ? bs : -1;
            if (go) ubs[bs] = __int_as_float(0x7fc00000);
        }
        __syncthreads();
        int s = sc[0];
        if (s < 0) break;
        int a = 16 * s;
        Best b = best_empty();
        if (L == 5) {
            // level-5 rows [a, a+16) from level 4, then their samples into the argmax
            vert_pairs<80>(Y, n4, 0, t5, a, a / 2, a / 2 + 8);
            __syncthreads();
            b = horiz_pairs_argmax<80>(t5, 16, a);
        } else {
            // level-6 rows [a, a+16) need level-5 rows [a/2 - 1, a/2 + 8] (clamped), produced
            // as whole pairs from level 4
            int lo5 = a / 2 - 1 < 0 ? 0 : a / 2 - 1, hi5 = a / 2 + 8 > 159 ? 159 : a / 2 + 8;
            int plo = lo5 / 2, phi = hi5 / 2, base5 = 2 * plo, nr5 = 2 * (phi - plo + 1);
            vert_pairs<80>(Y, n4, 0, t5, base5, plo, phi + 1);
            __syncthreads();
            horiz_pairs<80>(t5, lv5, nr5);
            __syncthreads();
            vert_pairs<160>(lv5, 160, base5, t6, a, a / 2, a / 2 + 8);
            __syncthreads();
            b = horiz_pairs_argmax<160>(t6, 16, a);
        }
        b = block_best(b, red);
        best = best_merge(best, b);
        __syncthreads();
    }
    // All strips NaN (or none processed): the reference's scan leaves (1, 1). Its value is
    // recomputed from the first strip's first sample, which is NaN in that case.
    if (best.idx == 0x7fffffff) {
        best.idx = 0;
        best.v = __int_as_float(0x7fc00000);
    }
    out.si = best.idx / nfin;
    out.sj = best.idx - out.si * nfin;
    out.val = best.v;
    return out;
}

}  // namespace ar
