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

// One full pyrUp step in LDS: src (n x n, row-major) -> dst (2n x 2n), via tmp (2n x n).
__device__ inline void pyrup_full(const float* src, float* tmp, float* dst, int n) {
    int th = 2 * n;
    for (int t = threadIdx.x; t < th * n; t += AR_WG) {
        int i = t / n, c = t - i * n;
        int m, c0, q, nt;
        taps(i, n, m, c0, q, nt);
        tmp[t] = nt == 3 ? tap3(src[m * n + c], src[c0 * n + c], src[q * n + c])
                         : tap2(src[c0 * n + c], src[q * n + c]);
    }
    __syncthreads();
    int tw = 2 * n;
    for (int t = threadIdx.x; t < th * tw; t += AR_WG) {
        int i = t / tw, j = t - i * tw;
        int m, c0, q, nt;
        taps(j, n, m, c0, q, nt);
        const float* row = tmp + i * n;
        dst[t] = nt == 3 ? tap3(row[m], row[c0], row[q]) : tap2(row[c0], row[q]);
    }
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
    int n = 5;
    for (int k = 1; k <= LM; k++) {
        float* dst = ((LM - k) & 1) ? X : Y;
        pyrup_full(cur, T, dst, n);
        cur = dst;
        n *= 2;
    }

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
            // tmp5 strip: vertical pass of level 4 for level-5 rows [a, a+16)
            for (int t = threadIdx.x; t < 16 * n4; t += AR_WG) {
                int ii = t / n4, c = t - ii * n4;
                int m, c0, q, nt;
                taps(a + ii, n4, m, c0, q, nt);
                t5[t] = nt == 3 ? tap3(Y[m * n4 + c], Y[c0 * n4 + c], Y[q * n4 + c])
                                : tap2(Y[c0 * n4 + c], Y[q * n4 + c]);
            }
            __syncthreads();
            for (int t = threadIdx.x; t < 16 * 160; t += AR_WG) {
                int ii = t / 160, j = t - ii * 160;
                int m, c0, q, nt;
                taps(j, n4, m, c0, q, nt);
                const float* row = t5 + ii * n4;
                float v = nt == 3 ? tap3(row[m], row[c0], row[q]) : tap2(row[c0], row[q]);
                int idx = (a + ii) * 160 + j;
                if (better(v, idx, b)) { b.v = v; b.idx = idx; }
            }
        } else {
            int lo5, hi5;
            feed_rows(a, a + 16, 160, lo5, hi5);
            int nr5 = hi5 - lo5 + 1;
            // tmp5 rows lo5..hi5 from level 4
            for (int t = threadIdx.x; t < nr5 * n4; t += AR_WG) {
                int ii = t / n4, c = t - ii * n4;
                int m, c0, q, nt;
                taps(lo5 + ii, n4, m, c0, q, nt);
                t5[t] = nt == 3 ? tap3(Y[m * n4 + c], Y[c0 * n4 + c], Y[q * n4 + c])
                                : tap2(Y[c0 * n4 + c], Y[q * n4 + c]);
            }
            __syncthreads();
            // level-5 rows lo5..hi5 (160 wide)
            for (int t = threadIdx.x; t < nr5 * 160; t += AR_WG) {
                int ii = t / 160, j = t - ii * 160;
                int m, c0, q, nt;
                taps(j, n4, m, c0, q, nt);
                const float* row = t5 + ii * n4;
                lv5[t] = nt == 3 ? tap3(row[m], row[c0], row[q]) : tap2(row[c0], row[q]);
            }
            __syncthreads();
            // tmp6 rows a..a+15 (160 wide) from level 5
            for (int t = threadIdx.x; t < 16 * 160; t += AR_WG) {
                int ii = t / 160, c = t - ii * 160;
                int m, c0, q, nt;
                taps(a + ii, 160, m, c0, q, nt);
                m -= lo5; c0 -= lo5; q -= lo5;
                t6[t] = nt == 3 ? tap3(lv5[m * 160 + c], lv5[c0 * 160 + c], lv5[q * 160 + c])
                                : tap2(lv5[c0 * 160 + c], lv5[q * 160 + c]);
            }
            __syncthreads();
            // final level-6 samples of the strip, straight into the argmax
            for (int t = threadIdx.x; t < 16 * 320; t += AR_WG) {
                int ii = t / 320, j = t - ii * 320;
                int m, c0, q, nt;
                taps(j, 160, m, c0, q, nt);
                const float* row = t6 + ii * 160;
                float v = nt == 3 ? tap3(row[m], row[c0], row[q]) : tap2(row[c0], row[q]);
                int idx = (a + ii) * 320 + j;
                if (better(v, idx, b)) { b.v = v; b.idx = idx; }
            }
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
