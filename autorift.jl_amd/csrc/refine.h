// The pyrUp subpixel cascade of src/peak.jl:196-522 (pyrUp steps and subpixel_peak), run by one
// wave per point with no workgroup barriers. Every upsampled sample is computed by the
// reference's formula in FP32 from the same inputs, so the result is bit-identical whatever the
// order in which samples are produced, and whichever of them are produced more than once; the
// argmax follows peak_index's rule (src/peak.jl:32-61).
//
// Cost control without approximation. Levels up to 40x40 (8x) are materialised in LDS. Beyond,
// the final level is cut into 16x16 tiles, each computed from the 40x40 level through its own
// small blocks of the intermediate levels. A tile is skipped only when a proven upper bound on
// every sample it could hold is strictly below a sample already found: first the tile holding
// the image of the 8x maximum is computed (its best sample is the threshold), then every other
// tile fed by an 8x sample whose bound reaches the threshold, each re-checked against the bound
// of its own intermediate blocks before going deeper.
//
// The bound. A pass forms tap3 = 0.75 b + 0.125 (a + c) or tap2 = 0.5 (b + c) in FP32. With
// M = max(a, b, c): fl(a + c) <= 2M and fl(b + c) <= 2M (rounding is monotone and 2M is
// representable), the power-of-two scalings are exact, fl(0.75 b) <= 0.75 M + 0.75 |M| u, so a
// pass exceeds M by at most 1.75 |M| u (u = 2^-24) plus subnormal rounding (< 2^-149 per
// operation). Six passes (8x -> 64x) stay below 6.3e-7 |M|; the bound adds 1e-6 |M| + 1e-30.
#pragma once
#include "track_common.h"

namespace ar {

#define AR_TILE 16        // final-level tile edge
#define AR_TILE_WS 416    // tile scratch: vertical tmp 16 x 16 + block 10 x 16
#define AR_CASC_HDR 48    // patch (5x5) and padding; the levels start 16-byte aligned

__host__ __device__ inline int casc_levels(int up) {
    int L = 0;
    while ((1 << L) < up) L++;
    return L;
}

// Floats of one wave's cascade region for upsampling `up` (power of two, 2..64): the patch, the
// last materialised level Y and the vertical-pass tmp T. The level before Y lives in Y's own
// bytes (it is dead before the last horizontal pass writes Y), and the tile scratch reuses T once
// Y is complete.
__host__ __device__ inline int cascade_floats(int up) {
    int L = casc_levels(up);
    int LM = L < 3 ? L : 3;
    int nM = 5 << LM;
    int nP = 5 << (LM > 0 ? LM - 1 : 0);
    int T = nM * nP;
    if (L > 3 && T < AR_TILE_WS) T = AR_TILE_WS;
    return AR_CASC_HDR + nM * nM + T;
}

struct CascadeOut {
    int si, sj;      // 0-based sample on the final level
    float val;
};

// Intra-wave LDS hand-off: the LDS operations of one wave complete in order, so ordering the
// compiler's accesses is enough.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Largest non-NaN value of the wave (-inf if none). A pass's output is NaN whenever one of its
// taps is, so its non-NaN outputs are bounded by the largest non-NaN input.
__device__ inline float wave_max(float v) { return wave_max_f32(v); }

// The wave's argmax under the peak_index rule: the largest value, then the smallest index; NaN
// never wins; (-inf, INT32_MAX) when no lane holds a value.
__device__ inline Best wave_best_dpp(Best b) {
    Best r;
    r.v = wave_max_f32(b.v);
    r.idx = wave_min_i32(b.v == r.v ? b.idx : 0x7fffffff);
    if (!(r.v > -INFINITY)) {
        r.v = -INFINITY;
        r.idx = 0x7fffffff;
    }
    return r;
}

__device__ inline float ub_margin(float m) {
    if (!(m > -INFINITY)) return m;   // no finite feeder: nothing can beat anything
    return m + fabsf(m) * 1e-6f + 1e-30f;
}

// Pair form of the taps: source sample k of an axis of n samples feeds outputs 2k and 2k+1,
//   out[2k]   = tap3(src[km], src[k], src[kq]),   out[2k+1] = tap2(src[k], src[kq]),
// with km = k == 0 ? 1 : k - 1 and kq = min(k + 1, n - 1) -- exactly taps(2k) and taps(2k+1).
// Three loads give two outputs.
__device__ inline void pair_src(int k, int n, int& km, int& kq) {
    km = k == 0 ? 1 : k - 1;
    kq = k + 1 < n - 1 ? k + 1 : n - 1;
}

// The passes below give each lane up to AR_U tasks at a time (t, t + 64, ...), issuing all
// loads before any arithmetic.
#define AR_U 1

// One full pyrUp step by one wave: src (N x N, row-major) -> dst (2N x 2N), via tmp (2N x N),
// vertical pass first; src may share dst's bytes (only the vertical pass reads it). Returns the
// lane's argmax over the samples it wrote (row-major index).
template <int N>
__device__ inline Best wave_pyrup(const float* src, float* tmp, float* dst) {
    const int lane = threadIdx.x & 63;
    constexpr int TV = N * N, TH = 2 * N * N;
    for (int t0 = lane; t0 < TV; t0 += 64 * AR_U) {
        float a[AR_U], b[AR_U], q[AR_U];
        int o[AR_U];
#pragma unroll
        for (int u = 0; u < AR_U; u++) {
            int t = t0 + 64 * u;
            t = t < TV ? t : TV - 1;
            int k = t / N, c = t - k * N;
            int km, kq;
            pair_src(k, N, km, kq);
            a[u] = src[km * N + c];
            b[u] = src[k * N + c];
            q[u] = src[kq * N + c];
            o[u] = 2 * k * N + c;
        }
#pragma unroll
        for (int u = 0; u < AR_U; u++) {
            if (t0 + 64 * u < TV) {
                tmp[o[u]] = tap3(a[u], b[u], q[u]);
                tmp[o[u] + N] = tap2(b[u], q[u]);
            }
        }
    }
    wave_sync();
    Best best = best_empty();
    for (int t0 = lane; t0 < TH; t0 += 64 * AR_U) {
        float a[AR_U], b[AR_U], q[AR_U];
        int o[AR_U];
#pragma unroll
        for (int u = 0; u < AR_U; u++) {
            int t = t0 + 64 * u;
            t = t < TH ? t : TH - 1;
            int r = t / N, k = t - r * N;
            int km, kq;
            pair_src(k, N, km, kq);
            const float* row = tmp + r * N;
            a[u] = row[km];
            b[u] = row[k];
            q[u] = row[kq];
            o[u] = r * 2 * N + 2 * k;
        }
#pragma unroll
        for (int u = 0; u < AR_U; u++) {
            if (t0 + 64 * u < TH) {
                float e = tap3(a[u], b[u], q[u]), d = tap2(b[u], q[u]);
                dst[o[u]] = e;
                dst[o[u] + 1] = d;
                if (better(e, o[u], best)) { best.v = e; best.idx = o[u]; }
                if (better(d, o[u] + 1, best)) { best.v = d; best.idx = o[u] + 1; }
            }
        }
    }
    wave_sync();
    return best;
}

// Source range [lo, hi] (0-based, inclusive) feeding outputs [a, b] of a pyrUp axis over n >= 2
// source samples: the smallest range holding every tap of every output in [a, b] (taps above:
// output i reads i/2 - 1, i/2, i/2 + 1 for even i, i/2 and i/2 + 1 for odd i, reflected at 0
// and clamped at n - 1).
__host__ __device__ inline void feed_range(int a, int b, int n, int& lo, int& hi) {
    int l = ((a + 1) >> 1) - 1;
    lo = l > 0 ? l : 0;
    int h = (b >> 1) + 1;
    hi = h < n - 1 ? h : n - 1;
}

// Rows (and columns) of the level-L block of a final-level tile: 16 on the final level F, then
// 10 and 7 -- each the widest source range feed_range gives for the level above.
__host__ __device__ constexpr int tile_rows(int L, int F) {
    return L == F ? AR_TILE : (L == F - 1 ? 10 : 7);
}

// Widest source range of w consecutive outputs of a pyrUp axis (feed_range): (w + 1) / 2 + 2.
__host__ __device__ constexpr int feed_width(int w) { return (w + 1) / 2 + 2; }

// The wave computes the final-level tile (ti, tj) of level F from the materialised 8x level Z
// (40 x 40, row-major), through its blocks of the levels in between (pitch 16 in `ws`). With
// `prune`, it stops after an intermediate block whose bound is below `lb`. Returns the tile's
// argmax under the peak_index rule (wave-uniform; empty if pruned or all NaN). Each pass issues
// all of a lane's loads before any arithmetic.
template <int F>
__device__ inline Best tile_run(const float* Z, float* ws, int ti, int tj, float lb, bool prune) {
    const int lane = threadIdx.x & 63;
    int rlo[F + 1], rhi[F + 1], clo[F + 1], chi[F + 1];
    rlo[F] = AR_TILE * ti;
    rhi[F] = rlo[F] + AR_TILE - 1;
    clo[F] = AR_TILE * tj;
    chi[F] = clo[F] + AR_TILE - 1;
#pragma unroll
    for (int L = F; L > 3; L--) {
        feed_range(rlo[L], rhi[L], 5 << (L - 1), rlo[L - 1], rhi[L - 1]);
        feed_range(clo[L], chi[L], 5 << (L - 1), clo[L - 1], chi[L - 1]);
    }
    float* tmp = ws;         // vertical pass of the level being built: rows x 16
    float* blk = ws + 256;   // last intermediate block: rows x 16
    const float* src = Z;
    int sr0 = 0, sc0 = 0, sp = 40;
    Best best = best_empty();
#pragma unroll
    for (int L = 4; L <= F; L++) {
        // Lanes cover each pass's own extent: the block is W x W (at most), its vertical pass
        // W rows by WS source columns. Blocks keep a pitch of 16 floats.
        const int W = tile_rows(L, F), WS = feed_width(W);
        const int NV = (W * WS + 63) / 64, NH = (W * W + 63) / 64;   // lane iterations per pass
        const int n = 5 << (L - 1);
        const int nr = rhi[L] - rlo[L] + 1;
        const int ncs = chi[L - 1] - clo[L - 1] + 1;
        float xa[4], xb[4], xq[4];
        bool x3[4];
        // vertical pass: tmp[r][c] = level-L row rlo + r over level-(L-1) column clo + c
#pragma unroll
        for (int u = 0; u < NV; u++) {
            int t = lane + 64 * u;
            int r = t / WS, c = t - r * WS;
            r = r < nr ? r : nr - 1;
            c = c < ncs ? c : ncs - 1;
            int m, cc, q, nt;
            taps(rlo[L] + r, n, m, cc, q, nt);
            const float* col = src + (clo[L - 1] + c - sc0);
            xa[u] = col[(m - sr0) * sp];
            xb[u] = col[(cc - sr0) * sp];
            xq[u] = col[(q - sr0) * sp];
            x3[u] = nt == 3;
        }
#pragma unroll
        for (int u = 0; u < NV; u++) {
            int t = lane + 64 * u;
            int r = t / WS, c = t - r * WS;
            if (r < nr && c < ncs)
                tmp[r * 16 + c] = x3[u] ? tap3(xa[u], xb[u], xq[u]) : tap2(xb[u], xq[u]);
        }
        wave_sync();
        const float* row0 = tmp - clo[L - 1];
        const int nc = chi[L] - clo[L] + 1;
        // horizontal pass: level-L samples (rlo + r, clo + c)
#pragma unroll
        for (int u = 0; u < NH; u++) {
            int t = lane + 64 * u;
            int r = t / W, c = t - r * W;
            r = r < nr ? r : nr - 1;
            c = c < nc ? c : nc - 1;
            int m, cc, q, nt;
            taps(clo[L] + c, n, m, cc, q, nt);
            const float* row = row0 + r * 16;
            xa[u] = row[m];
            xb[u] = row[cc];
            xq[u] = row[q];
            x3[u] = nt == 3;
        }
        if (L < F) {
            float mx = -INFINITY;
#pragma unroll
            for (int u = 0; u < NH; u++) {
                int t = lane + 64 * u;
                int r = t / W, c = t - r * W;
                if (r < nr && c < nc) {
                    float v = x3[u] ? tap3(xa[u], xb[u], xq[u]) : tap2(xb[u], xq[u]);
                    blk[r * 16 + c] = v;
                    mx = v > mx ? v : mx;
                }
            }
            const bool cut = prune && ub_margin(wave_max(mx)) < lb;
            wave_sync();
            if (cut) return best;
            src = blk;
            sr0 = rlo[L];
            sc0 = clo[L];
            sp = 16;
        } else {
            const int nf = 5 << F;
#pragma unroll
            for (int u = 0; u < NH; u++) {
                int t = lane + 64 * u;
                float v = x3[u] ? tap3(xa[u], xb[u], xq[u]) : tap2(xb[u], xq[u]);
                int idx = (rlo[F] + (t >> 4)) * nf + clo[F] + (t & 15);
                if (better(v, idx, best)) { best.v = v; best.idx = idx; }
            }
            best = wave_best_dpp(best);
            wave_sync();   // tmp is rewritten by the next tile
        }
    }
    return best;
}

// Final levels 16x..64x (F = 4..6) from the materialised 8x level Z; b3 is the wave's argmax
// of Z. ws is the tile scratch.
template <int F>
__device__ inline Best cascade_tiles(const float* Z, float* ws, Best b3) {
    const int lane = threadIdx.x & 63;
    constexpr int NT = (5 << F) / AR_TILE;
    int i3 = 20, j3 = 20;
    if (b3.idx != 0x7fffffff) {
        i3 = b3.idx / 40;
        j3 = b3.idx - 40 * i3;
    }
    // the tile holding the image of the 8x maximum, unpruned: its best is the threshold
    const int ti0 = (i3 << (F - 3)) / AR_TILE, tj0 = (j3 << (F - 3)) / AR_TILE;
    Best best = tile_run<F>(Z, ws, ti0, tj0, -INFINITY, false);
    const float LB = best.v;

    // Candidates: tiles fed by an 8x sample whose bound reaches LB (a tile's bound is that of its
    // largest feeder; NaN feeders never raise it). Lane k < NT holds the 8x range feeding tile
    // row (and column) k, and the bitmask of flagged tile columns of tile row k.
    int mylo = 1 << 20, myhi = -1;
    if (lane < NT) {
        int lo = AR_TILE * lane, hi = lo + AR_TILE - 1;
        for (int L = F; L > 3; L--) feed_range(lo, hi, 5 << (L - 1), lo, hi);
        mylo = lo;
        myhi = hi;
    }
    uint32_t rowflags = 0;
    const float4* Z4 = (const float4*)Z;
    for (int c0 = 0; c0 < 400; c0 += 64) {
        const int c4 = c0 + lane;
        float4 v4 = Z4[c4 < 400 ? c4 : 399];
        const float vv[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
        for (int e = 0; e < 4; e++) {
            uint64_t hm = __ballot(c4 < 400 && ub_margin(vv[e]) >= LB);
            while (hm) {
                const int src = __ffsll((unsigned long long)hm) - 1;
                hm &= hm - 1;
                const int idx = 4 * (c0 + src) + e, r3 = idx / 40, c3 = idx - 40 * r3;
                const uint64_t rm = __ballot(mylo <= r3 && r3 <= myhi);
                const uint64_t cm = __ballot(mylo <= c3 && c3 <= myhi);
                if ((rm >> lane) & 1) rowflags |= (uint32_t)cm;
            }
        }
    }
    if (lane == ti0) rowflags &= ~(1u << tj0);
    float lb = LB;
    for (int ti = 0; ti < NT; ti++) {
        uint32_t m = (uint32_t)__shfl((int)rowflags, ti);
        while (m) {
            const int tj = __ffs(m) - 1;
            m &= m - 1;
            Best r = tile_run<F>(Z, ws, ti, tj, lb, true);
            best = best_merge(best, r);
            lb = best.v > lb ? best.v : lb;
        }
    }
    return best;
}

// ---- nested blocks: the common case of a single sharp peak in one pass per level
//
// A source sample s of a pyrUp axis over n samples feeds exactly the outputs
// [max(2s - 2, 0), min(2s + 2, 2n - 1)] (taps above, reflection at 0 and clamping at n - 1
// included). Blocks of 16 x 16 samples are computed at levels 3 .. F, each centred on the
// previous block's argmax (and placed so that every sample feeding it lies in that block; level
// 3 is fed by the materialised 4x level). A sample of level L is interior when every sample it
// feeds lies in the level L + 1 block. Every final-level sample outside the final block is fed only
// by non-interior samples of the level before, and, level by level, every sample outside a block
// only by non-interior samples of the level below; so with M the largest non-NaN value among the
// non-interior samples of the 4x level and of the blocks below F, every final-level sample outside
// the final block is at most ub_margin(M) (at most eight passes: 8 * 1.75 u < 1e-6, and the
// subnormal terms stay below 1e-30). When that is strictly below the final block's maximum, the
// block's argmax is the level's. Otherwise the 8x level is materialised and the tile search runs.

// Outputs of level 2n fed by source s: [max(2s - 2, 0), min(2s + 2, 2n - 1)] within [lo, lo + 15].
__device__ inline bool feeds_inside(int s, int n, int lo) {
    const int a = 2 * s - 2 > 0 ? 2 * s - 2 : 0;
    const int b = 2 * s + 2 < 2 * n - 1 ? 2 * s + 2 : 2 * n - 1;
    return a >= lo && b <= lo + 15;
}

// Level-L samples (R0 .. R0 + 15) x (C0 .. C0 + 15) of a level of n samples a side, from the
// level below held at src (its sample (s, t) at src[(s - sr0) sp + t - sc0]) through tmp (16 x 16
// floats, pitch 16), into dst (pitch 16). Returns the lane's argmax under the peak_index rule, indexed
// r 16 + c within the block.
__device__ inline Best block16(const float* src, int sr0, int sc0, int sp, int n, int R0, int C0,
                               float* tmp, float* dst) {
    const int lane = threadIdx.x & 63, ns = n >> 1;
    int clo, chi;
    feed_range(C0, C0 + 15, ns, clo, chi);
    const int ncs = chi - clo + 1;   // at most feed_width(16) = 10
    float xa[4], xb[4], xq[4];
    bool x3[4];
#pragma unroll
    for (int u = 0; u < 3; u++) {   // vertical: 16 rows x ncs source columns
        int t = lane + 64 * u;
        int r = t / 10, c = t - r * 10;
        r = r < 16 ? r : 15;
        c = c < ncs ? c : ncs - 1;
        int m, cc, q, nt;
        taps(R0 + r, ns, m, cc, q, nt);
        const float* col = src + (clo + c - sc0);
        xa[u] = col[(m - sr0) * sp];
        xb[u] = col[(cc - sr0) * sp];
        xq[u] = col[(q - sr0) * sp];
        x3[u] = nt == 3;
    }
#pragma unroll
    for (int u = 0; u < 3; u++) {
        int t = lane + 64 * u;
        int r = t / 10, c = t - r * 10;
        if (r < 16 && c < ncs) tmp[r * 16 + c] = x3[u] ? tap3(xa[u], xb[u], xq[u]) : tap2(xb[u], xq[u]);
    }
    wave_sync();
#pragma unroll
    for (int u = 0; u < 4; u++) {   // horizontal: 16 x 16
        const int t = lane + 64 * u, r = t >> 4, c = t & 15;
        int m, cc, q, nt;
        taps(C0 + c, ns, m, cc, q, nt);
        const float* row = tmp + r * 16 - clo;
        xa[u] = row[m];
        xb[u] = row[cc];
        xq[u] = row[q];
        x3[u] = nt == 3;
    }
    Best best = best_empty();
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int t = lane + 64 * u;
        const float v = x3[u] ? tap3(xa[u], xb[u], xq[u]) : tap2(xb[u], xq[u]);
        if (dst) dst[t] = v;
        if (better(v, t, best)) { best.v = v; best.idx = t; }
    }
    wave_sync();
    return best;
}

// Start row of the level-2n block centred on row i of a level-n block at R0: every sample
// feeding it (feed_range) stays in R0 .. R0 + 15, and the block stays in the level.
__device__ inline int next_block(int R0, int i, int n) {
    int r = 2 * i - 7;
    const int lo = R0 == 0 ? 0 : 2 * R0 + 1;
    const int hi = R0 + 16 == n ? 2 * n - 16 : 2 * R0 + 14;
    r = r < lo ? lo : (r > hi ? hi : r);
    return r < 0 ? 0 : (r > 2 * n - 16 ? 2 * n - 16 : r);
}

// Largest non-NaN value among the non-interior samples of a 16 x 16 block at (R0, C0) of a level
// of n samples, against the next level's block at (R1, C1) (wave-uniform).
__device__ inline float block_ring_max(const float* blk, int R0, int C0, int n, int R1, int C1) {
    const int lane = threadIdx.x & 63;
    float mx = -INFINITY;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int t = lane + 64 * u, r = t >> 4, c = t & 15;
        const float v = blk[t];
        if (!(feeds_inside(R0 + r, n, R1) && feeds_inside(C0 + c, n, C1)) && v > mx) mx = v;
    }
    return wave_max(mx);
}

// Final levels 16x..64x (F = 4..6) by nested blocks from the materialised level BL (Z, NB x NB
// with NB = 5 * 2^BL) and its argmax bb; ws holds 768 floats. Returns false when the bound does
// not settle the argmax (then `out` is unset).
template <int F, int BL>
__device__ inline bool cascade_nested(const float* Z, float* ws, Best bb, Best& out) {
    constexpr int NB = 5 << BL, NQ = NB / 4;
    if (bb.idx == 0x7fffffff) return false;
    const int lane = threadIdx.x & 63;
    float* blk[2] = {ws, ws + 256};
    float* tmp = ws + 512;
    const int ib = bb.idx / NB, jb = bb.idx - NB * (bb.idx / NB);
    int R0 = 2 * ib - 7, C0 = 2 * jb - 7;
    R0 = R0 < 0 ? 0 : (R0 > 2 * NB - 16 ? 2 * NB - 16 : R0);
    C0 = C0 < 0 ? 0 : (C0 > 2 * NB - 16 ? 2 * NB - 16 : C0);
    // non-interior samples of level BL against the first block (rows of NQ float4)
    float mring = -INFINITY;
    {
        const float4* Z4 = (const float4*)Z;
        for (int k = lane; k < NB * NQ; k += 64) {
            const int r = k / NQ, c4 = 4 * (k - NQ * r);
            const bool rin = feeds_inside(r, NB, R0);
            const float4 v4 = Z4[k];
            const float vv[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (!(rin && feeds_inside(c4 + e, NB, C0)) && vv[e] > mring) mring = vv[e];
        }
        mring = wave_max(mring);
    }
    const float* src = Z;
    int sr0 = 0, sc0 = 0, sp = NB;
#pragma unroll
    for (int L = BL + 1; L <= F; L++) {
        const int n = 5 << L;
        float* dst = L < F ? blk[L & 1] : nullptr;
        const Best b = wave_best_dpp(block16(src, sr0, sc0, sp, n, R0, C0, tmp, dst));
        if (L == F) {
            if (!(b.idx != 0x7fffffff && ub_margin(mring) < b.v)) return false;
            out.v = b.v;
            out.idx = (R0 + (b.idx >> 4)) * n + C0 + (b.idx & 15);
            return true;
        }
        if (b.idx == 0x7fffffff) return false;
        const int R1 = next_block(R0, R0 + (b.idx >> 4), n);
        const int C1 = next_block(C0, C0 + (b.idx & 15), n);
        const float m = block_ring_max(dst, R0, C0, n, R1, C1);
        mring = m > mring ? m : mring;
        src = dst;
        sr0 = R0;
        sc0 = C0;
        sp = 16;
        R0 = R1;
        C0 = C1;
    }
    return false;
}

// The cascade for upsampling `up` (power of two, 2..64). `reg` is the wave's region
// (cascade_floats(up) floats, 16-byte aligned) with the 5x5 patch (row-major) at its front.
// Returns the final-level argmax under the peak_index rule, identical in every lane.
__device__ inline CascadeOut cascade(float* reg, int up) {
    const float* patch = reg;
    const int L = casc_levels(up);
    const int LM = L < 3 ? L : 3;
    const int nM = 5 << LM;
    float* Y = reg + AR_CASC_HDR;          // nM x nM
    float* T = Y + nM * nM;                // vertical-pass tmp (2n x n), then the tile scratch

    // Levels 1..LM, each in Y's bytes: a step's vertical pass reads the level before into T, its
    // horizontal pass writes the new level over it. The last step's argmax comes along.
    Best b = best_empty();
    if (LM >= 1) b = wave_pyrup<5>(patch, T, Y);
    if (LM >= 2) b = wave_pyrup<10>(Y, T, Y);
    int nf = 5 << LM;
    if (L >= 4) {
        // nested blocks from the 4x level; if they do not settle it, the 8x level and the tiles
        b = wave_best_dpp(b);
        Best nb;
        const bool done = L == 4 ? cascade_nested<4, 2>(Y, T, b, nb)
                        : L == 5 ? cascade_nested<5, 2>(Y, T, b, nb)
                                 : cascade_nested<6, 2>(Y, T, b, nb);
        if (done) {
            b = nb;
        } else {
            b = wave_best_dpp(wave_pyrup<20>(Y, T, Y));
            if (L == 4) b = cascade_tiles<4>(Y, T, b);
            else if (L == 5) b = cascade_tiles<5>(Y, T, b);
            else b = cascade_tiles<6>(Y, T, b);
        }
        nf = 5 << L;
    } else {
        if (LM >= 3) b = wave_pyrup<20>(Y, T, Y);
        b = wave_best_dpp(b);
    }
    CascadeOut out;
    // All samples NaN: the reference's scan leaves (1, 1), whose value is NaN.
    if (b.idx == 0x7fffffff) {
        b.idx = 0;
        b.v = __int_as_float(0x7fc00000);
    }
    out.si = b.idx / nf;
    out.sj = b.idx - out.si * nf;
    out.val = b.v;
    return out;
}

}  // namespace ar
