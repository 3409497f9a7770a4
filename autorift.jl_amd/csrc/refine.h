// Integer peak selection and the 64x pyrUp subpixel cascade, run by one 256-thread workgroup per
// point. Follows src/peak.jl:32-106 (argmax rule, offsets) and src/peak.jl:196-522 (pyrUp steps
// and subpixel_peak) element for element: every upsampled sample is computed by the reference's
// formula in FP32 from the same inputs, so the result is bit-identical whatever the order in which
// samples are produced, and whichever of them are produced more than once.
//
// Cost control without approximation. Levels up to 40x40 (8x) are materialised in LDS. Beyond,
// the final level is cut into 16x16 tiles, each computed by one wave from the 40x40 level
// through its own small blocks of the intermediate levels. A tile is skipped only when a
// proven upper bound on every sample it could hold is strictly below a sample already found:
// first the 2x2 tiles around the image of the 8x maximum are computed (that best sample is the
// threshold), then every other tile fed by an 8x sample whose bound reaches the threshold, each
// re-checked against the bound of its own intermediate blocks before going deeper.
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
#define AR_TILE_WS 416    // per-wave tile scratch: vertical tmp 16 x 16 + block 10 x 16

// Region layout (floats): patch [0, 32), scalars [32, 64), per-tile 8x row ranges lo [64, 96)
// and hi [96, 128), tile flags [128, 528) and tile list [528, 928) as ints, then the levels.
#define AR_CASC_HDR 928

__host__ __device__ inline int casc_levels(int up) {
    int L = 0;
    while ((1 << L) < up) L++;
    return L;
}

// Floats needed by the cascade region for upsampling `up` (power of two, 1..64).
__host__ __device__ inline int cascade_floats(int up) {
    if (up <= 1) return AR_CASC_HDR;
    // materialised levels: Y (last materialised level), X (previous), T (vertical tmp)
    int L = casc_levels(up);
    int LM = L < 3 ? L : 3;
    int nM = 5 << LM;
    int nP = 5 << (LM > 0 ? LM - 1 : 0);
    int Y = nM * nM;
    int X = nP * nP;
    int T = nM * nP;
    int tiles = L > 3 ? AR_NWAVES * AR_TILE_WS : 0;   // aliases X and T once Y is complete
    return AR_CASC_HDR + Y + (X + T > tiles ? X + T : tiles);
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

// Horizontal pass: `rows` rows of WS source columns -> 2 WS output columns. With ARGMAX, also
// returns the thread's argmax over the samples it wrote (row-major index of the output).
template <int WS, bool ARGMAX = false>
__device__ inline Best horiz_pairs(const float* src, float* dst, int rows) {
    Best best = best_empty();
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
                float e = tap3(a[u], b[u], q[u]), d = tap2(b[u], q[u]);
                dst[o[u]] = e;
                dst[o[u] + 1] = d;
                if (ARGMAX) {
                    if (better(e, o[u], best)) { best.v = e; best.idx = o[u]; }
                    if (better(d, o[u] + 1, best)) { best.v = d; best.idx = o[u] + 1; }
                }
            }
        }
    }
    return best;
}

// One full pyrUp step in LDS: src (N x N, row-major) -> dst (2N x 2N), via tmp (2N x N).
template <int N, bool ARGMAX = false>
__device__ inline Best pyrup_full(const float* src, float* tmp, float* dst) {
    vert_pairs<N>(src, N, 0, tmp, 0, 0, N);
    __syncthreads();
    Best b = horiz_pairs<N, ARGMAX>(tmp, dst, 2 * N);
    __syncthreads();
    return b;
}

__device__ inline float ub_margin(float m) {
    if (!(m > -INFINITY)) return m;   // no finite feeder: nothing can beat anything
    return m + fabsf(m) * 1e-6f + 1e-30f;
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

// Intra-wave LDS hand-off: the LDS operations of one wave complete in order, so ordering the
// compiler's accesses is enough.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) {
        float u = __shfl_xor(v, o);
        v = u > v ? u : v;
    }
    return v;
}

// Rows of the level-L block of a final-level tile: 16 on the final level F, then 10 and 7.
__host__ __device__ constexpr int tile_rows(int L, int F) {
    return L == F ? AR_TILE : (L == F - 1 ? 10 : 7);
}

// One wave computes the final-level tile (ti, tj) of level F from the materialised 8x level Z
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
        const int NIT = (tile_rows(L, F) * 16 + 63) / 64;   // lane iterations covering the block
        const int n = 5 << (L - 1);
        const int nr = rhi[L] - rlo[L] + 1;
        const int ncs = chi[L - 1] - clo[L - 1] + 1;
        float xa[4], xb[4], xq[4];
        bool x3[4];
        // vertical pass: tmp[r][c] = level-L row rlo + r over level-(L-1) column clo + c
#pragma unroll
        for (int u = 0; u < NIT; u++) {
            int t = lane + 64 * u;
            int r = t >> 4, c = t & 15;
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
        for (int u = 0; u < NIT; u++) {
            int t = lane + 64 * u;
            if ((t >> 4) < nr && (t & 15) < ncs)
                tmp[t] = x3[u] ? tap3(xa[u], xb[u], xq[u]) : tap2(xb[u], xq[u]);
        }
        wave_sync();
        const float* row0 = tmp - clo[L - 1];
        const int nc = chi[L] - clo[L] + 1;
        // horizontal pass: level-L samples (rlo + r, clo + c)
#pragma unroll
        for (int u = 0; u < NIT; u++) {
            int t = lane + 64 * u;
            int r = t >> 4, c = t & 15;
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
            for (int u = 0; u < NIT; u++) {
                int t = lane + 64 * u;
                if ((t >> 4) < nr && (t & 15) < nc) {
                    float v = x3[u] ? tap3(xa[u], xb[u], xq[u]) : tap2(xb[u], xq[u]);
                    blk[t] = v;
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
            for (int u = 0; u < NIT; u++) {
                int t = lane + 64 * u;
                float v = x3[u] ? tap3(xa[u], xb[u], xq[u]) : tap2(xb[u], xq[u]);
                int idx = (rlo[F] + (t >> 4)) * nf + clo[F] + (t & 15);
                if (better(v, idx, best)) { best.v = v; best.idx = idx; }
            }
            best = wave_best(best);
            wave_sync();   // tmp is rewritten by the wave's next tile
        }
    }
    return best;
}

// Set-up of the tile search, run before the materialised levels (their barriers publish it):
// per-tile 8x row ranges (the same serve columns), cleared flags and counter.
template <int F>
__device__ inline void tiles_init(float* reg) {
    const int tid = threadIdx.x;
    int* sc = (int*)(reg + 32);
    int* r3lo = (int*)(reg + 64);
    int* r3hi = (int*)(reg + 96);
    int* flags = (int*)(reg + 128);
    const int NT = (5 << F) / AR_TILE;
    if (tid < NT) {
        int lo = AR_TILE * tid, hi = lo + AR_TILE - 1;
        for (int L = F; L > 3; L--) feed_range(lo, hi, 5 << (L - 1), lo, hi);
        r3lo[tid] = lo;
        r3hi[tid] = hi;
    }
    for (int t = tid; t < NT * NT; t += AR_WG) flags[t] = 0;
    if (tid == 0) sc[0] = 0;
}

// Final levels 16x..64x (F = 4..6) from the materialised 8x level in Y; b3 is each thread's
// argmax over the 8x samples it produced.
template <int F>
__device__ inline CascadeOut cascade_tiles(float* reg, const float* Y, float* ws, Best b3,
                                           Best* red) {
    const int tid = threadIdx.x, w = tid >> 6;
    int* sc = (int*)(reg + 32);
    const int* r3lo = (const int*)(reg + 64);
    const int* r3hi = (const int*)(reg + 96);
    int* flags = (int*)(reg + 128);
    int* list = (int*)(reg + 528);
    const int nf = 5 << F, NT = nf / AR_TILE;

    // Phase 1: the 2x2 tiles around the image of the 8x maximum, one per wave, no pruning.
    b3 = block_best(b3, red);
    int i3 = 20, j3 = 20;
    if (b3.idx != 0x7fffffff) {
        i3 = b3.idx / 40;
        j3 = b3.idx - 40 * i3;
    }
    int ti0 = ((i3 << (F - 3)) + AR_TILE / 2) / AR_TILE - 1;
    int tj0 = ((j3 << (F - 3)) + AR_TILE / 2) / AR_TILE - 1;
    ti0 = ti0 < 0 ? 0 : (ti0 > NT - 2 ? NT - 2 : ti0);
    tj0 = tj0 < 0 ? 0 : (tj0 > NT - 2 ? NT - 2 : tj0);
    Best wb = tile_run<F>(Y, ws, ti0 + (w >> 1), tj0 + (w & 1), -INFINITY, false);
    const Best pb = block_best(wb, red);
    const float LB = pb.v;

    // Phase 2: list every other tile fed by an 8x sample whose bound reaches LB (a tile's bound
    // is that of its largest feeder; NaN feeders never raise it). The flags deduplicate.
    for (int t4 = tid; t4 < 400; t4 += AR_WG) {
        const float4 v4 = ((const float4*)Y)[t4];
        const float vv[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
        for (int e = 0; e < 4; e++) {
            if (!(ub_margin(vv[e]) >= LB)) continue;
            int t = 4 * t4 + e, r3 = t / 40, c3 = t - 40 * r3;
            int a0 = NT, a1 = -1, b0 = NT, b1 = -1;
            for (int k = 0; k < NT; k++) {
                int lo = r3lo[k], hi = r3hi[k];
                if (lo <= r3 && r3 <= hi) { a0 = k < a0 ? k : a0; a1 = k; }
                if (lo <= c3 && c3 <= hi) { b0 = k < b0 ? k : b0; b1 = k; }
            }
            for (int ti = a0; ti <= a1; ti++)
                for (int tj = b0; tj <= b1; tj++) {
                    bool first = (unsigned)(ti - ti0) < 2u && (unsigned)(tj - tj0) < 2u;
                    int f = ti * NT + tj;
                    if (!first && atomicExch(&flags[f], 1) == 0) list[atomicAdd(&sc[0], 1)] = f;
                }
        }
    }
    __syncthreads();
    const int cnt = sc[0];

    // Phase 3: the listed tiles, wave-strided, each pruned against the best value known.
    Best best = pb;
    float lb = LB;
    for (int k = w; k < cnt; k += AR_NWAVES) {
        int t = list[k];
        int ti = t / NT, tj = t - ti * NT;
        Best r = tile_run<F>(Y, ws, ti, tj, lb, true);
        best = best_merge(best, r);
        lb = best.v > lb ? best.v : lb;
    }
    best = block_best(best, red);

    CascadeOut out;
    // All samples NaN: the reference's scan leaves (1, 1), whose value is NaN.
    if (best.idx == 0x7fffffff) {
        best.idx = 0;
        best.v = __int_as_float(0x7fc00000);
    }
    out.si = best.idx / nf;
    out.sj = best.idx - out.si * nf;
    out.val = best.v;
    return out;
}

// The cascade. `patch` (5x5, row-major) is in LDS; `reg` is the cascade region
// (cascade_floats(up) floats, 16-byte aligned). Returns the final-level argmax under the
// peak_index rule; identical for every thread.
__device__ inline CascadeOut cascade(float* reg, int up, Best* red) {
    const float* patch = reg;
    float* body = reg + AR_CASC_HDR;
    const int L = casc_levels(up);
    const int LM = L < 3 ? L : 3;
    const int nM = 5 << LM;
    const int nP = 5 << (LM > 0 ? LM - 1 : 0);
    float* Y = body;                       // nM x nM
    float* X = Y + nM * nM;                // nP x nP
    float* T = X + nP * nP;                // vertical-pass tmp (2n x n)

    if (L == 4) tiles_init<4>(reg);
    else if (L == 5) tiles_init<5>(reg);
    else if (L == 6) tiles_init<6>(reg);

    // Levels 1..LM, ping-ponging so that level LM lands in Y.
    const float* cur = patch;
    Best b3 = best_empty();
    if (LM >= 1) { float* d = ((LM - 1) & 1) ? X : Y; pyrup_full<5>(cur, T, d); cur = d; }
    if (LM >= 2) { float* d = ((LM - 2) & 1) ? X : Y; pyrup_full<10>(cur, T, d); cur = d; }
    if (LM >= 3) { float* d = ((LM - 3) & 1) ? X : Y; b3 = pyrup_full<20, true>(cur, T, d); cur = d; }

    if (L >= 4) {
        float* ws = X + (threadIdx.x >> 6) * AR_TILE_WS;
        CascadeOut out;
        if (L == 4) out = cascade_tiles<4>(reg, Y, ws, b3, red);
        else if (L == 5) out = cascade_tiles<5>(reg, Y, ws, b3, red);
        else out = cascade_tiles<6>(reg, Y, ws, b3, red);
        __syncthreads();
        return out;
    }

    // final level materialised: plain argmax
    const int nfin = 5 << LM;
    Best b = best_empty();
    for (int t = threadIdx.x; t < nfin * nfin; t += AR_WG) {
        float v = cur[t];
        if (better(v, t, b)) { b.v = v; b.idx = t; }
    }
    b = block_best(b, red);
    if (b.idx == 0x7fffffff) b.idx = 0;
    CascadeOut out;
    out.si = b.idx / nfin;
    out.sj = b.idx - out.si * nfin;
    out.val = cur[b.idx];
    __syncthreads();
    return out;
}

}  // namespace ar
