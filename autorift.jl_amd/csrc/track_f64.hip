// Direct Float64 numerator for Float32 images (and UInt8 chips whose area is not a power of two):
// the reference's arithmetic in its own order, fast.
//
// A point's correlation is split over two kernels with a staging slot per point between them; the
// points go in chunks, and the preparation of the next chunk runs on a second stream while the
// current chunk is tracked (it is latency-bound, the tracker is bound by the FP64 pipes):
//
//  * f64_prep_kernel (four waves per point): the chip's Float64 mean (src/correlate.jl:437-441),
//    the mean-removed chip Float32(x - mean) written as Float64 for the second kernel, the chip
//    energy (the reference's sequential sum, evaluated exactly by whole-wave steps), and the window's summed-area
//    tables in the reference's recurrence (src/integral.jl:141-170), from which the box sums of
//    W and W^2 at every shift (src/integral.jl:244-247) go into the slot.
//
//  * track_f64_kernel (one lane per 8-row tile of shifts): the direct numerator of every shift
//    (src/correlate.jl:799-814) as eight independent Float64 FMA chains per lane. Each chain
//    adds its products in the reference's order -- chip column jj outer, chip row ii inner -- and
//    a product of two Float32 values is exact in Float64, so fma(c, w, acc) is the reference's
//    `acc += c * w`. The window sits in LDS as Float32; the chip values come from the slot through
//    scalar loads (they are uniform over the workgroup) and enter the FMAs as scalar operands.
//    The normalisation (src/correlate.jl:591-630), argmax (src/peak.jl:32-61) and the 5x5 patch
//    hand-off to the refinement kernel follow.
//
// The sum of the chip (the mean's numerator) is the only reduction whose order could matter in
// parallel. It is taken in parallel only when every partial sum of the reference's sequential loop
// is exactly representable, so that every order gives the same bits: all values are multiples
// of 2^emin (the lowest set bit over the chip) and |partial sums| <= n 2^(emax+1) < 2^(emin+53).
// UInt8 chips always qualify; otherwise one lane sums in the reference's order.
#include "layout.h"
#include "track_common.h"

namespace ar {

// A chunk's point k -> the block that handles it: the blocks of each XCD (dispatched round-robin
// over the 8 XCDs) take one contiguous eighth of the chunk, so overlapping windows of neighbouring
// points are read through the same L2. Returns -1 for the padding blocks of the grid.
__device__ inline int64_t f64_point_of_block(int64_t cnt) {
    const int64_t b = blockIdx.x, per = (cnt + 7) / 8;
    const int64_t k = (b % 8) * per + b / 8;
    return (b / 8 < per && k < cnt) ? k : -1;
}

__device__ inline float f64_pix(const TrackArgs& a, int is_u8, int64_t r, int64_t c) {
    const int64_t k = pix(a, r, c);
    if (k < 0) return 0.0f;
    return is_u8 ? (float)((const uint8_t*)a.ref)[k] : ((const float*)a.ref)[k];
}

// Lowest set bit exponent (value = m 2^e with m odd) and the exponent of the leading bit of a
// finite nonzero float.
__device__ inline void f32_bits(float x, int& elow, int& ehigh) {
    const uint32_t u = __float_as_uint(x) & 0x7fffffffu;
    const int be = (int)(u >> 23);
    const uint32_t m = u & 0x7fffffu;
    if (be == 0) {   // subnormal
        elow = -149 + __builtin_ctz(m);
        ehigh = -149 + 31 - __builtin_clz(m);
    } else {
        elow = be - 150 + __builtin_ctz(m | 0x800000u);
        ehigh = be - 127;
    }
}

// Stages `total` elements into LDS, element e = load(e): each thread issues 16 independent global
// loads before it stores any, so a staging loop waits for memory once per 16 elements instead of
// once per element.
template <typename L>
__device__ inline void f64_stage(float* dst, int total, int tid, int nthr, L load) {
    for (int e0 = tid; e0 < total; e0 += 16 * nthr) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int e = e0 + u * nthr;
            v[u] = e < total ? load(e) : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int e = e0 + u * nthr;
            if (e < total) dst[e] = v[u];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// kernel 1: chip statistics, mean-removed chip, box sums of every shift
// ---------------------------------------------------------------------------------------------

// Summed-area table of W (SQ = 0) or of W^2 (SQ = 1) in the reference's recurrence
// (src/integral.jl:141-170), split over `nw` waves (wave index w): phase 1 writes column j's running
// sums into table column j + 1, rows 1..winh; phase 2 (after a workgroup barrier) forms
// S[i, j + 1] = run(i, j) + S[i, j] along each row i.
template <int IS_U8, int SQ>
__device__ inline void f64_table_cols(const TrackArgs& b, const Geom& g, int winh, int winw,
                                      double* T, const float* Wl, int w, int nw, int lane) {
    const int ls = winh + 1;
    const bool inside = g.wr0 - b.pady >= 1 && g.wc0 - b.padx >= 1 &&
                        g.wr0 - b.pady + winh - 1 <= b.ny && g.wc0 - b.padx + winw - 1 <= b.nx;
    const int64_t base = (g.wr0 - b.pady - 1) + (g.wc0 - b.padx - 1) * b.ny;
    for (int j = w * 64 + lane; j <= winw; j += nw * 64) T[j * ls] = 0.0;
    for (int j = w * 64 + lane; j < winw; j += nw * 64) {
        double run = 0.0;
        for (int i0 = 0; i0 < winh; i0 += 16) {
            float v[16];
#pragma unroll
            for (int u = 0; u < 16; u++) {
                const int r = i0 + u;
                if (r >= winh) v[u] = 0.0f;
                else if (Wl) v[u] = Wl[r + j * winh];
                else if (inside) {
                    const int64_t q = base + r + (int64_t)j * b.ny;
                    v[u] = IS_U8 ? (float)((const uint8_t*)b.ref)[q] : ((const float*)b.ref)[q];
                } else {
                    v[u] = f64_pix(b, IS_U8, g.wr0 + r, g.wc0 + j);
                }
            }
#pragma unroll
            for (int u = 0; u < 16; u++) {
                if (i0 + u < winh) {
                    const double x = (double)v[u];
                    run += SQ ? x * x : x;
                    T[(i0 + u + 1) + (j + 1) * ls] = run;
                }
            }
        }
    }
}

__device__ inline void f64_table_rows(int winh, int winw, double* T, int w, int nw, int lane) {
    const int ls = winh + 1;
    for (int i = w * 64 + lane + 1; i <= winh; i += nw * 64) {
        double s = 0.0;   // S[i, 1] = 0
        int j = 1;
        for (; j + 8 <= winw + 1; j += 8) {
            double r[8];
#pragma unroll
            for (int u = 0; u < 8; u++) r[u] = T[i + (j + u) * ls];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                s = r[u] + s;
                T[i + (j + u) * ls] = s;
            }
        }
        for (; j <= winw; j++) {
            s = T[i + j * ls] + s;
            T[i + j * ls] = s;
        }
    }
    for (int i = w * 64 + lane; i <= winh; i += nw * 64) T[i] = 0.0;
}

// ---- exact sequential sums of non-negative terms by whole-wave steps

__device__ inline double dpp_f64(double v, int ctrl, int row_mask) {
    const uint64_t u = __double_as_longlong(v);
    int lo = (int)(uint32_t)u, hi = (int)(uint32_t)(u >> 32);
    switch (ctrl) {
        case 1: lo = __builtin_amdgcn_update_dpp(0, lo, 0x111, 0xF, 0xF, false);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x111, 0xF, 0xF, false); break;
        case 2: lo = __builtin_amdgcn_update_dpp(0, lo, 0x112, 0xF, 0xF, false);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x112, 0xF, 0xF, false); break;
        case 4: lo = __builtin_amdgcn_update_dpp(0, lo, 0x114, 0xF, 0xF, false);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x114, 0xF, 0xF, false); break;
        case 8: lo = __builtin_amdgcn_update_dpp(0, lo, 0x118, 0xF, 0xF, false);
                hi = __builtin_amdgcn_update_dpp(0, hi, 0x118, 0xF, 0xF, false); break;
        case 15: lo = __builtin_amdgcn_update_dpp(0, lo, 0x142, 0xA, 0xF, false);
                 hi = __builtin_amdgcn_update_dpp(0, hi, 0x142, 0xA, 0xF, false); break;
        default: lo = __builtin_amdgcn_update_dpp(0, lo, 0x143, 0xC, 0xF, false);
                 hi = __builtin_amdgcn_update_dpp(0, hi, 0x143, 0xC, 0xF, false); break;
    }
    (void)row_mask;
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Inclusive prefix sum over the 64 lanes (row_shr 1, 2, 4, 8 within rows of 16, then row_bcast
// 15 and 31). Used only on non-negative integers below 2^53, where every addition is exact.
__device__ inline double wave_scan_f64(double v) {
    v += dpp_f64(v, 1, 0xF);
    v += dpp_f64(v, 2, 0xF);
    v += dpp_f64(v, 4, 0xF);
    v += dpp_f64(v, 8, 0xF);
    v += dpp_f64(v, 15, 0xA);
    v += dpp_f64(v, 31, 0xC);
    return v;
}

__device__ inline double lane_f64(double v, int l) {
    const uint64_t u = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// s + t_0 + t_1 + ... + t_63 added left to right in Float64 (t_l >= 0 finite, this lane's t), the
// bits of the sequential loop, for a wave-uniform s >= 0. While the running sum stays in one
// binade [2^E, 2^(E+1)), every addition rounds to the grid u = 2^(E-52) that s lies on, so
// s_k = s + u (r_0 + ... + r_k) with r_l = t_l / u rounded to the nearest integer -- unless t_l / u
// is a tie (the even neighbour then depends on s) or the sum would leave the binade. The first
// lane where either happens is added on its own, the reference's way, and the scan resumes after
// it. Every step is exact: the r_l and their prefix sums are integers below 2^53.
__device__ inline double seq_sum64(double s, double t, int lane) {
    int start = 0;
    while (start < 64) {
        if (!(s >= 0x1p-900)) {   // zero or tiny: one term at a time
            s = s + lane_f64(t, start);
            start++;
            continue;
        }
        const int E = (int)((__double_as_longlong(s) >> 52) & 0x7ff) - 1023;
        const double u = __longlong_as_double((long long)(E - 52 + 1023) << 52);
        const double inv = __longlong_as_double((long long)(52 - E + 1023) << 52);
        const double q = lane >= start ? t * inv : 0.0;   // exact: scaling by a power of two
        const double fl = floor(q), fr = q - fl;
        const bool tie = fr == 0.5;
        const double r = fr > 0.5 ? fl + 1.0 : fl;
        const double P = wave_scan_f64(r);
        const double lim = __longlong_as_double((long long)(E + 1 + 1023) << 52) - u;
        const bool bad = lane >= start && (tie || !(s + P * u <= lim));
        const uint64_t bm = __ballot(bad);
        const int f = bm ? __builtin_ctzll(bm) : 64;
        if (f > start) s = s + lane_f64(P, f - 1) * u;   // lanes start .. f-1: exact
        if (f < 64) {
            s = s + lane_f64(t, f);
            start = f + 1;
        } else {
            start = 64;
        }
    }
    return s;
}

// The chip energy sum over elements [e0, e1) (src/correlate.jl:442-447: sq += d * d, d =
// x - mean, left to right), continuing sq, by whole-wave steps (seq_sum64). sq is wave-uniform.
__device__ inline double f64_energy(const float* chip, int e0, int e1, double mean, double sq,
                                    int lane) {
    for (int seg = e0; seg < e1; seg += 64) {
        const int e = seg + lane;
        double t = 0.0;
        if (e < e1) {
            const double d = (double)chip[e] - mean;
            t = d * d;
        }
        sq = seq_sum64(sq, t, lane);
    }
    return sq;
}

// One point per workgroup at a time (persistent grid). LDS: the chip (Float32); with MODE >= 1
// one summed-area table, built for W, read for the box sums of W, then rebuilt for W^2 (MODE 0:
// the table is the block's own region of S.tables); with MODE 2 also the window (Float32,
// column pitch winh), staged once with coalesced loads for both tables.
template <int IS_U8, int MODE>
__global__ __launch_bounds__(F64_PREP_WG) void f64_prep_kernel(TrackArgs a, F64Stage S) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* chip = (float*)smem;
    __shared__ double red_d[F64_PREP_WG / 64];
    __shared__ int red_i[2 * (F64_PREP_WG / 64)];
    __shared__ double stat[2];
    const size_t lchip = ((size_t)a.max_ch * a.max_cw * 4 + 15) / 16 * 16;
    const size_t lwin = MODE == 2 ? ((size_t)(a.max_ch + 2 * a.max_ry - 1) *
                                     (a.max_cw + 2 * a.max_rx - 1) * 4 + 15) / 16 * 16 : 0;
    float* Wl = MODE == 2 ? (float*)(smem + lchip) : nullptr;
    double* T = MODE >= 1 ? (double*)(smem + lchip + lwin)
                          : S.tables + (size_t)blockIdx.x * S.table_doubles;
    for (int64_t k = blockIdx.x; k < S.cnt; k += gridDim.x) {
        const int64_t p = a.order ? (int64_t)a.order[S.p0 + k] : S.p0 + k;
        // The argument block read again from the kernarg segment per point, and the lane index
        // hidden from the optimiser: what a point needs is formed for it instead of being held
        // in (spilled) registers across the loop.
        TrackArgs b = reload_args();
        if (b.pairs) use_pair(b, p);
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int lane = tid & 63, wave = tid >> 6;
        unsigned char* slot = S.slots + (size_t)k * S.slot_bytes;
        const Geom g = point_geom(b, p);
        bool go = g.status != 0;
        if (go) {
            // _any_valid (src/track.jl:416-421); any_valid() assumes a 256-thread workgroup
            int found = 0;
            if (b.ok == nullptr) {
                const int64_t r0 = g.cr0 - b.pady - 1, c0 = g.cc0 - b.padx - 1;
                found = r0 + g.ch - 1 >= 0 && r0 < b.ny && c0 + g.cw - 1 >= 0 && c0 < b.nx;
            } else {
                for (int e = tid; e < g.ch * g.cw; e += F64_PREP_WG) {
                    const int64_t q = pix(b, g.cr0 + e % g.ch, g.cc0 + e / g.ch);
                    if (q >= 0 && b.ok[q]) found = 1;
                }
            }
            go = __syncthreads_or(found) != 0;
        }
        if (!go) {
            write_skip(a, p, false);
            if (tid == 0) *(int32_t*)slot = 0;
            __syncthreads();
            continue;
        }
        const int ch = g.ch, cw = g.cw, n = ch * cw;
        const int NR = 2 * g.ry, NC = 2 * g.rx;
        const int winh = ch + NR - 1, winw = cw + NC - 1, ls = winh + 1;
        // the chip (from the secondary), Float32 in LDS, with the bit span of its values
        f64_stage(chip, n, tid, F64_PREP_WG, [&](int e) {
            const int64_t q = pix(b, g.cr0 + e % ch, g.cc0 + e / ch);
            if (q < 0) return 0.0f;
            return IS_U8 ? (float)((const uint8_t*)b.sec)[q] : ((const float*)b.sec)[q];
        });
        int elow = 1 << 20, ehigh = -(1 << 20);
        for (int e = tid; e < n; e += F64_PREP_WG) {
            const float v = chip[e];   // this thread's own stores
            if (v != 0.0f) {
                int lo, hi;
                f32_bits(v, lo, hi);
                elow = min(elow, lo);
                ehigh = max(ehigh, hi);
            }
        }
        if (MODE == 2) {
            // the window (from the reference), column pitch winh
            const bool inside = g.wr0 - b.pady >= 1 && g.wc0 - b.padx >= 1 &&
                                g.wr0 - b.pady + winh - 1 <= b.ny &&
                                g.wc0 - b.padx + winw - 1 <= b.nx;
            const int64_t base = (g.wr0 - b.pady - 1) + (g.wc0 - b.padx - 1) * b.ny;
            f64_stage(Wl, winh * winw, tid, F64_PREP_WG, [&](int e) {
                const int r = e % winh, c = e / winh;
                if (inside) {
                    const int64_t q = base + r + (int64_t)c * b.ny;
                    return IS_U8 ? (float)((const uint8_t*)b.ref)[q] : ((const float*)b.ref)[q];
                }
                return f64_pix(b, IS_U8, g.wr0 + r, g.wc0 + c);
            });
        }
        for (int o = 32; o > 0; o >>= 1) {
            elow = min(elow, __shfl_xor(elow, o));
            ehigh = max(ehigh, __shfl_xor(ehigh, o));
        }
        const int nwv = F64_PREP_WG / 64;
        if (lane == 0) {
            red_i[wave] = elow;
            red_i[nwv + wave] = ehigh;
        }
        __syncthreads();
        elow = red_i[0];
        ehigh = red_i[nwv];
        for (int w = 1; w < nwv; w++) {
            elow = min(elow, red_i[w]);
            ehigh = max(ehigh, red_i[nwv + w]);
        }
        // ceil(log2 n) + ehigh + 1 <= elow + 53: every partial sum exact, any order the same
        const int lg = 32 - __builtin_clz((unsigned)(n - 1) | 1u);
        const bool exact = ehigh < elow || lg + ehigh + 1 <= elow + 53;
        if (exact) {
            double s = 0.0;
            for (int e = tid; e < n; e += F64_PREP_WG) s += (double)chip[e];
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
            if (lane == 0) red_d[wave] = s;
            __syncthreads();
            if (tid == 0) {
                double t = red_d[0];
                for (int w = 1; w < nwv; w++) t += red_d[w];
                stat[0] = t / (double)n;
            }
        } else if (tid == 0) {
            // src/correlate.jl:437-440: column-major, left to right
            double s = 0.0;
            for (int e = 0; e < n; e++) s += (double)chip[e];
            stat[0] = s / (double)n;
        }
        __syncthreads();
        const double mean = stat[0];
        // Float32(x - mean), as Float64, column pitch chp (rows past ch zero)
        double* cd = (double*)(slot + S.off_chip);
        for (int e = tid; e < S.chp * cw; e += F64_PREP_WG) {
            const int i = e % S.chp, j = e / S.chp;
            cd[e] = i < ch ? (double)(float)((double)chip[i + j * ch] - mean) : 0.0;
        }
        // the table of W, then of W^2 (integral_both!, src/integral.jl:141-170), each built by
        // waves 0 and 2 and read for its box sums (src/integral.jl:244-247):
        // ((S[i+h, j+w] - S[i, j+w]) - S[i+h, j]) + S[i, j]. Meanwhile wave 1 sums the chip
        // energy, half during each phase of the first table.
        double sq = 0.0;
        const int tw = wave == 0 ? 0 : (wave == 2 ? 1 : -1);   // table waves
#pragma unroll
        for (int q = 0; q < 2; q++) {
            if (tw >= 0) {
                if (q == 0) f64_table_cols<IS_U8, 0>(b, g, winh, winw, T, Wl, tw, 2, lane);
                else f64_table_cols<IS_U8, 1>(b, g, winh, winw, T, Wl, tw, 2, lane);
            } else if (wave == 1 && q == 0) {
                sq = f64_energy(chip, 0, n / 2, mean, sq, lane);
            }
            __syncthreads();
            if (tw >= 0) f64_table_rows(winh, winw, T, tw, 2, lane);
            else if (wave == 1 && q == 0) sq = f64_energy(chip, n / 2, n, mean, sq, lane);
            __syncthreads();
            double* bq = (double*)(slot + (q == 0 ? S.off_b1 : S.off_b2));
            for (int e = tid; e < NR * NC; e += F64_PREP_WG) {
                const int i = e % NR + 1, j = e / NR + 1;
                bq[e] = T[(i + ch - 1) + (j + cw - 1) * ls] - T[(i - 1) + (j + cw - 1) * ls] -
                        T[(i + ch - 1) + (j - 1) * ls] + T[(i - 1) + (j - 1) * ls];
            }
            __syncthreads();
        }
        // ok = sq > eps(Float64) * ch * cw (src/correlate.jl:452)
        if (wave == 1 && lane == 0) {
            const bool ok = sq > 2.220446049250313e-16 * (double)ch * (double)cw;
            *(int32_t*)slot = ok ? 2 : 1;
            *(double*)(slot + 8) = sqrt(sq);
            stat[1] = ok ? 1.0 : 0.0;
        }
        __syncthreads();
        if (stat[1] == 0.0) write_skip(a, p, true);   // searched, no measurement
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------------
// kernel 2: numerators, normalisation, argmax, patch
// ---------------------------------------------------------------------------------------------

template <int IS_U8>
__global__ __launch_bounds__(1024) void track_f64_kernel(TrackArgs a, F64Stage S, int pitch,
                                                         const double* __restrict__ stage) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* W = (float*)smem;
    __shared__ Best red[16];
    const int64_t k = f64_point_of_block(S.cnt);
    if (k < 0) return;
    const int64_t p = a.order ? (int64_t)a.order[S.p0 + k] : S.p0 + k;
    const unsigned char* slot = S.slots + (size_t)k * S.slot_bytes;
    if (*(const int32_t*)slot != 2) return;   // skipped or degenerate: written by kernel 1
    TrackArgs b = a;
    if (a.pairs) use_pair(b, p);
    const Geom g = point_geom(b, p);
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int ch = g.ch, cw = g.cw, NR = 2 * g.ry, NC = 2 * g.rx;
    const int winh = ch + NR - 1, winw = cw + NC - 1;
    const int chp = (ch + 7) / 8 * 8;
    // the window, Float32, row r of column c at W[c * pitch + 1 + r]; zeros below it
    const bool inside = g.wr0 - b.pady >= 1 && g.wc0 - b.padx >= 1 &&
                        g.wr0 - b.pady + winh - 1 <= b.ny && g.wc0 - b.padx + winw - 1 <= b.nx;
    const int64_t base = (g.wr0 - b.pady - 1) + (g.wc0 - b.padx - 1) * b.ny;
    f64_stage(W, winw * pitch, tid, nthr, [&](int e) {
        const int c = e / pitch, r = e % pitch - 1;
        if (r < 0 || r >= winh) return 0.0f;
        if (inside) {
            const int64_t q = base + r + (int64_t)c * b.ny;
            return IS_U8 ? (float)((const uint8_t*)b.ref)[q] : ((const float*)b.ref)[q];
        }
        return f64_pix(b, IS_U8, g.wr0 + r, g.wc0 + c);
    });
    __syncthreads();
    // this lane's tile: shift rows i0 .. i0 + 7 of shift column j (lanes run along j, so the 8
    // lanes of a 128-byte LDS read fetch columns pitch = 4 mod 8 floats apart: distinct banks)
    const int ntr = (NR + 7) / 8;
    const int t = tid < ntr * NC ? tid : 0;   // idle lanes repeat tile 0 and drop it
    const int j = t % NC, i0 = 8 * (t / NC);
    // the chip, chp x cw Float64 (column-major, pitch S.chp), read through scalar loads: every
    // lane reads the same values, so they are scalar operands of the FMAs
    const double* __restrict__ cd = stage + ((size_t)k * S.slot_bytes + S.off_chip) / 8;
    double acc[8];
#pragma unroll
    for (int r = 0; r < 8; r++) acc[r] = 0.0;
    const int groups = chp / 8;   // groups of 8 chip rows per chip column (rows past ch are 0)
    const double* cp = cd;
    double c[8];
#pragma unroll
    for (int u = 0; u < 8; u++) c[u] = cp[u];
    for (int jj = 0; jj < cw; jj++) {
        const float* wc = W + (j + jj) * pitch + i0;   // wc[1 + m] = W[i0 + m, j + jj]
        double w[8];
        {
            const float4 u0 = *(const float4*)wc;
            const float4 u1 = *(const float4*)(wc + 4);
            w[0] = u0.y; w[1] = u0.z; w[2] = u0.w;
            w[3] = u1.x; w[4] = u1.y; w[5] = u1.z; w[6] = u1.w;
        }
        for (int gq = 0; gq < groups; gq++, wc += 8) {
            // the next group's chip values (past the last column: the slot's box sums, unused)
            const double* np = gq + 1 < groups ? cp + 8 : cd + (size_t)(jj + 1) * S.chp;
            // rows i0 + 8 gq + 7 .. i0 + 8 gq + 14 of the window column
            const float4 u0 = *(const float4*)(wc + 8);
            const float4 u1 = *(const float4*)(wc + 12);
            const float nx[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
            double cn[8];
#pragma unroll
            for (int u = 0; u < 8; u++) cn[u] = np[u];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                // step 8 gq + u: slot (u + 7) % 8 takes W[i0 + 8 gq + u + 7]
                w[(u + 7) & 7] = (double)nx[u];
#pragma unroll
                for (int r = 0; r < 8; r++) acc[r] = __builtin_fma(c[u], w[(u + r) & 7], acc[r]);
            }
#pragma unroll
            for (int u = 0; u < 8; u++) c[u] = cn[u];
            cp = np;
        }
    }
    __syncthreads();   // every lane is past the window: its LDS now holds the surface
    float* surf = (float*)smem;   // surf[i + j * NR]
    const double cnorm = *(const double*)(slot + 8);
    const double* b1 = (const double*)(slot + S.off_b1);
    const double* b2 = (const double*)(slot + S.off_b2);
    const double nn = (double)(ch * cw);
    Best best = best_empty();
    if (tid < ntr * NC) {
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int i = i0 + r;
            if (i < NR) {
                // src/correlate.jl:591-600 (ZNCC) and 619-628 (NCC)
                double den;
                if (a.measure == 1) {
                    const double wsq = b2[i + j * NR];
                    den = cnorm * sqrt(wsq < 0.0 ? 0.0 : wsq);
                } else {
                    const double s1 = b1[i + j * NR], s2 = b2[i + j * NR];
                    const double wv = s2 - s1 * s1 / nn;
                    den = cnorm * sqrt(wv < 0.0 ? 0.0 : wv);
                }
                const float v = den > 0 ? (float)(acc[r] / den) : 0.0f;
                surf[i + j * NR] = v;
                const int idx = i * NC + j;
                if (better(v, idx, best)) {
                    best.v = v;
                    best.idx = idx;
                }
            }
        }
    }
    // argmax under peak_index's rule over the workgroup
    best = wave_best(best);
    const int nw = (nthr + 63) >> 6;
    if ((tid & 63) == 0) red[tid >> 6] = best;
    __syncthreads();
    best = red[0];
    for (int w = 1; w < nw; w++) best = best_merge(best, red[w]);
    if (best.idx == 0x7fffffff) best.idx = 0;   // all NaN: (1, 1)
    const int pi = best.idx / NC + 1, pj = best.idx % NC + 1;
    const float pval = surf[(pi - 1) + (pj - 1) * NR];
    int pi0 = 1, pj0 = 1;
    const bool refine = a.up > 1 && patch_origin(pi, pj, NR, NC, pi0, pj0);
    float pv = 0.0f;
    if (refine && tid < 25) pv = surf[(pi0 - 1 + tid / 5) + (pj0 - 1 + tid % 5) * NR];
    finish_point(a, p, g, pi, pj, pval, pi0, pj0, refine, pv);
}

template __global__ void f64_prep_kernel<0, 0>(TrackArgs, F64Stage);
template __global__ void f64_prep_kernel<1, 0>(TrackArgs, F64Stage);
template __global__ void f64_prep_kernel<0, 1>(TrackArgs, F64Stage);
template __global__ void f64_prep_kernel<1, 1>(TrackArgs, F64Stage);
template __global__ void f64_prep_kernel<0, 2>(TrackArgs, F64Stage);
template __global__ void f64_prep_kernel<1, 2>(TrackArgs, F64Stage);
template __global__ void track_f64_kernel<0>(TrackArgs, F64Stage, int, const double*);
template __global__ void track_f64_kernel<1>(TrackArgs, F64Stage, int, const double*);

}  // namespace ar
