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
#include <type_traits>

#include "layout.h"
#include "refine.h"
#include "seqsum.h"
#include "track_common.h"

namespace ar {

typedef int v4i __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------------------------
// shared tail: peak, refinement, outputs
// ---------------------------------------------------------------------------------------------

// The pyrUp cascade of every point the correlation kernel handed over (src/peak.jl:456-522),
// with the offsets of src/peak.jl:515-521 and the prior. One wave per workgroup, one point at a
// time; no workgroup barriers, so many points are in flight on each CU.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void refine_kernel(
    TrackArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* reg = (float*)smem;
    const int lane = threadIdx.x;
    const PointRange pr = point_range(a.n);
    for (int64_t p = pr.first; p < pr.end; p += pr.step) {
        const int32_t m = a.rmeta[p];
        if (m < 0) continue;
        if (lane < 25) reg[lane] = a.rpatch[32 * p + lane];
        wave_sync();
        CascadeOut c = cascade(reg, a.up);
        const int i0 = m & 0xffff, j0 = m >> 16;
        Geom g;
        g.rx = (int)a.radius_x[p];
        g.ry = (int)a.radius_y[p];
        double factor = (double)a.up;
        double row = (double)c.si / factor + (double)(i0 - 1);
        double col = (double)c.sj / factor + (double)(j0 - 1);
        write_result(a, p, g, col - (double)g.rx, row - (double)g.ry, c.val);
        wave_sync();
    }
}

// ---------------------------------------------------------------------------------------------
// u8 path
// ---------------------------------------------------------------------------------------------

// Upper bound on |approximate - exact| for the FP32 screening score (see u8 epilogue).
#define AR_SCREEN_EPS 4e-6f
// Columns whose global loads a lane issues before storing any (chip staging).
#define AR_SB 4
// s_waitcnt immediate for vmcnt(0) with expcnt and lgkmcnt left at their maxima (gfx9 encoding:
// vmcnt [3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8]).
#define AR_WAIT_VMCNT0 0x0F70

struct U8Point {
    int64_t n;       // chip area (power of two)
    int64_t sumT;    // sum of chip samples
    double cnorm;    // sqrt(sum (x - mean)^2), src/correlate.jl:440-452
    // screening constants: n, 128 n - sumT and n (128 sumT - 16384 n) as Float64 (exact
    // integers), 1 / (cnorm sqrt(n)) as Float32
    double nd, cd, dd;
    float k;
    double invn;     // 1 / n (exact: n is a power of two)
};

// Every quantity below is an integer below 2^53, so the Float64 forms are exact:
//   sum W*T = acc + 128 (sumT + s1) - 16384 n      (|.| < 2^31)
//   n * num = n * sum W*T - sumT * s1
//   n * var = n * s2 - s1^2                         (ZNCC; NCC uses n * s2)
__device__ inline double n_times_num(const U8Point& P, int32_t acc, int32_t s1) {
    int32_t swt = acc + 128 * ((int32_t)P.sumT + s1) - 16384 * (int32_t)P.n;
    return (double)swt * (double)P.n - (double)P.sumT * (double)s1;
}

__device__ inline double n_times_var(const U8Point& P, int32_t s1, int32_t s2, int measure) {
    double ns2 = (double)P.n * (double)s2;
    return measure == 1 ? ns2 : ns2 - (double)s1 * (double)s1;
}

// The reference's normalisation (src/correlate.jl:591-600) on the exact sums.
__device__ inline float exact_u8(const U8Point& P, int32_t acc, int32_t s1, int32_t s2,
                                 int measure) {
    const double invn = P.invn;                     // exact: n = 2^k
    double num = n_times_num(P, acc, s1) * invn;
    double den;
    if (measure == 1) {
        den = P.cnorm * sqrt((double)s2);
    } else {
        double wvar = n_times_var(P, s1, s2, 0) * invn;
        den = P.cnorm * sqrt(wvar);
    }
    return den > 0 ? (float)(num / den) : 0.0f;
}

// FP32 screening score (measure M: 0 ZNCC, 1 NCC): (n num) / (cnorm sqrt(n) sqrt(n var)),
// 0 for a flat window. n num = n acc + (128 n - sumT) s1 + n (128 sumT - 16384 n) and n var are
// the exact integers of n_times_num / n_times_var (every partial sum of the fused forms is an
// integer below 2^53). Two conversions to Float32, v_rsq_f32 (1 ulp), the constant (sq n
// converted to Float32, then v_rsq_f32) and two products: relative error below 8 * 2^-24 < 5e-7.
// A flat window has n var = 0 and n num = 0, so flooring n var at 1e-30 (every other n var is
// >= 1) makes its score 0 without a branch.
template <int M>
__device__ inline float approx_u8(const U8Point& P, int32_t acc, int32_t s1, int32_t s2) {
    const double ds1 = (double)s1;
    const double ns2 = P.nd * (double)s2;
    const double nv = M == 1 ? ns2 : fma(-ds1, ds1, ns2);
    const double nn = fma(P.cd, ds1, fma(P.nd, (double)acc, P.dd));
    return (float)nn * P.k * __builtin_amdgcn_rsqf(fmaxf((float)nv, 1e-30f));
}

// Sum W*T of one batch of 4 row-shift tiles (ti0 .. ti0 + 3) of column-shift tile tj on the int8
// matrix cores, for band height NRB 16-row blocks. K is walked by quads of chip columns: lane
// group g16 carries chip column v = 4 m + g16 (columns cw, cw + 1 are zero), all groups the same
// band block rb. The B fragment of (v, rb) -- 16 chip bytes from row 16 rb - l16 -- does not
// depend on the tile, and the A fragment of window rows 16 rho .. 16 rho + 15 serves every
// (tile t, rb) with t + rb = rho, so each m loads NRB B fragments and NRB + 3 A fragments for
// 4 NRB MFMAs.
template <int NRB>
__device__ inline void mfma_batch(const int8_t* Wl, const int8_t* Tp, int ldw, int cpad, int cw,
                                  int tj, int ti0, int l16, int g16, v4i (&acc)[4]) {
    const int nmq = (cw + 3) >> 2;
    int aoff = (16 * tj + l16 + g16) * ldw + 16 * ti0;
    int boff = g16 * cpad + 16 - l16;
    const uint32_t sb = (uint32_t)(boff & 3);   // the same for every m and rb
    for (int m = 0; m < nmq; m++) {
        // B fragments realigned from five dwords (a byte-misaligned ds_read_b128 costs more)
        v4i bf[NRB];
#pragma unroll
        for (int rb = 0; rb < NRB; rb++) {
            const uint32_t* bp = (const uint32_t*)(Tp + ((boff + 16 * rb) & ~3));
            uint32_t d0 = bp[0], d1 = bp[1], d2 = bp[2], d3 = bp[3], d4 = bp[4];
            bf[rb] = (v4i){(int)__builtin_amdgcn_alignbyte(d1, d0, sb),
                           (int)__builtin_amdgcn_alignbyte(d2, d1, sb),
                           (int)__builtin_amdgcn_alignbyte(d3, d2, sb),
                           (int)__builtin_amdgcn_alignbyte(d4, d3, sb)};
        }
        const int8_t* ap = Wl + aoff;
#pragma unroll
        for (int rho = 0; rho < NRB + 3; rho++) {
            const v4i af = *(const v4i*)(ap + 16 * rho);
#pragma unroll
            for (int rb = 0; rb < NRB; rb++) {
                const int t = rho - rb;
                if (t >= 0 && t < 4)
                    acc[t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af, bf[rb], acc[t], 0, 0, 0);
            }
        }
        aoff += 4 * ldw;
        boff += 4 * cpad;
    }
}

#define AR_NRB_MAX 9   // band heights with a dedicated MFMA loop (chips up to 130 rows)

// The chip energy sum (x - mean)^2 in the reference's order (src/correlate.jl:440-447: column by
// column, left to right in Float64) over the staged chip (byte u of column v at Tp[v cpad + 16 + u],
// value ^ 0x80), by one wave. Kept out of line: only chips whose integer energy reaches 2^53 / n^2
// take it, and it should cost the tracker's main loop no registers.
__device__ __noinline__ double u8_energy_in_order(const int8_t* Tp, int cpad, int ch, int cw,
                                                  int64_t sumT, double invn, int lane) {
    const double mean = (double)sumT * invn;   // exact: n is a power of two
    const int n = ch * cw;
    double acc = 0.0;
    for (int e0 = 0; e0 < n; e0 += 64) {
        const int e = e0 + lane;
        double t = 0.0;
        if (e < n) {
            const int v = e / ch, u = e - v * ch;
            const double d = (double)(uint8_t)(Tp[v * cpad + 16 + u] ^ 0x80) - mean;
            t = d * d;
        }
        acc = seq_sum64_ties(acc, t, lane);
    }
    return acc;
}

// Four workgroups per CU (128 VGPRs) when the LDS layout allows it.
// BATCH: many pairs in one launch. Point k of the launch is order[k]; its images come from its
// pair's descriptor. The single-pair instance is the same code without the indirection.
// OCC: waves per SIMD the registers are budgeted for (5: 96 VGPRs, five workgroups per CU at
// chip 32 r 25, some spills; 4: 128; 2: 256, for launches whose LDS allows at most two
// workgroups per CU, with no spills).
// SG: the box sums S1 / S2 in the workgroup's global scratch slot (a.sg) instead of LDS.
template <bool BATCH, int OCC, bool SG>
__global__ __launch_bounds__(AR_WG) __attribute__((amdgpu_waves_per_eu(OCC))) void track_u8_kernel(
    TrackArgs a0) {
    TrackArgs a = a0;
    extern __shared__ __align__(16) unsigned char smem[];
    const U8Layout L = u8_layout(a.max_ch, a.max_cw, a.max_rx, a.max_ry, a.box1 != nullptr, SG);
    // one set of reduction slots per call site (see block_sum2_u32 / block_best1)
    __shared__ Best red_screen[AR_NWAVES], red_exact[AR_NWAVES];
    __shared__ uint32_t red_chip[2 * AR_NWAVES];
    int8_t* Wl = (int8_t*)(smem + L.off_w);
    int8_t* Tp = (int8_t*)(smem + L.off_t);
    int32_t* S1 = SG ? a.sg + (int64_t)blockIdx.x * a.sg_stride : (int32_t*)(smem + L.off_s1);
    int32_t* S2 = SG ? S1 + a.sg_stride / 2 : (int32_t*)(smem + L.off_s2);
    int32_t* colS = (int32_t*)(smem + L.off_tmp);
    // the wave index in a scalar register: loops over a wave's share are scalar loops
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    // floor(t / (cpad / 4)) as a multiply-high by ceil(2^32 / (cpad / 4)): exact for every
    // t below 2^32 / (cpad / 4)
    const uint32_t nd_magic = 0xffffffffu / (uint32_t)(L.cpad >> 2) + 1u;
    // chip band rows and trailing columns are zero; the fetched chip path writes only chip rows
    for (int t = threadIdx.x; t < (L.cols_t * L.cpad) >> 2; t += AR_WG) ((uint32_t*)Tp)[t] = 0;
    __syncthreads();
    const PointRange pr = point_range(a0.n);
    for (int64_t k = pr.first; k < pr.end; k += pr.step) {
        // The lane index hidden from the optimiser per point: addresses derived from it are
        // formed again for each point instead of being held across the loop and spilled (a
        // reload is a vector memory access, and waiting for it waits for the window DMA too).
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int lane = tid & 63;
        const int64_t p = BATCH ? a0.order[k] : k;
        // The argument block (and a batched point's pair) read again at each phase: a field is
        // then live within its phase only, which keeps the scalar registers from spilling.
        auto refresh = [&]() {
            a = reload_args();
            if (BATCH) use_pair(a, p);
        };
        refresh();
        const uint8_t* ref = (const uint8_t*)a.ref;
        const uint8_t* sec = (const uint8_t*)a.sec;
        Geom g = load_geom(a, p);
        if (!g.status) { write_skip(a, p, false); continue; }
        if (!any_valid(a, g)) { write_skip(a, p, false); continue; }
        const int ch = g.ch, cw = g.cw, NR = 2 * g.ry, NC = 2 * g.rx;
        const int winh = ch + NR - 1, winw = cw + NC - 1;
        const int nti = (NR + 15) / 16, ntj = (NC + 15) / 16;
        const int nrb = (ch + 15 + 15) / 16;          // band rows r - i in [0, ch), i < 16
        const int rows_used = 16 * (nti + nrb);

        // ---- S1 / S2 of every shift from the pass's box images (exact integers), loaded straight
        // into LDS (global_load_lds: no registers; the barrier of the chip sums below waits for
        // them). Wave w fills elements 64 (w + 4 k) + lane of the column-major arrays.
        // Addresses are a per-point base (scalar) plus a 32-bit byte offset per lane; the
        // dispatcher keeps every such offset below 2^32. Wave w takes shift columns w, w + 4, ...
        // with lane = shift row (in chunks of 64 rows): the lane offsets never change, the
        // column base moves in scalar registers.
        if (a.box1) {
            const int64_t q0 = (g.wr0 - 1) + (g.wc0 - 1) * a.box_ld;
            const char* b1 = uniform_ptr((const char*)(a.box1 + q0));
            const char* b2 = uniform_ptr((const char*)(a.box2 + q0));
            const uint32_t ldb = 4u * (uint32_t)a.box_ld;   // bytes between box-image columns
            for (int i0 = 0; i0 < NR; i0 += 64) {
                if (i0 + lane < NR) {
                    const uint32_t lo = 4u * (uint32_t)(i0 + lane);
                    for (int j = wave; j < NC; j += AR_NWAVES) {
                        const uint32_t co = (uint32_t)j * ldb;
                        __builtin_amdgcn_global_load_lds(b1 + co + lo, S1 + j * NR + i0, 4, 0, 0);
                        __builtin_amdgcn_global_load_lds(b2 + co + lo, S2 + j * NR + i0, 4, 0, 0);
                    }
                }
            }
        }
        // ---- stage window (int8, value ^ 0x80; virtual zero padding -> 0x80) and chip
        {
            // window origin in unpadded 0-based coordinates; interior windows skip the
            // per-pixel padding test and keep several independent loads in flight
            int64_t orow = g.wr0 - a.pady - 1, ocol = g.wc0 - a.padx - 1;
            bool inside = orow >= 0 && ocol >= 0 && orow + winh <= a.ny && ocol + winw <= a.nx;
            if (inside && a.refx) {
                // Straight into LDS (global_load_lds, a dword per lane): the window's columns
                // are one contiguous LDS region of ldw-byte columns, filled in dwords
                // d = 64 (wave + 4 k) + lane, column d / (ldw / 4). Each column reads from the
                // int8 copy of the reference in which its first row starts a dword. Rows past
                // the window bring image bytes that meet zero chip rows or dropped shifts.
                // column c of copy sh starts at base + sh (stride + 1) + c ny
                const int64_t off0 = orow + ocol * a.ny;
                const uint8_t* base = uniform_ptr(a.refx + off0);
                const uint32_t ny = (uint32_t)a.ny, st1 = (uint32_t)a.refx_stride + 1u;
                const int cq = L.ldw >> 2, tot = winw * cq, dc = 256 / cq, dr = 256 - dc * cq;
                const int c = (64 * wave + lane) / cq;
                int r4 = 64 * wave + lane - c * cq;
                uint32_t coff = (uint32_t)c * ny;
                const uint32_t dco = (uint32_t)dc * ny, lo0 = (uint32_t)off0;
                const int kc = 64 / cq;   // whole window columns one DMA instruction covers
                if ((ny & 3u) == 0 && kc >= 1) {
                    // Every column starts at the window's alignment, so one copy serves them all;
                    // each instruction fills kc whole columns (lane = column cin, dword r4 of
                    // it): the lane offsets never change, the column base moves in scalar
                    // registers.
                    const uint8_t* bs = uniform_ptr(base + (uint64_t)((0u - lo0) & 3u) * st1);
                    const int cin = lane / cq, r4l = lane - cin * cq;
                    if (cin < kc) {
                        const uint32_t lo = (uint32_t)cin * ny + 4u * (uint32_t)r4l;
                        for (int c0 = wave * kc; c0 < winw; c0 += AR_NWAVES * kc)
                            if (c0 + cin < winw)
                                __builtin_amdgcn_global_load_lds(bs + (uint64_t)c0 * ny + lo,
                                                                 Wl + c0 * L.ldw, 4, 0, 0);
                    }
                } else if ((ny & 3u) == 0) {
                    // columns longer than 64 dwords: dwords d = 64 (wave + 4 k) + lane in order
                    const uint8_t* bs = uniform_ptr(base + (uint64_t)((0u - lo0) & 3u) * st1);
                    uint32_t off = coff + 4u * (uint32_t)r4;
                    const uint32_t doff = dco + 4u * (uint32_t)dr, wrap = ny - 4u * (uint32_t)cq;
                    for (int d0 = 64 * wave; d0 < tot; d0 += 256) {
                        if (d0 + lane < tot) __builtin_amdgcn_global_load_lds(bs + off, Wl + 4 * d0, 4, 0, 0);
                        r4 += dr;
                        off += doff;
                        if (r4 >= cq) {
                            r4 -= cq;
                            off += wrap;
                        }
                    }
                } else {
                    for (int d0 = 64 * wave; d0 < tot; d0 += 256) {
                        if (d0 + lane < tot) {
                            const uint32_t sh = (0u - (lo0 + coff)) & 3u;
                            __builtin_amdgcn_global_load_lds(
                                base + (sh * st1 + coff + 4u * (uint32_t)r4), Wl + 4 * d0, 4, 0, 0);
                        }
                        r4 += dr;
                        coff += dco;
                        if (r4 >= cq) {
                            r4 -= cq;
                            coff += ny;
                        }
                    }
                }
            } else {
                for (int c = wave; c < winw; c += AR_NWAVES)
                    for (int r = lane; r < rows_used; r += 64) {
                        int8_t v = 0;
                        if (r < winh) {
                            int64_t k = pix(a, g.wr0 + r, g.wc0 + c);
                            v = (int8_t)((k >= 0 ? ref[k] : 0) ^ 0x80);
                        }
                        Wl[c * L.ldw + r] = v;
                    }
            }
        }
        int64_t st = 0, st2 = 0;
        if (L.acc_cols_t > 0 && (ch & 3) == 0) {
            // The last point's accumulators ran over the first chip columns: their band rows
            // (and whole trailing columns) are zero again before the chip is staged. The paths
            // that stage whole columns write the same zeros; the fetched path writes chip rows
            // only.
            const int nd = L.cpad >> 2, c0 = ch >> 2;   // dwords per column; rows 16 + u: u < ch
            for (int t = tid; t < L.acc_cols_t * nd; t += AR_WG) {
                const int v = (int)__umulhi((uint32_t)t, nd_magic), d = t - v * nd;
                if (v >= cw || d < 4 || d >= 4 + c0) ((uint32_t*)Tp)[t] = 0u;
            }
        }
        {
            int64_t orow = g.cr0 - a.pady - 1, ocol = g.cc0 - a.padx - 1;
            bool inside = orow >= 0 && ocol >= 0 && orow + ch <= a.ny && ocol + cw <= a.nx;
            if (inside && (ch & 3) == 0 && a.box1 && a.secx) {
                // One chip size in the pass, so the zero band rows and trailing columns written
                // at kernel start stay valid: only the chip's rows are fetched, straight into LDS
                // (a dword per lane from the int8 copy of the secondary in which the column
                // starts a dword); wave w takes columns w, w + 4, ... and, once its loads have
                // landed, sums x and x^2 over them.
                const int64_t off0 = orow + ocol * a.ny;
                const uint8_t* base = uniform_ptr(a.secx + off0);
                const uint32_t ny = (uint32_t)a.ny, st1 = (uint32_t)a.refx_stride + 1u;
                const uint32_t lo0 = (uint32_t)off0;
                const int cd = ch >> 2;
                for (int v = wave; v < cw; v += AR_NWAVES) {
                    const uint32_t coff = (uint32_t)v * ny;
                    const uint32_t sh = (0u - (lo0 + coff)) & 3u;
                    for (int x0 = 0; x0 < cd; x0 += 64)
                        if (x0 + lane < cd)
                            __builtin_amdgcn_global_load_lds(
                                base + (sh * st1 + coff + 4u * (uint32_t)(x0 + lane)),
                                Tp + v * L.cpad + 16 + 4 * x0, 4, 0, 0);
                }
                __builtin_amdgcn_s_waitcnt(AR_WAIT_VMCNT0);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);   // keep the LDS reads below the wait
                uint32_t acc1 = 0, acc2 = 0;
                for (int v = wave; v < cw; v += AR_NWAVES)
                    for (int x = lane; x < cd; x += 64) {
                        const uint32_t raw =
                            *(const uint32_t*)(Tp + v * L.cpad + 16 + 4 * x) ^ 0x80808080u;
                        acc1 = __builtin_amdgcn_sad_u8(raw, 0u, acc1);
                        acc2 = __builtin_amdgcn_udot4(raw, raw, acc2, false);
                    }
                st = acc1;
                st2 = acc2;
            } else if (inside && (ch & 3) == 0) {
                // dword path: chip rows u = x - 16 for x in [16, 16 + ch), realigned as above;
                // the zero band rows and the two trailing zero columns are written as zeros.
                // Half-waves take columns hw, hw + 8, ...; loads of AR_SB columns go first.
                const uint8_t* gb = sec + orow + ocol * a.ny;
                const int nx4 = L.cpad >> 2;
                const int hw = tid >> 5, l32 = lane & 31;
                uint32_t acc1 = 0, acc2 = 0;
                for (int x4 = l32; x4 < nx4; x4 += 32) {
                    const int u = 4 * x4 - 16;
                    const bool rowok = u >= 0 && u < ch;
                    for (int v0 = hw; v0 < cw + 2; v0 += 8 * AR_SB) {
                        uint32_t lo[AR_SB], hi[AR_SB], sh[AR_SB];
#pragma unroll
                        for (int k = 0; k < AR_SB; k++) {
                            const int v = v0 + 8 * k;
                            const uintptr_t ad = (uintptr_t)(gb + (int64_t)v * a.ny + u);
                            sh[k] = (uint32_t)(ad & 3);
                            lo[k] = 0;
                            hi[k] = 0;
                            if (rowok && v < cw) {
                                const uint32_t* al = (const uint32_t*)(ad & ~(uintptr_t)3);
                                lo[k] = al[0];
                                hi[k] = al[1];
                            }
                        }
#pragma unroll
                        for (int k = 0; k < AR_SB; k++) {
                            const int v = v0 + 8 * k;
                            if (v < cw + 2) {
                                uint32_t w = 0;
                                if (rowok && v < cw) {
                                    uint32_t raw = __builtin_amdgcn_alignbyte(hi[k], lo[k], sh[k]);
                                    acc1 = __builtin_amdgcn_sad_u8(raw, 0u, acc1);
                                    acc2 = __builtin_amdgcn_udot4(raw, raw, acc2, false);
                                    w = raw ^ 0x80808080u;
                                }
                                *(uint32_t*)(Tp + v * L.cpad + 4 * x4) = w;
                            }
                        }
                    }
                }
                st = acc1;
                st2 = acc2;
            } else {
                for (int v = wave; v < cw + 2; v += AR_NWAVES)   // two trailing zero columns
                    for (int x = lane; x < L.cpad; x += 64) {
                        int u = x - 16;
                        int8_t b = 0;
                        if (u >= 0 && u < ch && v < cw) {
                            int64_t k = pix(a, g.cr0 + u, g.cc0 + v);
                            uint8_t raw = k >= 0 ? sec[k] : 0;
                            st += raw;
                            st2 += (int64_t)raw * raw;
                            b = (int8_t)(raw ^ 0x80);
                        }
                        Tp[v * L.cpad + x] = b;
                    }
            }
        }
        U8Point P;
        P.n = (int64_t)ch * cw;
        // chip energy, exact in int64: sum (n x - sumT)^2 = n^2 sum x^2 - n sumT^2, over n^2
        // (sum x and sum x^2 of at most 16384 UInt8 samples fit in uint32)
        uint32_t sum1, sum2;
        block_sum2_u32((uint32_t)st, (uint32_t)st2, red_chip, sum1, sum2);
        P.sumT = sum1;
        const int64_t sumx2 = sum2;
        const int64_t sumk2 = P.n * P.n * sumx2 - P.n * P.sumT * P.sumT;
        // n = 2^lg: 1 / n and the division by n^2 are exact scalings
        P.invn = ldexp(1.0, -(int)__builtin_ctzll((unsigned long long)P.n));
        double sq = (double)sumk2 * P.invn * P.invn;
        if (sumk2 >= ((int64_t)1 << 53)) {
            // The reference sums (x - mean)^2 left to right (src/correlate.jl:440-447). Each term is
            // a multiple of 1/n^2 and the partial sums grow, so the sum is exact, and equal to the
            // integer form above, while it stays below 2^53 / n^2; past that (128 x 128 chips whose
            // variance exceeds 2048) its additions round: one wave replays them.
            __syncthreads();   // every wave has read the chip sums out of red_chip
            if (wave == 0) {
                const double acc = u8_energy_in_order(Tp, L.cpad, ch, cw, P.sumT, P.invn, lane);
                if (lane == 0) {
                    const uint64_t b = (uint64_t)__double_as_longlong(acc);
                    red_chip[0] = (uint32_t)b;
                    red_chip[1] = (uint32_t)(b >> 32);
                }
            }
            __syncthreads();
            sq = __longlong_as_double(
                (long long)(((uint64_t)red_chip[1] << 32) | (uint64_t)red_chip[0]));
            __syncthreads();   // red_chip is the next point's reduction slot
        }
        bool ok = sq > 2.220446049250313e-16 * (double)ch * (double)cw;
        if (!ok) { write_skip(a, p, true); __syncthreads(); continue; }
        P.cnorm = sqrt(sq);
        P.nd = (double)P.n;
        P.cd = (double)(128 * P.n - P.sumT);
        P.dd = (double)(P.n * (128 * P.sumT - 16384 * P.n));
        P.k = __builtin_amdgcn_rsqf((float)(sq * P.nd));   // 1 / (cnorm sqrt(n)), sq n exact

        // ---- integer box sums S1 (W) and S2 (W^2) over every shift (unless read from the box
        // images above): sliding sums, with each window column split into KC row chunks and each
        // shift row into KR column chunks; W first, then W^2, through one buffer of column sums
        if (!a.box1) {
            constexpr int KC = 3, KR = 5;
            int32_t* cs = colS;
            const int lc = (NR + KC - 1) / KC, lr = (NC + KR - 1) / KR;
#pragma unroll 1
            for (int sq = 0; sq < 2; sq++) {
                for (int t = tid; t < winw * KC; t += AR_WG) {
                    int c = t / KC, k = t - c * KC;
                    int r0 = k * lc, r1 = r0 + lc < NR ? r0 + lc : NR;
                    if (r0 >= r1) continue;
                    const uint8_t* col = (const uint8_t*)(Wl + c * L.ldw);
                    int32_t acc = 0;
                    for (int u = 0; u < ch; u++) {
                        int32_t w = col[r0 + u] ^ 0x80;
                        acc += sq ? w * w : w;
                    }
                    for (int r = r0; r < r1; r++) {
                        cs[c * NR + r] = acc;
                        int32_t wo = col[r] ^ 0x80, wi = col[r + ch] ^ 0x80;
                        acc += sq ? wi * wi - wo * wo : wi - wo;
                    }
                }
                __syncthreads();
                int32_t* Sx = sq ? S2 : S1;
                for (int t = tid; t < NR * KR; t += AR_WG) {
                    int i = t / KR, k = t - i * KR;
                    int j0 = k * lr, j1 = j0 + lr < NC ? j0 + lr : NC;
                    if (j0 >= j1) continue;
                    int32_t acc = 0;
                    for (int v = 0; v < cw; v++) acc += cs[(j0 + v) * NR + i];
                    for (int j = j0; j < j1; j++) {
                        Sx[j * NR + i] = acc;
                        if (j + 1 < j1) acc += cs[(j + cw) * NR + i] - cs[j * NR + i];
                    }
                }
                __syncthreads();
            }
        }

        refresh();
        // ---- numerator on int8 MFMA, stored per shift (column-major o = j NR + i)
        int32_t* accL = (int32_t*)(smem + L.off_acc);   // int32 [NR*NC]
        const int g16 = lane >> 4, l16 = lane & 15;
        // the shifts of one tile batch, from a wave's accumulators
        auto store_batch = [&](const v4i (&acc)[4], int tj, int ti0) {
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int i = 16 * (ti0 + t) + l16;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int j = 16 * tj + 4 * g16 + r;
                    if (i < NR && j < NC) accL[j * NR + i] = acc[t][r];   // i < NR: tile kept
                }
            }
        };
        if (nrb <= AR_NRB_MAX) {
            // batches of 4 row-shift tiles of one column-shift tile; tiles past nti in the last
            // batch read finite LDS data and are dropped. With the accumulators in the window's
            // bytes, a wave has at most one batch and stores it once every wave is past the
            // window.
            const int nbt = (nti + 3) >> 2, nbatch = ntj * nbt;
            v4i acc[4];
            int btj = 0, bti0 = 0;
            for (int bt = wave; bt < nbatch; bt += AR_NWAVES) {
                const int tj = bt / nbt, ti0 = 4 * (bt - tj * nbt);
                btj = tj;
                bti0 = ti0;
#pragma unroll
                for (int k = 0; k < 4; k++) acc[k] = (v4i){0, 0, 0, 0};
                switch (nrb) {
                    case 2: mfma_batch<2>(Wl, Tp, L.ldw, L.cpad, cw, tj, ti0, l16, g16, acc); break;
                    case 3: mfma_batch<3>(Wl, Tp, L.ldw, L.cpad, cw, tj, ti0, l16, g16, acc); break;
                    case 4: mfma_batch<4>(Wl, Tp, L.ldw, L.cpad, cw, tj, ti0, l16, g16, acc); break;
                    case 5: mfma_batch<5>(Wl, Tp, L.ldw, L.cpad, cw, tj, ti0, l16, g16, acc); break;
                    case 6: mfma_batch<6>(Wl, Tp, L.ldw, L.cpad, cw, tj, ti0, l16, g16, acc); break;
                    case 7: mfma_batch<7>(Wl, Tp, L.ldw, L.cpad, cw, tj, ti0, l16, g16, acc); break;
                    case 8: mfma_batch<8>(Wl, Tp, L.ldw, L.cpad, cw, tj, ti0, l16, g16, acc); break;
                    default: mfma_batch<9>(Wl, Tp, L.ldw, L.cpad, cw, tj, ti0, l16, g16, acc); break;
                }
                if (!L.acc_in_w) store_batch(acc, tj, ti0);
            }
            if (L.acc_in_w) {
                __syncthreads();
                if (wave < nbatch) store_batch(acc, btj, bti0);
            }
        } else {
            const int ntiles = nti * ntj;
            const int nq = cw * nrb;                       // 16-wide k-blocks
            const int nm = (nq + 3) / 4;                   // MFMAs per tile
            // Taller bands: K is walked in 16-wide blocks q = 4m + g16 = (chip column v, band row
            // block rb). Blocks past the last (nq) read the zero chip columns cw, cw + 1, so the
            // loop has no branches. Each wave takes batches of 4 tiles (t = tb + 4k); a batch past
            // the last tile repeats the last tile and its results are dropped.
            for (int tb = wave; tb < ntiles; tb += 16) {
                int tt[4], ti[4], tj[4], abase[4];
    #pragma unroll
                for (int k = 0; k < 4; k++) {
                    tt[k] = tb + 4 * k;
                    int t = tt[k] < ntiles ? tt[k] : ntiles - 1;
                    ti[k] = t % nti;
                    tj[k] = t / nti;
                    abase[k] = (16 * tj[k] + l16) * L.ldw + 16 * ti[k];
                }
                v4i acc[4];
    #pragma unroll
                for (int k = 0; k < 4; k++) acc[k] = (v4i){0, 0, 0, 0};
                int v = g16 / nrb, rb = g16 - v * nrb;
                for (int m = 0; m < nm; m++) {
                    // B: 16 bytes of the zero-padded chip column v from row 16 rb - l16, realigned
                    // from five dwords
                    int off = v * L.cpad + 16 + 16 * rb - l16;
                    const uint32_t* bp = (const uint32_t*)(Tp + (off & ~3));
                    uint32_t d0 = bp[0], d1 = bp[1], d2 = bp[2], d3 = bp[3], d4 = bp[4];
                    uint32_t sb = (uint32_t)(off & 3);
                    v4i bf = (v4i){(int)__builtin_amdgcn_alignbyte(d1, d0, sb),
                                   (int)__builtin_amdgcn_alignbyte(d2, d1, sb),
                                   (int)__builtin_amdgcn_alignbyte(d3, d2, sb),
                                   (int)__builtin_amdgcn_alignbyte(d4, d3, sb)};
                    const int aoff = v * L.ldw + 16 * rb;
                    v4i af[4];
    #pragma unroll
                    for (int k = 0; k < 4; k++) af[k] = *(const v4i*)(Wl + abase[k] + aoff);
    #pragma unroll
                    for (int k = 0; k < 4; k++)
                        acc[k] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[k], bf, acc[k], 0, 0, 0);
                    rb += 4;
                    if (rb >= nrb) { rb -= nrb; v++; }
                    if (rb >= nrb) { rb -= nrb; v++; }
                }
                for (int k = 0; k < 4; k++) {
                    if (tt[k] >= ntiles) break;
                    int i = 16 * ti[k] + l16;
                    for (int r = 0; r < 4; r++) {
                        int j = 16 * tj[k] + 4 * g16 + r;
                        if (i < NR && j < NC) {
                            int o = j * NR + i;
                            accL[o] = acc[k][r];
                        }
                    }
                }
            }
        }
        __syncthreads();

        refresh();
        // ---- integer peak: screen every shift in FP32 (all lanes busy, the matrix-core tiles'
        // padding left out), then exact values for the contenders. The screening best of a lane
        // is the first maximum it meets, indexed by o: ties are settled by the exact values. It also
        // keeps its second-largest score, so that a lane with one contender needs no rescan.
        const int nsh = NR * NC;
        Best lb = best_empty();
        float lb2 = -INFINITY;
        // (lb2 <= lb.v throughout, so the median of the two and the new score is the new
        // second-largest; scores are never NaN)
        auto screen = [&](auto mc) {
#pragma unroll 2
            for (int o = tid; o < nsh; o += AR_WG) {
                const float sv = approx_u8<decltype(mc)::value>(P, accL[o], S1[o], S2[o]);
                const bool up = sv > lb.v;
                lb2 = __builtin_amdgcn_fmed3f(lb.v, lb2, sv);
                lb.v = up ? sv : lb.v;
                lb.idx = up ? o : lb.idx;
            }
        };
        if (a.measure == 1) screen(std::integral_constant<int, 1>());
        else screen(std::integral_constant<int, 0>());
        const Best b = block_best1<true>(lb, red_screen);
        const float thr = b.v - 2.0f * AR_SCREEN_EPS;
        // the patch around the screening peak, evaluated while the contenders are
        int i0 = 1, j0 = 1;
        const int bo = b.idx < nsh ? b.idx : 0;
        const bool srefine = a.up > 1 && patch_origin(bo % NR + 1, bo / NR + 1, NR, NC, i0, j0);
        float pv = 0.0f;
        if (srefine && tid < 25) {
            const int o = (j0 - 1 + tid % 5) * NR + (i0 - 1 + tid / 5);
            pv = exact_u8(P, accL[o], S1[o], S2[o], a.measure);
        }
        // contenders: only lanes whose own best reaches the threshold hold any; they score their
        // shifts again (the same function of the same integers)
        Best e2 = best_empty();
        auto contend = [&](auto mc) {
            for (int o = tid; o < nsh; o += AR_WG)
                if (approx_u8<decltype(mc)::value>(P, accL[o], S1[o], S2[o]) >= thr) {
                    const int j = o / NR, i = o - j * NR;
                    const float v = exact_u8(P, accL[o], S1[o], S2[o], a.measure);
                    if (better(v, i * NC + j, e2)) { e2.v = v; e2.idx = i * NC + j; }
                }
        };
        if (lb.v >= thr) {
            if (lb2 < thr) {
                // the lane's only contender is its best shift
                const int o = lb.idx, j = o / NR, i = o - j * NR;
                e2.v = exact_u8(P, accL[o], S1[o], S2[o], a.measure);
                e2.idx = i * NC + j;
            } else if (a.measure == 1) {
                contend(std::integral_constant<int, 1>());
            } else {
                contend(std::integral_constant<int, 0>());
            }
        }
        e2 = block_best1<false>(e2, red_exact);
        const int pi = e2.idx / NC + 1, pj = e2.idx % NC + 1;
        const float pval = e2.v;
        const int si0 = i0, sj0 = j0;
        const bool refine = a.up > 1 && patch_origin(pi, pj, NR, NC, i0, j0);
        if (refine && (i0 != si0 || j0 != sj0) && tid < 25) {
            const int o = (j0 - 1 + tid % 5) * NR + (i0 - 1 + tid / 5);
            pv = exact_u8(P, accL[o], S1[o], S2[o], a.measure);
        }
        refresh();
        finish_point(a, p, g, pi, pj, pval, i0, j0, refine, pv);
        __syncthreads();
    }
}

template __global__ void track_u8_kernel<false, 5, false>(TrackArgs);
template __global__ void track_u8_kernel<true, 4, false>(TrackArgs);
template __global__ void track_u8_kernel<true, 2, false>(TrackArgs);
template __global__ void track_u8_kernel<true, 4, true>(TrackArgs);
template __global__ void track_u8_kernel<true, 2, true>(TrackArgs);

// ---------------------------------------------------------------------------------------------
// generic path (FP64, reference order)
// ---------------------------------------------------------------------------------------------

__device__ inline float load_pix_f(const TrackArgs& a, const void* img, int is_u8, int64_t r,
                                   int64_t c) {
    int64_t k = pix(a, r, c);
    if (k < 0) return 0.0f;
    return is_u8 ? (float)((const uint8_t*)img)[k] : ((const float*)img)[k];
}

// _rotate_chip (src/track.jl:382-409): chip0 (ch x cw, column-major) rotated about its centre,
// bilinear in Float64, Float32 out, zero where a source corner leaves the chip; s, c are the sine
// and cosine of the angle.
__device__ inline void rotate_chip(const float* chip0, float* out, int ch, int cw, double s,
                                   double c) {
    const double yc = (double)(ch + 1) / 2.0, xc = (double)(cw + 1) / 2.0;
    for (int e = threadIdx.x; e < ch * cw; e += AR_WG) {
        const int i = e % ch + 1, j = e / ch + 1;
        const double dyy = (double)i - yc, dxx = (double)j - xc;
        const double sy = yc + c * dyy + s * dxx;
        const double sx = xc - s * dyy + c * dxx;
        const int64_t i0 = jfloor(sy), j0 = jfloor(sx);
        if (i0 < 1 || j0 < 1 || i0 + 1 > ch || j0 + 1 > cw) {
            out[e] = 0.0f;
            continue;
        }
        const double fy = sy - (double)i0, fx = sx - (double)j0;
        const double v00 = (double)chip0[(i0 - 1) + (j0 - 1) * ch];
        const double v10 = (double)chip0[i0 + (j0 - 1) * ch];
        const double v01 = (double)chip0[(i0 - 1) + j0 * ch];
        const double v11 = (double)chip0[i0 + j0 * ch];
        out[e] = (float)((1 - fy) * ((1 - fx) * v00 + fx * v01) + fy * ((1 - fx) * v10 + fx * v11));
    }
}

// Julia's maximum over a Float32 surface: NaN-propagating.
__device__ inline float nan_max(float m, float v) { return (v != v || v > m) ? v : m; }

template <int IS_U8>
__global__ __launch_bounds__(AR_WG) void track_generic_kernel(TrackArgs a) {
    __shared__ Best red[AR_NWAVES];
    __shared__ double shd[4];
    __shared__ float redm[AR_NWAVES];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const GenScratch G = gen_scratch(a.max_ch, a.max_cw, a.max_rx, a.max_ry);
    unsigned char* base = a.scratch + (size_t)blockIdx.x * a.scratch_per_block;
    float* chip = (float*)(base + G.off_chip);
    float* win = (float*)(base + G.off_win);
    double* S = (double*)(base + G.off_s);
    double* S2 = (double*)(base + G.off_s2);
    float* surf = (float*)(base + G.off_surf);
    float* chip0 = (float*)(base + G.off_chip0);
    float* bsurf = (float*)(base + G.off_best);

    const PointRange pr = point_range(a.n);
    for (int64_t p = pr.first; p < pr.end; p += pr.step) {
        Geom g = point_geom(a, p);
        if (!g.status) { write_skip(a, p, false); continue; }
        if (!any_valid(a, g)) { write_skip(a, p, false); continue; }
        const int ch = g.ch, cw = g.cw, NR = 2 * g.ry, NC = 2 * g.rx;
        const int winh = ch + NR - 1, winw = cw + NC - 1, ls = winh + 1;
        for (int c = wave; c < winw; c += AR_NWAVES)
            for (int r = lane; r < winh; r += 64)
                win[c * winh + r] = load_pix_f(a, a.ref, IS_U8, g.wr0 + r, g.wc0 + c);
        for (int v = wave; v < cw; v += AR_NWAVES)
            for (int u = lane; u < ch; u += 64)
                chip0[v * ch + u] = load_pix_f(a, a.sec, IS_U8, g.cr0 + u, g.cc0 + v);
        __syncthreads();
        // integral_both!: column running sums, then the row recurrence, in the reference order
        // (the window is the same for every chip rotation)
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

        // one correlation, or one per chip rotation keeping the surface with the highest
        // maximum (_correlate_rotations!, src/track.jl:316-351: an earlier angle wins a tie;
        // every angle degenerate makes the point degenerate)
        const int nang = a.nrot > 0 ? a.nrot : 1;
        float best = -INFINITY;
        bool found = false;
        for (int k = 0; k < nang; k++) {
            if (a.nrot > 0 && !((a.rot_zero >> k) & 1u))
                rotate_chip(chip0, chip, ch, cw, a.rot_s[k], a.rot_c[k]);
            else
                for (int e = tid; e < ch * cw; e += AR_WG) chip[e] = chip0[e];
            __syncthreads();
            // prepare_chip!: sequential Float64 mean and energy in column-major order
            if (tid == 0) {
                double s = 0.0;
                for (int q = 0; q < ch * cw; q++) s += (double)chip[q];
                double mean = s / (double)(ch * cw);
                double sq = 0.0;
                for (int q = 0; q < ch * cw; q++) {
                    double d = (double)chip[q] - mean;
                    sq += d * d;
                }
                shd[0] = mean;
                shd[1] = sq;
            }
            __syncthreads();
            const double mean = shd[0], sq = shd[1];
            const bool ok = sq > 2.220446049250313e-16 * (double)ch * (double)cw;
            __syncthreads();
            if (!ok) continue;
            const double cnorm = sqrt(sq);
            for (int q = tid; q < ch * cw; q += AR_WG) chip[q] = (float)((double)chip[q] - mean);
            __syncthreads();
            // direct numerator (src/correlate.jl:799-814) and the normalisation, per shift
            const double n = (double)(ch * cw);
            float m = -INFINITY;
            for (int j = wave; j < NC; j += AR_NWAVES)
            for (int i = lane; i < NR; i += 64) {
                int o = j * NR + i;
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
                const float v = den > 0 ? (float)(acc / den) : 0.0f;
                surf[o] = v;
                m = nan_max(m, v);
            }
            if (a.nrot == 0) {
                found = true;
                break;
            }
            // the surface's maximum, and the best surface so far
            for (int o = 32; o > 0; o >>= 1) m = nan_max(m, __shfl_xor(m, o));
            __syncthreads();
            if (lane == 0) redm[wave] = m;
            __syncthreads();
            float pk = redm[0];
            for (int w = 1; w < AR_NWAVES; w++) pk = nan_max(pk, redm[w]);
            if (pk > best) {
                best = pk;
                found = true;
                for (int o = tid; o < NR * NC; o += AR_WG) bsurf[o] = surf[o];
            }
            __syncthreads();
        }
        if (!found) { write_skip(a, p, true); __syncthreads(); continue; }
        const float* fs = a.nrot > 0 ? bsurf : surf;
        __syncthreads();
        Best b = best_empty();
        for (int j = wave; j < NC; j += AR_NWAVES)
            for (int i = lane; i < NR; i += 64) {
                float v = fs[j * NR + i];
                int idx = i * NC + j;
                if (better(v, idx, b)) { b.v = v; b.idx = idx; }
            }
        b = block_best(b, red);
        if (b.idx == 0x7fffffff) b.idx = 0;   // all NaN: (1, 1)
        int pi = b.idx / NC + 1, pj = b.idx % NC + 1;
        float pval = fs[(pj - 1) * NR + (pi - 1)];
        int i0 = 1, j0 = 1;
        bool refine = a.up > 1 && patch_origin(pi, pj, NR, NC, i0, j0);
        float pv = 0.0f;
        if (refine && tid < 25) pv = fs[(j0 - 1 + tid % 5) * NR + (i0 - 1 + tid / 5)];
        finish_point(a, p, g, pi, pj, pval, i0, j0, refine, pv);
        __syncthreads();
    }
}

template __global__ void track_generic_kernel<0>(TrackArgs);
template __global__ void track_generic_kernel<1>(TrackArgs);

}  // namespace ar
