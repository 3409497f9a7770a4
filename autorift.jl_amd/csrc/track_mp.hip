// Mixed-precision tracker for Float32 prepared images (quantize = :none).
//
// The correlation contraction sum(chip' * W) runs on the matrix cores with BF16 or FP16 operands
// and FP32 accumulation (v_mfma_f32_16x16x32_{bf16,f16}); everything around it keeps the
// reference's precision:
//   * the chip mean and energy are Float64 sums and the chip is Float32(x - mean)
//     (src/correlate.jl:423-453) before it is rounded to the 16-bit type;
//   * the box sums S1 = sum W and S2 = sum W^2 come from Float64 box images of the pass
//     (box_f64_colpass_kernel: direct column sums; box_f64_rowpass_kernel: sliding row sums
//     restarted every 64 columns) and the window variance s2 - s1^2 / n is formed in Float64;
//     the screening score of each shift divides the FP32 numerator by cnorm sqrt(variance) in
//     FP32, and the values reported (peak and 5x5 patch) use the reference's Float64
//     normalisation (src/correlate.jl:575-630);
//   * the integer peak's value and the 5x5 patch the subpixel fit starts from are recomputed
//     with the Float64 direct numerator (src/correlate.jl:799-814), so the pyrUp cascade
//     (src/peak.jl:456-522) starts from full-precision samples.
// The Float64 sums (chip mean and energy, the 26 exact numerators) are block reductions in a
// fixed order -- deterministic, but not the reference's left-to-right order, so a value can
// differ from the exact path's in its last Float32 bit when it lies within ~1e-16 of a rounding
// boundary. The argmax over the low-precision surface follows peak_index's rule
// (src/peak.jl:32-61).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "ar_internal.h"
#include "layout.h"
#include "track_common.h"

namespace ar {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

// Operand type traits: Float32 -> 16-bit bits (round to nearest even) and the MFMA.
template <int PREC> struct MpType;
template <> struct MpType<AR_PREC_BF16> {
    static __device__ inline uint16_t bits(float v) {
        return __builtin_bit_cast(uint16_t, (__bf16)v);
    }
    static __device__ inline v4f mfma(v4i a, v4i b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a),
                                                      __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
    }
};
template <> struct MpType<AR_PREC_FP16> {
    static __device__ inline uint16_t bits(float v) {
        return __builtin_bit_cast(uint16_t, (_Float16)v);
    }
    static __device__ inline v4f mfma(v4i a, v4i b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(v8h, a),
                                                     __builtin_bit_cast(v8h, b), c, 0, 0, 0);
    }
};


// sum chip' * W of one batch of 4 row-shift tiles (ti0 .. ti0 + 3) of column-shift tile tj.
// Output rows are column shifts j, output columns row shifts i; K = 32 is four chip columns
// (lane group g16 carries chip column v = 4 m + g16) times 8 window rows. The B fragment of
// (v, band block rb) -- chip rows 8 rb - l16 .. 8 rb - l16 + 7 -- does not depend on the tile,
// and the A fragment of window rows 16 ti0 + 8 rho .. + 7 serves every (tile t, rb) with
// rho = 2 t + rb: per chip-column quad NRB B and NRB + 6 A fragments feed 4 NRB MFMAs.
template <int PREC, int NRB>
__device__ inline void mp_batch(const uint16_t* Wl, const uint16_t* Cl, int ldw, int cpad, int cw,
                                int tj, int ti0, int l16, int g16, v4f (&acc)[4]) {
    const int nmq = (cw + 3) >> 2;
    int aoff = (16 * tj + l16 + g16) * ldw + 16 * ti0;   // elements
    int boff = g16 * cpad + 16 - l16;                    // elements
    const uint32_t sb = (uint32_t)(boff & 1) * 2;        // byte shift, the same for every m, rb
    for (int m = 0; m < nmq; m++) {
        v4i bf[NRB];
#pragma unroll
        for (int rb = 0; rb < NRB; rb++) {
            const uint32_t* bp = (const uint32_t*)(Cl + ((boff + 8 * rb) & ~1));
            const uint32_t d0 = bp[0], d1 = bp[1], d2 = bp[2], d3 = bp[3], d4 = bp[4];
            bf[rb] = (v4i){(int)__builtin_amdgcn_alignbyte(d1, d0, sb),
                           (int)__builtin_amdgcn_alignbyte(d2, d1, sb),
                           (int)__builtin_amdgcn_alignbyte(d3, d2, sb),
                           (int)__builtin_amdgcn_alignbyte(d4, d3, sb)};
        }
        const uint16_t* ap = Wl + aoff;
#pragma unroll
        for (int rho = 0; rho < NRB + 6; rho++) {
            const v4i af = *(const v4i*)(ap + 8 * rho);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int rb = rho - 2 * t;
                if (rb >= 0 && rb < NRB) acc[t] = MpType<PREC>::mfma(af, bf[rb], acc[t]);
            }
        }
        aoff += 4 * ldw;
        boff += 4 * cpad;
    }
}


template <int PREC, int NRB>
__global__ __launch_bounds__(AR_WG) void track_mp_kernel(TrackArgs a0) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red_d[AR_NWAVES];
    __shared__ float pvals[32];
    __shared__ Best red[AR_NWAVES];
    TrackArgs a = a0;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const MpLayout L = mp_layout(a.max_ch, a.max_cw, a.max_rx, a.max_ry, NRB);
    uint16_t* Wl = (uint16_t*)(smem + L.off_w);
    float* num = (float*)(smem + L.off_num);
    uint16_t* Cl = (uint16_t*)(smem + L.off_c);
    float* chipf = (float*)(smem + L.off_w);      // raw Float32 chip, before the window
    float* pwin = (float*)(smem + L.off_w);       // after the argmax: window block under the patch
    float* pchip = (float*)(smem + L.off_pc);     // and the mean-removed chip

    const PointRange pr = point_range(a.n);
    for (int64_t p = pr.first; p < pr.end; p += pr.step) {
        // The argument block is read again from the kernarg segment at each phase and the lane
        // index is hidden from the optimiser per point: values the loop does not need across
        // phases are formed where they are used instead of being held in (spilled) registers.
        a = reload_args();
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int lane = tid & 63, g16 = lane >> 4, l16 = lane & 15;
        const float* ref = (const float*)a.ref;
        const float* sec = (const float*)a.sec;
        const Geom g = point_geom(a, p);
        if (!g.status || !any_valid(a, g)) { write_skip(a, p, false); continue; }
        const int ch = g.ch, cw = g.cw, NR = 2 * g.ry, NC = 2 * g.rx;
        const int winh = ch + NR - 1, winw = cw + NC - 1;

        // ---- chip: Float32 samples, then the reference's sequential Float64 mean and energy
#pragma unroll 4
        for (int v = wave; v < cw; v += AR_NWAVES)
            for (int u = lane; u < ch; u += 64) chipf[u + v * ch] = mp_pix(a, sec, g.cr0 + u, g.cc0 + v);
        __syncthreads();
        // chip mean and energy: Float64 block sums in a fixed order
        const int nn = ch * cw;
        double s = 0.0;
        for (int k = tid; k < nn; k += AR_WG) s += (double)chipf[k];
        const double mean = block_sum(s, red_d) / (double)nn;
        double q = 0.0;
        for (int k = tid; k < nn; k += AR_WG) {
            const double d = (double)chipf[k] - mean;
            q += d * d;
        }
        const double sq = block_sum(q, red_d);
        if (!(sq > 2.220446049250313e-16 * (double)ch * (double)cw)) {   // degenerate chip
            write_skip(a, p, true);
            __syncthreads();
            continue;
        }
        const double cnorm = sqrt(sq);
        // the mean-removed chip Float32(x - mean), scaled, as a 16-bit copy in full padded
        // columns
        for (int v = wave; v < L.cw4; v += AR_NWAVES)
            for (int r = lane; r < L.cpad; r += 64) {
                const int u = r - 16;
                float c = 0.0f;
                if (v < cw && u >= 0 && u < ch) c = (float)((double)chipf[u + v * ch] - mean);
                Cl[v * L.cpad + r] = MpType<PREC>::bits(c * a.cscale);
            }
        __syncthreads();
        // ---- window, scaled to the 16-bit type; rows and columns past it zero. A window inside
        // the image is read through one base pointer with no per-pixel padding test.
        {
            const int64_t orow = g.wr0 - a.pady - 1, ocol = g.wc0 - a.padx - 1;
            const bool inside = orow >= 0 && ocol >= 0 && orow + winh <= a.ny &&
                                ocol + winw <= a.nx;
            const float* wb = ref + (inside ? orow + ocol * a.ny : 0);
#pragma unroll 4
            for (int c = wave; c < L.cols; c += AR_NWAVES) {
                const float* col = wb + (int64_t)c * a.ny;
                for (int r = lane; r < L.ldw; r += 64) {
                    float w = 0.0f;
                    if (c < winw && r < winh)
                        w = inside ? col[r] : mp_pix(a, ref, g.wr0 + r, g.wc0 + c);
                    Wl[c * L.ldw + r] = MpType<PREC>::bits(w * a.wscale);
                }
            }
        }
        __syncthreads();

        a = reload_args();
        // ---- numerators on the matrix cores
        const int nti = (NR + 15) >> 4, ntj = (NC + 15) >> 4, nbt = (nti + 3) >> 2;
        auto store = [&](const v4f (&acc)[4], int tj, int ti0) {
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int i = 16 * (ti0 + t) + l16;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int j = 16 * tj + 4 * g16 + r;
                    if (i < NR && j < NC) num[j * NR + i] = acc[t][r];
                }
            }
        };
        if (L.num_in_w) {
            // at most one tile batch per wave: its accumulators wait in registers until every
            // wave is past the window, then go over it
            v4f acc[4];
#pragma unroll
            for (int t = 0; t < 4; t++) acc[t] = (v4f){0.0f, 0.0f, 0.0f, 0.0f};
            const bool mine = wave < ntj * nbt;
            const int tj = wave % ntj, ti0 = 4 * (wave / ntj);
            if (mine) mp_batch<PREC, NRB>(Wl, Cl, L.ldw, L.cpad, cw, tj, ti0, l16, g16, acc);
            __syncthreads();
            if (mine) store(acc, tj, ti0);
        } else {
            for (int item = wave; item < ntj * nbt; item += AR_NWAVES) {
                const int tj = item % ntj, ti0 = 4 * (item / ntj);
                v4f acc[4];
#pragma unroll
                for (int t = 0; t < 4; t++) acc[t] = (v4f){0.0f, 0.0f, 0.0f, 0.0f};
                mp_batch<PREC, NRB>(Wl, Cl, L.ldw, L.cpad, cw, tj, ti0, l16, g16, acc);
                store(acc, tj, ti0);
            }
        }
        __syncthreads();

        a = reload_args();
        // ---- normalisation per shift and the argmax (peak_index rule): the window variance
        // s2 - s1^2 / n formed in Float64 by the box pass (where it cancels) and read as Float32,
        // the rest of the 16-bit surface's score in FP32 (the exact values of the chosen peak and
        // patch are Float64, below)
        const double n = (double)(ch * cw);
        const int64_t bb = (g.wr0 - 1) + (g.wc0 - 1) * a.box_ld;
        const float inv_cn = (float)(1.0 / cnorm);
        const float nsc = (float)a.numscale;
        Best b = best_empty();
        // the variances of four shift columns (two row blocks of 64 each) loaded before any is
        // used
        for (int jb = wave; jb < NC; jb += 4 * AR_NWAVES)
            for (int ib = 0; ib < NR; ib += 128) {
                float den[4][2];
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int j = jb + (e >> 1) * AR_NWAVES, i = ib + lane + 64 * (e & 1);
                    den[e >> 1][e & 1] =
                        j < NC && i < NR ? a.fwv[bb + i + (int64_t)j * a.box_ld] : 0.0f;
                }
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    const int j = jb + (e >> 1) * AR_NWAVES, i = ib + lane + 64 * (e & 1);
                    if (j >= NC || i >= NR) continue;
                    const float d = den[e >> 1][e & 1];
                    const float v = d > 0.0f
                                        ? num[j * NR + i] * nsc * inv_cn * __builtin_amdgcn_rsqf(d)
                                        : 0.0f;
                    const int idx = i * NC + j;
                    if (better(v, idx, b)) { b.v = v; b.idx = idx; }
                }
            }
        b = block_best(b, red);
        if (b.idx == 0x7fffffff) b.idx = 0;
        const int pi = b.idx / NC + 1, pj = b.idx % NC + 1;
        int i0 = 1, j0 = 1;
        const bool refine = a.up > 1 && patch_origin(pi, pj, NR, NC, i0, j0);

        // ---- exact Float64 values: the 5x5 patch (threads 0..24, row-major) or the peak (25).
        // The window block under them (top-left shift (bi, bj)) and the mean-removed chip are
        // staged in LDS first (the 16-bit window and the scores are dead by now).
        const int bi = refine ? i0 - 1 : pi - 1, bj = refine ? j0 - 1 : pj - 1;
        const int ph = ch + 4, pw = cw + 4;
#pragma unroll 4
        for (int c = wave; c < pw; c += AR_NWAVES)
            for (int r = lane; r < ph; r += 64)
                pwin[r + c * ph] = (bi + r < winh && bj + c < winw)
                                       ? mp_pix(a, ref, g.wr0 + bi + r, g.wc0 + bj + c)
                                       : 0.0f;
#pragma unroll 4
        for (int v = wave; v < cw; v += AR_NWAVES)
            for (int u = lane; u < ch; u += 64)
                pchip[u + v * ch] = (float)((double)mp_pix(a, sec, g.cr0 + u, g.cc0 + v) - mean);
        __syncthreads();
        // value v (0..24 the patch, row-major; 25 the peak alone) is wave v % 4's. Lanes map to
        // chip rows and column groups: for ch <= 64, lane = jg ch + ii takes rows ii of columns
        // jg, jg + G, ... (G = 64 / ch); taller chips take rows lane and lane + 64 of every column
        const int G = ch <= 64 ? 64 / ch : 1;
        const int ii0 = ch <= 64 ? lane % ch : lane, jg = ch <= 64 ? lane / ch : 0;
        const bool active = jg < G;
        for (int v = wave; v < 26; v += AR_NWAVES) {
            if ((v < 25) != refine) continue;
            const int di = v < 25 ? v / 5 : 0, dj = v < 25 ? v % 5 : 0;
            const float* wb = pwin + dj * ph + di;
            double acc = 0.0;
            if (active)
                for (int jj = jg; jj < cw; jj += G) {
                    const float* cc = pchip + jj * ch;
                    const float* wc = wb + jj * ph;
                    for (int ii = ii0; ii < ch; ii += 64) acc += (double)cc[ii] * (double)wc[ii];
                }
            acc = wave_sum(acc);
            if (lane == 0) {
                const int64_t q = bb + (bi + di) + (int64_t)(bj + dj) * a.box_ld;
                pvals[v] = mp_normalise(acc, a.fbox1[q], a.fbox2[q], n, cnorm, a.measure);
            }
        }
        __syncthreads();
        const float pv = tid < 25 ? pvals[tid] : 0.0f;
        a = reload_args();
        finish_point(a, p, g, pi, pj, pvals[25], i0, j0, refine, pv);
        __syncthreads();
    }
}

#define AR_MP_INST(P, N) template __global__ void track_mp_kernel<P, N>(TrackArgs);
AR_MP_INST(AR_PREC_BF16, 4) AR_MP_INST(AR_PREC_BF16, 6) AR_MP_INST(AR_PREC_BF16, 8)
AR_MP_INST(AR_PREC_BF16, 10) AR_MP_INST(AR_PREC_BF16, 14) AR_MP_INST(AR_PREC_BF16, 18)
AR_MP_INST(AR_PREC_FP16, 4) AR_MP_INST(AR_PREC_FP16, 6) AR_MP_INST(AR_PREC_FP16, 8)
AR_MP_INST(AR_PREC_FP16, 10) AR_MP_INST(AR_PREC_FP16, 14) AR_MP_INST(AR_PREC_FP16, 18)
#undef AR_MP_INST

// ---------------------------------------------------------------------------------------------
// Float64 box images of the zero-padded Float32 reference over every chip position
// ---------------------------------------------------------------------------------------------

// Padded 0-based (r, c) -> pixel value, 0 in the padding.
__device__ inline double boxf_pix(const float* img, int64_t ny, int64_t nx, int64_t pady,
                                  int64_t padx, int64_t r, int64_t c) {
    const int64_t orow = r - pady, ocol = c - padx;
    if (orow < 0 || orow >= ny || ocol < 0 || ocol >= nx) return 0.0;
    return (double)img[orow + ocol * ny];
}

#define AR_BOXF_BLK 64
// V1 (r, c) = sum_{u < ch} W(r + u, c), V2 likewise of W^2, r < HB, every padded column c. One
// thread per output row and column (workgroup = 256 rows of one column, nrb workgroups per
// column): consecutive lanes take consecutive rows, so every load and store is coalesced, and the
// ch terms of a sum are added directly in order (the overlapping loads of neighbouring lanes hit
// in the caches), eight loads in flight at a time.
__global__ __launch_bounds__(256) void box_f64_colpass_kernel(const float* __restrict__ img,
                                                              int64_t ny, int64_t nx, int64_t pady,
                                                              int64_t padx, int ch, int64_t HB,
                                                              int nrb, double* __restrict__ V1,
                                                              double* __restrict__ V2) {
    const int64_t c = blockIdx.x / (unsigned)nrb;
    const int64_t r = (int64_t)(blockIdx.x - (unsigned)c * (unsigned)nrb) * 256 + threadIdx.x;
    if (r >= HB) return;
    double s1 = 0.0, s2 = 0.0;
    const int64_t ocol = c - padx, o0 = r - pady;
    if (ocol >= 0 && ocol < nx) {
        const float* col = img + ocol * ny;
        if (o0 >= 0 && o0 + ch <= ny) {
            const float* q = col + o0;
            int u = 0;
            for (; u + 8 <= ch; u += 8) {
                float w[8];
#pragma unroll
                for (int k = 0; k < 8; k++) w[k] = q[u + k];
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const double d = (double)w[k];
                    s1 += d;
                    s2 += d * d;
                }
            }
            for (; u < ch; u++) {
                const double d = (double)q[u];
                s1 += d;
                s2 += d * d;
            }
        } else {
            for (int u = 0; u < ch; u++) {
                const int64_t orow = o0 + u;
                const double d = (orow >= 0 && orow < ny) ? (double)col[orow] : 0.0;
                s1 += d;
                s2 += d * d;
            }
        }
    }
    V1[c * HB + r] = s1;
    V2[c * HB + r] = s2;
}

// B (r, c) = sum_{v < cw} V(r, c + v), c < WB, and FW (r, c) = Float32(max(wv, 0)) with the
// window variance wv = s2 - s1 s1 (1 / n) (ZNCC) or s2 (NCC), formed as the trackers' screening
// forms it. One thread per (row, block of 64 columns):
// the block's first sum in order, then a sliding sum; consecutive lanes take consecutive rows
// (workgroup = 256 rows of one column block, nrb workgroups per block).
__global__ __launch_bounds__(256) void box_f64_rowpass_kernel(const double* __restrict__ V1,
                                                              const double* __restrict__ V2,
                                                              int cw, int64_t HB, int64_t WB,
                                                              int nrb, double inv_n, int measure,
                                                              double* __restrict__ B1,
                                                              double* __restrict__ B2,
                                                              float* __restrict__ FW) {
    const int64_t blk = blockIdx.x / (unsigned)nrb;
    const int64_t r = (int64_t)(blockIdx.x - (unsigned)blk * (unsigned)nrb) * 256 + threadIdx.x;
    if (r >= HB) return;
    const int64_t c0 = blk * AR_BOXF_BLK;
    const int64_t c1 = c0 + AR_BOXF_BLK < WB ? c0 + AR_BOXF_BLK : WB;
    const double* v1 = V1 + r;
    const double* v2 = V2 + r;
    double s1 = 0.0, s2 = 0.0;
#pragma unroll 8
    for (int v = 0; v < cw; v++) {
        s1 += v1[(c0 + v) * HB];
        s2 += v2[(c0 + v) * HB];
    }
#pragma unroll 8
    for (int64_t c = c0; c < c1; c++) {
        B1[r + c * HB] = s1;
        B2[r + c * HB] = s2;
        const double wv = measure == 1 ? s2 : s2 - s1 * s1 * inv_n;
        FW[r + c * HB] = (float)(wv < 0.0 ? 0.0 : wv);
        if (c + 1 < c1) {
            s1 = (s1 + v1[(c + cw) * HB]) - v1[c * HB];
            s2 = (s2 + v2[(c + cw) * HB]) - v2[c * HB];
        }
    }
}

// Largest |value| of a Float32 image (non-finite values excluded), as the bits of a
// non-negative float (ordered like unsigned integers) in *out, which starts at 0.
__global__ __launch_bounds__(256) void absmax_f32_kernel(const float* img, int64_t n,
                                                         unsigned int* out) {
    unsigned int m = 0;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        const float v = fabsf(img[k]);
        if (v <= 3.402823466e38f) m = max(m, __float_as_uint(v));
    }
    for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned int)__shfl_xor((int)m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

}  // namespace ar

using namespace ar;

static int64_t mp_grid(int64_t work) {
    int64_t b = (work + 255) / 256;
    return b < 1 ? 1 : (b > 65536 ? 65536 : b);
}
// (grid-stride loops: a grid of at most 65 536 workgroups covers any work count)

// Set-up of a mixed-precision pass: operand scales, the pass's Float64 box images and the
// kernel instance. `a` already holds the pass geometry (max extents, padding).
ar_status ar_mp_setup(ar_context* ctx, TrackArgs& a, void (**kern)(TrackArgs), size_t* lds) {
    const int prec = ctx->params.precision;
    const int nrb = mp_nrb(a.max_ch);
    if (nrb == 0) return AR_ERR_UNSUPPORTED;
    const MpLayout L = mp_layout(a.max_ch, a.max_cw, a.max_rx, a.max_ry, nrb);
    if (L.total > 160 * 1024 - 2048) return AR_ERR_UNSUPPORTED;
    *lds = L.total;
    // operand scales: FP16 keeps |W| <= 2^14 and |chip'| <= 2^14 (|chip'| <= 2 max|W|)
    float wscale = 1.0f, cscale = 1.0f;
    double numscale = 1.0;
    if (prec == AR_PREC_FP16) {
        if (ctx->absmax_gen != ctx->ref_gen) {
            if (ar_status st = ctx->flags.reserve(ctx, 64)) return st;
            if (hipMemsetAsync(ctx->flags.ptr, 0, 4, ctx->stream) != hipSuccess) return AR_ERR_DEVICE;
            const int64_t N = ctx->ny * ctx->nx;
            absmax_f32_kernel<<<mp_grid(N), 256, 0, ctx->stream>>>((const float*)ctx->pref.ptr, N,
                                                                   (unsigned int*)ctx->flags.ptr);
            absmax_f32_kernel<<<mp_grid(N), 256, 0, ctx->stream>>>((const float*)ctx->psec.ptr, N,
                                                                   (unsigned int*)ctx->flags.ptr);
            if (hipGetLastError() != hipSuccess) return AR_ERR_DEVICE;
            unsigned int bits = 0;
            if (hipMemcpyAsync(&bits, ctx->flags.ptr, 4, hipMemcpyDeviceToHost, ctx->stream) !=
                    hipSuccess ||
                hipStreamSynchronize(ctx->stream) != hipSuccess)
                return AR_ERR_DEVICE;
            float m;
            std::memcpy(&m, &bits, 4);
            ctx->absmax = m;
            ctx->absmax_gen = ctx->ref_gen;
        }
        int e = 0;
        if (ctx->absmax > 0.0f) {
            std::frexp(ctx->absmax, &e);   // absmax < 2^e
            e = e > 14 ? e - 14 : 0;
        }
        wscale = std::ldexp(1.0f, -e);
        cscale = std::ldexp(1.0f, -(e + 1));
        numscale = std::ldexp(1.0, 2 * e + 1);
    }
    a.wscale = wscale;
    a.cscale = cscale;
    a.numscale = numscale;
    if (ar_status st = ar_f64_box_images(ctx, a)) return st;
#define AR_MP_PICK(P)                                                          \
    *kern = nrb == 4 ? track_mp_kernel<P, 4> : nrb == 6 ? track_mp_kernel<P, 6> \
          : nrb == 8 ? track_mp_kernel<P, 8> : nrb == 10 ? track_mp_kernel<P, 10> \
          : nrb == 14 ? track_mp_kernel<P, 14> : track_mp_kernel<P, 18>;
    if (prec == AR_PREC_BF16) { AR_MP_PICK(AR_PREC_BF16) }
    else { AR_MP_PICK(AR_PREC_FP16) }
#undef AR_MP_PICK
    return AR_OK;
}

// Float64 box images of W and W^2 over the padded Float32 reference's chip positions for the
// pass's chip size (a.max_ch x a.max_cw), and the Float32 window variances the trackers screen
// with: a.fbox1, a.fbox2, a.fwv, a.box_ld.
ar_status ar_f64_box_images(ar_context* ctx, TrackArgs& a) {
    const int64_t NY = ctx->ny + 2 * a.pady, NX = ctx->nx + 2 * a.padx;
    const int64_t HB = NY - a.max_ch + 1, WB = NX - a.max_cw + 1;
    if (HB < 1 || WB < 1) return AR_ERR_UNSUPPORTED;
    const size_t vbytes = (size_t)HB * NX * 8, bbytes = (size_t)HB * WB * 8;
    if (ar_status st = ctx->box.reserve(ctx, 2 * vbytes + 2 * bbytes + bbytes / 2)) return st;
    ctx->box_key.ptr = nullptr;   // the integer box images are overwritten
    double* V1 = (double*)ctx->box.ptr;
    double* V2 = (double*)((char*)ctx->box.ptr + vbytes);
    double* B1 = (double*)((char*)ctx->box.ptr + 2 * vbytes);
    double* B2 = (double*)((char*)ctx->box.ptr + 2 * vbytes + bbytes);
    float* FW = (float*)((char*)ctx->box.ptr + 2 * vbytes + 2 * bbytes);
    const int64_t nrb = (HB + 255) / 256, nblk = (WB + AR_BOXF_BLK - 1) / AR_BOXF_BLK;
    if (nrb * NX > 0x7fffffff || nrb * nblk > 0x7fffffff) return AR_ERR_UNSUPPORTED;
    box_f64_colpass_kernel<<<(unsigned)(nrb * NX), 256, 0, ctx->stream>>>(
        (const float*)ctx->pref.ptr, ctx->ny, ctx->nx, a.pady, a.padx, a.max_ch, HB, (int)nrb, V1,
        V2);
    const double inv_n = 1.0 / (double)((int64_t)a.max_ch * a.max_cw);
    box_f64_rowpass_kernel<<<(unsigned)(nrb * nblk), 256, 0, ctx->stream>>>(
        V1, V2, a.max_cw, HB, WB, (int)nrb, inv_n, a.measure, B1, B2, FW);
    if (hipGetLastError() != hipSuccess) return AR_ERR_DEVICE;
    a.fbox1 = B1;
    a.fbox2 = B2;
    a.fwv = FW;
    a.box_ld = HB;
    return AR_OK;
}
