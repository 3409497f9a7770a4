// Complex (single-look-complex radar) pairs: preparation (isfinite masks, Deramp, NoQuantize)
// and the Coherence tracker.
//
// Coherence (src/correlate.jl:632-683, src/types.jl:154-200) is the complex analogue of ZNCC:
// |sum conj(T') W| / (|T'| sqrt(sum |W|^2 - |sum W|^2 / n)) with T' the chip less its complex
// mean. The pass keeps the reference's operation order so that results are bit-identical to a
// sequential restatement: ComplexF64 chip mean and energy in column-major order, complex
// summed-area tables built column by column, and per shift a ComplexF64 accumulation of the
// direct numerator in column-major order (_cnumerators_direct!, src/correlate.jl:712-730;
// the reference transforms instead, whose rounding no other program reproduces).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/autorift_hip.h"
#include "ar_internal.h"
#include "layout.h"
#include "refine.h"
#include "track_common.h"

// Every product and sum below rounds on its own, as the reference's scalar Float64 code does
// (abs2 of a complex sum is two rounded squares and a rounded add, not a fused form).
#pragma clang fp contract(off)

namespace ar {
namespace {

constexpr int kRampBlocks = 1024;

__global__ void cplx_finite_mask_kernel(const float2* z, uint8_t* m, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m[i] = (isfinite(z[i].x) && isfinite(z[i].y)) ? 1 : 0;
}

// Per-block partial sums of conj(z_k) z_l over valid neighbour pairs, l one row down (slot 0,
// phase_y) or one column right (slot 1, phase_x): each thread a fixed stride, the block in a
// fixed tree, so the result is the same on every run.
__global__ __launch_bounds__(256) void ramp_partial_kernel(const float2* z, const uint8_t* m,
                                                           int64_t nr, int64_t nc,
                                                           double* part) {
    __shared__ double red[4][256];
    const int64_t N = nr * nc;
    double yr = 0.0, yi = 0.0, xr = 0.0, xi = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < N;
         k += (int64_t)gridDim.x * 256) {
        if (!m[k]) continue;
        const int64_t j = k / nr, i = k - j * nr;
        const double zr = (double)z[k].x, zi = (double)z[k].y;
        if (i + 1 < nr && m[k + 1]) {
            const double ur = (double)z[k + 1].x, ui = (double)z[k + 1].y;
            yr += zr * ur + zi * ui;
            yi += zr * ui - zi * ur;
        }
        if (j + 1 < nc && m[k + nr]) {
            const double ur = (double)z[k + nr].x, ui = (double)z[k + nr].y;
            xr += zr * ur + zi * ui;
            xi += zr * ui - zi * ur;
        }
    }
    red[0][threadIdx.x] = yr;
    red[1][threadIdx.x] = yi;
    red[2][threadIdx.x] = xr;
    red[3][threadIdx.x] = xi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 4; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[threadIdx.x * kRampBlocks + blockIdx.x] = red[threadIdx.x][0];
}

// The block partials in block order, then the phases (phase_y, phase_x); an axis not asked for
// keeps phase 0 (src/preprocess.jl:184-231).
__global__ void ramp_final_kernel(const double* part, int axis, double* phase) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s[4];
    for (int q = 0; q < 4; q++) {
        double acc = 0.0;
        for (int b = 0; b < kRampBlocks; b++) acc += part[q * kRampBlocks + b];
        s[q] = acc;
    }
    phase[0] = axis != 1 ? atan2(s[1], s[0]) : 0.0;
    phase[1] = axis != 2 ? atan2(s[3], s[2]) : 0.0;
}

// deramp's product (ComplexF64(z) * cis(-((i-1) phase_y + (j-1) phase_x)), narrowed), then
// NoQuantize for complex values: a value with a non-finite component becomes zero.
__global__ void deramp_apply_kernel(const float2* z, int64_t nr, int64_t nc, const double* phase,
                                    int deramp, float2* out) {
    const int64_t N = nr * nc;
    const double py = deramp ? phase[0] : 0.0, px = deramp ? phase[1] : 0.0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < N;
         k += (int64_t)gridDim.x * blockDim.x) {
        float2 v = z[k];
        if (deramp) {
            const int64_t j = k / nr, i = k - j * nr;
            const double th = -((double)i * py + (double)j * px);
            double s, c;
            sincos(th, &s, &c);
            const double zr = (double)v.x, zi = (double)v.y;
            v = make_float2((float)(zr * c - zi * s), (float)(zr * s + zi * c));
        }
        if (!(isfinite(v.x) && isfinite(v.y))) v = make_float2(0.0f, 0.0f);
        out[k] = v;
    }
}

__host__ __device__ inline size_t rupz(size_t x, size_t m) { return (x + m - 1) / m * m; }

struct CohScratch {
    size_t off_chip, off_win, off_sr, off_si, off_s2, off_surf, total;
};

__host__ __device__ inline CohScratch coh_scratch(int ch, int cw, int rx, int ry) {
    CohScratch C;
    const size_t winh = ch + 2 * ry - 1, winw = cw + 2 * rx - 1;
    const size_t tab = rupz((winh + 1) * (winw + 1) * 8, 256);
    C.off_chip = 0;
    C.off_win = rupz((size_t)ch * cw * 8, 256);
    C.off_sr = C.off_win + rupz(winh * winw * 8, 256);
    C.off_si = C.off_sr + tab;
    C.off_s2 = C.off_si + tab;
    C.off_surf = C.off_s2 + tab;
    C.total = C.off_surf + rupz((size_t)4 * rx * ry * 4, 256);
    return C;
}

__device__ inline float2 load_c(const TrackArgs& a, const void* img, int64_t r, int64_t c) {
    const int64_t k = pix(a, r, c);
    if (k < 0) return make_float2(0.0f, 0.0f);
    return ((const float2*)img)[k];
}

}  // namespace

// LDS: the window and the chip sit in dynamic LDS (when they fit) instead of the workgroup's
// global scratch; the arithmetic and its order are the same.
template <bool LDS>
__global__ __launch_bounds__(AR_WG) void track_coherence_kernel(TrackArgs a) {
    __shared__ Best red[AR_NWAVES];
    __shared__ double shd[4];
    extern __shared__ float2 dyn[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const CohScratch G = coh_scratch(a.max_ch, a.max_cw, a.max_rx, a.max_ry);
    unsigned char* base = a.scratch + (size_t)blockIdx.x * a.scratch_per_block;
    const int wmax = (a.max_ch + 2 * a.max_ry - 1) * (a.max_cw + 2 * a.max_rx - 1);
    float2* chip = LDS ? dyn + wmax : (float2*)(base + G.off_chip);
    float2* win = LDS ? dyn : (float2*)(base + G.off_win);
    double* Sr = (double*)(base + G.off_sr);
    double* Si = (double*)(base + G.off_si);
    double* S2 = (double*)(base + G.off_s2);
    float* surf = (float*)(base + G.off_surf);

    const PointRange pr = point_range(a.n);
    for (int64_t p = pr.first; p < pr.end; p += pr.step) {
        Geom g = point_geom(a, p);
        if (!g.status) { write_skip(a, p, false); continue; }
        if (!any_valid(a, g)) { write_skip(a, p, false); continue; }
        const int ch = g.ch, cw = g.cw, NR = 2 * g.ry, NC = 2 * g.rx;
        const int winh = ch + NR - 1, winw = cw + NC - 1, ls = winh + 1;
        for (int c = wave; c < winw; c += AR_NWAVES)
            for (int r = lane; r < winh; r += 64)
                win[c * winh + r] = load_c(a, a.ref, g.wr0 + r, g.wc0 + c);
        for (int v = wave; v < cw; v += AR_NWAVES)
            for (int u = lane; u < ch; u += 64)
                chip[v * ch + u] = load_c(a, a.sec, g.cr0 + u, g.cc0 + v);
        __syncthreads();
        // prepare_chip! for complex input (src/correlate.jl:480-502), sequential
        if (tid == 0) {
            double sr = 0.0, si = 0.0;
            for (int q = 0; q < ch * cw; q++) {
                sr += (double)chip[q].x;
                si += (double)chip[q].y;
            }
            const double n = (double)(ch * cw);
            const double mr = sr / n, mi = si / n;
            double sq = 0.0;
            for (int q = 0; q < ch * cw; q++) {
                const double dr = (double)chip[q].x - mr, di = (double)chip[q].y - mi;
                sq += dr * dr + di * di;
            }
            shd[0] = mr;
            shd[1] = mi;
            shd[2] = sq;
        }
        __syncthreads();
        const double mr = shd[0], mi = shd[1], sq = shd[2];
        const bool ok = sq > 2.220446049250313e-16 * (double)ch * (double)cw;
        if (!ok) { write_skip(a, p, true); __syncthreads(); continue; }
        const double cnorm = sqrt(sq);
        for (int q = tid; q < ch * cw; q += AR_WG)
            chip[q] = make_float2((float)((double)chip[q].x - mr), (float)((double)chip[q].y - mi));
        // integral_both! for complex input (src/integral.jl:172-200)
        for (int j = tid; j <= winw; j += AR_WG) { Sr[j * ls] = 0.0; Si[j * ls] = 0.0; S2[j * ls] = 0.0; }
        for (int i = tid; i <= winh; i += AR_WG) { Sr[i] = 0.0; Si[i] = 0.0; S2[i] = 0.0; }
        for (int j = tid; j < winw; j += AR_WG) {
            double rr = 0.0, ri = 0.0, runsq = 0.0;
            for (int i = 0; i < winh; i++) {
                const float2 v = win[i + j * winh];
                const double vr = (double)v.x, vi = (double)v.y;
                rr += vr;
                ri += vi;
                runsq += vr * vr + vi * vi;
                Sr[(i + 1) + (j + 1) * ls] = rr;
                Si[(i + 1) + (j + 1) * ls] = ri;
                S2[(i + 1) + (j + 1) * ls] = runsq;
            }
        }
        __syncthreads();
        for (int i = tid + 1; i <= winh; i += AR_WG) {
            for (int j = 1; j <= winw; j++) {
                Sr[i + j * ls] = Sr[i + j * ls] + Sr[i + (j - 1) * ls];
                Si[i + j * ls] = Si[i + j * ls] + Si[i + (j - 1) * ls];
                S2[i + j * ls] = S2[i + j * ls] + S2[i + (j - 1) * ls];
            }
        }
        __syncthreads();
        // per shift: the direct complex numerator, |num| (Julia's Float32 hypot: squares and sum
        // in Float64, one rounding to Float32) and the normalisation
        const double n = (double)(ch * cw);
        // shifts spread over all 256 lanes (a surface column is shorter than a wave)
        for (int o = tid; o < NR * NC; o += AR_WG) {
                const int i = o % NR, j = o / NR;
                double ar = 0.0, ai = 0.0;
                for (int jj = 0; jj < cw; jj++) {
                    const float2* wc = win + (j + jj) * winh + i;
                    const float2* cc = chip + jj * ch;
                    for (int ii = 0; ii < ch; ii++) {
                        const double cr = (double)cc[ii].x, ci = (double)cc[ii].y;
                        const double wr = (double)wc[ii].x, wi = (double)wc[ii].y;
                        ar += cr * wr + ci * wi;
                        ai += cr * wi - ci * wr;
                    }
                }
                const float n32r = (float)ar, n32i = (float)ai;
                const float mag = (float)sqrt((double)n32r * (double)n32r + (double)n32i * (double)n32i);
                const int I = i + 1, J = j + 1;
#define BOX(T) (T[(I + ch - 1) + (J + cw - 1) * ls] - T[(I - 1) + (J + cw - 1) * ls] - \
                T[(I + ch - 1) + (J - 1) * ls] + T[(I - 1) + (J - 1) * ls])
                const double s1r = BOX(Sr), s1i = BOX(Si), s2 = BOX(S2);
#undef BOX
                const double a2 = s1r * s1r + s1i * s1i;
                const double wv = s2 - a2 / n;
                const double den = cnorm * sqrt(wv < 0.0 ? 0.0 : wv);
                surf[j * NR + i] = den > 0 ? (float)((double)mag / den) : 0.0f;
            }
        __syncthreads();
        Best b = best_empty();
        for (int j = wave; j < NC; j += AR_NWAVES)
            for (int i = lane; i < NR; i += 64) {
                const float v = surf[j * NR + i];
                const int idx = i * NC + j;
                if (better(v, idx, b)) { b.v = v; b.idx = idx; }
            }
        b = block_best(b, red);
        if (b.idx == 0x7fffffff) b.idx = 0;   // all NaN: (1, 1)
        const int pi = b.idx / NC + 1, pj = b.idx % NC + 1;
        const float pval = surf[(pj - 1) * NR + (pi - 1)];
        int i0 = 1, j0 = 1;
        const bool refine = a.up > 1 && patch_origin(pi, pj, NR, NC, i0, j0);
        float pv = 0.0f;
        if (refine && tid < 25) pv = surf[(j0 - 1 + tid % 5) * NR + (i0 - 1 + tid / 5)];
        finish_point(a, p, g, pi, pj, pval, i0, j0, refine, pv);
        __syncthreads();
    }
}

template __global__ void track_coherence_kernel<false>(TrackArgs);
template __global__ void track_coherence_kernel<true>(TrackArgs);

}  // namespace ar

namespace ar {

size_t ar_coherence_scratch_impl(int ch, int cw, int rx, int ry) {
    return coh_scratch(ch, cw, rx, ry).total;
}

}  // namespace ar

size_t ar_coherence_scratch(int ch, int cw, int rx, int ry) {
    return ar::ar_coherence_scratch_impl(ch, cw, rx, ry);
}

// Dynamic LDS bytes of the window and chip of the largest point, or 0 when they do not fit
// beside the kernel's static LDS.
size_t ar_coherence_lds(int ch, int cw, int rx, int ry) {
    const size_t b = ((size_t)(ch + 2 * ry - 1) * (cw + 2 * rx - 1) + (size_t)ch * cw) * 8;
    return b <= AR_LDS_LIMIT ? b : 0;
}

using namespace ar;

// One complex image into dst (ComplexF32) and its validity into dmask: isfinite unless a mask
// is given, then Deramp when asked for, then NoQuantize (src/preprocess.jl:74-79, 184-231,
// 511-517).
ar_status ar_prepare_complex(ar_context* ctx, const void* img, const uint8_t* mask, int64_t ny,
                             int64_t nx, void* dst, uint8_t* dmask) {
    const int64_t N = ny * nx;
    ar_status st;
    if ((st = ctx->raw.reserve(ctx, N * 8)) ||
        (st = ctx->scratch.reserve(ctx, 4 * kRampBlocks * 8 + 64)))
        return st;
    auto chk = [&](hipError_t e) {
        return e == hipSuccess ? AR_OK
                               : (e == hipErrorOutOfMemory ? AR_ERR_OUT_OF_MEMORY : AR_ERR_DEVICE);
    };
    if ((st = chk(hipMemcpyAsync(ctx->raw.ptr, img, N * 8, hipMemcpyDefault, ctx->stream))))
        return st;
    const int blocks = (int)std::min<int64_t>((N + 255) / 256, 8192);
    if (mask) {
        if ((st = chk(hipMemcpyAsync(dmask, mask, N, hipMemcpyDefault, ctx->stream)))) return st;
    } else {
        cplx_finite_mask_kernel<<<blocks, 256, 0, ctx->stream>>>((const float2*)ctx->raw.ptr,
                                                                 dmask, N);
        if ((st = chk(hipGetLastError()))) return st;
    }
    const int deramp = ctx->params.preprocess == AR_PRE_DERAMP;
    double* part = (double*)ctx->scratch.ptr;
    double* phase = part + 4 * kRampBlocks;
    if (deramp) {
        ramp_partial_kernel<<<kRampBlocks, 256, 0, ctx->stream>>>((const float2*)ctx->raw.ptr,
                                                                  dmask, ny, nx, part);
        ramp_final_kernel<<<1, 64, 0, ctx->stream>>>(part, ctx->params.deramp_axis, phase);
        if ((st = chk(hipGetLastError()))) return st;
    }
    deramp_apply_kernel<<<blocks, 256, 0, ctx->stream>>>((const float2*)ctx->raw.ptr, ny, nx,
                                                         phase, deramp, (float2*)dst);
    if ((st = chk(hipGetLastError()))) return st;
    return chk(hipStreamSynchronize(ctx->stream));
}
