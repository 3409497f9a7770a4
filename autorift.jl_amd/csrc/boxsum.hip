// Box images for a tracking pass with one chip size: for every chip position (r, c) of the
// zero-padded UInt8 reference image P (padded 0-based coordinates),
//   B1(r, c) = sum_{u < ch, v < cw} P(r + u, c + v),   B2(r, c) = sum P(r + u, c + v)^2,
// exact int32 (ch <= 256 rows of 255 and areas <= 16384 keep both below 2^31). The tracker's
// S1 / S2 for shift (i, j) of a point whose window starts at padded (wr0, wc0) (1-based) are
// B(wr0 - 1 + i, wc0 - 1 + j): the integer box sums of src/correlate.jl:574-600 (summed-area
// tables, src/integral.jl:141-170) read instead of rebuilt per point, identical as integers.
//
// One wave per (band of 64 RPL padded rows, chunk of output columns, its width chosen per launch
// so that the chunks fill the resident waves in one round); lane l holds rows
// RPL l .. RPL l + RPL - 1 of the band. The wave walks the chunk's columns keeping each row's
// horizontal running sums over cw columns in registers (loads of one column are coalesced
// across lanes), then forms the vertical box of ch rows from an inclusive prefix over the band
// (in-lane, then across lanes) as S(r + ch - 1) - S(r - 1), giving 64 RPL - ch + 1 output rows
// per column, stored coalesced (column-major, leading dimension HB). Sums wrap modulo 2^32;
// differences of box size are exact.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/autorift_hip.h"
#include "ar_internal.h"

namespace ar {

// rows per lane for chip height ch: the band of 64 RPL rows yields at least half as many outputs
__host__ __device__ inline int box_rpl(int ch) { return ch <= 128 ? 4 : 8; }

// Inclusive prefix sum over the 64 lanes by DPP: shifts by 1, 2, 4, 8 within each row of 16
// lanes, then the last lane of rows 0 and 2 into rows 1 and 3 (row_bcast:15) and the last lane
// of row 1 into rows 2 and 3 (row_bcast:31). Lanes a shift leaves without a source add 0.
__device__ inline uint32_t wave_scan_u32(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);
    return v;
}

template <int RPL>
__global__ __launch_bounds__(256) void box_u8_kernel(const uint8_t* img, int64_t ny, int64_t nx,
                                                    int64_t pady, int64_t padx, int ch, int cw,
                                                    int64_t HB, int64_t WB, int chunk_w,
                                                    int32_t* B1, int32_t* B2) {
    // columns per group of loads: 16 loaded dwords per lane in flight whatever the band height
    constexpr int BX_G = 32 / RPL;
    __shared__ __align__(16) uint32_t sS[4][2][64 * RPL];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t* s1 = sS[wave][0];
    uint32_t* s2 = sS[wave][1];
    const int OUT = 64 * RPL - ch + 1;   // output rows per band
    const int64_t nb = (HB + OUT - 1) / OUT, nchunks = (WB + chunk_w - 1) / chunk_w;
    const int64_t nw = (int64_t)gridDim.x * 4;
    for (int64_t w = (int64_t)blockIdx.x * 4 + wave; w < nb * nchunks; w += nw) {
        const int64_t band = w % nb, chunk = w / nb;
        const int64_t R0 = band * OUT, C0 = chunk * chunk_w;
        const int64_t C1 = C0 + chunk_w < WB ? C0 + chunk_w : WB;
        // image rows of this lane (padded rows R0 + RPL lane + k); the column walk is wave-uniform
        const int64_t ir0 = R0 + (int64_t)RPL * lane - pady;
        uint32_t rowok = 0;
#pragma unroll
        for (int k = 0; k < RPL; k++)
            rowok |= (ir0 + k >= 0 && ir0 + k < ny) ? 1u << k : 0u;
        const bool full = rowok == (1u << RPL) - 1;
        // the lane's RPL bytes of padded column pc (zero outside the image), packed 4 per dword
        auto load = [&](int64_t pc, uint32_t* w) {
            const int64_t col = pc - padx;
#pragma unroll
            for (int q = 0; q < RPL / 4; q++) w[q] = 0;
            if (col < 0 || col >= nx) return;
            const uint8_t* base = img + col * ny + ir0;
            if (full) {
#pragma unroll
                for (int q = 0; q < RPL / 4; q++) {
                    const uintptr_t ad = (uintptr_t)(base + 4 * q);
                    const uint32_t* al = (const uint32_t*)(ad & ~(uintptr_t)3);
                    w[q] = __builtin_amdgcn_alignbyte(al[1], al[0], (uint32_t)(ad & 3));
                }
            } else {
#pragma unroll
                for (int k = 0; k < RPL; k++)
                    if ((rowok >> k) & 1) w[k >> 2] |= (uint32_t)base[k] << (8 * (k & 3));
            }
        };
        auto byte = [](const uint32_t* w, int k) -> uint32_t { return (w[k >> 2] >> (8 * (k & 3))) & 0xffu; };
        uint32_t h1[RPL], h2[RPL];
#pragma unroll
        for (int k = 0; k < RPL; k++) h1[k] = h2[k] = 0;
        // horizontal sums of the first column, BX_G column loads in flight
        for (int v0 = 0; v0 < cw; v0 += BX_G) {
            uint32_t w[BX_G][RPL / 4];
#pragma unroll
            for (int g = 0; g < BX_G; g++) {
                if (v0 + g < cw) {
                    load(C0 + v0 + g, w[g]);
                } else {
#pragma unroll
                    for (int q = 0; q < RPL / 4; q++) w[g][q] = 0;
                }
            }
#pragma unroll
            for (int g = 0; g < BX_G; g++)
#pragma unroll
                for (int k = 0; k < RPL; k++) {
                    uint32_t x = byte(w[g], k);
                    h1[k] += x;
                    h2[k] += x * x;
                }
        }
        // columns leaving (c) and entering (c + cw) the horizontal window after output column c
        auto load_group = [&](int64_t cg, uint32_t (*wo)[RPL / 4], uint32_t (*wi)[RPL / 4]) {
#pragma unroll
            for (int g = 0; g < BX_G; g++) {
                if (cg + g < C1) {
                    load(cg + g, wo[g]);
                    load(cg + g + cw, wi[g]);
                } else {
#pragma unroll
                    for (int q = 0; q < RPL / 4; q++) wo[g][q] = wi[g][q] = 0;
                }
            }
        };
        // output column c from the current horizontal sums, then slide them by one column
        auto step = [&](int64_t c, const uint32_t* wo, const uint32_t* wi) {
            // inclusive prefix over the band: in-lane, then an exclusive scan of lane totals
            uint32_t p1[RPL], p2[RPL];
            p1[0] = h1[0];
            p2[0] = h2[0];
#pragma unroll
            for (int k = 1; k < RPL; k++) {
                p1[k] = p1[k - 1] + h1[k];
                p2[k] = p2[k - 1] + h2[k];
            }
            const uint32_t t1 = wave_scan_u32(p1[RPL - 1]), t2 = wave_scan_u32(p2[RPL - 1]);
            const uint32_t e1 = t1 - p1[RPL - 1], e2 = t2 - p2[RPL - 1];
#pragma unroll
            for (int q = 0; q < RPL / 4; q++) {
                ((uint4*)s1)[(RPL / 4) * lane + q] =
                    make_uint4(e1 + p1[4 * q], e1 + p1[4 * q + 1], e1 + p1[4 * q + 2], e1 + p1[4 * q + 3]);
                ((uint4*)s2)[(RPL / 4) * lane + q] =
                    make_uint4(e2 + p2[4 * q], e2 + p2[4 * q + 1], e2 + p2[4 * q + 2], e2 + p2[4 * q + 3]);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // output rows lane, lane + 64, ...: each store covers 64 consecutive rows
#pragma unroll
            for (int k = 0; k < RPL; k++) {
                const int rb = lane + 64 * k;
                const int64_t R = R0 + rb;
                if (rb < OUT && R < HB) {
                    const uint32_t lo1 = rb > 0 ? s1[rb - 1] : 0u, lo2 = rb > 0 ? s2[rb - 1] : 0u;
                    B1[R + c * HB] = (int32_t)(s1[rb + ch - 1] - lo1);
                    B2[R + c * HB] = (int32_t)(s2[rb + ch - 1] - lo2);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int k = 0; k < RPL; k++) {
                const uint32_t xo = byte(wo, k), xi = byte(wi, k);
                h1[k] += xi - xo;
                h2[k] += xi * xi - xo * xo;
            }
        };
        // two register sets: the loads of the next group are in flight while a group is used
        uint32_t ao[BX_G][RPL / 4], ai[BX_G][RPL / 4], bo[BX_G][RPL / 4], bi[BX_G][RPL / 4];
        load_group(C0, ao, ai);
        for (int64_t cg = C0; cg < C1; cg += 2 * BX_G) {
            load_group(cg + BX_G, bo, bi);
#pragma unroll
            for (int g = 0; g < BX_G; g++)
                if (cg + g < C1) step(cg + g, ao[g], ai[g]);
            load_group(cg + 2 * BX_G, ao, ai);
#pragma unroll
            for (int g = 0; g < BX_G; g++)
                if (cg + BX_G + g < C1) step(cg + BX_G + g, bo[g], bi[g]);
        }
    }
}

}  // namespace ar

using namespace ar;

ar_status ar_box_images(ar_context* ctx, int ch, int cw, int64_t pady, int64_t padx, bool reuse,
                        int32_t** b1, int32_t** b2, int64_t* ld) {
    const int64_t NY = ctx->ny + 2 * pady, NX = ctx->nx + 2 * padx;
    const int64_t HB = NY - ch + 1, WB = NX - cw + 1;
    if (HB < 1 || WB < 1) return AR_ERR_UNSUPPORTED;
    const size_t each = (size_t)HB * WB * 4;
    ar_status st = ctx->box.reserve(ctx, 2 * each);
    if (st) return st;
    int32_t* p1 = (int32_t*)ctx->box.ptr;
    int32_t* p2 = (int32_t*)((char*)ctx->box.ptr + each);
    auto& k = ctx->box_key;
    if (reuse && k.ptr == ctx->box.ptr && k.ch == ch && k.cw == cw && k.pady == pady && k.padx == padx &&
        k.gen == ctx->ref_gen) {
        *b1 = p1;   // made by an earlier pass over the same pair, chip and padding
        *b2 = p2;
        *ld = HB;
        return AR_OK;
    }
    k.ptr = nullptr;
    const int rpl = box_rpl(ch);
    const int OUT = 64 * rpl - ch + 1;
    auto kern = rpl == 4 ? box_u8_kernel<4> : (rpl == 8 ? box_u8_kernel<8> : box_u8_kernel<16>);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0) != hipSuccess ||
        per_cu < 1)
        return AR_ERR_DEVICE;
    // Column chunks sized so that the work items (row band, column chunk) fill the resident
    // waves in one round: with a fixed width the last round ran mostly idle. Chunks of at least
    // 64 columns keep the horizontal warm-up (cw columns per chunk) a minor share.
    const int64_t nb = (HB + OUT - 1) / OUT, resident = (int64_t)ctx->ncu * per_cu * 4;
    int64_t nchunks = resident / nb;
    if (nchunks < 1) nchunks = 1;
    int64_t chunk_w = (WB + nchunks - 1) / nchunks;
    if (chunk_w < 64) chunk_w = 64;
    if (chunk_w > 1 << 20) chunk_w = 1 << 20;
    const int64_t waves = nb * ((WB + chunk_w - 1) / chunk_w);
    int64_t blocks = (int64_t)ctx->ncu * per_cu;
    if (blocks > (waves + 3) / 4) blocks = (waves + 3) / 4;
    const uint8_t* img = (const uint8_t*)ctx->pref.ptr;
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, img, ctx->ny,
                       ctx->nx, pady, padx, ch, cw, HB, WB, (int)chunk_w, p1, p2);
    if (hipGetLastError() != hipSuccess) return AR_ERR_DEVICE;
    k.ch = ch;
    k.cw = cw;
    k.pady = pady;
    k.padx = padx;
    k.gen = ctx->ref_gen;
    k.ptr = ctx->box.ptr;
    *b1 = p1;
    *b2 = p2;
    *ld = HB;
    return AR_OK;
}
