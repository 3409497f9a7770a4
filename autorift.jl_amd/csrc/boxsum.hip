// Box images for a tracking pass with one chip size: for every chip position (r, c) of the
// zero-padded UInt8 reference image P (padded 0-based coordinates),
//   B1(r, c) = sum_{u < ch, v < cw} P(r + u, c + v),   B2(r, c) = sum P(r + u, c + v)^2,
// exact int32 (ch <= 256 rows of 255 and areas <= 16384 keep both below 2^31). The tracker's
// S1 / S2 for shift (i, j) of a point whose window starts at padded (wr0, wc0) (1-based) are
// B(wr0 - 1 + i, wc0 - 1 + j): the integer box sums of src/correlate.jl:574-600 (summed-area
// tables, src/integral.jl:141-170) read instead of rebuilt per point, identical as integers.
//
// One workgroup per 32 x 64 tile of positions: the tile's input bytes go to LDS (coalesced
// column reads), then vertical sliding sums of 8 rows per task, then horizontal sliding sums of
// 8 columns per task, written column-major with leading dimension HB.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/autorift_hip.h"
#include "ar_internal.h"

namespace ar {

#define BX_TR 32
#define BX_TC 64

__host__ __device__ inline int bx_rup(int x, int m) { return (x + m - 1) / m * m; }

__host__ __device__ inline size_t box_lds_bytes(int ch, int cw) {
    const int IR = BX_TR + ch - 1, IC = BX_TC + cw - 1;
    return (size_t)bx_rup(IC * bx_rup(IR, 4), 16) + (size_t)BX_TR * IC * 8;
}

__global__ __launch_bounds__(256) void box_u8_kernel(const uint8_t* img, int64_t ny, int64_t nx,
                                                    int64_t pady, int64_t padx, int ch, int cw,
                                                    int64_t HB, int64_t WB, int32_t* B1,
                                                    int32_t* B2) {
    extern __shared__ __align__(16) unsigned char sm[];
    const int IR = BX_TR + ch - 1, IC = BX_TC + cw - 1;
    const int ldi = bx_rup(IR, 4);
    uint8_t* in = sm;
    int32_t* V1 = (int32_t*)(sm + bx_rup(IC * ldi, 16));   // V[c * BX_TR + r]
    int32_t* V2 = V1 + BX_TR * IC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t ntr = (HB + BX_TR - 1) / BX_TR, ntc = (WB + BX_TC - 1) / BX_TC;
    for (int64_t tile = blockIdx.x; tile < ntr * ntc; tile += gridDim.x) {
        const int64_t r0 = (tile % ntr) * BX_TR, c0 = (tile / ntr) * BX_TC;
        // input bytes of the tile, zero outside the real image
        for (int c = wave; c < IC; c += 4) {
            const int64_t col = c0 + c - padx;
            const bool colok = col >= 0 && col < nx;
            for (int r = lane; r < IR; r += 64) {
                const int64_t row = r0 + r - pady;
                in[c * ldi + r] = (colok && row >= 0 && row < ny) ? img[row + col * ny] : 0;
            }
        }
        __syncthreads();
        // vertical sums over ch rows: task (column c, rows 8k .. 8k + 7)
        for (int t = tid; t < IC * (BX_TR / 8); t += 256) {
            const int c = t / (BX_TR / 8), k = t - c * (BX_TR / 8);
            const uint8_t* col = in + c * ldi + 8 * k;
            int32_t a1 = 0, a2 = 0;
            for (int u = 0; u < ch; u++) {
                int32_t w = col[u];
                a1 += w;
                a2 += w * w;
            }
            int32_t* v1 = V1 + c * BX_TR + 8 * k;
            int32_t* v2 = V2 + c * BX_TR + 8 * k;
            for (int r = 0; r < 8; r++) {
                v1[r] = a1;
                v2[r] = a2;
                if (r + 1 < 8) {
                    int32_t wo = col[r], wi = col[r + ch];
                    a1 += wi - wo;
                    a2 += wi * wi - wo * wo;
                }
            }
        }
        __syncthreads();
        // horizontal sums over cw columns: task (row r, columns 8k .. 8k + 7)
        {
            const int r = tid & (BX_TR - 1), k = tid / BX_TR;   // 256 tasks = 32 rows x 8
            int32_t a1 = 0, a2 = 0;
            for (int v = 0; v < cw; v++) {
                a1 += V1[(8 * k + v) * BX_TR + r];
                a2 += V2[(8 * k + v) * BX_TR + r];
            }
            const int64_t R = r0 + r;
            for (int c = 0; c < 8; c++) {
                const int64_t C = c0 + 8 * k + c;
                if (R < HB && C < WB) {
                    B1[R + C * HB] = a1;
                    B2[R + C * HB] = a2;
                }
                if (c + 1 < 8) {
                    a1 += V1[(8 * k + c + cw) * BX_TR + r] - V1[(8 * k + c) * BX_TR + r];
                    a2 += V2[(8 * k + c + cw) * BX_TR + r] - V2[(8 * k + c) * BX_TR + r];
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace ar

using namespace ar;

ar_status ar_box_images(ar_context* ctx, int ch, int cw, int64_t pady, int64_t padx,
                        int32_t** b1, int32_t** b2, int64_t* ld) {
    const int64_t NY = ctx->ny + 2 * pady, NX = ctx->nx + 2 * padx;
    const int64_t HB = NY - ch + 1, WB = NX - cw + 1;
    if (HB < 1 || WB < 1) return AR_ERR_UNSUPPORTED;
    const size_t each = (size_t)HB * WB * 4;
    ar_status st = ctx->box.reserve(ctx, 2 * each);
    if (st) return st;
    int32_t* p1 = (int32_t*)ctx->box.ptr;
    int32_t* p2 = (int32_t*)((char*)ctx->box.ptr + each);
    const size_t lds = box_lds_bytes(ch, cw);
    if (hipFuncSetAttribute((const void*)box_u8_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
        return AR_ERR_DEVICE;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, box_u8_kernel, 256, lds) !=
            hipSuccess ||
        per_cu < 1)
        return AR_ERR_DEVICE;
    const int64_t tiles = ((HB + BX_TR - 1) / BX_TR) * ((WB + BX_TC - 1) / BX_TC);
    int64_t blocks = (int64_t)ctx->ncu * per_cu;
    if (blocks > tiles) blocks = tiles;
    hipLaunchKernelGGL(box_u8_kernel, dim3((unsigned)blocks), dim3(256), lds, ctx->stream,
                       (const uint8_t*)ctx->pref.ptr, ctx->ny, ctx->nx, pady, padx, ch, cw, HB, WB,
                       p1, p2);
    if (hipGetLastError() != hipSuccess) return AR_ERR_DEVICE;
    *b1 = p1;
    *b2 = p2;
    *ld = HB;
    return AR_OK;
}
