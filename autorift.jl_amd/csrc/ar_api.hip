// Host side of the C ABI: context lifecycle, device-resident image pairs, pass geometry and the
// dispatch of the per-point tracker. See include/autorift_hip.h for the contract of each entry.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/autorift_hip.h"
#include "ar_internal.h"
#include "layout.h"
#include "refine.h"
#include "track_common.h"

namespace ar {

// Defined in track.hip.
template <bool BATCH, int OCC, bool SG>
__global__ void track_u8_kernel(TrackArgs a);
__global__ void refine_kernel(TrackArgs a);
__global__ void track_fft_kernel(TrackArgs a);
template <int IS_U8>
__global__ void track_generic_kernel(TrackArgs a);
template <bool LDS>
__global__ void track_coherence_kernel(TrackArgs a);
// Defined in track_f64.hip.
template <int IS_U8, int MODE>
__global__ void f64_prep_kernel(TrackArgs a, F64Stage S);
template <int IS_U8>
__global__ void track_f64_kernel(TrackArgs a, F64Stage S, int pitch, const double* stage);

}  // namespace ar

using namespace ar;

static std::mutex g_err_mu;
static std::string g_err;   // errors with no context (create)

static const size_t kLdsLimit = AR_LDS_LIMIT;

static ar_status fail(ar_context* ctx, ar_status st, const std::string& msg) {
    if (ctx) {
        ctx->err = msg;
    } else {
        std::lock_guard<std::mutex> lk(g_err_mu);
        g_err = msg;
    }
    return st;
}

#define HIPCHK(ctx, expr)                                                                    \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess)                                                                \
            return fail((ctx), _e == hipErrorOutOfMemory ? AR_ERR_OUT_OF_MEMORY : AR_ERR_DEVICE, \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                  \
    } while (0)

// ---------------------------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------------------------

ar_status DevBuf::reserve(ar_context* ctx, size_t n) {
    if (n <= cap) return AR_OK;
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    // 256 bytes of slack: kernels may read a few bytes past the last element with aligned
    // dword loads (never used, never faulting)
    size_t want = (n < 256 ? 256 : n) + 256;
    HIPCHK(ctx, hipMalloc(&ptr, want));
    cap = want - 256;
    return AR_OK;
}

void DevBuf::release() {
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

// ---------------------------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------------------------

// Combined validity (src/preprocess.jl:106) and an "every pixel valid" flag.
__global__ void combine_masks_kernel(const uint8_t* a, const uint8_t* b, uint8_t* out, int64_t n,
                                     int* all_valid) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int bad = 0;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint8_t v = (a[i] && b[i]) ? 1 : 0;
        out[i] = v;
        bad |= !v;
    }
    if (__syncthreads_or(bad) && threadIdx.x == 0) atomicAnd(all_valid, 0);
}

// isfinite mask of a float image (src/preprocess.jl:74-79).
__global__ void finite_mask_kernel(const float* img, uint8_t* out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) out[i] = isfinite(img[i]) ? 1 : 0;
}

__global__ void fill_u8_kernel(uint8_t* out, uint8_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) out[i] = v;
}

// _pass_geometry (src/track.jl:176-197) plus the facts the dispatcher needs. Ten results, each
// in its own 256-byte line of g (slot k at g[AR_GEOM_SLOT * k]) so that the blocks' atomics on
// different results do not queue behind each other: 0..5 = cx, cy, rx, ry, px, py (max), 6..7 =
// the smallest even chip extents (min), 8 = every window fits, 9 = every searchable chip area a
// power of two (and). Each block reduces its points first and issues one atomic per result.
#define AR_GEOM_SLOT 32
__global__ __launch_bounds__(256) void pass_geometry_kernel(ar_points pts, int64_t n, int64_t ny,
                                                            int64_t nx, unsigned long long* g,
                                                            const PairDesc* pairs,
                                                            const int32_t* pair_of) {
    __shared__ unsigned long long red[4][10];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long v[10] = {0, 0, 0, 0, 0, 0, ~0ULL, ~0ULL, 1, 1};
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t prx = pts.radius_x[i], pry = pts.radius_y[i];
        if (!(prx > 0 && pry > 0)) continue;
        int64_t csx = pts.chip_size_x[i], csy = pts.chip_size_y[i];
        v[0] = max(v[0], (unsigned long long)(csx < 0 ? 0 : csx));
        v[1] = max(v[1], (unsigned long long)(csy < 0 ? 0 : csy));
        v[2] = max(v[2], (unsigned long long)prx);
        v[3] = max(v[3], (unsigned long long)pry);
        int64_t qx = csx / 2 + prx + (int64_t)ceil(fabs(pts.dx_prior[i]));
        int64_t qy = csy / 2 + pry + (int64_t)ceil(fabs(pts.dy_prior[i]));
        v[4] = max(v[4], (unsigned long long)(qx < 0 ? 0 : qx));
        v[5] = max(v[5], (unsigned long long)(qy < 0 ? 0 : qy));
        v[6] = min(v[6], (unsigned long long)(csx < 0 ? 0 : 2 * (csx / 2)));
        v[7] = min(v[7], (unsigned long long)(csy < 0 ? 0 : 2 * (csy / 2)));
        int64_t hx = csx / 2, hy = csy / 2;
        int64_t fx = (int64_t)floor(pts.x[i]), fy = (int64_t)floor(pts.y[i]);
        int64_t r0 = fy - hy - pry, r1 = fy + hy + pry - 2;
        int64_t c0 = fx - hx - prx, c1 = fx + hx + prx - 2;
        const int64_t pny = pairs ? pairs[pair_of[i]].ny : ny;   // batched: the point's pair
        const int64_t pnx = pairs ? pairs[pair_of[i]].nx : nx;
        if (!(r0 >= 1 && r1 <= pny && c0 >= 1 && c1 <= pnx)) v[8] = 0;
        int64_t area = (2 * hy) * (2 * hx);
        if (!(area > 0 && (area & (area - 1)) == 0)) v[9] = 0;
    }
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int k = 0; k < 10; k++) {
            const unsigned long long u = __shfl_xor(v[k], o);
            v[k] = k < 6 ? max(v[k], u) : (k < 8 ? min(v[k], u) : (v[k] & u));
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 10; k++) red[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 10) {
        const int k = threadIdx.x;
        unsigned long long r = red[0][k];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++)
            r = k < 6 ? max(r, red[w][k]) : (k < 8 ? min(r, red[w][k]) : (r & red[w][k]));
        unsigned long long* slot = g + AR_GEOM_SLOT * k;
        if (k < 6) atomicMax(slot, r);
        else if (k < 8) atomicMin(slot, r);
        else if (r == 0) atomicAnd(slot, 0ULL);
    }
}

// Per-point geometry records of a pass (point_geom once per point instead of once per wave).
__global__ __launch_bounds__(256) void point_prep_kernel(TrackArgs a) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; p < a.n; p += (int64_t)gridDim.x * blockDim.x) {
        int32_t r[AR_PREC];
        if (a.pairs) {   // batched pass: bounds against the point's own pair
            TrackArgs b = a;
            use_pair(b, p);
            point_record(point_geom(b, p), r);
        } else {
            point_record(point_geom(a, p), r);
        }
        int4* q = (int4*)(a.prec + AR_PREC * p);
        q[0] = make_int4(r[0], r[1], r[2], r[3]);
        q[1] = make_int4(r[4], r[5], r[6], r[7]);
    }
}

__global__ void skip_all_kernel(float* dx, float* dy, float* corr, uint8_t* s, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        dx[i] = __int_as_float(0x7fc00000);
        dy[i] = __int_as_float(0x7fc00000);
        corr[i] = __int_as_float(0x7fc00000);
        s[i] = 0;
    }
}

// AR_BOX_IMAGES=always|never overrides the choice between box images and per-point box sums
// (both exact); unset or anything else: automatic.
static int box_images_mode() {
    const char* e = std::getenv("AR_BOX_IMAGES");
    if (!e) return 0;
    if (std::strcmp(e, "always") == 0) return 1;
    if (std::strcmp(e, "never") == 0) return -1;
    return 0;
}

// AR_F64_DIRECT=never sends the passes the direct Float64 tracker would take to the one-kernel
// generic tracker instead (the same results; kept for comparisons). Unset or anything else: the
// direct tracker wherever its window fits in LDS.
static bool f64_direct_enabled() {
    const char* e = std::getenv("AR_F64_DIRECT");
    return !(e && std::strcmp(e, "never") == 0);
}

// With precision BF16 / FP16, passes of fewer multiply-adds per point than this take the exact
// direct Float64 tracker instead: it measured faster there on MI355X (chip 16 r 8: 29 against 21
// M points/s; chip 24 r 12: 16.5 against 15.8; chip 16 r 25: 27 against 31 ms for the 10000^2
// search), and exact results are within any tolerance of the 16-bit form. AR_MP_MIN_MACS
// (environment, read per pass) sets the threshold; 0 sends every such pass to the 16-bit tracker.
//
// The 16-bit pass also builds Float64 box images over every chip position of the padded
// reference (about 26 ps a position: 2.6 ms at 10000^2), which the exact tracker does without.
// On references of at least 4 M positions the pass's work must repay them: measured at chip 32,
// r 25 on the 10000^2 pair, the 16-bit contraction saves about 54 fs per multiply-add, so it
// needs searched points x multiply-adds >= 512 x positions (a multi-chip level holding a few
// hundred unfinished points of a 10000^2 search fails this; the full 168 100-point pass clears
// it eightfold). Setting AR_MP_MIN_MACS turns this second test off. The FFT tracker builds the
// same images and saves about as much per multiply-add (45 fs), so its passes take the same test.
static bool box_images_repaid(double macs, int64_t nsearch, double positions) {
    return positions < 4194304.0 || (double)nsearch * macs >= 512.0 * positions;
}
static bool mp_worth_it(double macs, int64_t nsearch, double positions) {
    const char* e = std::getenv("AR_MP_MIN_MACS");
    if (e && *e) {
        char* end = nullptr;
        const double v = std::strtod(e, &end);
        if (end && *end == 0 && v >= 0) return macs >= v;
    }
    return macs >= 8.0e5 && box_images_repaid(macs, nsearch, positions);
}

// Staging bytes the direct Float64 tracker uses per chunk of points.
static const size_t kF64StageBytes = (size_t)2 << 30;

// The direct Float64 tracker's launch shape for a pass: window LDS bytes, column pitch and
// threads per workgroup (one lane per 8-row tile of shifts); false when it does not apply.
static bool f64_shape(const TrackArgs& a, size_t* lds, int* pitch, int* threads) {
    const int p = f64_pitch(a.max_ch, a.max_ry);
    const int winw = a.max_cw + 2 * a.max_rx - 1;
    const size_t wl = (size_t)winw * p * 4;
    const int tiles = (2 * a.max_ry + 7) / 8 * (2 * a.max_rx);
    const int t = (tiles + 63) / 64 * 64;
    const size_t surf = (size_t)(2 * a.max_ry) * (2 * a.max_rx) * 4;   // reuses the window's LDS
    *lds = wl > surf ? wl : surf;
    *pitch = p;
    *threads = t;
    return *lds <= kLdsLimit && t <= 1024;
}

// Four int8 copies of a UInt8 image (value ^ 0x80), copy k shifted k bytes later
// (copy_k[y] = ref[y - k] ^ 0x80, int8 zero outside), so any image offset starts a dword in one of
// them.
// Copy k (at dst + k stride; stride a multiple of 16) holds src[y - k] ^ 0x80 at position y, and
// 0x80 where y - k falls outside [0, n). A thread forms 16 bytes of every copy from one 16-byte
// load and the dword before it (copy k's dword d is bytes 4 - k .. 7 - k of dwords d - 1, d of
// the source block, one alignbyte); the first block and blocks reaching past n take their bytes
// one at a time.
__global__ void refx_kernel(const uint8_t* src, uint8_t* dst, int64_t n, int64_t stride) {
    const int64_t nv = stride >> 4;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t y0 = v << 4;
        uint32_t o[4][4];
        if (y0 >= 16 && y0 + 16 <= n) {
            const uint4 q = *(const uint4*)(src + y0);
            const uint32_t c[5] = {*(const uint32_t*)(src + y0 - 4), q.x, q.y, q.z, q.w};
#pragma unroll
            for (int d = 0; d < 4; d++) {
                o[0][d] = c[d + 1] ^ 0x80808080u;
#pragma unroll
                for (int k = 1; k < 4; k++)
                    o[k][d] = __builtin_amdgcn_alignbyte(c[d + 1], c[d], 4 - k) ^ 0x80808080u;
            }
        } else {
            for (int k = 0; k < 4; k++)
                for (int d = 0; d < 4; d++) {
                    uint32_t w = 0;
                    for (int i = 0; i < 4; i++) {
                        const int64_t y = y0 + 4 * d + i;
                        const uint32_t b = (y >= k && y - k < n) ? (uint32_t)(src[y - k] ^ 0x80)
                                                                 : 0x80u;
                        w |= b << (8 * i);
                    }
                    o[k][d] = w;
                }
        }
#pragma unroll
        for (int k = 0; k < 4; k++)
            *(uint4*)(dst + k * stride + y0) = make_uint4(o[k][0], o[k][1], o[k][2], o[k][3]);
    }
}

// Every kernel that takes dynamic LDS gets the largest allowance once per process (the
// attribute is process-wide, so lowering it for one context could break another's launch), and
// its occupancy for a given size is cached per context: both calls cost host time on every pass
// otherwise.
static std::mutex g_attr_mu;
static std::set<const void*> g_attr_set;

static ar_status lds_allowance(ar_context* ctx, const void* kern) {
    std::lock_guard<std::mutex> lk(g_attr_mu);
    if (!g_attr_set.count(kern)) {
        HIPCHK(ctx, hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)kLdsLimit));
        g_attr_set.insert(kern);
    }
    return AR_OK;
}

static ar_status kernel_occupancy(ar_context* ctx, const void* kern, int threads, size_t lds,
                                  int* per_cu) {
    const auto key = std::make_tuple(kern, threads, lds);
    auto it = ctx->occupancy.find(key);
    if (it != ctx->occupancy.end()) {
        *per_cu = it->second;
        return AR_OK;
    }
    if (ar_status st = lds_allowance(ctx, kern)) return st;
    if (lds > kLdsLimit) {
        *per_cu = 0;
        return AR_OK;
    }
    HIPCHK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, kern, threads, lds));
    ctx->occupancy[key] = *per_cu;
    return AR_OK;
}

static int grid_for(int64_t n, int per_block = 256) {
    int64_t b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > 4096) b = 4096;
    return (int)b;
}

// ---------------------------------------------------------------------------------------------
// ABI
// ---------------------------------------------------------------------------------------------

extern "C" {

void ar_params_default(ar_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->preprocess = AR_PRE_HIGHPASS;
    p->filter_width = 5;
    p->quantize = AR_QUANT_UINT8;
    p->measure = AR_ZNCC;
    p->upsampling = 64;
    p->outliers = 1;
    p->gardner_window = 5;
    p->gardner_iterations = 3;
    p->min_agree_fraction = 8.0 / 25.0;
    p->agree_tolerance = 0.2;
    p->mad_scale = 4.0;
    p->chip_size_base = 32;
    p->chip_size_min = 32;
    p->chip_size_max = 128;
    p->chip_aspect = 1.0;
    p->grid_spacing = 32;
    p->search_radius_x = 25;
    p->search_radius_y = 25;
    p->min_search_radius = 6;
    p->coarse_stride = 4;
    p->coarse_buffer = 8;
    p->min_coarse_valid_fraction = 0.01;
    p->dx_prior = 0.0;
    p->dy_prior = 0.0;
    p->fill_window = 3;
}

const char* ar_version(void) { return "autorift_hip 0.1.0 (gfx950)"; }

const char* ar_last_error(const ar_context* ctx) {
    if (ctx) return ctx->err.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_err.c_str();
}

ar_status ar_context_create(const ar_params* p, int32_t device, ar_context** out) {
    if (!out) return fail(nullptr, AR_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, AR_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= ndev)
        return fail(nullptr, AR_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        return fail(nullptr, AR_ERR_DEVICE, "hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, AR_ERR_NO_DEVICE,
                    std::string("device is ") + prop.gcnArchName + ", this build targets gfx950");
    ar_context* c = new ar_context();
    if (p) c->params = *p;
    else ar_params_default(&c->params);
    c->device = device;
    c->ncu = prop.multiProcessorCount;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->evm) != hipSuccess || hipEventCreate(&c->evb) != hipSuccess) {
        delete c;
        return fail(nullptr, AR_ERR_DEVICE, "stream/event creation failed");
    }
    *out = c;
    return AR_OK;
}

void ar_context_destroy(ar_context* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (DevBuf* b : ctx->all_bufs()) b->release();
    if (ctx->stream2) {
        hipStreamSynchronize(ctx->stream2);
        for (int k = 0; k < 5; k++) hipEventDestroy(ctx->evf[k]);
        hipStreamDestroy(ctx->stream2);
    }
    if (ctx->pin) hipHostFree(ctx->pin);
    hipEventDestroy(ctx->ev0);
    hipEventDestroy(ctx->ev1);
    if (ctx->evm) hipEventDestroy(ctx->evm);
    if (ctx->evb) hipEventDestroy(ctx->evb);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // extern "C"

ar_status ar_pinned(ar_context* ctx, size_t n, void** out) {
    if (n > ctx->pin_cap) {
        if (ctx->pin) {
            if (hipStreamSynchronize(ctx->stream) != hipSuccess)
                return fail(ctx, AR_ERR_DEVICE, "stream synchronisation failed");
            hipHostFree(ctx->pin);
            ctx->pin = nullptr;
            ctx->pin_cap = 0;
        }
        const size_t cap = (n + 65535) & ~(size_t)65535;
        if (hipHostMalloc(&ctx->pin, cap, hipHostMallocDefault) != hipSuccess) {
            ctx->pin = nullptr;
            return fail(ctx, AR_ERR_OUT_OF_MEMORY, "page-locked staging allocation failed");
        }
        ctx->pin_cap = cap;
    }
    *out = ctx->pin;
    return AR_OK;
}

// Device bytes the context holds.
size_t ar_context_bytes(ar_context* ctx) {
    size_t t = 0;
    for (DevBuf* b : ctx->all_bufs()) t += b->cap;
    return t;
}

// Gives every device buffer back; the context keeps its stream and parameters and needs a new
// pair before the next pass.
void ar_context_release(ar_context* ctx) {
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (DevBuf* b : ctx->all_bufs()) b->release();
    ctx->has_pair = false;
    ctx->refx_gen = ~0ULL;
    ctx->box_key.ptr = nullptr;
    ctx->absmax_gen = ~0ULL;
}

// Upload masks (or derive them) and combine; sets ctx->all_valid.
__global__ void set_flag_kernel(int* f, int v) {
    if (threadIdx.x == 0) *f = v;
}

ar_status ar_upload_masks(ar_context* ctx, const uint8_t* ref_valid, const uint8_t* sec_valid,
                          int64_t N) {
    ar_status st;
    if ((st = ctx->rmask.reserve(ctx, N)) || (st = ctx->smask.reserve(ctx, N)) ||
        (st = ctx->okmask.reserve(ctx, N)) || (st = ctx->flags.reserve(ctx, 64)))
        return st;
    // ref/sec validity come from the caller, else from the prepared images (already written to
    // rmask/smask by the preparation step when the caller passed NULL).
    if (ref_valid) HIPCHK(ctx, hipMemcpyAsync(ctx->rmask.ptr, ref_valid, N, hipMemcpyHostToDevice, ctx->stream));
    if (sec_valid) HIPCHK(ctx, hipMemcpyAsync(ctx->smask.ptr, sec_valid, N, hipMemcpyHostToDevice, ctx->stream));
    set_flag_kernel<<<1, 64, 0, ctx->stream>>>((int*)ctx->flags.ptr, 1);
    combine_masks_kernel<<<grid_for(N), 256, 0, ctx->stream>>>(
        (const uint8_t*)ctx->rmask.ptr, (const uint8_t*)ctx->smask.ptr, (uint8_t*)ctx->okmask.ptr, N,
        (int*)ctx->flags.ptr);
    HIPCHK(ctx, hipGetLastError());
    void* pin = nullptr;
    if ((st = ar_pinned(ctx, 256, &pin))) return st;
    HIPCHK(ctx, hipMemcpyAsync(pin, ctx->flags.ptr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->all_valid = *(const int*)pin != 0;
    return AR_OK;
}

extern "C" ar_status ar_set_pair(ar_context* ctx, const void* ref, const void* sec,
                                 const uint8_t* ref_valid, const uint8_t* sec_valid, int64_t ny,
                                 int64_t nx, int32_t dtype) {
    if (!ctx) return fail(nullptr, AR_ERR_INVALID_ARGUMENT, "context is NULL");
    if (!ref || !sec) return fail(ctx, AR_ERR_INVALID_ARGUMENT, "both images are required");
    if (ny <= 0 || nx <= 0) return fail(ctx, AR_ERR_INVALID_ARGUMENT, "image size must be positive");
    if (dtype != AR_U8 && dtype != AR_F32)
        return fail(ctx, AR_ERR_UNSUPPORTED, "prepared pairs must be UInt8 or Float32");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int64_t N = ny * nx;
    size_t es = dtype == AR_U8 ? 1 : 4;
    ar_status st;
    if ((st = ctx->pref.reserve(ctx, N * es)) || (st = ctx->psec.reserve(ctx, N * es)) ||
        (st = ctx->rmask.reserve(ctx, N)) || (st = ctx->smask.reserve(ctx, N)))
        return st;
    HIPCHK(ctx, hipMemcpyAsync(ctx->pref.ptr, ref, N * es, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->psec.ptr, sec, N * es, hipMemcpyHostToDevice, ctx->stream));
    ctx->ref_gen++;
    ctx->ny = ny;
    ctx->nx = nx;
    ctx->pdtype = dtype;
    ctx->has_pair = true;
    // default masks: isfinite (all true for integer images)
    if (!ref_valid) {
        if (dtype == AR_U8) fill_u8_kernel<<<grid_for(N), 256, 0, ctx->stream>>>((uint8_t*)ctx->rmask.ptr, 1, N);
        else finite_mask_kernel<<<grid_for(N), 256, 0, ctx->stream>>>((const float*)ctx->pref.ptr, (uint8_t*)ctx->rmask.ptr, N);
    }
    if (!sec_valid) {
        if (dtype == AR_U8) fill_u8_kernel<<<grid_for(N), 256, 0, ctx->stream>>>((uint8_t*)ctx->smask.ptr, 1, N);
        else finite_mask_kernel<<<grid_for(N), 256, 0, ctx->stream>>>((const float*)ctx->psec.ptr, (uint8_t*)ctx->smask.ptr, N);
    }
    HIPCHK(ctx, hipGetLastError());
    return ar_upload_masks(ctx, ref_valid, sec_valid, N);
}

extern "C" ar_status ar_get_pair(ar_context* ctx, void* ref, void* sec, uint8_t* ref_valid,
                                 uint8_t* sec_valid, int32_t* dtype) {
    if (!ctx) return fail(nullptr, AR_ERR_INVALID_ARGUMENT, "context is NULL");
    if (!ctx->has_pair) return fail(ctx, AR_ERR_INVALID_ARGUMENT, "no image pair set");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    int64_t N = ctx->ny * ctx->nx;
    size_t es = ctx->pdtype == AR_U8 ? 1 : (ctx->pdtype == AR_C64 ? 8 : 4);
    if (ref) HIPCHK(ctx, hipMemcpyAsync(ref, ctx->pref.ptr, N * es, hipMemcpyDeviceToHost, ctx->stream));
    if (sec) HIPCHK(ctx, hipMemcpyAsync(sec, ctx->psec.ptr, N * es, hipMemcpyDeviceToHost, ctx->stream));
    if (ref_valid) HIPCHK(ctx, hipMemcpyAsync(ref_valid, ctx->rmask.ptr, N, hipMemcpyDeviceToHost, ctx->stream));
    if (sec_valid) HIPCHK(ctx, hipMemcpyAsync(sec_valid, ctx->smask.ptr, N, hipMemcpyDeviceToHost, ctx->stream));
    if (dtype) *dtype = ctx->pdtype;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return AR_OK;
}

// Launch plan of the direct Float64 tracker over a.n points (a.order, when set, lists them):
// kernels, LDS sizes, chunking, and the staging buffer, reserved here so that f64_enqueue makes no
// allocation and no runtime query (it may run under graph capture). The points go in chunks whose
// staging slots fit half of kF64StageBytes; each chunk is a preparation launch (persistent, one
// point per workgroup at a time) and a tracking launch (one workgroup per point). With `overlap`
// and more than one chunk the preparations run on the context's second stream, one chunk ahead of
// the tracking launches, over two alternating halves of the staging buffer; events order the two
// streams.
struct F64Plan {
    void (*fprep)(TrackArgs, F64Stage);
    void (*ftrack)(TrackArgs, F64Stage, int, const double*);
    size_t lprep, wl;
    int pitch, threads;
    int64_t chunk, nprep;
    bool overlap;
    int nbuf;
    size_t half, tbytes;
    F64Slot sl;
    size_t td;
};

// Kernels, LDS sizes and slot layout of the plan (no runtime call).
static void f64_plan_shape(const TrackArgs& a, bool u8, F64Plan* P) {
    f64_shape(a, &P->wl, &P->pitch, &P->threads);
    const int winh = a.max_ch + 2 * a.max_ry - 1, winw = a.max_cw + 2 * a.max_rx - 1;
    P->td = (size_t)(winh + 1) * (winw + 1);
    // the preparation kernel's LDS: chip, then the window (mode 2) and the table (modes 1, 2)
    const size_t lchip = ((size_t)a.max_ch * a.max_cw * 4 + 15) / 16 * 16;
    const size_t lwin = ((size_t)winh * winw * 4 + 15) / 16 * 16;
    const size_t td8 = P->td * 8;
    const int mode = lchip + lwin + td8 <= kLdsLimit ? 2 : (lchip + td8 <= kLdsLimit ? 1 : 0);
    if (mode == 2) P->fprep = u8 ? f64_prep_kernel<1, 2> : f64_prep_kernel<0, 2>;
    else if (mode == 1) P->fprep = u8 ? f64_prep_kernel<1, 1> : f64_prep_kernel<0, 1>;
    else P->fprep = u8 ? f64_prep_kernel<1, 0> : f64_prep_kernel<0, 0>;
    P->lprep = lchip + (mode == 2 ? lwin : 0) + (mode >= 1 ? td8 : 0);
    P->ftrack = u8 ? track_f64_kernel<1> : track_f64_kernel<0>;
    P->sl = f64_slot(a.max_ch, a.max_cw, a.max_rx, a.max_ry);
    P->tbytes = mode >= 1 ? 0 : td8;   // one block's table; scaled by the launch plan
}

ar_status ar_stream2(ar_context* ctx) {
    if (ctx->stream2) return AR_OK;
    HIPCHK(ctx, hipStreamCreate(&ctx->stream2));
    for (int k = 0; k < 5; k++)
        HIPCHK(ctx, hipEventCreateWithFlags(&ctx->evf[k], hipEventDisableTiming));
    return AR_OK;
}

static ar_status f64_plan(ar_context* ctx, const TrackArgs& a, bool u8, bool allow_overlap,
                          F64Plan* P) {
    ar_status st;
    f64_plan_shape(a, u8, P);
    const size_t table1 = P->tbytes;
    int64_t chunk = (int64_t)(kF64StageBytes / 2 / P->sl.bytes);
    if (chunk < 1) chunk = 1;
    P->overlap = allow_overlap && a.n > chunk;
    if (chunk > a.n) chunk = a.n;
    P->chunk = chunk;
    int pa = 0, pb = 0;
    if ((st = kernel_occupancy(ctx, (const void*)P->fprep, F64_PREP_WG, P->lprep, &pa))) return st;
    if ((st = kernel_occupancy(ctx, (const void*)P->ftrack, P->threads, P->wl, &pb))) return st;
    if (pa < 1 || pb < 1) return fail(ctx, AR_ERR_UNSUPPORTED, "tracker kernel does not fit on a CU");
    // Running beside the tracker, the preparation takes up to three workgroups per CU: its
    // phases are latency-bound, so co-resident workgroups hide each other's waits. Measured on the
    // 10000^2 search at chip 16, r 25 (preparation-bound): 26.9 ms with one, 24.3 with two, 23.1
    // with three or more; at chip 32, r 25 (tracker-bound) 34.8 ms throughout.
    P->nprep = (int64_t)ctx->ncu * (P->overlap ? std::min(3, pa) : pa);
    if (P->nprep > chunk) P->nprep = chunk;
    P->nbuf = P->overlap ? 2 : 1;
    P->half = (size_t)chunk * P->sl.bytes;
    P->tbytes = (size_t)P->nbuf * P->nprep * table1;
    if ((st = ctx->scratch.reserve(ctx, P->nbuf * P->half + P->tbytes))) return st;
    if (P->overlap) return ar_stream2(ctx);
    return AR_OK;
}

static ar_status f64_enqueue(ar_context* ctx, const TrackArgs& a, const F64Plan& P,
                             unsigned char* base) {
    F64Stage S;
    S.slot_bytes = P.sl.bytes;
    S.off_chip = P.sl.off_chip;
    S.off_b1 = P.sl.off_b1;
    S.off_b2 = P.sl.off_b2;
    S.chp = P.sl.chp;
    S.table_doubles = P.td;
    hipStream_t ps = P.overlap ? ctx->stream2 : ctx->stream;
    if (P.overlap) {
        // evf[0]: everything before the pass; evf[1 + b]: preparation into half b done;
        // evf[3 + b]: tracking from half b done
        HIPCHK(ctx, hipEventRecord(ctx->evf[0], ctx->stream));
        HIPCHK(ctx, hipStreamWaitEvent(ctx->stream2, ctx->evf[0], 0));
    }
    int64_t c = 0;
    for (int64_t p0 = 0; p0 < a.n; p0 += P.chunk, c++) {
        const int h = (int)(c & 1) % P.nbuf;
        S.p0 = p0;
        S.cnt = a.n - p0 < P.chunk ? a.n - p0 : P.chunk;
        S.slots = base + (size_t)h * P.half;
        S.tables = (double*)(base + P.nbuf * P.half + (size_t)h * (P.tbytes / P.nbuf));
        if (P.overlap && c >= 2)
            HIPCHK(ctx, hipStreamWaitEvent(ctx->stream2, ctx->evf[3 + h], 0));
        const int64_t gp = P.nprep < S.cnt ? P.nprep : S.cnt;
        hipLaunchKernelGGL(P.fprep, dim3((unsigned)gp), dim3(F64_PREP_WG), P.lprep, ps, a, S);
        HIPCHK(ctx, hipGetLastError());
        if (P.overlap) {
            HIPCHK(ctx, hipEventRecord(ctx->evf[1 + h], ctx->stream2));
            HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, ctx->evf[1 + h], 0));
        }
        const int64_t gt = 8 * ((S.cnt + 7) / 8);
        hipLaunchKernelGGL(P.ftrack, dim3((unsigned)gt), dim3(P.threads), P.wl, ctx->stream, a, S,
                           P.pitch, (const double*)S.slots);
        HIPCHK(ctx, hipGetLastError());
        if (P.overlap) HIPCHK(ctx, hipEventRecord(ctx->evf[3 + h], ctx->stream));
    }
    return AR_OK;
}

// The pass on device-resident points. All pointers device pointers.
ar_status ar_track_pass(ar_context* ctx, const ar_points& dp, int64_t n, int64_t up,
                        int32_t measure, float* dx, float* dy, float* corr, uint8_t* srch,
                        const ar_pass_hint* hint, int64_t searchable_max, bool reuse_box) {
    if (!ctx->has_pair) return fail(ctx, AR_ERR_INVALID_ARGUMENT, "no image pair set");
    if (n < 0) return fail(ctx, AR_ERR_INVALID_ARGUMENT, "negative point count");
    if (up < 1 || up > 64 || (up & (up - 1)) != 0)
        return fail(ctx, AR_ERR_UNSUPPORTED,
                    "upsampling must be a power of two in 1..64 on the device path");
    if (measure != AR_ZNCC && measure != AR_NCC && measure != AR_COHERENCE)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "measure must be ZNCC, NCC or Coherence");
    // Coherence exploits the phase of a complex pair; the real measures need real images
    // (src/correlate.jl:739-748)
    const bool cplx = ctx->pdtype == AR_C64;
    if (measure == AR_COHERENCE && !cplx)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT,
                    "`Coherence` needs complex input, but the images are real. Coherence exploits "
                    "the phase of an SLC pair; a real image has none, so there is nothing for it "
                    "to measure. Use ZNCC for real imagery, or pass the complex data.");
    if (measure != AR_COHERENCE && cplx)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT,
                    "ZNCC and NCC need real images; use Coherence for a complex pair");
    if (cplx && ctx->params.rotation_count > 0)
        return fail(ctx, AR_ERR_UNSUPPORTED, "rotation search needs real images");
    if (n == 0) return AR_OK;
    ar_status st;
    unsigned long long g[10];
    if (hint) {
        // the caller's bound: one chip size, every point's chip area that of the hint
        const int64_t ex = 2 * (hint->cx / 2), ey = 2 * (hint->cy / 2), area = ex * ey;
        g[0] = (unsigned long long)hint->cx;
        g[1] = (unsigned long long)hint->cy;
        g[2] = (unsigned long long)hint->rx;
        g[3] = (unsigned long long)hint->ry;
        g[4] = (unsigned long long)hint->px;
        g[5] = (unsigned long long)hint->py;
        g[6] = (unsigned long long)ex;
        g[7] = (unsigned long long)ey;
        g[8] = hint->fits ? 1 : 0;
        g[9] = (area > 0 && (area & (area - 1)) == 0) ? 1 : 0;
    } else {
        unsigned long long gs[10 * AR_GEOM_SLOT] = {0};
        if ((st = ctx->geom.reserve(ctx, sizeof(gs)))) return st;
        for (int k = 6; k < 10; k++) gs[AR_GEOM_SLOT * k] = k < 8 ? ~0ULL : 1ULL;
        HIPCHK(ctx, hipMemcpyAsync(ctx->geom.ptr, gs, sizeof(gs), hipMemcpyHostToDevice,
                                   ctx->stream));
        pass_geometry_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(
            dp, n, ctx->ny, ctx->nx, (unsigned long long*)ctx->geom.ptr, nullptr, nullptr);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipMemcpyAsync(gs, ctx->geom.ptr, sizeof(gs), hipMemcpyDeviceToHost,
                                   ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        for (int k = 0; k < 10; k++) g[k] = gs[AR_GEOM_SLOT * k];
    }
    int64_t cx = g[0], cy = g[1], rx = g[2], ry = g[3], px = g[4], py = g[5];
    // one even chip extent for every searchable point
    const bool uniform = g[6] == (unsigned long long)(2 * (cx / 2)) &&
                         g[7] == (unsigned long long)(2 * (cy / 2));
    bool fits = g[8] != 0, pow2 = g[9] != 0;
    if (cx == 0) {
        skip_all_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(dx, dy, corr, srch, n);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        return AR_OK;
    }
    if (cx > 512 || cy > 512 || rx > 256 || ry > 256)
        return fail(ctx, AR_ERR_UNSUPPORTED, "chip size above 512 or search radius above 256");

    TrackArgs a;
    std::memset(&a, 0, sizeof(a));
    a.ref = ctx->pref.ptr;
    a.sec = ctx->psec.ptr;
    a.ok = ctx->all_valid ? nullptr : (const uint8_t*)ctx->okmask.ptr;
    a.ny = ctx->ny;
    a.nx = ctx->nx;
    a.padx = fits ? 0 : px + 2;
    a.pady = fits ? 0 : py + 2;
    a.x = dp.x; a.y = dp.y; a.dx_prior = dp.dx_prior; a.dy_prior = dp.dy_prior;
    a.radius_x = dp.radius_x; a.radius_y = dp.radius_y;
    a.chip_size_x = dp.chip_size_x; a.chip_size_y = dp.chip_size_y;
    a.n = n;
    a.up = (int32_t)up;
    a.measure = measure;
    a.dx = dx; a.dy = dy; a.corr = corr; a.searched = srch;
    // even chip extents actually used (chip_bounds), src/points.jl:358-366
    a.max_ch = (int32_t)(2 * (cy / 2));
    a.max_cw = (int32_t)(2 * (cx / 2));
    a.max_rx = (int32_t)rx;
    a.max_ry = (int32_t)ry;

    // RotationSearch: rotated chips are Float32, so the pass takes the generic Float64 path;
    // sines and cosines of deg * (pi / 180) are computed here, once
    const int nrot = ctx->params.rotation_count;
    if (nrot < 0 || nrot > 8)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "rotation_count must be in 0..8");
    a.nrot = nrot;
    for (int k = 0; k < nrot; k++) {
        const double deg = ctx->params.rotation_angles[k];
        if (!std::isfinite(deg)) return fail(ctx, AR_ERR_INVALID_ARGUMENT, "rotation angles must be finite");
        const double rad = deg * (M_PI / 180.0);
        a.rot_s[k] = std::sin(rad);
        a.rot_c[k] = std::cos(rad);
        if (deg == 0.0) a.rot_zero |= 1u << k;
    }
    // the u8 kernel's integer identities hold in int32 for chip areas up to 128 x 128, and its
    // uint16 column sums of W for chip heights up to 256
    bool use_f64 = false;   // direct Float64 tracker (track_f64.hip)
    bool use_u8 = nrot == 0 && ctx->pdtype == AR_U8 && pow2 &&
                  (int64_t)a.max_ch * a.max_cw <= 16384 && a.max_ch <= 256;
    if (use_u8) {
        // the layout with per-point box sums is the larger one
        U8Layout L = u8_layout(a.max_ch, a.max_cw, a.max_rx, a.max_ry, false);
        if (L.total > kLdsLimit) use_u8 = false;
        else a.lds_bytes = L.total;
    }
    void (*kern)(TrackArgs);
    if (use_u8) {
        kern = track_u8_kernel<false, 5, false>;
        // the int8 copy of the reference the matrix cores read (made once per prepared image)
        const int64_t N = ctx->ny * ctx->nx;
        const int64_t stride = (N + 1024 + 255) / 256 * 256;   // room for a window column's tail
        // the kernel addresses a window or chip column of copy k at a 32-bit byte offset
        // k (stride + 1) + column * ny + row from the window's origin in copy 0
        const int64_t reach = 3 * (stride + 1) +
                              (int64_t)(a.max_cw + 2 * a.max_rx + 2) * ctx->ny + 4096;
        const bool copies = reach < ((int64_t)1 << 32);
        if (copies && ctx->refx_gen != ctx->ref_gen) {
            if ((st = ctx->refx.reserve(ctx, 8 * stride))) return st;
            refx_kernel<<<grid_for(stride >> 4), 256, 0, ctx->stream>>>(
                (const uint8_t*)ctx->pref.ptr, (uint8_t*)ctx->refx.ptr, N, stride);
            HIPCHK(ctx, hipGetLastError());
            refx_kernel<<<grid_for(stride >> 4), 256, 0, ctx->stream>>>(
                (const uint8_t*)ctx->psec.ptr, (uint8_t*)ctx->refx.ptr + 4 * stride, N, stride);
            HIPCHK(ctx, hipGetLastError());
            ctx->refx_gen = ctx->ref_gen;
        }
        if (copies) {
            a.refx = (const uint8_t*)ctx->refx.ptr;
            a.secx = (const uint8_t*)ctx->refx.ptr + 4 * stride;
            a.refx_stride = stride;
        }
    }
    const int prec = ctx->params.precision;
    if (prec < AR_PREC_EXACT || prec > AR_PREC_FFT)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "precision must be EXACT, BF16, FP16 or FFT");
    const double macs = (double)(2 * a.max_ry) * (2 * a.max_rx) * a.max_ch * a.max_cw;
    // points that can be searched: all of them, or fewer when the caller knows (a multi-chip
    // level whose points were mostly finished at a smaller chip)
    const int64_t nsearch = searchable_max >= 0 && searchable_max < n ? searchable_max : n;
    const double box_positions = (double)(a.ny + 2 * a.pady - a.max_ch + 1) *
                                 (double)(a.nx + 2 * a.padx - a.max_cw + 1);
    // FFT requested: the reference's dispatch (src/correlate.jl:785-797) on the pass's largest
    // point; the transform kernel carves its LDS for one chip size per pass and takes transforms
    // up to 128 per side, the exact direct tracker everything else, and passes too small to
    // repay the transform pass's box images on a large reference
    bool use_fft = !use_u8 && nrot == 0 && !cplx && ctx->pdtype == AR_F32 && prec == AR_PREC_FFT &&
                   uniform && macs > 20000.0 && box_images_repaid(macs, nsearch, box_positions);
    bool mp_box_timed = false;
    if (use_fft) {
        size_t lds = 0;
        HIPCHK(ctx, hipEventRecord(ctx->evb, ctx->stream));   // the box images count in the pass
        st = ar_fft_setup(ctx, a, &kern, &lds);
        if (st == AR_ERR_UNSUPPORTED) {
            use_fft = false;
        } else if (st == AR_ERR_OUT_OF_MEMORY) {
            return fail(ctx, st, "out of device memory (box images)");
        } else if (st) {
            return fail(ctx, st, "FFT set-up failed");
        } else {
            mp_box_timed = true;
            a.lds_bytes = lds;
        }
    }
    // Float32 images with a 16-bit contraction requested: the mixed-precision kernel, which
    // carves its LDS for one chip size per pass
    const bool use_mp = !use_u8 && nrot == 0 && ctx->pdtype == AR_F32 &&
                        (prec == AR_PREC_BF16 || prec == AR_PREC_FP16) &&
                        mp_worth_it(macs, nsearch, box_positions);
    if (use_fft) {
        // set up above
    } else if (use_mp) {
        if (!uniform)
            return fail(ctx, AR_ERR_UNSUPPORTED,
                        "the BF16 / FP16 tracker needs one chip size per pass");
        size_t lds = 0;
        HIPCHK(ctx, hipEventRecord(ctx->evb, ctx->stream));   // the box images count in the pass
        mp_box_timed = true;
        st = ar_mp_setup(ctx, a, &kern, &lds);
        if (st == AR_ERR_UNSUPPORTED)
            return fail(ctx, st, "chip or search window too large for the BF16 / FP16 tracker");
        if (st == AR_ERR_OUT_OF_MEMORY) return fail(ctx, st, "out of device memory (box images)");
        if (st) return fail(ctx, st, "mixed-precision set-up failed");
        a.lds_bytes = lds;
    } else if (cplx) {
        a.lds_bytes = ar_coherence_lds(a.max_ch, a.max_cw, a.max_rx, a.max_ry);
        kern = a.lds_bytes ? track_coherence_kernel<true> : track_coherence_kernel<false>;
    } else if (!use_u8) {
        a.lds_bytes = 0;
        kern = ctx->pdtype == AR_U8 ? track_generic_kernel<1> : track_generic_kernel<0>;
        size_t wl = 0;
        int pitch = 0, threads = 0;
        use_f64 = nrot == 0 && f64_direct_enabled() && f64_shape(a, &wl, &pitch, &threads);
    }
    bool box_timed = mp_box_timed;
    const int kwg = use_fft ? AR_FFT_WG : AR_WG;   // threads per workgroup of the tracker kernel
    const int bmode = box_images_mode();
    // Box images of W and W^2 replace the per-point box sums when every point has the same chip,
    // the searchable points' shifts cover at least a quarter of the image's chip positions and
    // there are at least 2^22 positions (below that the extra launch costs more than it saves).
    // (The kernel reads a point's S1 / S2 at 32-bit byte offsets 4 (i + j ld) from its first
    // shift, ld = ny + 2 pady - ch + 1.)
    const int64_t box_reach = 4 * (2 * a.max_ry + 2 * a.max_rx * (ctx->ny + 2 * a.pady));
    if (bmode >= 0 && use_u8 && uniform && a.max_ch > 0 && a.max_cw > 0 &&
        box_reach < ((int64_t)1 << 32)) {
        const double positions = (double)(a.ny + 2 * a.pady - a.max_ch + 1) *
                                 (double)(a.nx + 2 * a.padx - a.max_cw + 1);
        const double shifts = (double)nsearch * (2.0 * a.max_rx) * (2.0 * a.max_ry);
        if (positions >= 1 &&
            (bmode == 1 || (positions >= 4194304.0 && 4.0 * shifts >= positions))) {
            int32_t *b1 = nullptr, *b2 = nullptr;
            int64_t ld = 0;
            HIPCHK(ctx, hipEventRecord(ctx->evb, ctx->stream));
            box_timed = true;
            st = ar_box_images(ctx, a.max_ch, a.max_cw, a.pady, a.padx, reuse_box, &b1, &b2,
                               &ld);
            if (st == AR_ERR_OUT_OF_MEMORY) return fail(ctx, st, "out of device memory (box images)");
            if (st == AR_ERR_DEVICE) return fail(ctx, st, "box image kernel failed to launch");
            if (st == AR_OK) {
                a.box1 = b1;
                a.box2 = b2;
                a.box_ld = ld;
            }
        }
    }
    if (use_u8 && a.box1)
        a.lds_bytes = u8_layout(a.max_ch, a.max_cw, a.max_rx, a.max_ry, true).total;
    // refinement hand-off: 32 floats of patch and one packed origin per point
    if ((st = ctx->refine.reserve(ctx, (size_t)n * 132))) return st;
    a.rpatch = (float*)ctx->refine.ptr;
    a.rmeta = (int32_t*)((char*)ctx->refine.ptr + (size_t)n * 128);
    int per_cu = 0;
    if (cplx && a.lds_bytes) {
        // the Coherence tracker's LDS instance, unless its window and chip leave no room on a CU
        if ((st = kernel_occupancy(ctx, (const void*)kern, AR_WG, a.lds_bytes, &per_cu))) return st;
        if (per_cu < 1) {
            kern = track_coherence_kernel<false>;
            a.lds_bytes = 0;
        }
    }
    if (!use_f64) {
        if ((st = kernel_occupancy(ctx, (const void*)kern, kwg, a.lds_bytes, &per_cu))) return st;
        if (per_cu < 1) return fail(ctx, AR_ERR_UNSUPPORTED, "tracker kernel does not fit on a CU");
    }
    int64_t nblocks = (int64_t)ctx->ncu * per_cu;
    if (nblocks > n) nblocks = n;
    if (!use_u8 && !use_mp && !use_fft && !use_f64) {
        const size_t per = cplx ? ar_coherence_scratch(a.max_ch, a.max_cw, a.max_rx, a.max_ry)
                                : gen_scratch(a.max_ch, a.max_cw, a.max_rx, a.max_ry).total;
        a.scratch_per_block = per;
        if ((st = ctx->scratch.reserve(ctx, per * nblocks))) return st;
        a.scratch = (unsigned char*)ctx->scratch.ptr;
    }
    // refinement kernel: one wave per workgroup with its own cascade region
    const size_t rlds = (size_t)cascade_floats((int)up) * 4;
    int64_t rblocks = 0;
    if (up > 1) {
        int rpc = 0;
        if ((st = kernel_occupancy(ctx, (const void*)refine_kernel, 64, rlds, &rpc))) return st;
        if (rpc < 1) return fail(ctx, AR_ERR_UNSUPPORTED, "refinement kernel does not fit on a CU");
        rblocks = (int64_t)ctx->ncu * rpc;
        if (rblocks > n) rblocks = n;
    }
    if (use_u8) {
        if ((st = ctx->prec.reserve(ctx, (size_t)n * AR_PREC * 4))) return st;
        a.prec = (const int32_t*)ctx->prec.ptr;
        point_prep_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(a);
        HIPCHK(ctx, hipGetLastError());
    }
    if (use_f64) {
        F64Plan P;
        if ((st = f64_plan(ctx, a, ctx->pdtype == AR_U8, true, &P))) return st;
        HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        if ((st = f64_enqueue(ctx, a, P, (unsigned char*)ctx->scratch.ptr))) return st;
    } else {
        HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(kwg), a.lds_bytes, ctx->stream, a);
        HIPCHK(ctx, hipGetLastError());
    }
    HIPCHK(ctx, hipEventRecord(ctx->evm, ctx->stream));
    if (up > 1) {
        hipLaunchKernelGGL(refine_kernel, dim3((unsigned)rblocks), dim3(64), rlds, ctx->stream, a);
        HIPCHK(ctx, hipGetLastError());
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    ctx->timing_pending = true;
    ctx->pending_box = box_timed;
    ctx->pending_refine = up > 1;
    ctx->last_path = use_u8 ? 0 : (use_mp ? 2 : (use_fft ? 4 : (use_f64 ? 3 : 1)));
    if (hint) return AR_OK;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ar_resolve_timing(ctx);
    return AR_OK;
}

void ar_resolve_timing(ar_context* ctx) {
    if (!ctx->timing_pending) return;
    ctx->timing_pending = false;
    if (hipEventSynchronize(ctx->ev1) != hipSuccess) return;
    float ms = 0, ms_t = 0, ms_r = 0, ms_b = 0;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    hipEventElapsedTime(&ms_t, ctx->ev0, ctx->evm);
    hipEventElapsedTime(&ms_r, ctx->evm, ctx->ev1);
    if (ctx->pending_box) hipEventElapsedTime(&ms_b, ctx->evb, ctx->ev0);
    ctx->last_ms = ms;
    ctx->last_ms_box = ms_b;
    ctx->last_ms_track = ms_t;
    ctx->last_ms_refine = ctx->pending_refine ? ms_r : 0.0;
}

extern "C" ar_status ar_track_device(ar_context* ctx, const ar_points* dev_pts, int64_t n,
                                     int64_t upsampling, int32_t measure, float* dev_dx,
                                     float* dev_dy, float* dev_corr, uint8_t* dev_searched) {
    if (!ctx) return fail(nullptr, AR_ERR_INVALID_ARGUMENT, "context is NULL");
    if (!dev_pts || !dev_dx || !dev_dy || !dev_corr || !dev_searched)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    return ar_track_pass(ctx, *dev_pts, n, upsampling, measure, dev_dx, dev_dy, dev_corr,
                         dev_searched);
}

// Uploads host point fields into the context's point buffers; returns device views.
ar_status ar_upload_points(ar_context* ctx, const ar_points* pts, int64_t n, ar_points* dp,
                           int64_t csx_fill, int64_t csy_fill) {
    ar_status st;
    size_t d = sizeof(double) * n, l = sizeof(int64_t) * n;
    if ((st = ctx->pts.reserve(ctx, 4 * d + 4 * l))) return st;
    char* b = (char*)ctx->pts.ptr;
    double* x = (double*)b; double* y = (double*)(b + d);
    double* pdx = (double*)(b + 2 * d); double* pdy = (double*)(b + 3 * d);
    int64_t* rx = (int64_t*)(b + 4 * d); int64_t* ry = (int64_t*)(b + 4 * d + l);
    int64_t* cx = (int64_t*)(b + 4 * d + 2 * l); int64_t* cy = (int64_t*)(b + 4 * d + 3 * l);
    HIPCHK(ctx, hipMemcpyAsync(x, pts->x, d, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(y, pts->y, d, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(pdx, pts->dx_prior, d, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(pdy, pts->dy_prior, d, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(rx, pts->radius_x, l, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ry, pts->radius_y, l, hipMemcpyHostToDevice, ctx->stream));
    if (pts->chip_size_x && pts->chip_size_y) {
        HIPCHK(ctx, hipMemcpyAsync(cx, pts->chip_size_x, l, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(cy, pts->chip_size_y, l, hipMemcpyHostToDevice, ctx->stream));
    } else {
        std::vector<int64_t> fx(n, csx_fill), fy(n, csy_fill);
        HIPCHK(ctx, hipMemcpyAsync(cx, fx.data(), l, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(cy, fy.data(), l, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    dp->x = x; dp->y = y; dp->dx_prior = pdx; dp->dy_prior = pdy;
    dp->radius_x = rx; dp->radius_y = ry; dp->chip_size_x = cx; dp->chip_size_y = cy;
    return AR_OK;
}

extern "C" ar_status ar_track(ar_context* ctx, const ar_points* pts, int64_t n,
                              int64_t upsampling, int32_t measure, float* dx, float* dy,
                              float* corr, uint8_t* searched) {
    if (!ctx) return fail(nullptr, AR_ERR_INVALID_ARGUMENT, "context is NULL");
    if (!pts || !dx || !dy || !corr || !searched)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    if (!pts->x || !pts->y || !pts->radius_x || !pts->radius_y || !pts->dx_prior ||
        !pts->dy_prior || !pts->chip_size_x || !pts->chip_size_y)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "NULL point field");
    if (n == 0) return AR_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    ar_status st;
    ar_points dp;
    if ((st = ar_upload_points(ctx, pts, n, &dp, 0, 0))) return st;
    if ((st = ctx->out.reserve(ctx, 13 * n))) return st;
    float* ddx = (float*)ctx->out.ptr;
    float* ddy = ddx + n;
    float* dc = ddy + n;
    uint8_t* ds = (uint8_t*)(dc + n);
    if ((st = ar_track_pass(ctx, dp, n, upsampling, measure, ddx, ddy, dc, ds))) return st;
    HIPCHK(ctx, hipMemcpyAsync(dx, ddx, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(dy, ddy, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(corr, dc, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(searched, ds, n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return AR_OK;
}

extern "C" ar_status ar_last_kernel(const ar_context* ctx, double* ms, int32_t* path) {
    if (!ctx) return AR_ERR_INVALID_ARGUMENT;
    ar_resolve_timing(const_cast<ar_context*>(ctx));
    if (ms) *ms = ctx->last_ms;
    if (path) *path = ctx->last_path;
    return AR_OK;
}

extern "C" ar_status ar_last_timing(const ar_context* ctx, double* box_ms, double* track_ms,
                                    double* refine_ms) {
    if (!ctx) return AR_ERR_INVALID_ARGUMENT;
    ar_resolve_timing(const_cast<ar_context*>(ctx));
    if (box_ms) *box_ms = ctx->last_ms_box;
    if (track_ms) *track_ms = ctx->last_ms_track;
    if (refine_ms) *refine_ms = ctx->last_ms_refine;
    return AR_OK;
}

extern "C" ar_status ar_correlate_point(ar_context* ctx, const void* search, const void* chip,
                                        int32_t dtype, int64_t ch, int64_t cw, int64_t rx,
                                        int64_t ry, int64_t upsampling, int32_t measure,
                                        double* dx, double* dy, float* corr) {
    if (!ctx) return fail(nullptr, AR_ERR_INVALID_ARGUMENT, "context is NULL");
    if (!search || !chip || !dx || !dy || !corr)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    if (ch <= 0 || cw <= 0 || (ch & 1) || (cw & 1) || rx <= 0 || ry <= 0)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "chip extents must be even and radii positive");
    if (dtype != AR_U8 && dtype != AR_F32)
        return fail(ctx, AR_ERR_UNSUPPORTED, "dtype must be UInt8 or Float32");
    // Lay the window out as the reference image and the chip at offset (ry, rx) of an
    // equally sized secondary, so one grid point with zero prior reproduces correlate!.
    int64_t wh = ch + 2 * ry - 1, ww = cw + 2 * rx - 1;
    size_t es = dtype == AR_U8 ? 1 : 4;
    std::vector<unsigned char> sec(wh * ww * es, 0);
    for (int64_t j = 0; j < cw; j++)
        std::memcpy(&sec[((ry) + (j + rx) * wh) * es], (const unsigned char*)chip + j * ch * es,
                    ch * es);
    std::vector<uint8_t> ones(wh * ww, 1);
    // keep the caller's pair, if any, intact
    ar_context* c = ctx;
    bool had = c->has_pair;
    DevBuf sref = c->pref, ssec = c->psec, srm = c->rmask, ssm = c->smask, sok = c->okmask;
    int64_t sny = c->ny, snx = c->nx;
    int32_t sdt = c->pdtype;
    bool sall = c->all_valid;
    c->pref = DevBuf(); c->psec = DevBuf(); c->rmask = DevBuf(); c->smask = DevBuf();
    c->okmask = DevBuf();
    ar_status st = ar_set_pair(c, search, sec.data(), ones.data(), ones.data(), wh, ww, dtype);
    if (st == AR_OK) {
        double x = (double)(1 + cw / 2 + rx), y = (double)(1 + ch / 2 + ry), z = 0.0;
        int64_t r_x = rx, r_y = ry, c_x = cw, c_y = ch;
        ar_points p = {&x, &y, &r_x, &r_y, &z, &z, &c_x, &c_y};
        float fdx, fdy, fc;
        uint8_t s;
        st = ar_track(c, &p, 1, upsampling, measure, &fdx, &fdy, &fc, &s);
        *dx = fdx;
        *dy = fdy;
        *corr = fc;
    }
    c->pref.release(); c->psec.release(); c->rmask.release(); c->smask.release();
    c->okmask.release();
    c->pref = sref; c->psec = ssec; c->rmask = srm; c->smask = ssm; c->okmask = sok;
    c->ny = sny; c->nx = snx; c->pdtype = sdt; c->all_valid = sall; c->has_pair = had;
    c->ref_gen++;   // the restored pair is not the one the int8 copies were made from
    return st;
}

// ---------------------------------------------------------------------------------------------
// ar_correlate_point_noalloc: ar_correlate_point without any device allocation
// ---------------------------------------------------------------------------------------------

// The device memory of ar_correlate_point_noalloc: the window, the chip's secondary, the point,
// its outputs, the refinement hand-off and the direct Float64 tracker's staging slot (and its
// summed-area table when that does not fit in LDS). One array per device, for the life of the
// process. Host code never uses the name: the device address comes from hipGetSymbolAddress.
#define AR_NOALLOC_BYTES ((size_t)1 << 20)
static __device__ __align__(256) unsigned char scratch[AR_NOALLOC_BYTES];

static std::mutex g_noalloc_mu;
static unsigned char* g_noalloc_ptr[64];   // per device ordinal, nullptr until looked up

// The array's address on the context's device (current), looked up once per device.
static ar_status noalloc_scratch(ar_context* ctx, unsigned char** out) {
    if (ctx->device < 0 || ctx->device >= 64)
        return fail(ctx, AR_ERR_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[ctx->device]) {
        void* ptr = nullptr;
        HIPCHK(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(ctx, AR_ERR_DEVICE, "hipGetSymbolAddress returned no address");
        g_noalloc_ptr[ctx->device] = (unsigned char*)ptr;
    }
    *out = g_noalloc_ptr[ctx->device];
    return AR_OK;
}

static size_t up256(size_t x) { return (x + 255) / 256 * 256; }

extern "C" ar_status ar_correlate_point_noalloc(ar_context* ctx, const void* search,
                                                const void* chip, int32_t dtype, int64_t ch,
                                                int64_t cw, int64_t rx, int64_t ry,
                                                int64_t upsampling, int32_t measure, double* dx,
                                                double* dy, float* corr) {
    if (!ctx) return fail(nullptr, AR_ERR_INVALID_ARGUMENT, "context is NULL");
    if (!search || !chip || !dx || !dy || !corr)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    if (ch <= 0 || cw <= 0 || (ch & 1) || (cw & 1) || rx <= 0 || ry <= 0)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "chip extents must be even and radii positive");
    if (dtype != AR_U8 && dtype != AR_F32)
        return fail(ctx, AR_ERR_UNSUPPORTED, "dtype must be UInt8 or Float32");
    if (upsampling < 1 || upsampling > 64 || (upsampling & (upsampling - 1)) != 0)
        return fail(ctx, AR_ERR_UNSUPPORTED,
                    "upsampling must be a power of two in 1..64 on the device path");
    if (measure != AR_ZNCC && measure != AR_NCC)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "measure must be ZNCC or NCC");
    if (ch > 512 || cw > 512 || rx > 256 || ry > 256)
        return fail(ctx, AR_ERR_UNSUPPORTED, "chip size above 512 or search radius above 256");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    // the window is the reference image and the chip sits at offset (ry, rx) of an equally sized
    // secondary, as in ar_correlate_point: one point with zero prior reproduces correlate!
    const int64_t wh = ch + 2 * ry - 1, ww = cw + 2 * rx - 1;
    const size_t es = dtype == AR_U8 ? 1 : 4;
    TrackArgs a;
    std::memset(&a, 0, sizeof(a));
    a.max_ch = (int32_t)ch;
    a.max_cw = (int32_t)cw;
    a.max_rx = (int32_t)rx;
    a.max_ry = (int32_t)ry;
    a.n = 1;
    a.up = (int32_t)upsampling;
    a.measure = measure;
    a.ny = wh;
    a.nx = ww;
    size_t wl = 0;
    int pitch = 0, threads = 0;
    if (!f64_shape(a, &wl, &pitch, &threads))
        return fail(ctx, AR_ERR_UNSUPPORTED, "chip or search window too large for the device path");
    F64Plan P;
    f64_plan_shape(a, dtype == AR_U8, &P);
    P.chunk = 1;
    P.overlap = false;
    P.nbuf = 1;
    P.nprep = 1;
    P.half = up256(P.sl.bytes);
    // regions of the array
    const size_t img = up256((size_t)wh * ww * es);
    const size_t o_ref = 0, o_sec = o_ref + img, o_pts = o_sec + img;
    const size_t o_out = o_pts + 256, o_ref_h = o_out + 256, o_slot = o_ref_h + 256;
    const size_t need = o_slot + P.half + up256(P.tbytes);
    if (need > AR_NOALLOC_BYTES)
        return fail(ctx, AR_ERR_UNSUPPORTED,
                    "chip " + std::to_string(ch) + "x" + std::to_string(cw) + " at radius " +
                        std::to_string(ry) + "x" + std::to_string(rx) + " needs " +
                        std::to_string(need) + " bytes of the no-alloc scratch, which holds " +
                        std::to_string(AR_NOALLOC_BYTES));
    unsigned char* base = nullptr;
    ar_status st;
    if ((st = noalloc_scratch(ctx, &base))) return st;
    if ((st = lds_allowance(ctx, (const void*)P.fprep))) return st;
    if ((st = lds_allowance(ctx, (const void*)P.ftrack))) return st;
    hipStream_t s = ctx->stream;
    // inputs: the window, then the chip into a zeroed secondary (a fill kernel, not a memset)
    HIPCHK(ctx, hipMemcpyAsync(base + o_ref, search, (size_t)wh * ww * es, hipMemcpyHostToDevice, s));
    fill_u8_kernel<<<grid_for(wh * ww * (int64_t)es), 256, 0, s>>>(base + o_sec, 0,
                                                                   wh * ww * (int64_t)es);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpy2DAsync(base + o_sec + (ry + rx * wh) * es, wh * es, chip, ch * es,
                                 ch * es, cw, hipMemcpyHostToDevice, s));
    struct {
        double x, y, pdx, pdy;
        int64_t rx, ry, cx, cy;
    } pt = {(double)(1 + cw / 2 + rx), (double)(1 + ch / 2 + ry), 0.0, 0.0, rx, ry, cw, ch};
    HIPCHK(ctx, hipMemcpyAsync(base + o_pts, &pt, sizeof(pt), hipMemcpyHostToDevice, s));
    double* dpts = (double*)(base + o_pts);
    int64_t* ipts = (int64_t*)(base + o_pts + 32);
    a.ref = base + o_ref;
    a.sec = base + o_sec;
    a.ok = nullptr;
    a.x = dpts; a.y = dpts + 1; a.dx_prior = dpts + 2; a.dy_prior = dpts + 3;
    a.radius_x = ipts; a.radius_y = ipts + 1; a.chip_size_x = ipts + 2; a.chip_size_y = ipts + 3;
    float* dout = (float*)(base + o_out);
    a.dx = dout; a.dy = dout + 1; a.corr = dout + 2;
    a.searched = (uint8_t*)(dout + 3);
    a.rpatch = (float*)(base + o_ref_h);
    a.rmeta = (int32_t*)(base + o_ref_h + 128);
    if ((st = f64_enqueue(ctx, a, P, base + o_slot))) return st;
    if (upsampling > 1) {
        const size_t rlds = (size_t)cascade_floats((int)upsampling) * 4;
        hipLaunchKernelGGL(refine_kernel, dim3(1), dim3(64), rlds, s, a);
        HIPCHK(ctx, hipGetLastError());
    }
    float o[3];
    HIPCHK(ctx, hipMemcpyAsync(o, dout, sizeof(o), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    *dx = o[0];
    *dy = o[1];
    *corr = o[2];
    return AR_OK;
}

// ---------------------------------------------------------------------------------------------
// Batched passes: many prepared UInt8 pairs resident at once, and one tracking pass over the
// points of all of them (each point names its pair). The pairs keep their own buffers (prepared
// images, validity, int8 copies); the pass shares one virtual zero padding, large enough for
// every point, which leaves each point's result as in a pass over its own pair alone (the
// reference pads every window into bounds, src/track.jl:121-138, and padding is invalid).
// ---------------------------------------------------------------------------------------------

#include <map>
#include <tuple>

namespace {

struct BatchPair {
    DevBuf ref, sec, ok, refx;
    int64_t ny = 0, nx = 0, stride = 0;
    bool all_valid = true;
    bool copies = false;
};

struct GraphKey {
    const void* p[11];
    int64_t n, up;
    int32_t measure, split, exact;
    bool operator==(const GraphKey& o) const {
        return std::memcmp(p, o.p, sizeof(p)) == 0 && n == o.n && up == o.up &&
               measure == o.measure && split == o.split && exact == o.exact;
    }
};

}  // namespace

struct ar_batch {
    ar_context* ctx = nullptr;          // preparation, stream, pass scratch, error text
    std::vector<BatchPair*> pairs;
    DevBuf desc, lists, counts, pidx;   // descriptor table, class lists and counts, pair indices
    DevBuf sbuf;                        // global box sums of the largest radius class
    bool desc_dirty = true;
    // chips whose area is not a power of two go to the direct Float64 tracker
    bool exact = false;
    // radius split chosen by the tuner, per pass shape (max ch, cw, rx, ry, points / 4096,
    // exact); also recorded in the tuning cache under the device's name
    std::map<std::tuple<int, int, int, int, int64_t, int>, int> tuned;
    std::string tune_device;            // "<arch>-<CUs>cu"
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    GraphKey gkey;
    bool has_graph = false;
    int graph_launches = 0, graph_split = 0;
    // device buffers the captured pass addresses; a replay needs every one where it was (a later
    // pass that grew one of them moved it)
    const void* gbufs[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    void graph_bufs(const void** out) const {
        out[0] = desc.ptr; out[1] = lists.ptr; out[2] = counts.ptr; out[3] = ctx->refine.ptr;
        out[4] = ctx->prec.ptr; out[5] = ctx->scratch.ptr; out[6] = sbuf.ptr;
    }
    bool graph_bufs_same() const {
        const void* b[7];
        graph_bufs(b);
        for (int k = 0; k < 7; k++)
            if (b[k] != gbufs[k]) return false;
        return true;
    }
    double last_ms = 0;
    int last_split = 0, last_launches = 0, last_replayed = 0;
};

// Radius classes of a batched pass: class 0 max(rx, ry) <= R1, class 1 <= R2 (R2 > R1, or 0:
// no such class), class 2 the rest, and, with `exact`, class 3 every searchable point whose chip
// area is not a power of two (the direct Float64 tracker takes those); each class list in any
// order. One atomic
// per wave and class (ballot, then each lane's rank among the wave's lanes of its class).
__global__ __launch_bounds__(256) void batch_classify_kernel(ar_points pts, int64_t n, int R1,
                                                             int R2, int exact, int32_t* lists,
                                                             uint32_t* counts) {
    const int lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n; i0 += stride) {
        const int64_t i = i0 + threadIdx.x;
        int c = -1;
        if (i < n) {
            const int64_t r = max(pts.radius_x[i], pts.radius_y[i]);
            c = r <= R1 ? 0 : ((R2 > 0 && r <= R2) ? 1 : 2);
            // a searchable chip whose area is not a power of two: the direct Float64 tracker
            const int64_t area = (2 * (pts.chip_size_y[i] / 2)) * (2 * (pts.chip_size_x[i] / 2));
            if (exact && pts.radius_x[i] > 0 && pts.radius_y[i] > 0 && area > 0 &&
                (area & (area - 1)) != 0)
                c = 3;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint64_t m = __ballot(c == k);
            if (m == 0) continue;
            const int leader = __ffsll((long long)m) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(counts + k, (uint32_t)__popcll(m));
            base = (uint32_t)__shfl((int)base, leader);
            if (c == k) lists[k * n + base + (uint32_t)__popcll(m & below)] = (int32_t)i;
        }
    }
}

__global__ void batch_iota_kernel(int32_t* out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (int32_t)i;
}

__global__ void batch_zero_kernel(uint32_t* c, int k) {
    if ((int)threadIdx.x < k) c[threadIdx.x] = 0u;
}

// Counts pair indices outside [0, npairs).
__global__ __launch_bounds__(256) void batch_check_kernel(const int32_t* pair_of, int64_t n,
                                                          int64_t npairs, uint32_t* bad) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (pair_of[i] < 0 || pair_of[i] >= npairs) atomicAdd(bad, 1u);
}

static ar_status batch_sync_desc(ar_batch* B) {
    ar_context* ctx = B->ctx;
    if (!B->desc_dirty) return AR_OK;
    std::vector<PairDesc> d(B->pairs.size());
    for (size_t k = 0; k < B->pairs.size(); k++) {
        BatchPair* p = B->pairs[k];
        d[k].ref = (const uint8_t*)p->ref.ptr;
        d[k].sec = (const uint8_t*)p->sec.ptr;
        d[k].refx = p->copies ? (const uint8_t*)p->refx.ptr : nullptr;
        d[k].secx = p->copies ? (const uint8_t*)p->refx.ptr + 4 * p->stride : nullptr;
        d[k].ok = p->all_valid ? nullptr : (const uint8_t*)p->ok.ptr;
        d[k].ny = p->ny;
        d[k].nx = p->nx;
        d[k].refx_stride = p->stride;
    }
    ar_status st;
    if ((st = B->desc.reserve(ctx, d.size() * sizeof(PairDesc) + 64))) return st;
    HIPCHK(ctx, hipMemcpyAsync(B->desc.ptr, d.data(), d.size() * sizeof(PairDesc),
                               hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    B->desc_dirty = false;
    return AR_OK;
}

// Kernel and LDS of one class launch of the batched tracker (largest chip, the class's radius
// cap). A class keeps its box sums S1 / S2 in global scratch when that admits more workgroups
// per CU than holding them in LDS; launches whose LDS leaves room for at most two workgroups per
// CU take the instance with twice the registers (no spills).
struct BatchKern {
    void (*kern)(TrackArgs);
    size_t lds;
    bool sg;
    int64_t sg_ints;   // int32 per workgroup slot (S1 and S2)
};
static BatchKern batch_kernel(const TrackArgs& a, int rcap) {
    const int rx = a.max_rx < rcap ? a.max_rx : rcap, ry = a.max_ry < rcap ? a.max_ry : rcap;
    BatchKern k;
    k.lds = u8_layout(a.max_ch, a.max_cw, rx, ry, false).total;
    k.sg = false;
    k.sg_ints = 0;
    // workgroups per CU the LDS admits (at most 4: the register budget of the 4-wave instance)
    auto fit = [](size_t l) { const size_t f = kLdsLimit / (l > 0 ? l : 1); return f < 4 ? f : 4; };
    const size_t l2 = u8_layout(a.max_ch, a.max_cw, rx, ry, false, true).total;
    if (fit(l2) > fit(k.lds)) {
        k.lds = l2;
        k.sg = true;
        k.sg_ints = 2 * (int64_t)(2 * rx) * (2 * ry);
    }
    const bool wide = k.lds > (160 * 1024) / 3;
    if (k.sg) k.kern = wide ? track_u8_kernel<true, 2, true> : track_u8_kernel<true, 4, true>;
    else k.kern = wide ? track_u8_kernel<true, 2, false> : track_u8_kernel<true, 4, false>;
    return k;
}

// One class launch of the batched tracker: the points in `list` (count m), LDS carved for the
// class's largest radius.
static ar_status batch_launch(ar_batch* B, TrackArgs a, const int32_t* list, int64_t m, int rcap) {
    ar_context* ctx = B->ctx;
    if (m == 0) return AR_OK;
    const BatchKern bk = batch_kernel(a, rcap);
    a.order = list;
    a.n = m;
    a.max_rx = a.max_rx < rcap ? a.max_rx : rcap;
    a.max_ry = a.max_ry < rcap ? a.max_ry : rcap;
    a.lds_bytes = bk.lds;
    int per_cu = 0;
    if (ar_status st = kernel_occupancy(ctx, (const void*)bk.kern, AR_WG, a.lds_bytes, &per_cu))
        return st;
    if (per_cu < 1) return fail(ctx, AR_ERR_UNSUPPORTED, "batched tracker does not fit on a CU");
    int64_t nb = (int64_t)ctx->ncu * per_cu;
    if (nb > m) nb = m;
    if (bk.sg) {
        // reserved before any capture (batch_pass sizes it for every class)
        if (B->sbuf.cap < (size_t)nb * bk.sg_ints * 4)
            return fail(ctx, AR_ERR_DEVICE, "batched tracker box-sum scratch not reserved");
        a.sg = (int32_t*)B->sbuf.ptr;
        a.sg_stride = bk.sg_ints;
    }
    auto kern = bk.kern;
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(AR_WG), a.lds_bytes, ctx->stream, a);
    HIPCHK(ctx, hipGetLastError());
    B->last_launches++;
    return AR_OK;
}

// A radius split read back from the tuning cache: R1 | R2 << 8 with both in 0..255.
static bool batch_split_ok(int32_t s) { return s >= 0 && s < (1 << 16); }

// The batched pass on device-resident points, pair indices and outputs. split: radius threshold
// of the small class (0: one launch over every point), or -1: the tuner's choice.
static ar_status batch_pass(ar_batch* B, const ar_points& dp, const int32_t* dpair, int64_t n,
                            int64_t up, int32_t measure, float* dx, float* dy, float* corr,
                            uint8_t* srch, int split, bool replay) {
    ar_context* ctx = B->ctx;
    if (B->pairs.empty()) return fail(ctx, AR_ERR_INVALID_ARGUMENT, "the batch holds no pair");
    if (n < 0) return fail(ctx, AR_ERR_INVALID_ARGUMENT, "negative point count");
    if (n > 0x7fffffffLL) return fail(ctx, AR_ERR_UNSUPPORTED, "more than 2^31 - 1 points");
    if (up < 1 || up > 64 || (up & (up - 1)) != 0)
        return fail(ctx, AR_ERR_UNSUPPORTED,
                    "upsampling must be a power of two in 1..64 on the device path");
    if (measure != AR_ZNCC && measure != AR_NCC)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "measure must be ZNCC or NCC");
    B->last_launches = 0;
    B->last_replayed = 0;
    if (n == 0) return AR_OK;
    ar_status st;
    if ((st = batch_sync_desc(B))) return st;
    GraphKey key;
    const void* ps[11] = {dp.x, dp.y, dp.dx_prior, dp.dy_prior, dp.radius_x, dp.radius_y,
                          dp.chip_size_x, dp.chip_size_y, dpair, dx, (const void*)srch};
    std::memcpy(key.p, ps, sizeof(ps));
    key.n = n;
    key.up = up;
    key.measure = measure;
    key.split = split;
    key.exact = B->exact;
    if (replay && B->has_graph && B->gkey == key && B->graph_bufs_same()) {
        // the captured pass over the same buffers (contents may differ only in values)
        HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        HIPCHK(ctx, hipGraphLaunch(B->gexec, ctx->stream));
        HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        float ms = 0;
        hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
        B->last_ms = ms;
        B->last_replayed = 1;
        B->last_launches = B->graph_launches;
        B->last_split = B->graph_split;
        return AR_OK;
    }
    // pair indices in range
    if ((st = B->counts.reserve(ctx, 256))) return st;
    uint32_t* cnt = (uint32_t*)B->counts.ptr;
    batch_zero_kernel<<<1, 64, 0, ctx->stream>>>(cnt, 4);
    batch_check_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(dpair, n, (int64_t)B->pairs.size(),
                                                             cnt + 2);
    HIPCHK(ctx, hipGetLastError());
    uint32_t bad = 0;
    HIPCHK(ctx, hipMemcpyAsync(&bad, cnt + 2, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (bad) return fail(ctx, AR_ERR_INVALID_ARGUMENT, "pair index outside the batch");
    // pass geometry over every pair
    unsigned long long gs[10 * AR_GEOM_SLOT] = {0};
    if ((st = ctx->geom.reserve(ctx, sizeof(gs)))) return st;
    for (int k = 6; k < 10; k++) gs[AR_GEOM_SLOT * k] = k < 8 ? ~0ULL : 1ULL;
    HIPCHK(ctx, hipMemcpyAsync(ctx->geom.ptr, gs, sizeof(gs), hipMemcpyHostToDevice, ctx->stream));
    pass_geometry_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(
        dp, n, 0, 0, (unsigned long long*)ctx->geom.ptr, (const PairDesc*)B->desc.ptr, dpair);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpyAsync(gs, ctx->geom.ptr, sizeof(gs), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    unsigned long long g[10];
    for (int k = 0; k < 10; k++) g[k] = gs[AR_GEOM_SLOT * k];
    const int64_t cx = g[0], cy = g[1], rx = g[2], ry = g[3], px = g[4], py = g[5];
    const bool fits = g[8] != 0, pow2 = g[9] != 0;
    if (cx == 0) {
        skip_all_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(dx, dy, corr, srch, n);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        return AR_OK;
    }
    TrackArgs a;
    std::memset(&a, 0, sizeof(a));
    a.padx = fits ? 0 : px + 2;
    a.pady = fits ? 0 : py + 2;
    a.x = dp.x; a.y = dp.y; a.dx_prior = dp.dx_prior; a.dy_prior = dp.dy_prior;
    a.radius_x = dp.radius_x; a.radius_y = dp.radius_y;
    a.chip_size_x = dp.chip_size_x; a.chip_size_y = dp.chip_size_y;
    a.n = n;
    a.up = (int32_t)up;
    a.measure = measure;
    a.dx = dx; a.dy = dy; a.corr = corr; a.searched = srch;
    a.max_ch = (int32_t)(2 * (cy / 2));
    a.max_cw = (int32_t)(2 * (cx / 2));
    a.max_rx = (int32_t)rx;
    a.max_ry = (int32_t)ry;
    a.pairs = (const PairDesc*)B->desc.ptr;
    a.pair_of = dpair;
    if ((int64_t)a.max_ch * a.max_cw > 16384 || a.max_ch > 256 ||
        u8_layout(a.max_ch, a.max_cw, a.max_rx, a.max_ry, false).total > kLdsLimit)
        return fail(ctx, AR_ERR_UNSUPPORTED, "chip or search window too large for the batched tracker");
    if ((st = ctx->refine.reserve(ctx, (size_t)n * 132))) return st;
    a.rpatch = (float*)ctx->refine.ptr;
    a.rmeta = (int32_t*)((char*)ctx->refine.ptr + (size_t)n * 128);
    if ((st = ctx->prec.reserve(ctx, (size_t)n * AR_PREC * 4))) return st;
    a.prec = (const int32_t*)ctx->prec.ptr;
    if ((st = B->lists.reserve(ctx, (size_t)n * 16))) return st;
    int32_t* lists = (int32_t*)B->lists.ptr;
    // the tuner: time each candidate split once per pass shape, keep the fastest
    const int rmax = (int)(rx > ry ? rx : ry);
    if (split < 0) {
        auto tk = std::make_tuple(a.max_ch, a.max_cw, a.max_rx, a.max_ry, n >> 12, (int)B->exact);
        auto it = B->tuned.find(tk);
        char ckey[96];
        std::snprintf(ckey, sizeof ckey, "batch c%dx%d r%dx%d n%lld e%d", a.max_ch, a.max_cw,
                      a.max_rx, a.max_ry, (long long)(n >> 12), (int)B->exact);
        int32_t cached = 0;
        if (it != B->tuned.end()) {
            split = it->second;
        } else if (ar_tune_cache_lookup(B->tune_device.c_str(), ckey, &cached) &&
                   batch_split_ok(cached)) {
            B->tuned[tk] = cached;
            split = cached;
        } else {
            int best = 0;
            double best_ms = 1e30;
            // one launch; two classes split at 16, 20, 24, 28 or 32; three classes
            const int cands[11] = {0, 16, 20, 24, 28, 32, 12 | (24 << 8), 16 | (28 << 8),
                                   16 | (32 << 8), 20 | (32 << 8), 24 | (32 << 8)};
            for (int c : cands) {
                const int c1 = c & 0xff, c2 = c >> 8;
                if (c != 0 && (c1 >= rmax || (c2 > 0 && c2 >= rmax))) continue;
                if ((st = batch_pass(B, dp, dpair, n, up, measure, dx, dy, corr, srch, c, false)))
                    return st;
                if (B->last_ms < best_ms) {
                    best_ms = B->last_ms;
                    best = c;
                }
            }
            B->tuned[tk] = best;
            split = best;
            ar_tune_cache_record(B->tune_device.c_str(), ckey, best);   // dropped if it fails
        }
        key.split = split;
    }
    int R1 = split & 0xff, R2 = (split >> 8) & 0xff;
    if (R2 > 0 && (R2 <= R1 || R2 >= rmax)) R2 = 0;
    if (R1 >= rmax) R1 = R2 = 0;
    split = R1 | (R2 << 8);
    // occupancy queries (and the once-per-process LDS allowances) happen before a capture
    // begins: every class's size is queried here
    {
        int pc = 0;
        const int caps0[3] = {R1 > 0 ? R1 : 1 << 30, R2 > 0 ? R2 : 1 << 30, 1 << 30};
        size_t sbytes = 0;
        for (int k = 0; k < 3; k++) {
            const BatchKern bk = batch_kernel(a, caps0[k]);
            if ((st = kernel_occupancy(ctx, (const void*)bk.kern, AR_WG, bk.lds, &pc))) return st;
            if (bk.sg) {
                const size_t need = (size_t)ctx->ncu * pc * bk.sg_ints * 4;
                if (need > sbytes) sbytes = need;
            }
        }
        if (sbytes && (st = B->sbuf.reserve(ctx, sbytes))) return st;
    }
    const size_t rlds = (size_t)cascade_floats((int)up) * 4;
    int64_t rblocks = 0;
    if (up > 1) {
        int rpc = 0;
        if ((st = kernel_occupancy(ctx, (const void*)refine_kernel, 64, rlds, &rpc))) return st;
        if (rpc < 1) return fail(ctx, AR_ERR_UNSUPPORTED, "refinement kernel does not fit on a CU");
        rblocks = (int64_t)ctx->ncu * rpc;
        if (rblocks > n) rblocks = n;
    }
    // class sizes (the class lists are rebuilt inside the pass, so a replay needs the same points)
    int64_t m[4] = {n, 0, 0, 0};
    const bool classify = R1 > 0 || (!pow2 && B->exact);
    if (classify) {
        batch_zero_kernel<<<1, 64, 0, ctx->stream>>>(cnt, 4);
        batch_classify_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(dp, n, R1, R2, B->exact,
                                                                    lists, cnt);
        HIPCHK(ctx, hipGetLastError());
        uint32_t c4[4] = {0, 0, 0, 0};
        HIPCHK(ctx, hipMemcpyAsync(c4, cnt, 16, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        for (int k = 0; k < 4; k++) m[k] = c4[k];
    }
    // the non-power-of-two chips' launches are planned (occupancy, staging) before any capture
    F64Plan fp;
    TrackArgs af = a;
    af.order = lists + 3 * n;
    af.n = m[3];
    if (m[3] > 0) {
        size_t wl = 0;
        int pitch = 0, threads = 0;
        if (!f64_shape(af, &wl, &pitch, &threads))
            return fail(ctx, AR_ERR_UNSUPPORTED,
                        "chip or search window too large for the batched Float64 tracker");
        if ((st = f64_plan(ctx, af, true, false, &fp))) return st;
    }
    const bool capture = replay;
    if (capture) {
        if (B->has_graph) {
            hipGraphExecDestroy(B->gexec);
            hipGraphDestroy(B->graph);
            B->has_graph = false;
        }
        HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    } else {
        HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    }
    B->last_launches = 0;
    // ---- the pass: point records, class lists, the tracker per class, the refinement
    point_prep_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(a);
    if (classify) {
        batch_zero_kernel<<<1, 64, 0, ctx->stream>>>(cnt, 4);
        batch_classify_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(dp, n, R1, R2, B->exact,
                                                                    lists, cnt);
    } else {
        batch_iota_kernel<<<grid_for(n), 256, 0, ctx->stream>>>(lists, n);
    }
    // class caps: R1, then R2 (or the largest radius when there is no middle class), then rmax
    const int caps[3] = {R1 > 0 ? R1 : rmax, R2 > 0 ? R2 : rmax, rmax};
    ar_status lst = AR_OK;
    for (int k = 0; k < 3 && lst == AR_OK; k++)
        if (m[k] > 0) lst = batch_launch(B, a, lists + k * n, m[k], caps[k]);
    if (lst == AR_OK && m[3] > 0) {
        lst = f64_enqueue(ctx, af, fp, (unsigned char*)ctx->scratch.ptr);
        B->last_launches += 2 * (int)((m[3] + fp.chunk - 1) / fp.chunk);
    }
    if (lst == AR_OK && up > 1)
        hipLaunchKernelGGL(refine_kernel, dim3((unsigned)rblocks), dim3(64), rlds, ctx->stream, a);
    if (capture) {
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &gr);
        if (lst != AR_OK) {
            if (gr) hipGraphDestroy(gr);
            return lst;
        }
        if (e != hipSuccess) return fail(ctx, AR_ERR_DEVICE, "graph capture of the batched pass failed");
        hipGraphExec_t ex = nullptr;
        if (hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) != hipSuccess) {
            hipGraphDestroy(gr);
            return fail(ctx, AR_ERR_DEVICE, "graph instantiation of the batched pass failed");
        }
        B->graph = gr;
        B->gexec = ex;
        B->gkey = key;
        B->graph_bufs(B->gbufs);
        B->has_graph = true;
        B->graph_launches = B->last_launches;
        B->graph_split = split;
        const int launches = B->last_launches;
        // first run of the captured pass
        HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        HIPCHK(ctx, hipGraphLaunch(ex, ctx->stream));
        B->last_launches = launches;
    } else if (lst != AR_OK) {
        return lst;
    }
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    B->last_ms = ms;
    B->last_split = split;
    return AR_OK;
}

extern "C" {

ar_status ar_batch_create(const ar_params* p, int32_t device, ar_batch** out) {
    if (!out) return fail(nullptr, AR_ERR_INVALID_ARGUMENT, "NULL output handle");
    *out = nullptr;
    ar_params q;
    if (p) q = *p;
    else ar_params_default(&q);
    if (q.quantize != AR_QUANT_UINT8)
        return fail(nullptr, AR_ERR_UNSUPPORTED, "the batched tracker takes UInt8-quantised pairs");
    if (q.rotation_count != 0)
        return fail(nullptr, AR_ERR_UNSUPPORTED, "the batched tracker does not take a rotation search");
    ar_context* c = nullptr;
    ar_status st = ar_context_create(&q, device, &c);
    if (st) return st;
    ar_batch* B = new ar_batch();
    B->ctx = c;
    B->tune_device = "gfx950-" + std::to_string(c->ncu) + "cu";
    *out = B;
    return AR_OK;
}

ar_status ar_batch_set_exact(ar_batch* B, int32_t exact) {
    if (!B) return fail(nullptr, AR_ERR_INVALID_ARGUMENT, "batch is NULL");
    B->exact = exact != 0;
    return AR_OK;
}

void ar_batch_destroy(ar_batch* B) {
    if (!B) return;
    hipSetDevice(B->ctx->device);
    hipStreamSynchronize(B->ctx->stream);
    if (B->has_graph) {
        hipGraphExecDestroy(B->gexec);
        hipGraphDestroy(B->graph);
    }
    for (BatchPair* p : B->pairs) {
        p->ref.release();
        p->sec.release();
        p->ok.release();
        p->refx.release();
        delete p;
    }
    B->desc.release();
    B->lists.release();
    B->counts.release();
    B->pidx.release();
    B->sbuf.release();
    ar_context_destroy(B->ctx);
    delete B;
}

ar_status ar_batch_add_pair(ar_batch* B, const void* ref, const void* sec,
                            const uint8_t* ref_valid, const uint8_t* sec_valid, int64_t ny,
                            int64_t nx, int32_t dtype, int32_t* index) {
    if (!B) return fail(nullptr, AR_ERR_INVALID_ARGUMENT, "batch is NULL");
    ar_context* ctx = B->ctx;
    if (!ref || !sec) return fail(ctx, AR_ERR_INVALID_ARGUMENT, "a batched pair needs both images");
    if (B->pairs.size() >= 0x7fffffff) return fail(ctx, AR_ERR_UNSUPPORTED, "too many pairs");
    ar_status st = ar_set_images(ctx, ref, sec, ref_valid, sec_valid, ny, nx, dtype);
    if (st) return st;
    if (ctx->pdtype != AR_U8) return fail(ctx, AR_ERR_UNSUPPORTED, "prepared pair is not UInt8");
    BatchPair* bp = new BatchPair();
    const int64_t N = ny * nx;
    bp->ny = ny;
    bp->nx = nx;
    bp->all_valid = ctx->all_valid;
    bp->stride = (N + 1024 + 255) / 256 * 256;
    // the tracker addresses the copies with 32-bit offsets (chips up to 256, radii up to 256)
    bp->copies = 3 * (bp->stride + 1) + (int64_t)(256 + 2 * 256 + 2) * ny + 4096 < ((int64_t)1 << 32);
    if ((st = bp->ref.reserve(ctx, N)) || (st = bp->sec.reserve(ctx, N)) ||
        (!bp->all_valid && (st = bp->ok.reserve(ctx, N))) ||
        (bp->copies && (st = bp->refx.reserve(ctx, 8 * bp->stride)))) {
        bp->ref.release(); bp->sec.release(); bp->ok.release(); bp->refx.release();
        delete bp;
        return st;
    }
    HIPCHK(ctx, hipMemcpyAsync(bp->ref.ptr, ctx->pref.ptr, N, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(bp->sec.ptr, ctx->psec.ptr, N, hipMemcpyDeviceToDevice, ctx->stream));
    if (!bp->all_valid)
        HIPCHK(ctx, hipMemcpyAsync(bp->ok.ptr, ctx->okmask.ptr, N, hipMemcpyDeviceToDevice,
                                   ctx->stream));
    if (bp->copies) {
        refx_kernel<<<grid_for(bp->stride >> 4), 256, 0, ctx->stream>>>(
            (const uint8_t*)ctx->pref.ptr, (uint8_t*)bp->refx.ptr, N, bp->stride);
        refx_kernel<<<grid_for(bp->stride >> 4), 256, 0, ctx->stream>>>(
            (const uint8_t*)ctx->psec.ptr, (uint8_t*)bp->refx.ptr + 4 * bp->stride, N, bp->stride);
        HIPCHK(ctx, hipGetLastError());
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    B->pairs.push_back(bp);
    B->desc_dirty = true;
    if (index) *index = (int32_t)(B->pairs.size() - 1);
    return AR_OK;
}

ar_status ar_batch_track_device(ar_batch* B, const ar_points* dev_pts, const int32_t* dev_pair,
                                int64_t n, int64_t upsampling, int32_t measure, float* dev_dx,
                                float* dev_dy, float* dev_corr, uint8_t* dev_searched,
                                int32_t flags) {
    if (!B) return fail(nullptr, AR_ERR_INVALID_ARGUMENT, "batch is NULL");
    ar_context* ctx = B->ctx;
    if (!dev_pts || !dev_pair || !dev_dx || !dev_dy || !dev_corr || !dev_searched)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int split = (flags & AR_BATCH_SPLIT_SET) ? (flags >> 8) & 0xffff : -1;
    return batch_pass(B, *dev_pts, dev_pair, n, upsampling, measure, dev_dx, dev_dy, dev_corr,
                      dev_searched, split, (flags & AR_BATCH_GRAPH) != 0);
}

ar_status ar_batch_track(ar_batch* B, const ar_points* pts, const int32_t* pair, int64_t n,
                         int64_t upsampling, int32_t measure, float* dx, float* dy, float* corr,
                         uint8_t* searched) {
    if (!B) return fail(nullptr, AR_ERR_INVALID_ARGUMENT, "batch is NULL");
    ar_context* ctx = B->ctx;
    if (!pts || !pair || !dx || !dy || !corr || !searched)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    if (!pts->x || !pts->y || !pts->radius_x || !pts->radius_y || !pts->dx_prior ||
        !pts->dy_prior || !pts->chip_size_x || !pts->chip_size_y)
        return fail(ctx, AR_ERR_INVALID_ARGUMENT, "NULL point field");
    if (n == 0) return AR_OK;
    for (int64_t i = 0; i < n; i++)
        if (pair[i] < 0 || pair[i] >= (int64_t)B->pairs.size())
            return fail(ctx, AR_ERR_INVALID_ARGUMENT, "pair index outside the batch");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    ar_status st;
    ar_points dp;
    if ((st = ar_upload_points(ctx, pts, n, &dp, 0, 0))) return st;
    if ((st = B->pidx.reserve(ctx, 4 * (size_t)n))) return st;
    HIPCHK(ctx, hipMemcpyAsync(B->pidx.ptr, pair, 4 * n, hipMemcpyHostToDevice, ctx->stream));
    if ((st = ctx->out.reserve(ctx, 13 * n))) return st;
    float* ddx = (float*)ctx->out.ptr;
    float* ddy = ddx + n;
    float* dc = ddy + n;
    uint8_t* ds = (uint8_t*)(dc + n);
    if ((st = batch_pass(B, dp, (const int32_t*)B->pidx.ptr, n, upsampling, measure, ddx, ddy,
                         dc, ds, -1, false)))
        return st;
    HIPCHK(ctx, hipMemcpyAsync(dx, ddx, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(dy, ddy, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(corr, dc, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(searched, ds, n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return AR_OK;
}

ar_status ar_batch_info(const ar_batch* B, double* ms, int32_t* info) {
    if (!B) return AR_ERR_INVALID_ARGUMENT;
    if (ms) *ms = B->last_ms;
    if (info) {
        info[0] = (int32_t)B->pairs.size();
        info[1] = B->last_split;
        info[2] = B->last_launches;
        info[3] = B->last_replayed;
    }
    return AR_OK;
}

const char* ar_batch_last_error(const ar_batch* B) {
    return B ? B->ctx->err.c_str() : "batch is NULL";
}

}  // extern "C"
