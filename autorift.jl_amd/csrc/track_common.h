// Shared device-side pieces of the per-point tracker: pass arguments, point geometry,
// block reductions, the OpenCV-exact argmax rule and the pyrUp refinement cascade.
//
// Arithmetic that must match the CPU path bit for bit is written without FMA contraction
// (the library is compiled with -ffp-contract=off) and in the reference's operation order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define AR_WG 256
#define AR_NWAVES (AR_WG / 64)

namespace ar {

// One prepared UInt8 pair of a batch (batched pass): what a single-pair pass reads from
// TrackArgs' image fields.
struct PairDesc {
    const uint8_t* ref;
    const uint8_t* sec;
    const uint8_t* refx;
    const uint8_t* secx;
    const uint8_t* ok;        // nullptr: every pixel valid
    int64_t ny, nx, refx_stride;
};

struct TrackArgs {
    const void* ref;          // prepared reference image, column-major ny x nx
    const uint8_t* refx;      // u8 path: 4 int8 copies of the reference (value ^ 0x80), copy k
    const uint8_t* secx;      // shifted k bytes later, at refx + k * refx_stride; secx likewise
    int64_t refx_stride;      // for the secondary (both nullptr for images too large for the
                              // kernel's 32-bit offsets)
    const void* sec;          // prepared secondary image
    const uint8_t* ok;        // valid(ref) & valid(sec); nullptr when every pixel is valid
    const int32_t* prec;      // u8 path: per-point geometry records (point_record), 8 per point
    int64_t ny, nx;           // image size
    int64_t padx, pady;       // virtual zero padding of the pass (0 when every window fits)
    const double* x;          // point fields (src/points.jl:50-74), 1-based coordinates
    const double* y;
    const double* dx_prior;
    const double* dy_prior;
    const int64_t* radius_x;
    const int64_t* radius_y;
    const int64_t* chip_size_x;
    const int64_t* chip_size_y;
    int64_t n;
    int32_t up;               // upsampling; 1 = integer peak only
    int32_t measure;          // 0 ZNCC, 1 NCC
    float* dx;                // outputs, one per point
    float* dy;
    float* corr;
    uint8_t* searched;
    // refinement hand-off between the correlation and refinement kernels: the 5x5 patch of
    // point p at rpatch[32 p ..], its 1-based origin packed as i0 | j0 << 16 in rmeta[p]
    // (-1: nothing to refine; the correlation kernel has written the result)
    float* rpatch;
    int32_t* rmeta;
    // box images of W and W^2 for a pass with one chip size (u8 path), or nullptr: S1 / S2 of
    // shift (i, j) of a point whose window starts at padded (wr0, wc0) are at
    // (wr0 - 1 + i) + (wc0 - 1 + j) * box_ld
    const int32_t* box1;
    const int32_t* box2;
    int64_t box_ld;
    // batched tracker with the box sums in global memory: S1 then S2 of workgroup b at
    // sg + b * sg_stride (sg_stride / 2 int32 each)
    int32_t* sg;
    int64_t sg_stride;
    // mixed-precision path: Float64 box images of W and W^2 (same indexing as box1 / box2) and
    // the power-of-two operand scales (window, chip) with the factor that undoes them
    const double* fbox1;
    const double* fbox2;
    const float* fwv;           // Float32(max(window variance, 0)) per chip position (the
                                // measure's: s2 - s1^2 / n for ZNCC, s2 for NCC), leading dim box_ld
    float wscale, cscale;
    double numscale;
    // FFT path: transform extents (rows, columns) of the pass
    int32_t fft_fy, fft_fx;
    // per-pass maxima the LDS layout is carved for
    int32_t max_ch, max_cw, max_rx, max_ry;
    size_t lds_bytes;
    // generic path: per-block global scratch
    unsigned char* scratch;
    size_t scratch_per_block;
    // RotationSearch (generic path): nrot chip rotations (0: none); per angle its sine and cosine
    // (computed on the host) and, in bit k of rot_zero, whether angle k is exactly 0 (the chip is
    // then used unrotated, as the reference does)
    int32_t nrot;
    uint32_t rot_zero;
    double rot_s[8];
    double rot_c[8];
    // batched pass (many pairs in one launch): the pairs' descriptors, each point's pair, and the
    // points this launch visits (order[k], k < n); nullptr for a single-pair pass
    const PairDesc* pairs;
    const int32_t* pair_of;
    const int32_t* order;
};

// The kernel's argument block read again from the kernarg segment through a pointer the compiler
// cannot follow: a field used after such a read is loaded there (scalar loads) instead of being
// held in scalar registers, or spilled to vector lanes, since the previous one.
__device__ inline TrackArgs reload_args() {
    typedef const __attribute__((address_space(4))) TrackArgs* KArgs;
    KArgs p = (KArgs)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return *(const TrackArgs*)p;
}

// A batched pass reads point p's images from its pair's descriptor (uniform per workgroup).
__device__ inline void use_pair(TrackArgs& a, int64_t p) {
    const PairDesc d = a.pairs[a.pair_of[p]];
    a.ref = d.ref;
    a.sec = d.sec;
    a.refx = d.refx;
    a.secx = d.secx;
    a.ok = d.ok;
    a.ny = d.ny;
    a.nx = d.nx;
    a.refx_stride = d.refx_stride;
}

// Geometry of one point in padded coordinates (1-based, inclusive), src/points.jl:358-395 and
// the pass shift of src/track.jl:207-210.
struct Geom {
    int64_t cr0, cc0;   // chip top-left
    int64_t wr0, wc0;   // window top-left
    int ch, cw, rx, ry; // chip extent (even) and radii
    double pdx, pdy;
    int status;         // 0 skip (not searchable / out of bounds), 1 go
};

__device__ inline int64_t jfloor(double v) { return (int64_t)floor(v); }

__device__ inline Geom point_geom(const TrackArgs& a, int64_t p) {
    Geom g;
    g.status = 0;
    int64_t prx = a.radius_x[p], pry = a.radius_y[p];
    g.rx = (int)prx;
    g.ry = (int)pry;
    if (!(prx > 0 && pry > 0)) return g;
    double X = a.x[p] + ((double)a.padx + 0.5);
    double Y = a.y[p] + ((double)a.pady + 0.5);
    g.pdx = a.dx_prior[p];
    g.pdy = a.dy_prior[p];
    int64_t hx = a.chip_size_x[p] / 2, hy = a.chip_size_y[p] / 2;
    int64_t ccx = jfloor(X - g.pdx), ccy = jfloor(Y - g.pdy);
    g.cr0 = ccy - hy;
    g.cc0 = ccx - hx;
    int64_t cr1 = ccy + hy - 1, cc1 = ccx + hx - 1;
    int64_t wcx = jfloor(X), wcy = jfloor(Y);
    g.wr0 = wcy - hy - pry;
    g.wc0 = wcx - hx - prx;
    int64_t wr1 = wcy + hy + pry - 2, wc1 = wcx + hx + prx - 2;
    g.ch = (int)(2 * hy);
    g.cw = (int)(2 * hx);
    int64_t NY = a.ny + 2 * a.pady, NX = a.nx + 2 * a.padx;
    bool chip_in = (cr1 < g.cr0 || cc1 < g.cc0) ||
                   (g.cr0 >= 1 && cr1 <= NY && g.cc0 >= 1 && cc1 <= NX);
    bool win_in = (wr1 < g.wr0 || wc1 < g.wc0) ||
                  (g.wr0 >= 1 && wr1 <= NY && g.wc0 >= 1 && wc1 <= NX);
    if (!chip_in || !win_in) return g;
    if (g.ch == 0 || g.cw == 0) return g;   // empty chip: _any_valid is false
    g.status = 1;
    return g;
}

// The geometry of a point as 8 int32 (written once per pass by point_prep_kernel, read by the
// u8 tracker with one scalar load): status, cr0, cc0, wr0, wc0, ch | cw << 16, rx | ry << 16, 0.
// A searchable point lies inside the padded image, so its coordinates fit in int32.
#define AR_PREC 8
__device__ inline void point_record(const Geom& g, int32_t* r) {
    r[0] = g.status;
    r[1] = g.status ? (int32_t)g.cr0 : 0;
    r[2] = g.status ? (int32_t)g.cc0 : 0;
    r[3] = g.status ? (int32_t)g.wr0 : 0;
    r[4] = g.status ? (int32_t)g.wc0 : 0;
    r[5] = g.status ? (g.ch | (g.cw << 16)) : 0;
    r[6] = (g.rx & 0xffff) | (g.ry << 16);
    r[7] = 0;
}

// Geom from a point's record (the priors are read where the result is written).
__device__ inline Geom load_geom(const TrackArgs& a, int64_t p) {
    const int4* q = (const int4*)(a.prec + AR_PREC * p);
    const int4 u = q[0], v = q[1];
    Geom g;
    g.status = u.x;
    g.cr0 = u.y;
    g.cc0 = u.z;
    g.wr0 = u.w;
    g.wc0 = v.x;
    g.ch = v.y & 0xffff;
    g.cw = v.y >> 16;
    g.rx = v.z & 0xffff;
    g.ry = v.z >> 16;
    g.pdx = 0.0;
    g.pdy = 0.0;
    return g;
}

// Persistent point schedule. Workgroups are dispatched round-robin over the 8 XCDs, each with
// its own L2; the workgroups of one XCD take one contiguous eighth of the points, so the
// overlapping windows of neighbouring points are fetched into the same L2. Every point is
// visited exactly once (for fewer than 8 workgroups, a plain stride).
struct PointRange {
    int64_t first, end, step;
};

__device__ inline PointRange point_range(int64_t n) {
    PointRange r;
    const int64_t g = gridDim.x, b = blockIdx.x;
    if (g < 8) {
        r.first = b;
        r.end = n;
        r.step = g;
        return r;
    }
    const int64_t x = b % 8, slot = b / 8, cnt = (g - x + 7) / 8;
    const int64_t chunk = (n + 7) / 8;
    int64_t start = x * chunk, end = start + chunk < n ? start + chunk : n;
    r.first = start + slot;
    r.end = end;
    r.step = cnt;
    return r;
}

// Padded (1-based) pixel -> linear index into the unpadded image, or -1 in the padding.
__device__ inline int64_t pix(const TrackArgs& a, int64_t r, int64_t c) {
    int64_t orow = r - a.pady - 1, ocol = c - a.padx - 1;
    if (orow < 0 || orow >= a.ny || ocol < 0 || ocol >= a.nx) return -1;
    return orow + ocol * a.ny;
}

// _any_valid (src/track.jl:416-421) over the chip rectangle. Block-uniform result.
__device__ inline bool any_valid(const TrackArgs& a, const Geom& g) {
    if (a.ok == nullptr) {
        // every real pixel valid: the chip has a valid pixel iff it overlaps the image
        int64_t r0 = g.cr0 - a.pady - 1, c0 = g.cc0 - a.padx - 1;
        return r0 + g.ch - 1 >= 0 && r0 < a.ny && c0 + g.cw - 1 >= 0 && c0 < a.nx;
    }
    int found = 0;
    for (int v = threadIdx.x >> 6; v < g.cw; v += AR_NWAVES)
        for (int u = threadIdx.x & 63; u < g.ch; u += 64) {
            int64_t k = pix(a, g.cr0 + u, g.cc0 + v);
            if (k >= 0 && a.ok[k]) found = 1;
        }
    return __syncthreads_or(found) != 0;
}

// A pointer every lane holds the same value of, moved to scalar registers (so that loads through
// it can use the scalar-base + 32-bit offset addressing form).
template <typename T>
__device__ inline T* uniform_ptr(T* p) {
    const uint64_t u = (uint64_t)(uintptr_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    return (T*)(uintptr_t)(((uint64_t)hi << 32) | lo);
}

// ---------------------------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------------------------

template <typename T>
__device__ inline T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Block sum; every thread gets the total. `red` holds AR_NWAVES elements.
template <typename T>
__device__ inline T block_sum(T v, T* red) {
    v = wave_sum(v);
    int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    T s = red[0];
    for (int k = 1; k < AR_NWAVES; k++) s += red[k];
    return s;
}

// Two block sums in one exchange; `red` holds 2 AR_NWAVES elements.
template <typename T>
__device__ inline void block_sum2(T a, T b, T* red, T& sa, T& sb) {
    a = wave_sum(a);
    b = wave_sum(b);
    int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[w] = a;
        red[AR_NWAVES + w] = b;
    }
    __syncthreads();
    sa = red[0];
    sb = red[AR_NWAVES];
    for (int k = 1; k < AR_NWAVES; k++) {
        sa += red[k];
        sb += red[AR_NWAVES + k];
    }
}

// Argmax key under peak_index's rule (src/peak.jl:32-61): larger value wins; equal values go to
// the smaller row-major index; NaN never wins.
struct Best {
    float v;
    int idx;   // row-major linear index, INT32_MAX when empty
};

__device__ inline bool better(float v, int idx, const Best& b) {
    return v > b.v || (v == b.v && idx < b.idx);
}

__device__ inline Best best_merge(Best a, Best b) { return better(b.v, b.idx, a) ? b : a; }

__device__ inline Best wave_best(Best b) {
    for (int o = 32; o > 0; o >>= 1) {
        Best c;
        c.v = __shfl_xor(b.v, o);
        c.idx = __shfl_xor(b.idx, o);
        b = best_merge(b, c);
    }
    return b;
}

__device__ inline Best block_best(Best b, Best* red) {
    b = wave_best(b);
    int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = b;
    __syncthreads();
    Best r = red[0];
    for (int k = 1; k < AR_NWAVES; k++) r = best_merge(r, red[k]);
    return r;
}

// ---- one-barrier block reductions for a full workgroup (all lanes active). A row of 16 lanes
// is reduced by DPP (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror, row_mirror: afterwards
// every lane of the row holds the row's result), the four rows by readlane, and the waves through
// `slots` (AR_NWAVES entries owned by one call site, so no barrier has to guard them before the
// write: the call site is reached again only after later barriers).
__device__ inline int dpp_row(int v, int ctrl) {
    switch (ctrl) {
        case 0: return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
        case 1: return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
        case 2: return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);
        default: return __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);
    }
}

__device__ inline uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int k = 0; k < 4; k++) v += (uint32_t)dpp_row((int)v, k);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 0) + (uint32_t)__builtin_amdgcn_readlane((int)v, 16) +
           (uint32_t)__builtin_amdgcn_readlane((int)v, 32) + (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
}

__device__ inline float wave_max_f32(float v) {
#pragma unroll
    for (int k = 0; k < 4; k++) v = fmaxf(v, __int_as_float(dpp_row(__float_as_int(v), k)));
    const float a = fmaxf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0)),
                          __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16)));
    const float b = fmaxf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32)),
                          __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48)));
    return fmaxf(a, b);
}

__device__ inline int wave_min_i32(int v) {
#pragma unroll
    for (int k = 0; k < 4; k++) v = min(v, dpp_row(v, k));
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// Sums of two uint32 over the workgroup (each total below 2^32).
__device__ inline void block_sum2_u32(uint32_t a, uint32_t b, uint32_t* slots, uint32_t& sa,
                                      uint32_t& sb) {
    a = wave_sum_u32(a);
    b = wave_sum_u32(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        slots[w] = a;
        slots[AR_NWAVES + w] = b;
    }
    __syncthreads();
    sa = slots[0];
    sb = slots[AR_NWAVES];
    for (int k = 1; k < AR_NWAVES; k++) {
        sa += slots[k];
        sb += slots[AR_NWAVES + k];
    }
}

// Workgroup argmax under `better` (exact: the peak_index rule) or, with ANY, the largest value
// at any one of the lanes that hold it (ties left open; NaN never wins either way).
template <bool ANY>
__device__ inline Best block_best1(Best b, Best* slots) {
    const float m = wave_max_f32(b.v);
    Best r;
    r.v = m;
    if (ANY) {
        const uint64_t hit = __ballot(b.v == m);
        r.idx = hit ? __builtin_amdgcn_readlane(b.idx, __ffsll((unsigned long long)hit) - 1)
                    : 0x7fffffff;
    } else {
        r.idx = wave_min_i32(b.v == m ? b.idx : 0x7fffffff);
    }
    if (!(m > -INFINITY)) {   // empty (or all NaN) wave: the empty key, which any value beats
        r.v = -INFINITY;
        r.idx = 0x7fffffff;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) slots[w] = r;
    __syncthreads();
    Best q = slots[0];
    for (int k = 1; k < AR_NWAVES; k++) q = best_merge(q, slots[k]);
    return q;
}

__device__ inline Best best_empty() {
    Best b;
    b.v = -INFINITY;
    b.idx = 0x7fffffff;
    return b;
}

// ---------------------------------------------------------------------------------------------
// pyrUp (src/peak.jl:196-337): taps in upsampled coordinates with BORDER_REFLECT_101
// ---------------------------------------------------------------------------------------------

// 0-based output position k on an axis with n >= 2 source samples -> 0-based source taps.
// Closed form of src/peak.jl:312-324 (reflection in upsampled coordinates, BORDER_REFLECT_101):
// only three positions reflect -- k = 0 (minus tap of sample 0 reflects onto sample 1) and the
// last two positions (plus tap clamps onto sample n-1).
__host__ __device__ inline void taps(int k, int n, int& m, int& c, int& q, int& nt) {
    c = k >> 1;
    q = c + 1 < n - 1 ? c + 1 : n - 1;
    if ((k & 1) == 0) {
        m = c == 0 ? 1 : c - 1;
        nt = 3;
    } else {
        m = c;
        nt = 2;
    }
}

#define AR_WC (6.0f / 8)
#define AR_WS (1.0f / 8)
#define AR_WH (4.0f / 8)

__device__ inline float tap3(float a, float b, float c) { return AR_WC * b + AR_WS * (a + c); }
__device__ inline float tap2(float b, float c) { return AR_WH * (b + c); }

// Prepared Float32 pixel at padded 1-based (r, c); 0 in the padding.
__device__ inline float mp_pix(const TrackArgs& a, const float* img, int64_t r, int64_t c) {
    const int64_t k = pix(a, r, c);
    return k < 0 ? 0.0f : img[k];
}

// The reference's normalisation (src/correlate.jl:591-600 ZNCC, :618-627 NCC) of a numerator.
__device__ inline float mp_normalise(double num, double s1, double s2, double n, double cnorm,
                                     int measure) {
    double den;
    if (measure == 1) {
        den = cnorm * sqrt(s2 < 0.0 ? 0.0 : s2);
    } else {
        const double wv = s2 - s1 * s1 / n;
        den = cnorm * sqrt(wv < 0.0 ? 0.0 : wv);
    }
    return den > 0 ? (float)(num / den) : 0.0f;
}

// ---- shared tail of the tracker kernels: outputs, refinement hand-off, patch origin

__device__ inline void write_skip(const TrackArgs& a, int64_t p, bool searched) {
    if (threadIdx.x == 0) {
        a.dx[p] = __int_as_float(0x7fc00000);
        a.dy[p] = __int_as_float(0x7fc00000);
        a.corr[p] = __int_as_float(0x7fc00000);
        a.searched[p] = searched ? 1 : 0;
        a.rmeta[p] = -1;
    }
}

// Offsets of src/peak.jl:102-106 and src/peak.jl:515-521, plus the prior (src/track.jl:293-295).
__device__ inline void write_result(const TrackArgs& a, int64_t p, const Geom& g, double ddx,
                                    double ddy, float val) {
    if (threadIdx.x == 0) {
        a.dx[p] = (float)(ddx + a.dx_prior[p]);
        a.dy[p] = (float)(ddy + a.dy_prior[p]);
        a.corr[p] = val;
        a.searched[p] = 1;
    }
}

// After the integer peak (pi, pj) (1-based) and its exact value are known: without refinement,
// write the point's result; with it, hand the 5x5 patch (exact values; thread t < 25 holds
// element t, row-major) and its origin to the refinement kernel.
__device__ inline void finish_point(const TrackArgs& a, int64_t p, const Geom& g, int pi, int pj,
                                    float pval, int i0, int j0, bool refine, float pv) {
    if (!refine) {
        write_result(a, p, g, (double)(pj - g.rx - 1), (double)(pi - g.ry - 1), pval);
        if (threadIdx.x == 0) a.rmeta[p] = -1;
        return;
    }
    if (threadIdx.x < 25) a.rpatch[32 * p + threadIdx.x] = pv;
    if (threadIdx.x == 0) {
        a.rmeta[p] = i0 | (j0 << 16);
        a.searched[p] = 1;
    }
}

// Patch origin of src/peak.jl:473-478 (1-based); returns false when the surface is smaller than
// the patch, in which case the integer peak is reported.
__device__ inline bool patch_origin(int pi, int pj, int nr, int nc, int& i0, int& j0) {
    int hr = nr - 5 + 1 > 1 ? nr - 5 + 1 : 1;
    int hc = nc - 5 + 1 > 1 ? nc - 5 + 1 : 1;
    i0 = pi - 2 < 1 ? 1 : (pi - 2 > hr ? hr : pi - 2);
    j0 = pj - 2 < 1 ? 1 : (pj - 2 > hc ? hc : pj - 2);
    return !(nr < 5 || nc < 5);
}


}  // namespace ar
