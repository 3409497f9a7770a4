// The multi-chip-size search on the device (src/multichip.jl:106-422): per level, a decimated
// integer-peak pass gated by the relaxed outlier filter, exact Euclidean dilation and nearest
// resampling of the coarse mask, the subpixel pass, the Gardner filter (src/outliers.jl:90-173),
// median hole filling and the merge. The host only sequences the launches and reads back the
// few counts the reference branches on; every grid-sized array stays in HBM.
//
// Every grid kernel computes a point from its neighbourhood in the reference's arithmetic
// (Float32 agreement counts, insertion-sort medians, Float64 tolerances and distance transform),
// so the result does not depend on the launch order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/autorift_hip.h"
#include "ar_internal.h"

namespace ar {

#define GK 256
static inline unsigned gblocks(int64_t n) { return (unsigned)((n + GK - 1) / GK); }

__device__ inline float qnan() { return __int_as_float(0x7fc00000); }

__global__ void fill_result_kernel(float* dx, float* dy, float* corr, uint16_t* cs, uint8_t* it,
                                   int64_t n) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    dx[k] = qnan(); dy[k] = qnan(); corr[k] = qnan(); cs[k] = 0; it[k] = 0;
}

// counts of f0[k] != 0 and f1[k] != 0 into out[0], out[1]
__global__ void count2_kernel(const uint8_t* f0, const uint8_t* f1, int64_t n,
                              unsigned long long* out) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long a = __syncthreads_count(k < n && f0[k] ? 1 : 0);
    const unsigned long long b = __syncthreads_count(k < n && f1[k] ? 1 : 0);
    if (threadIdx.x == 0) {
        if (a) atomicAdd(out, a);
        if (b) atomicAdd(out + 1, b);
    }
}

// _level_points + sanitize! (src/multichip.jl:193-213, src/points.jl:332-346)
// Also writes wanted = chip_size == 0 (src/multichip.jl:118-136).
__global__ void level_points_kernel(const int64_t* grx, const int64_t* gry, const uint16_t* rcs,
                                    int64_t n, int64_t csx, int64_t csy, int64_t minr,
                                    int64_t* rx, int64_t* ry, int64_t* cx, int64_t* cy,
                                    uint8_t* srch, uint8_t* want) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    bool wanted = rcs[k] == 0;
    want[k] = wanted ? 1 : 0;
    int64_t a = wanted ? grx[k] : 0, b = wanted ? gry[k] : 0;
    if (a <= 0 || b <= 0) { a = 0; b = 0; }
    else { a = a < minr ? minr : a; b = b < minr ? minr : b; }
    rx[k] = a; ry[k] = b; cx[k] = csx; cy[k] = csy;
    srch[k] = (a > 0 && b > 0) ? 1 : 0;
}

// pts[stride:stride:nr, stride:stride:nc] with _cell_max_radius! (src/multichip.jl:224-306)
__global__ void coarse_points_kernel(const double* x, const double* y, const double* pdx,
                                     const double* pdy, const int64_t* rx, const int64_t* ry,
                                     int64_t nr, int64_t nc, int64_t stride, int64_t nrc,
                                     int64_t ncc, int64_t csx, int64_t csy, double* ox, double* oy,
                                     double* opdx, double* opdy, int64_t* orx, int64_t* ory,
                                     int64_t* ocx, int64_t* ocy, uint8_t* srch,
                                     unsigned long long* frac) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) { frac[0] = 0ull; frac[1] = 0ull; }   // the level's valid-fraction counters
    if (t >= nrc * ncc) return;
    int64_t jo = t / nrc, io = t - jo * nrc;
    int64_t i = (io + 1) * stride, j = (jo + 1) * stride;   // 1-based grid indices
    int64_t src = (i - 1) + (j - 1) * nr;
    ox[t] = x[src]; oy[t] = y[src]; opdx[t] = pdx[src]; opdy[t] = pdy[src];
    int64_t lo = stride / 2, hi = stride - 1 - lo;
    int64_t mx = 0, my = 0;
    int64_t j0 = j - lo > 1 ? j - lo : 1, j1 = j + hi < nc ? j + hi : nc;
    int64_t i0 = i - lo > 1 ? i - lo : 1, i1 = i + hi < nr ? i + hi : nr;
    for (int64_t jj = j0; jj <= j1; jj++)
        for (int64_t ii = i0; ii <= i1; ii++) {
            int64_t k = (ii - 1) + (jj - 1) * nr;
            mx = rx[k] > mx ? rx[k] : mx;
            my = ry[k] > my ? ry[k] : my;
        }
    orx[t] = mx; ory[t] = my; ocx[t] = csx; ocy[t] = csy;
    srch[t] = (mx > 0 && my > 0) ? 1 : 0;
}

__global__ void measured_kernel(const float* dx, int64_t n, uint8_t* m) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) m[k] = isnan(dx[k]) ? 0 : 1;
}

// nx = Float32(dx / r) (Float32 arithmetic; NaN where r == 0), src/outliers.jl:96-102
// The valid mask is the measured points (!isnan(dx)), also written to meas.
__global__ void normalise_kernel(const float* dx, const float* dy, const int64_t* rx,
                                 const int64_t* ry, uint8_t* meas, int64_t n, float* nx,
                                 float* ny, uint8_t* keep) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint8_t v = isnan(dx[k]) ? 0 : 1;
    meas[k] = v;
    nx[k] = rx[k] > 0 ? dx[k] / (float)rx[k] : qnan();
    ny[k] = ry[k] > 0 ? dy[k] / (float)ry[k] : qnan();
    keep[k] = v;
}

struct Win {
    int lx, ly, rx, ry;
};

__device__ inline float masked(const float* a, const uint8_t* keep, int64_t k) {
    return keep[k] ? a[k] : qnan();
}

__device__ inline void isort(float* b, int n) {
    for (int k = 1; k < n; k++) {
        float v = b[k];
        int m = k - 1;
        while (m >= 0 && b[m] > v) { b[m + 1] = b[m]; m--; }
        b[m + 1] = v;
    }
}

__device__ inline float smedian(const float* b, int n) {
    return (n & 1) ? b[(n + 1) / 2 - 1] : (b[n / 2 - 1] + b[n / 2]) / 2;
}

#define MAXWIN 81   // window^2 for windows up to 9

// ---- small windows (side <= WM, WM * WM <= 32) held in registers ------------------------------
// The reference's stable insertion sort (src/window.jl:547-558) puts value i at the position
// rank_i = #{j : v_j < v_i, or v_j == v_i and j before i}; selecting by that rank gives the same
// element, bit for bit (equal floats differ only as -0 / +0, and the earlier one keeps its
// place). All compares run unrolled on registers instead of a sort through private memory.

// Gathers the window around (i, j) in the reference's column-major order into slot
// (jj - j0) * WM + (ii - i0); ok marks the non-NaN (kept) values. Returns their count.
template <int WM, bool KEEP>
__device__ inline int gather_small(const float* a, const uint8_t* kin, int64_t i, int64_t j,
                                   int64_t nr, int64_t nc, Win w, float (&v)[WM * WM],
                                   uint32_t& ok) {
    ok = 0;
    const int64_t j0 = j - w.lx, i0 = i - w.ly;
#pragma unroll
    for (int cj = 0; cj < WM; cj++)
#pragma unroll
        for (int ci = 0; ci < WM; ci++) {
            const int slot = cj * WM + ci;
            const int64_t jj = j0 + cj, ii = i0 + ci;
            float x = qnan();
            if (cj <= w.lx + w.rx && ci <= w.ly + w.ry && jj >= 0 && jj < nc && ii >= 0 &&
                ii < nr) {
                const int64_t k = ii + jj * nr;
                x = KEEP ? masked(a, kin, k) : a[k];
            }
            v[slot] = x;
            if (!isnan(x)) ok |= 1u << slot;
        }
    return __builtin_popcount(ok);
}

// The median of the kept values in stable-sort order (src/window.jl:561-564): the value at
// position (n - 1) / 2, and for even n the Float32 mean of positions n / 2 - 1 and n / 2.
template <int WM>
__device__ inline float median_small(const float (&v)[WM * WM], uint32_t ok, int n) {
    const int k1 = (n - 1) / 2, k2 = n / 2;
    float a = qnan(), b = qnan();
#pragma unroll
    for (int q = 0; q < WM * WM; q++) {
        int r = 0;
#pragma unroll
        for (int t = 0; t < WM * WM; t++) {
            const bool before = v[t] < v[q] || (t < q && v[t] == v[q]);
            r += (((ok >> t) & 1u) && before) ? 1 : 0;
        }
        const bool live = (ok >> q) & 1u;
        if (live && r == k1) a = v[q];
        if (live && r == k2) b = v[q];
    }
    return (n & 1) ? a : (a + b) / 2;
}

// _medmad_of! (src/window.jl:535-545) on a small window; false when every value is NaN.
template <int WM>
__device__ inline bool medmad_small(const float* a, const uint8_t* kin, int64_t i, int64_t j,
                                    int64_t nr, int64_t nc, Win w, float& med, float& mad) {
    float v[WM * WM];
    uint32_t ok;
    const int n = gather_small<WM, true>(a, kin, i, j, nr, nc, w, v, ok);
    if (n == 0) return false;
    const float m = median_small<WM>(v, ok, n);
#pragma unroll
    for (int q = 0; q < WM * WM; q++) v[q] = fabsf(v[q] - m);
    med = m;
    mad = median_small<WM>(v, ok, n);
    return true;
}

// Stage 1 (agreement counts, src/window.jl:594-619) on values masked by keep_in. WM > 0: windows
// of side at most WM gathered into registers; WM == 0: any window.
template <int WM>
__global__ void gardner_stage1_kernel(const float* nx, const float* ny, const uint8_t* kin,
                                      uint8_t* kout, int64_t nr, int64_t nc, Win w, float tol,
                                      float needed) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr * nc) return;
    int64_t j = k / nr, i = k - j * nr;
    float cx = masked(nx, kin, k), cy = masked(ny, kin, k);
    float ax = 0.0f, ay = 0.0f;
    if (WM > 0) {
        constexpr int S = WM > 0 ? WM : 1;
        float v[S * S];
        uint32_t ok;
        if (!isnan(cx)) {
            gather_small<S, true>(nx, kin, i, j, nr, nc, w, v, ok);
            int c = 0;
#pragma unroll
            for (int q = 0; q < S * S; q++) c += (((ok >> q) & 1u) && fabsf(v[q] - cx) < tol) ? 1 : 0;
            ax = (float)c;
        }
        if (!isnan(cy)) {
            gather_small<S, true>(ny, kin, i, j, nr, nc, w, v, ok);
            int c = 0;
#pragma unroll
            for (int q = 0; q < S * S; q++) c += (((ok >> q) & 1u) && fabsf(v[q] - cy) < tol) ? 1 : 0;
            ay = (float)c;
        }
        kout[k] = (ax >= needed && ay >= needed) ? 1 : 0;
        return;
    }
    int64_t j0 = j - w.lx > 0 ? j - w.lx : 0, j1 = j + w.rx < nc - 1 ? j + w.rx : nc - 1;
    int64_t i0 = i - w.ly > 0 ? i - w.ly : 0, i1 = i + w.ry < nr - 1 ? i + w.ry : nr - 1;
    if (!isnan(cx)) {
        int c = 0;
        for (int64_t jj = j0; jj <= j1; jj++)
            for (int64_t ii = i0; ii <= i1; ii++) {
                float v = masked(nx, kin, ii + jj * nr);
                if (fabsf(v - cx) < tol) c++;
            }
        ax = (float)c;
    }
    if (!isnan(cy)) {
        int c = 0;
        for (int64_t jj = j0; jj <= j1; jj++)
            for (int64_t ii = i0; ii <= i1; ii++) {
                float v = masked(ny, kin, ii + jj * nr);
                if (fabsf(v - cy) < tol) c++;
            }
        ay = (float)c;
    }
    kout[k] = (ax >= needed && ay >= needed) ? 1 : 0;
}

// ---- 5 x 5 windows, one lane per window slot --------------------------------------------------
// Half a wave (32 lanes, 25 used) per grid point: lane s holds slot s = (jj - j0) * 5 + (ii - i0)
// of the window in the reference's column-major order, NaN where the window is clamped or the
// value is masked. Ranks follow the stable insertion sort as above (slot t before s wins a tie).
#define W5_PTS 8   // grid points per 256-thread workgroup

struct Slot5 {
    int64_t k;       // the point's linear index (or -1 past the end)
    int64_t kk;      // this lane's window element (or -1)
    int s, half;     // slot 0..31 (25..31 unused), half-wave
};

__device__ inline Slot5 slot5(int64_t n, int64_t nr, int64_t nc, Win w) {
    Slot5 q;
    const int lane = threadIdx.x & 63;
    q.half = lane >> 5;
    q.s = lane & 31;
    const int64_t k = (int64_t)blockIdx.x * W5_PTS + (threadIdx.x >> 5);
    q.k = k < n ? k : -1;
    q.kk = -1;
    if (q.k >= 0 && q.s < 25) {
        const int64_t j = k / nr, i = k - j * nr;
        const int cj = q.s / 5, ci = q.s - 5 * cj;
        const int64_t jj = j - w.lx + cj, ii = i - w.ly + ci;
        if (cj <= w.lx + w.rx && ci <= w.ly + w.ry && jj >= 0 && jj < nc && ii >= 0 && ii < nr)
            q.kk = ii + jj * nr;
    }
    return q;
}

// stable-sort rank of this lane's value among the non-NaN values of its half-wave: the
// half-wave's 32 values go to its slots in LDS (sh) and are read back as broadcasts (a wave's
// LDS operations run in order, so no barrier is needed between its own write and reads)
__device__ inline int rank5(float v, const Slot5& q, float* sh) {
    sh[q.s] = v;
    __builtin_amdgcn_wave_barrier();
    int r = 0;
#pragma unroll
    for (int t = 0; t < 25; t++) {
        const float x = sh[t];
        r += (t < q.s ? x <= v : x < v) ? 1 : 0;
    }
    __builtin_amdgcn_wave_barrier();
    return r;
}

// the value at sorted positions k1 and k2 of the half-wave (every lane gets them)
__device__ inline float median5(float v, const Slot5& q, int n, float* sh) {
    const int r = rank5(v, q, sh);
    const bool live = !isnan(v);
    const int k1 = (n - 1) / 2, k2 = n / 2;
    const uint64_t b1 = __ballot(live && r == k1), b2 = __ballot(live && r == k2);
    const uint32_t h1 = (uint32_t)(b1 >> (32 * q.half)), h2 = (uint32_t)(b2 >> (32 * q.half));
    const float a = __shfl(v, (q.half << 5) + (__ffs(h1) - 1));
    const float b = __shfl(v, (q.half << 5) + (__ffs(h2) - 1));
    return (n & 1) ? a : (a + b) / 2;
}

// median and MAD of one plane's window; n = 0 leaves both NaN
__device__ inline void medmad5(float v, const Slot5& q, float& med, float& mad, float* sh) {
    const uint64_t bl = __ballot(!isnan(v));
    const int n = __popc((uint32_t)(bl >> (32 * q.half)));
    med = qnan();
    mad = qnan();
    if (n == 0) return;
    const float m = median5(v, q, n, sh);
    med = m;
    mad = median5(fabsf(v - m), q, n, sh);
}

// Stage 1 on 5 x 5 windows: the agreement count is a ballot.
__global__ __launch_bounds__(256) void gardner_stage1_w5_kernel(const float* nx, const float* ny,
                                                                const uint8_t* kin, uint8_t* kout,
                                                                int64_t nr, int64_t nc, Win w,
                                                                float tol, float needed) {
    const Slot5 q = slot5(nr * nc, nr, nc, w);
    // A point that is not kept reads NaN at its centre and counts no agreement, so with a
    // positive threshold it is rejected whatever its window: a wave whose two points are both
    // such writes them and leaves (levels with few measured points are mostly such waves).
    if (needed > 0.0f && __ballot(q.k >= 0 && kin[q.k]) == 0) {
        if (q.k >= 0 && q.s == 0) kout[q.k] = 0;
        return;
    }
    const float vx = q.kk >= 0 ? masked(nx, kin, q.kk) : qnan();
    const float vy = q.kk >= 0 ? masked(ny, kin, q.kk) : qnan();
    const float cx = q.k >= 0 ? masked(nx, kin, q.k) : qnan();
    const float cy = q.k >= 0 ? masked(ny, kin, q.k) : qnan();
    const uint64_t bx = __ballot(fabsf(vx - cx) < tol), by = __ballot(fabsf(vy - cy) < tol);
    const float ax = isnan(cx) ? 0.0f : (float)__popc((uint32_t)(bx >> (32 * q.half)));
    const float ay = isnan(cy) ? 0.0f : (float)__popc((uint32_t)(by >> (32 * q.half)));
    if (q.k >= 0 && q.s == 0) kout[q.k] = (ax >= needed && ay >= needed) ? 1 : 0;
}

// Stage 2 on 5 x 5 windows (src/outliers.jl:140-157).
__global__ __launch_bounds__(256) void gardner_stage2_w5_kernel(
    const float* nx, const float* ny, const uint8_t* kin, uint8_t* kout, const int64_t* rxa,
    const int64_t* rya, int64_t nr, int64_t nc, Win w, double scale, int64_t up) {
    const Slot5 q = slot5(nr * nc, nr, nc, w);
    // A point that is not kept is rejected whatever its window: a wave whose two points are both
    // such writes them and leaves (levels with few measured points are mostly such waves).
    if (__ballot(q.k >= 0 && kin[q.k]) == 0) {
        if (q.k >= 0 && q.s == 0) kout[q.k] = 0;
        return;
    }
    const float vx = q.kk >= 0 ? masked(nx, kin, q.kk) : qnan();
    const float vy = q.kk >= 0 ? masked(ny, kin, q.kk) : qnan();
    __shared__ float sh[256];   // each half-wave's window values, for the ranks
    float* hs = sh + (threadIdx.x & ~31u);
    float medx, madx, medy, mady;
    medmad5(vx, q, medx, madx, hs);
    medmad5(vy, q, medy, mady, hs);
    if (q.k < 0 || q.s != 0) return;
    const int64_t k = q.k;
    if (!kin[k]) { kout[k] = 0; return; }
    const int64_t rx = rxa[k], ry = rya[k];
    double floor_x = rx > 0 ? 2.0 / (double)(up * rx) : INFINITY;
    double floor_y = ry > 0 ? 2.0 / (double)(up * ry) : INFINITY;
    double tx = scale * (double)madx, ty = scale * (double)mady;
    double tolx = (isnan(tx) || isnan(floor_x)) ? NAN : (tx < floor_x ? floor_x : tx);
    double toly = (isnan(ty) || isnan(floor_y)) ? NAN : (ty < floor_y ? floor_y : ty);
    kout[k] = ((double)fabsf(nx[k] - medx) <= tolx && (double)fabsf(ny[k] - medy) <= toly) ? 1 : 0;
}

// median and MAD of the non-NaN masked values around (i, j) (src/window.jl:513-563)
__device__ inline bool medmad(const float* a, const uint8_t* kin, int64_t i, int64_t j,
                              int64_t nr, int64_t nc, Win w, float& med, float& mad) {
    float buf[MAXWIN];
    int n = 0;
    int64_t j0 = j - w.lx > 0 ? j - w.lx : 0, j1 = j + w.rx < nc - 1 ? j + w.rx : nc - 1;
    int64_t i0 = i - w.ly > 0 ? i - w.ly : 0, i1 = i + w.ry < nr - 1 ? i + w.ry : nr - 1;
    for (int64_t jj = j0; jj <= j1; jj++)
        for (int64_t ii = i0; ii <= i1; ii++) {
            float v = masked(a, kin, ii + jj * nr);
            if (isnan(v)) continue;
            buf[n++] = v;
        }
    if (n == 0) return false;
    isort(buf, n);
    float m = smedian(buf, n);
    for (int k = 0; k < n; k++) buf[k] = fabsf(buf[k] - m);
    isort(buf, n);
    med = m;
    mad = smedian(buf, n);
    return true;
}

// Stage 2 (median-deviation test with the quantisation floor), src/outliers.jl:140-157
// WM > 0: windows of side at most WM in registers; WM == 0: any window up to MAXWIN values.
template <int WM>
__global__ void gardner_stage2_kernel(const float* nx, const float* ny, const uint8_t* kin,
                                      uint8_t* kout, const int64_t* rxa, const int64_t* rya,
                                      int64_t nr, int64_t nc, Win w, double scale, int64_t up) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr * nc) return;
    if (!kin[k]) { kout[k] = 0; return; }
    int64_t j = k / nr, i = k - j * nr;
    float medx = qnan(), madx = qnan(), medy = qnan(), mady = qnan();
    if (WM > 0) {
        constexpr int S = WM > 0 ? WM : 1;
        medmad_small<S>(nx, kin, i, j, nr, nc, w, medx, madx);
        medmad_small<S>(ny, kin, i, j, nr, nc, w, medy, mady);
    } else {
        medmad(nx, kin, i, j, nr, nc, w, medx, madx);
        medmad(ny, kin, i, j, nr, nc, w, medy, mady);
    }
    int64_t rx = rxa[k], ry = rya[k];
    double floor_x = rx > 0 ? 2.0 / (double)(up * rx) : INFINITY;
    double floor_y = ry > 0 ? 2.0 / (double)(up * ry) : INFINITY;
    double tx = scale * (double)madx, ty = scale * (double)mady;
    double tolx = (isnan(tx) || isnan(floor_x)) ? NAN : (tx < floor_x ? floor_x : tx);
    double toly = (isnan(ty) || isnan(floor_y)) ? NAN : (ty < floor_y ? floor_y : ty);
    kout[k] = ((double)fabsf(nx[k] - medx) <= tolx && (double)fabsf(ny[k] - medy) <= toly) ? 1 : 0;
}

// keep &= measured (src/multichip.jl:259-262), then numer/denom of the valid-fraction gate
// (265-276)
__global__ void fraction_kernel(const int64_t* rx, const uint8_t* meas, uint8_t* keep,
                                int64_t n, unsigned long long* out) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int d = 0, u = 0;
    if (k < n) {
        const uint8_t kp = (keep[k] && meas[k]) ? 1 : 0;
        keep[k] = kp;
        if (rx[k] > 0 && meas[k]) { d = 1; u = kp; }
    }
    unsigned long long cd = __syncthreads_count(d), cu = __syncthreads_count(u);
    if (threadIdx.x == 0) {
        if (cd) atomicAdd(&out[0], cd);
        if (cu) atomicAdd(&out[1], cu);
    }
}

// dilate_within (src/resample.jl:211-219) directly: a cell is set when a kept cell lies within
// squared distance r2 of it. The squared distances are exact integers in Float64, so this is the
// distance-transform form below cell for cell (no kept cell at all: every distance infinite,
// nothing set). A workgroup takes a 16 x 16 tile of cells, stages the tile and its margin of
// R = floor(sqrt(r2)) cells in LDS, and each thread ORs the kept flags of its disc without
// branches; the pipeline takes it for radii up to kDiscMax, where it is many times faster than
// the two serial transform passes (one thread per row or column).
constexpr int kDiscMax = 24, kDiscTile = 16;
__global__ __launch_bounds__(kDiscTile * kDiscTile) void dilate_disc_kernel(
    const uint8_t* m, int64_t nr, int64_t nc, int R, double r2, uint8_t* out) {
    constexpr int S = kDiscTile + 2 * kDiscMax;
    __shared__ uint8_t tile[S * S];
    const int w = kDiscTile + 2 * R;
    const int64_t i0 = (int64_t)blockIdx.x * kDiscTile - R, j0 = (int64_t)blockIdx.y * kDiscTile - R;
    for (int e = threadIdx.x; e < w * w; e += kDiscTile * kDiscTile) {
        const int a = e % w, b = e / w;
        const int64_t ii = i0 + a, jj = j0 + b;
        tile[e] = (ii >= 0 && ii < nr && jj >= 0 && jj < nc) ? m[ii + jj * nr] : 0;
    }
    __syncthreads();
    const int li = threadIdx.x % kDiscTile, lj = threadIdx.x / kDiscTile;
    const int64_t i = i0 + R + li, j = j0 + R + lj;
    if (i >= nr || j >= nc) return;
    uint32_t hit = 0;
    for (int dj = -R; dj <= R; dj++) {
        const uint8_t* col = tile + (lj + R + dj) * w + li + R;
        for (int di = -R; di <= R; di++)
            if ((double)(di * di + dj * dj) <= r2) hit |= col[di];
    }
    out[i + j * nr] = hit ? 1 : 0;
}

// Exact squared Euclidean distance transform (src/resample.jl:225-325): column pass.
__global__ void edt_col_kernel(const uint8_t* m, int64_t nr, int64_t nc, double* d2) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nc) return;
    double run = INFINITY;
    for (int64_t i = 0; i < nr; i++) {
        run = m[i + j * nr] ? 0.0 : run + 1.0;
        d2[i + j * nr] = run;
    }
    run = INFINITY;
    for (int64_t i = nr - 1; i >= 0; i--) {
        run = m[i + j * nr] ? 0.0 : run + 1.0;
        double c = d2[i + j * nr];
        d2[i + j * nr] = c < run ? c : run;
    }
    for (int64_t i = 0; i < nr; i++) {
        double v = d2[i + j * nr];
        d2[i + j * nr] = isinf(v) ? INFINITY : v * v;
    }
}

// Row pass: lower envelope of parabolas (Felzenszwalb-Huttenlocher), 1-based as the reference.
// Writes the dilation d2 <= r2 directly. Scratch: v (nc int64), z (nc+1 double), dd (nc double)
// per row.
__global__ void edt_row_dilate_kernel(const double* d2, int64_t nr, int64_t nc, double r2,
                                      int64_t* vs, double* zs, double* dds, uint8_t* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nr) return;
    int64_t* v = vs + i * nc;
    double* z = zs + i * (nc + 1);
    double* dd = dds + i * nc;
#define F(q) d2[i + ((q) - 1) * nr]
    bool allinf = true;
    for (int64_t q = 1; q <= nc; q++)
        if (!isinf(F(q))) { allinf = false; break; }
    if (allinf) {
        for (int64_t q = 0; q < nc; q++) out[i + q * nr] = INFINITY <= r2 ? 1 : 0;
        return;
    }
    int64_t k = 1;
    v[0] = 1;
    z[0] = -INFINITY;
    z[1] = INFINITY;
    for (int64_t q = 2; q <= nc; q++) {
        if (isinf(F(q))) continue;
        double s = 0.0;
        for (;;) {
            int64_t p = v[k - 1];
            s = ((F(q) + (double)(q * q)) - (F(p) + (double)(p * p))) / (double)(2 * q - 2 * p);
            if (s > z[k - 1]) break;
            k -= 1;
            if (k == 0) break;
        }
        if (k == 0) {
            k = 1;
            v[0] = q;
            z[0] = -INFINITY;
            z[1] = INFINITY;
        } else {
            k += 1;
            v[k - 1] = q;
            z[k - 1] = s;
            z[k] = INFINITY;
        }
    }
    k = 1;
    for (int64_t q = 1; q <= nc; q++) {
        while (z[k] < (double)q) k += 1;
        int64_t p = v[k - 1];
        dd[q - 1] = (double)((q - p) * (q - p)) + F(p);
    }
    for (int64_t q = 0; q < nc; q++) out[i + q * nr] = dd[q] <= r2 ? 1 : 0;
#undef F
}

// resample(grown, (nr, nc), Nearest()) .> 0.5 (src/resample.jl:56-71), then zero the radii
// outside the coarse gate (src/multichip.jl:171-176).
// frac = {denom, numer} of the valid-fraction gate: with denom == 0 or numer / denom below
// min_fraction the coarse pass returns `nothing` (src/multichip.jl:251-276) and the level has no
// searchable point, which leaves every result as the reference's skipped level does.
__global__ void nearest_gate_kernel(const uint8_t* grown, int64_t sr, int64_t sc, int64_t dr,
                                    int64_t dc, int64_t* rx, int64_t* ry, uint8_t* srch,
                                    const unsigned long long* frac, double min_fraction) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= dr * dc) return;
    const unsigned long long den = frac[0], num = frac[1];
    if (den == 0 || (double)num / (double)den < min_fraction) {
        rx[k] = 0;
        ry[k] = 0;
        srch[k] = 0;
        return;
    }
    int64_t j = k / dr, i = k - j * dr;
    double ys = (double)sr / (double)dr, xs = (double)sc / (double)dc;
    int64_t sj = (int64_t)floor(((double)(j + 1) - 0.5) * xs) + 1;
    sj = sj < 1 ? 1 : (sj > sc ? sc : sj);
    int64_t si = (int64_t)floor(((double)(i + 1) - 0.5) * ys) + 1;
    si = si < 1 ? 1 : (si > sr ? sr : si);
    float g = grown[(si - 1) + (sj - 1) * sr] ? 1.0f : 0.0f;
    if (!(g > 0.5f)) { rx[k] = 0; ry[k] = 0; }
    srch[k] = (rx[k] > 0 && ry[k] > 0) ? 1 : 0;
}

// _reject_and_fill! (src/multichip.jl:318-330): rejected points become NaN; clears `filled`.
__global__ void reject_apply_kernel(const uint8_t* keep, float* dx, float* dy, float* corr,
                                    uint8_t* filled, int64_t n) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    filled[k] = 0;
    if (!keep[k]) { dx[k] = qnan(); dy[k] = qnan(); corr[k] = qnan(); }
}

// _fill_holes! phase one (src/multichip.jl:347-390): pending medians for NaN points with at
// least `needed` measured neighbours, from the field as it stood before this pass.
__global__ void fill_pending_kernel(const float* dx, const float* dy, int64_t nr, int64_t nc,
                                    int lo, int hi, int needed, float* px, float* py,
                                    uint8_t* pend) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr * nc) return;
    pend[k] = 0;
    if (!isnan(dx[k])) return;
    int64_t j = k / nr, i = k - j * nr;
    float bx[MAXWIN], by[MAXWIN];
    int n = 0;
    int64_t j0 = j - lo > 0 ? j - lo : 0, j1 = j + hi < nc - 1 ? j + hi : nc - 1;
    int64_t i0 = i - lo > 0 ? i - lo : 0, i1 = i + hi < nr - 1 ? i + hi : nr - 1;
    for (int64_t jj = j0; jj <= j1; jj++)
        for (int64_t ii = i0; ii <= i1; ii++) {
            float v = dx[ii + jj * nr];
            if (isnan(v)) continue;
            bx[n] = v;
            by[n] = dy[ii + jj * nr];
            n++;
        }
    if (n < needed) return;
    isort(bx, n);
    isort(by, n);
    px[k] = smedian(bx, n);
    py[k] = smedian(by, n);
    pend[k] = 1;
}

__global__ void fill_apply_kernel(float* dx, float* dy, const float* px, const float* py,
                                  const uint8_t* pend, uint8_t* filled, int64_t n) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && pend[k]) { dx[k] = px[k]; dy[k] = py[k]; filled[k] = 1; }
}

// _merge_level! (src/multichip.jl:401-422)
__global__ void merge_kernel(float* rdx, float* rdy, float* rcorr, uint16_t* rcs, uint8_t* rit,
                             const float* dx, const float* dy, const float* corr,
                             const uint8_t* filled, uint16_t cs, int64_t n) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (rcs[k] != 0) return;
    if (isnan(dx[k])) return;
    rdx[k] = dx[k]; rdy[k] = dy[k]; rcorr[k] = corr[k]; rcs[k] = cs; rit[k] = filled[k];
}

__global__ void zero_u64_kernel(unsigned long long* c, int k) {
    if ((int)threadIdx.x < k) c[threadIdx.x] = 0ull;
}

}  // namespace ar

using namespace ar;

namespace {

// Carves the driver's planes out of one grow-only context buffer: a first pass (base == nullptr)
// only adds up the sizes, the second hands out 256-byte-aligned pointers. No allocation per call.
struct Carve {
    char* base = nullptr;
    size_t off = 0;
    template <typename T>
    T* get(int64_t n) {
        const size_t b = (((size_t)(n > 0 ? n : 1) * sizeof(T)) + 255) & ~(size_t)255;
        T* p = base ? (T*)(base + off) : nullptr;
        off += b;
        return p;
    }
};

}  // namespace

static ar_status mfail(ar_context* ctx, ar_status st, const std::string& m) {
    ctx->err = m;
    return st;
}

#define MCHK(expr)                                                                            \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            return mfail(ctx, _e == hipErrorOutOfMemory ? AR_ERR_OUT_OF_MEMORY : AR_ERR_DEVICE, \
                         std::string(#expr) + ": " + hipGetErrorString(_e));                  \
    } while (0)

// Readback of the first `words` counters into page-locked staging (rb), then into out.
static ar_status read_counts(ar_context* ctx, const unsigned long long* cnt, int words,
                             unsigned long long* rb, unsigned long long* out) {
    MCHK(hipMemcpyAsync(rb, cnt, 8 * words, hipMemcpyDeviceToHost, ctx->stream));
    MCHK(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < words; k++) out[k] = rb[k];
    return AR_OK;
}

// Counts of two flag arrays of n bytes with one readback.
static ar_status count_flags2(ar_context* ctx, const uint8_t* f0, const uint8_t* f1, int64_t n,
                              unsigned long long* cnt, unsigned long long* rb,
                              unsigned long long* out) {
    zero_u64_kernel<<<1, 64, 0, ctx->stream>>>(cnt, 2);
    count2_kernel<<<gblocks(n), GK, 0, ctx->stream>>>(f0, f1, n, cnt);
    MCHK(hipGetLastError());
    return read_counts(ctx, cnt, 2, rb, out);
}

// reject_outliers with the Gardner filter (or the measured mask when outliers are off) over the
// measured points (!isnan(dx)), which are also written to meas.
static ar_status reject(ar_context* ctx, const float* dx, const float* dy, const int64_t* rx,
                        const int64_t* ry, uint8_t* meas, int64_t nr, int64_t nc,
                        int64_t iterations, float* nxb, float* nyb, uint8_t* ka, uint8_t* kb,
                        uint8_t** keep_out) {
    const ar_params& p = ctx->params;
    int64_t n = nr * nc;
    if (!p.outliers) {
        measured_kernel<<<gblocks(n), GK, 0, ctx->stream>>>(dx, n, meas);
        MCHK(hipGetLastError());
        *keep_out = meas;
        return AR_OK;
    }
    int64_t w = p.gardner_window;
    int64_t wx = w < nc ? w : nc, wy = w < nr ? w : nr;
    Win win;
    win.lx = (int)(wx / 2);
    win.ly = (int)(wy / 2);
    win.rx = (int)(wx - 1 - win.lx);
    win.ry = (int)(wy - 1 - win.ly);
    normalise_kernel<<<gblocks(n), GK, 0, ctx->stream>>>(dx, dy, rx, ry, meas, n, nxb, nyb, ka);
    float needed = (float)(p.min_agree_fraction * (double)(w * w));
    float tol = (float)p.agree_tolerance;
    uint8_t *kin = ka, *kout = kb;
    for (int64_t it = 0; it < iterations; it++) {
        if (wx <= 5 && wy <= 5)
            gardner_stage1_w5_kernel<<<(unsigned)((n + W5_PTS - 1) / W5_PTS), 256, 0, ctx->stream>>>(
                nxb, nyb, kin, kout, nr, nc, win, tol, needed);
        else
            gardner_stage1_kernel<0><<<gblocks(n), GK, 0, ctx->stream>>>(nxb, nyb, kin, kout, nr,
                                                                         nc, win, tol, needed);
        uint8_t* t = kin; kin = kout; kout = t;
    }
    int64_t n2 = iterations - 1 > 1 ? iterations - 1 : 1;
    for (int64_t it = 0; it < n2; it++) {
        if (wx <= 5 && wy <= 5)
            gardner_stage2_w5_kernel<<<(unsigned)((n + W5_PTS - 1) / W5_PTS), 256, 0, ctx->stream>>>(
                nxb, nyb, kin, kout, rx, ry, nr, nc, win, p.mad_scale, p.upsampling);
        else
            gardner_stage2_kernel<0><<<gblocks(n), GK, 0, ctx->stream>>>(
                nxb, nyb, kin, kout, rx, ry, nr, nc, win, p.mad_scale, p.upsampling);
        uint8_t* t = kin; kin = kout; kout = t;
    }
    MCHK(hipGetLastError());
    *keep_out = kin;
    return AR_OK;
}

// The grid's fields copied into page-locked staging (x, y, dx_prior, dy_prior, radius_x,
// radius_y, each a contiguous plane of N 8-byte values) in one pass that also gathers what the
// level planning needs: whether every coordinate and prior is finite and below 1e15 in
// magnitude, the coordinate extents, the largest prior and the largest radii of the searchable
// points. Large grids are split over up to 8 threads when `threads` is set.
struct GridScan {
    bool ok = true;
    double xlo = INFINITY, xhi = -INFINITY, ylo = INFINITY, yhi = -INFINITY, pmax = 0.0;
    int64_t rx = 0, ry = 0;
};

static void grid_scan_range(const ar_points* g, char* pin, int64_t N, int64_t a, int64_t b,
                            GridScan& r) {
    const void* src[6] = {g->x, g->y, g->dx_prior, g->dy_prior, g->radius_x, g->radius_y};
    for (int f = 0; f < 6; f++)
        std::memcpy(pin + (size_t)f * 8 * N + 8 * (size_t)a, (const char*)src[f] + 8 * (size_t)a,
                    8 * (size_t)(b - a));
    bool ok = true;
    double xlo = r.xlo, xhi = r.xhi, ylo = r.ylo, yhi = r.yhi, pmax = r.pmax;
    int64_t rx = r.rx, ry = r.ry;
    for (int64_t k = a; k < b; k++) {
        const double x = g->x[k], y = g->y[k];
        const double ax = std::fabs(g->dx_prior[k]), ay = std::fabs(g->dy_prior[k]);
        ok = ok && std::fabs(x) <= 1e15 && std::fabs(y) <= 1e15 && ax <= 1e15 && ay <= 1e15;
        xlo = x < xlo ? x : xlo;
        xhi = x > xhi ? x : xhi;
        ylo = y < ylo ? y : ylo;
        yhi = y > yhi ? y : yhi;
        const double m = ax > ay ? ax : ay;
        pmax = m > pmax ? m : pmax;
        const int64_t gx = g->radius_x[k], gy = g->radius_y[k];
        if (gx > 0 && gy > 0) {
            rx = gx > rx ? gx : rx;
            ry = gy > ry ? gy : ry;
        }
    }
    r.ok = r.ok && ok;
    r.xlo = xlo; r.xhi = xhi; r.ylo = ylo; r.yhi = yhi; r.pmax = pmax; r.rx = rx; r.ry = ry;
}

static GridScan grid_scan(const ar_points* g, char* pin, int64_t N, bool threads) {
    int nt = threads ? (int)std::min<int64_t>(8, N / 32768) : 1;
    if (nt < 1) nt = 1;
    std::vector<GridScan> part(nt);
    std::vector<std::thread> ts;
    const int64_t per = (N + nt - 1) / nt;
    for (int t = 1; t < nt; t++) {
        const int64_t a = t * per, b = std::min(N, a + per);
        if (a < b) ts.emplace_back([=, &part] { grid_scan_range(g, pin, N, a, b, part[t]); });
    }
    grid_scan_range(g, pin, N, 0, std::min(N, per), part[0]);
    for (std::thread& t : ts) t.join();
    GridScan r = part[0];
    for (int t = 1; t < nt; t++) {
        const GridScan& q = part[t];
        r.ok = r.ok && q.ok;
        r.xlo = std::min(r.xlo, q.xlo); r.xhi = std::max(r.xhi, q.xhi);
        r.ylo = std::min(r.ylo, q.ylo); r.yhi = std::max(r.yhi, q.yhi);
        r.pmax = std::max(r.pmax, q.pmax);
        r.rx = std::max(r.rx, q.rx); r.ry = std::max(r.ry, q.ry);
    }
    return r;
}

extern "C" ar_status ar_run_multichip(ar_context* ctx, const ar_points* grid, int64_t nr,
                                      int64_t nc, float* dx, float* dy, float* corr,
                                      uint16_t* chip_size, uint8_t* interpolated) {
    if (!ctx) return AR_ERR_INVALID_ARGUMENT;
    if (!grid || !dx || !dy || !corr || !chip_size || !interpolated || !grid->x || !grid->y ||
        !grid->radius_x || !grid->radius_y || !grid->dx_prior || !grid->dy_prior)
        return mfail(ctx, AR_ERR_INVALID_ARGUMENT, "NULL pointer argument");
    if (nr <= 0 || nc <= 0) return mfail(ctx, AR_ERR_INVALID_ARGUMENT, "empty grid");
    if (!ctx->has_pair) return mfail(ctx, AR_ERR_INVALID_ARGUMENT, "no image pair set");
    const ar_params& p = ctx->params;
    if (p.gardner_window < 1 || p.gardner_window * p.gardner_window > MAXWIN ||
        p.fill_window < 1 || p.fill_window * p.fill_window > MAXWIN)
        return mfail(ctx, AR_ERR_UNSUPPORTED, "outlier and fill windows above 9 are not supported");
    if (p.coarse_stride < 1) return mfail(ctx, AR_ERR_INVALID_ARGUMENT, "coarse_stride must be positive");
    MCHK(hipSetDevice(ctx->device));
    const int64_t N = nr * nc;
    hipStream_t s = ctx->stream;
    ctx->box_key.ptr = nullptr;   // box images are shared within this run only
    // planes: the grid's six input fields and the five result planes each contiguous (one copy
    // in, one copy out through page-locked staging), then the level and coarse-grid planes
    int64_t stride = p.coarse_stride;
    int64_t nrc = nr / stride, ncc = nc / stride, M = nrc * ncc;
    double *gx, *gy, *gpdx, *gpdy;
    int64_t *grx, *gry, *lrx, *lry, *lcx, *lcy;
    float *rdx, *rdy, *rc, *fdx, *fdy, *fc, *nxb, *nyb, *px, *py;
    uint16_t* rcs;
    uint8_t *rit, *fs, *meas, *ka, *kb, *pend, *filled, *flags;
    unsigned long long* cnt;
    double *cxp, *cyp, *cpdx, *cpdy, *d2, *ez, *edd;
    int64_t *crx, *cry, *ccx, *ccy, *ev;
    float *cdx, *cdy, *ccr, *cnx, *cny;
    uint8_t *cs8, *cmeas, *cka, *ckb, *grown;
    char *gin, *res;
    auto carve = [&](Carve& C) {
        gin = C.get<char>(48 * N);
        gx = (double*)gin; gy = gx + N; gpdx = gy + N; gpdy = gpdx + N;
        grx = (int64_t*)(gpdy + N); gry = grx + N;
        res = C.get<char>(15 * N);
        rdx = (float*)res; rdy = rdx + N; rc = rdy + N;
        rcs = (uint16_t*)(rc + N); rit = (uint8_t*)(rcs + N);
        lrx = C.get<int64_t>(N); lry = C.get<int64_t>(N);
        lcx = C.get<int64_t>(N); lcy = C.get<int64_t>(N);
        fdx = C.get<float>(N); fdy = C.get<float>(N); fc = C.get<float>(N);
        fs = C.get<uint8_t>(N); meas = C.get<uint8_t>(N);
        ka = C.get<uint8_t>(N); kb = C.get<uint8_t>(N);
        nxb = C.get<float>(N); nyb = C.get<float>(N);
        px = C.get<float>(N); py = C.get<float>(N);
        pend = C.get<uint8_t>(N); filled = C.get<uint8_t>(N);
        flags = C.get<uint8_t>(N);
        cnt = C.get<unsigned long long>(4);
        cxp = C.get<double>(M); cyp = C.get<double>(M);
        cpdx = C.get<double>(M); cpdy = C.get<double>(M);
        crx = C.get<int64_t>(M); cry = C.get<int64_t>(M);
        ccx = C.get<int64_t>(M); ccy = C.get<int64_t>(M);
        cdx = C.get<float>(M); cdy = C.get<float>(M); ccr = C.get<float>(M);
        cs8 = C.get<uint8_t>(M); cmeas = C.get<uint8_t>(M);
        cka = C.get<uint8_t>(M); ckb = C.get<uint8_t>(M);
        cnx = C.get<float>(M); cny = C.get<float>(M);
        d2 = C.get<double>(M); ev = C.get<int64_t>(M);
        ez = C.get<double>(M + nrc); edd = C.get<double>(M);
        grown = C.get<uint8_t>(M);
    };
    ar_status st;
    Carve dry;
    carve(dry);
    if ((st = ctx->mc.reserve(ctx, dry.off))) return st;
    Carve C;
    C.base = (char*)ctx->mc.ptr;
    carve(C);
    // page-locked staging: the grid fields on the way in, the results on the way out, and the
    // count readbacks (after both)
    void* pinv = nullptr;
    const size_t rb_off = ((size_t)48 * N + 255) & ~(size_t)255;
    if ((st = ar_pinned(ctx, rb_off + 256, &pinv))) return st;
    char* pin = (char*)pinv;
    unsigned long long* rb = (unsigned long long*)(pin + rb_off);
    // (a server's slot contexts scan on their worker's thread alone: the slots run side by side)
    const GridScan gs = grid_scan(grid, pin, N, ctx->side_prep);
    MCHK(hipMemcpyAsync(gin, pin, 48 * (size_t)N, hipMemcpyHostToDevice, s));
    fill_result_kernel<<<gblocks(N), GK, 0, s>>>(rdx, rdy, rc, rcs, rit, N);
    MCHK(hipGetLastError());

    // What the host knows of every pass's geometry without asking the device: the grid's
    // extent, its largest radius and prior. A level's points (coarse and fine) sit at grid
    // positions with at most that radius, so when every grid position's window fits the image at
    // that radius the passes need no padding and their geometry is known here.
    // (floor is monotone: the floors of the extremes are the extremes of the floors)
    const bool grid_known = gs.ok;
    const double fx_lo = grid_known ? std::floor(gs.xlo) : 0;
    const double fx_hi = grid_known ? std::floor(gs.xhi) : 0;
    const double fy_lo = grid_known ? std::floor(gs.ylo) : 0;
    const double fy_hi = grid_known ? std::floor(gs.yhi) : 0;
    const double prior_max = grid_known ? gs.pmax : 0;
    const int64_t r0x = gs.rx, r0y = gs.ry;

    if (p.level_measure_count < 0 || p.level_measure_count > 8)
        return mfail(ctx, AR_ERR_INVALID_ARGUMENT, "level_measure_count must be in 0..8");
    {
        // _check_measures (src/types.jl:899-906): a tuple longer than the level list would
        // leave measures that never run
        int64_t nlevels = 0;
        for (int64_t cs = p.chip_size_base; cs > 0 && cs <= p.chip_size_max; cs *= 2)
            if (cs >= p.chip_size_min) nlevels++;
        if (p.level_measure_count > nlevels)
            return mfail(ctx, AR_ERR_INVALID_ARGUMENT,
                         "`similarity` names " + std::to_string(p.level_measure_count) +
                             " measures but there are only " + std::to_string(nlevels) +
                             " chip-size levels. Widen the `chip_size_min`/`chip_size_max` range, "
                             "or name fewer measures — the last one applies to every remaining "
                             "level.");
    }
    int level = 0;
    for (int64_t cs = p.chip_size_base; cs <= p.chip_size_max; cs *= 2) {
        if (cs < p.chip_size_min) continue;
        // measure_at (src/types.jl:842-870): the level's entry of the similarity tuple, the last
        // one repeating
        const int32_t measure =
            p.level_measure_count > 0
                ? p.level_measures[level < p.level_measure_count ? level : p.level_measure_count - 1]
                : p.measure;
        level++;
        unsigned long long c2[2] = {0, 0};
        int64_t csx = cs, csy = 2 * (int64_t)std::nearbyint((double)cs * p.chip_aspect / 2.0);
        level_points_kernel<<<gblocks(N), GK, 0, s>>>(grx, gry, rcs, N, csx, csy,
                                                      p.min_search_radius, lrx, lry, lcx, lcy, flags,
                                                      pend);
        if ((st = count_flags2(ctx, pend, flags, N, cnt, rb, c2))) return st;
        if (c2[0] == 0) break;
        if (c2[1] == 0) continue;
        ar_pass_hint hint{};
        const ar_pass_hint* hp = nullptr;
        if (grid_known && r0x > 0 && r0y > 0) {
            const int64_t Rx = std::max(r0x, p.min_search_radius);
            const int64_t Ry = std::max(r0y, p.min_search_radius);
            const int64_t hx = csx / 2, hy = csy / 2;
            const bool fits = fy_lo - hy - Ry >= 1 && fy_hi + hy + Ry - 2 <= ctx->ny &&
                              fx_lo - hx - Rx >= 1 && fx_hi + hx + Rx - 2 <= ctx->nx;
            if (fits) {
                const int64_t q = (int64_t)std::ceil(prior_max);
                hint = ar_pass_hint{csx, csy, Rx, Ry, hx + Rx + q, hy + Ry + q, true};
                hp = &hint;
            }
        }

        // ---- coarse pass
        Win dummy;
        (void)dummy;
        int64_t fw = p.outliers ? p.gardner_window : 0;
        bool gate_all = nrc < fw || ncc < fw;
        if (!gate_all) {
            // Every decision of the coarse pass that ends the level early (nothing searchable or
            // measured, too small a coherent fraction, an empty gate) is made on the device: the
            // gate then leaves no searchable point, and the fine pass, rejection, filling and
            // merge change nothing, as for the reference's skipped level.
            coarse_points_kernel<<<gblocks(M), GK, 0, s>>>(gx, gy, gpdx, gpdy, lrx, lry, nr, nc, stride,
                                                           nrc, ncc, csx, csy, cxp, cyp, cpdx, cpdy,
                                                           crx, cry, ccx, ccy, cs8, cnt + 2);
            MCHK(hipGetLastError());
            ar_points cp = {cxp, cyp, crx, cry, cpdx, cpdy, ccx, ccy};
            // the level's points bound the searchable coarse points (c2[1], read back above)
            if ((st = ar_track_pass(ctx, cp, M, 1, measure, cdx, cdy, ccr, cs8, hp, (int64_t)c2[1],
                                     true)))
                return st;
            uint8_t* keep = nullptr;
            int64_t iters = p.gardner_iterations - 1 > 1 ? p.gardner_iterations - 1 : 1;
            if ((st = reject(ctx, cdx, cdy, crx, cry, cmeas, nrc, ncc, iters, cnx, cny, cka, ckb,
                             &keep)))
                return st;
            fraction_kernel<<<gblocks(M), GK, 0, s>>>(crx, cmeas, keep, M, cnt + 2);
            double r2 = (double)p.coarse_buffer * (double)p.coarse_buffer;
            const int64_t rdisc = (int64_t)std::floor(std::sqrt(r2));
            if (rdisc <= kDiscMax && ncc <= (int64_t)65535 * kDiscTile) {
                const dim3 tiles((unsigned)((nrc + kDiscTile - 1) / kDiscTile),
                                 (unsigned)((ncc + kDiscTile - 1) / kDiscTile));
                dilate_disc_kernel<<<tiles, kDiscTile * kDiscTile, 0, s>>>(keep, nrc, ncc,
                                                                           (int)rdisc, r2, grown);
            } else {
                edt_col_kernel<<<gblocks(ncc), GK, 0, s>>>(keep, nrc, ncc, d2);
                edt_row_dilate_kernel<<<gblocks(nrc), GK, 0, s>>>(d2, nrc, ncc, r2, ev, ez, edd,
                                                                   grown);
            }
            nearest_gate_kernel<<<gblocks(N), GK, 0, s>>>(grown, nrc, ncc, nr, nc, lrx, lry, flags,
                                                          cnt + 2, p.min_coarse_valid_fraction);
            MCHK(hipGetLastError());
        }

        // ---- fine pass, rejection, hole filling, merge
        ar_points lp = {gx, gy, lrx, lry, gpdx, gpdy, lcx, lcy};
        if ((st = ar_track_pass(ctx, lp, N, p.upsampling, measure, fdx, fdy, fc, fs, hp,
                                (int64_t)c2[1], true)))
            return st;
        uint8_t* keep = nullptr;
        if ((st = reject(ctx, fdx, fdy, lrx, lry, meas, nr, nc, p.gardner_iterations, nxb, nyb, ka,
                         kb, &keep)))
            return st;
        reject_apply_kernel<<<gblocks(N), GK, 0, s>>>(keep, fdx, fdy, fc, filled, N);
        int fwn = (int)p.fill_window;
        int lo = fwn / 2, hi = fwn - 1 - lo, needed = 2 * fwn * fwn / 3;
        // three passes without counting: a pass with nothing pending changes nothing, so the
        // reference's early exit is an optimisation only
        for (int pass = 0; pass < 3; pass++) {
            fill_pending_kernel<<<gblocks(N), GK, 0, s>>>(fdx, fdy, nr, nc, lo, hi, needed, px, py,
                                                          pend);
            fill_apply_kernel<<<gblocks(N), GK, 0, s>>>(fdx, fdy, px, py, pend, filled, N);
        }
        merge_kernel<<<gblocks(N), GK, 0, s>>>(rdx, rdy, rc, rcs, rit, fdx, fdy, fc, filled,
                                               (uint16_t)cs, N);
        MCHK(hipGetLastError());
    }
    MCHK(hipMemcpyAsync(pin, res, 15 * (size_t)N, hipMemcpyDeviceToHost, s));
    MCHK(hipStreamSynchronize(s));
    std::memcpy(dx, pin, 4 * (size_t)N);
    std::memcpy(dy, pin + 4 * (size_t)N, 4 * (size_t)N);
    std::memcpy(corr, pin + 8 * (size_t)N, 4 * (size_t)N);
    std::memcpy(chip_size, pin + 12 * (size_t)N, 2 * (size_t)N);
    std::memcpy(interpolated, pin + 14 * (size_t)N, (size_t)N);
    return AR_OK;
}
