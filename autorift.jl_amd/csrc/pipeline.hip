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

// count of flags[k] != 0 into *out (atomic; counts are exact integers)
__global__ void count_kernel(const uint8_t* f, int64_t n, unsigned long long* out) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int v = k < n && f[k] ? 1 : 0;
    unsigned long long c = __syncthreads_count(v);
    if (threadIdx.x == 0 && c) atomicAdd(out, c);
}

// _level_points + sanitize! (src/multichip.jl:193-213, src/points.jl:332-346)
__global__ void level_points_kernel(const int64_t* grx, const int64_t* gry, const uint16_t* rcs,
                                    int64_t n, int64_t csx, int64_t csy, int64_t minr,
                                    int64_t* rx, int64_t* ry, int64_t* cx, int64_t* cy,
                                    uint8_t* srch) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    bool wanted = rcs[k] == 0;
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
                                     int64_t* ocx, int64_t* ocy, uint8_t* srch) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
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
__global__ void normalise_kernel(const float* dx, const float* dy, const int64_t* rx,
                                 const int64_t* ry, const uint8_t* valid, int64_t n, float* nx,
                                 float* ny, uint8_t* keep) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    nx[k] = rx[k] > 0 ? dx[k] / (float)rx[k] : qnan();
    ny[k] = ry[k] > 0 ? dy[k] / (float)ry[k] : qnan();
    keep[k] = valid[k] ? 1 : 0;
}

struct Win {
    int lx, ly, rx, ry;
};

__device__ inline float masked(const float* a, const uint8_t* keep, int64_t k) {
    return keep[k] ? a[k] : qnan();
}

// Stage 1 (agreement counts, src/window.jl:594-619) on values masked by keep_in.
__global__ void gardner_stage1_kernel(const float* nx, const float* ny, const uint8_t* kin,
                                      uint8_t* kout, int64_t nr, int64_t nc, Win w, float tol,
                                      float needed) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr * nc) return;
    int64_t j = k / nr, i = k - j * nr;
    float cx = masked(nx, kin, k), cy = masked(ny, kin, k);
    float ax = 0.0f, ay = 0.0f;
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
__global__ void gardner_stage2_kernel(const float* nx, const float* ny, const uint8_t* kin,
                                      uint8_t* kout, const int64_t* rxa, const int64_t* rya,
                                      int64_t nr, int64_t nc, Win w, double scale, int64_t up) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr * nc) return;
    if (!kin[k]) { kout[k] = 0; return; }
    int64_t j = k / nr, i = k - j * nr;
    float medx = qnan(), madx = qnan(), medy = qnan(), mady = qnan();
    medmad(nx, kin, i, j, nr, nc, w, medx, madx);
    medmad(ny, kin, i, j, nr, nc, w, medy, mady);
    int64_t rx = rxa[k], ry = rya[k];
    double floor_x = rx > 0 ? 2.0 / (double)(up * rx) : INFINITY;
    double floor_y = ry > 0 ? 2.0 / (double)(up * ry) : INFINITY;
    double tx = scale * (double)madx, ty = scale * (double)mady;
    double tolx = (isnan(tx) || isnan(floor_x)) ? NAN : (tx < floor_x ? floor_x : tx);
    double toly = (isnan(ty) || isnan(floor_y)) ? NAN : (ty < floor_y ? floor_y : ty);
    kout[k] = ((double)fabsf(nx[k] - medx) <= tolx && (double)fabsf(ny[k] - medy) <= toly) ? 1 : 0;
}

__global__ void and_kernel(uint8_t* a, const uint8_t* b, int64_t n) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = (a[k] && b[k]) ? 1 : 0;
}

// numer/denom of the valid-fraction gate (src/multichip.jl:265-276)
__global__ void fraction_kernel(const int64_t* rx, const uint8_t* meas, const uint8_t* keep,
                                int64_t n, unsigned long long* out) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int d = 0, u = 0;
    if (k < n && rx[k] > 0 && meas[k]) { d = 1; u = keep[k] ? 1 : 0; }
    unsigned long long cd = __syncthreads_count(d), cu = __syncthreads_count(u);
    if (threadIdx.x == 0) {
        if (cd) atomicAdd(&out[0], cd);
        if (cu) atomicAdd(&out[1], cu);
    }
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
__global__ void nearest_gate_kernel(const uint8_t* grown, int64_t sr, int64_t sc, int64_t dr,
                                    int64_t dc, int64_t* rx, int64_t* ry, uint8_t* srch) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= dr * dc) return;
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

__global__ void reject_apply_kernel(const uint8_t* keep, float* dx, float* dy, float* corr,
                                    int64_t n) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && !keep[k]) { dx[k] = qnan(); dy[k] = qnan(); corr[k] = qnan(); }
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

__global__ void wanted_kernel(const uint16_t* rcs, int64_t n, uint8_t* w) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) w[k] = rcs[k] == 0 ? 1 : 0;
}

}  // namespace ar

using namespace ar;

namespace {

struct Alloc {
    ar_context* ctx;
    std::vector<void*> ptrs;
    hipError_t err = hipSuccess;
    template <typename T>
    T* get(int64_t n) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, (size_t)(n > 0 ? n : 1) * sizeof(T));
        if (e != hipSuccess) { err = e; return nullptr; }
        ptrs.push_back(p);
        return (T*)p;
    }
    ~Alloc() {
        for (void* p : ptrs) (void)hipFree(p);
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

// Counts of two flag arrays of n bytes with one readback.
static ar_status count_flags2(ar_context* ctx, const uint8_t* f0, const uint8_t* f1, int64_t n,
                              unsigned long long* cnt, unsigned long long* out) {
    MCHK(hipMemsetAsync(cnt, 0, 16, ctx->stream));
    count_kernel<<<gblocks(n), GK, 0, ctx->stream>>>(f0, n, cnt);
    count_kernel<<<gblocks(n), GK, 0, ctx->stream>>>(f1, n, cnt + 1);
    MCHK(hipGetLastError());
    MCHK(hipMemcpyAsync(out, cnt, 16, hipMemcpyDeviceToHost, ctx->stream));
    MCHK(hipStreamSynchronize(ctx->stream));
    return AR_OK;
}

static ar_status count_flags(ar_context* ctx, const uint8_t* f, int64_t n,
                             unsigned long long* cnt, unsigned long long* out) {
    MCHK(hipMemsetAsync(cnt, 0, 8, ctx->stream));
    count_kernel<<<gblocks(n), GK, 0, ctx->stream>>>(f, n, cnt);
    MCHK(hipGetLastError());
    MCHK(hipMemcpyAsync(out, cnt, 8, hipMemcpyDeviceToHost, ctx->stream));
    MCHK(hipStreamSynchronize(ctx->stream));
    return AR_OK;
}

// reject_outliers with the Gardner filter (or the measured mask when outliers are off).
static ar_status reject(ar_context* ctx, const float* dx, const float* dy, const int64_t* rx,
                        const int64_t* ry, const uint8_t* valid, int64_t nr, int64_t nc,
                        int64_t iterations, float* nxb, float* nyb, uint8_t* ka, uint8_t* kb,
                        uint8_t** keep_out) {
    const ar_params& p = ctx->params;
    int64_t n = nr * nc;
    if (!p.outliers) {
        MCHK(hipMemcpyAsync(ka, valid, n, hipMemcpyDeviceToDevice, ctx->stream));
        *keep_out = ka;
        return AR_OK;
    }
    int64_t w = p.gardner_window;
    int64_t wx = w < nc ? w : nc, wy = w < nr ? w : nr;
    Win win;
    win.lx = (int)(wx / 2);
    win.ly = (int)(wy / 2);
    win.rx = (int)(wx - 1 - win.lx);
    win.ry = (int)(wy - 1 - win.ly);
    normalise_kernel<<<gblocks(n), GK, 0, ctx->stream>>>(dx, dy, rx, ry, valid, n, nxb, nyb, ka);
    float needed = (float)(p.min_agree_fraction * (double)(w * w));
    float tol = (float)p.agree_tolerance;
    uint8_t *kin = ka, *kout = kb;
    for (int64_t it = 0; it < iterations; it++) {
        gardner_stage1_kernel<<<gblocks(n), GK, 0, ctx->stream>>>(nxb, nyb, kin, kout, nr, nc, win,
                                                                  tol, needed);
        uint8_t* t = kin; kin = kout; kout = t;
    }
    int64_t n2 = iterations - 1 > 1 ? iterations - 1 : 1;
    for (int64_t it = 0; it < n2; it++) {
        gardner_stage2_kernel<<<gblocks(n), GK, 0, ctx->stream>>>(
            nxb, nyb, kin, kout, rx, ry, nr, nc, win, p.mad_scale, p.upsampling);
        uint8_t* t = kin; kin = kout; kout = t;
    }
    MCHK(hipGetLastError());
    *keep_out = kin;
    return AR_OK;
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
    Alloc A{ctx};
    double* gx = A.get<double>(N); double* gy = A.get<double>(N);
    double* gpdx = A.get<double>(N); double* gpdy = A.get<double>(N);
    int64_t* grx = A.get<int64_t>(N); int64_t* gry = A.get<int64_t>(N);
    int64_t* lrx = A.get<int64_t>(N); int64_t* lry = A.get<int64_t>(N);
    int64_t* lcx = A.get<int64_t>(N); int64_t* lcy = A.get<int64_t>(N);
    float* rdx = A.get<float>(N); float* rdy = A.get<float>(N); float* rc = A.get<float>(N);
    uint16_t* rcs = A.get<uint16_t>(N); uint8_t* rit = A.get<uint8_t>(N);
    float* fdx = A.get<float>(N); float* fdy = A.get<float>(N); float* fc = A.get<float>(N);
    uint8_t* fs = A.get<uint8_t>(N); uint8_t* meas = A.get<uint8_t>(N);
    uint8_t* ka = A.get<uint8_t>(N); uint8_t* kb = A.get<uint8_t>(N);
    float* nxb = A.get<float>(N); float* nyb = A.get<float>(N);
    float* px = A.get<float>(N); float* py = A.get<float>(N);
    uint8_t* pend = A.get<uint8_t>(N); uint8_t* filled = A.get<uint8_t>(N);
    uint8_t* flags = A.get<uint8_t>(N);
    unsigned long long* cnt = A.get<unsigned long long>(4);
    // coarse grid buffers
    int64_t stride = p.coarse_stride;
    int64_t nrc = nr / stride, ncc = nc / stride, M = nrc * ncc;
    double* cxp = A.get<double>(M); double* cyp = A.get<double>(M);
    double* cpdx = A.get<double>(M); double* cpdy = A.get<double>(M);
    int64_t* crx = A.get<int64_t>(M); int64_t* cry = A.get<int64_t>(M);
    int64_t* ccx = A.get<int64_t>(M); int64_t* ccy = A.get<int64_t>(M);
    float* cdx = A.get<float>(M); float* cdy = A.get<float>(M); float* ccr = A.get<float>(M);
    uint8_t* cs8 = A.get<uint8_t>(M); uint8_t* cmeas = A.get<uint8_t>(M);
    uint8_t* cka = A.get<uint8_t>(M); uint8_t* ckb = A.get<uint8_t>(M);
    float* cnx = A.get<float>(M); float* cny = A.get<float>(M);
    double* d2 = A.get<double>(M); int64_t* ev = A.get<int64_t>(M);
    double* ez = A.get<double>(M + nrc); double* edd = A.get<double>(M);
    uint8_t* grown = A.get<uint8_t>(M);
    if (A.err != hipSuccess)
        return mfail(ctx, AR_ERR_OUT_OF_MEMORY, std::string("multichip buffers: ") + hipGetErrorString(A.err));

    MCHK(hipMemcpyAsync(gx, grid->x, 8 * N, hipMemcpyHostToDevice, s));
    MCHK(hipMemcpyAsync(gy, grid->y, 8 * N, hipMemcpyHostToDevice, s));
    MCHK(hipMemcpyAsync(gpdx, grid->dx_prior, 8 * N, hipMemcpyHostToDevice, s));
    MCHK(hipMemcpyAsync(gpdy, grid->dy_prior, 8 * N, hipMemcpyHostToDevice, s));
    MCHK(hipMemcpyAsync(grx, grid->radius_x, 8 * N, hipMemcpyHostToDevice, s));
    MCHK(hipMemcpyAsync(gry, grid->radius_y, 8 * N, hipMemcpyHostToDevice, s));
    fill_result_kernel<<<gblocks(N), GK, 0, s>>>(rdx, rdy, rc, rcs, rit, N);
    MCHK(hipGetLastError());

    // What the host knows of every pass's geometry without asking the device: the grid's
    // extent, its largest radius and prior. A level's points (coarse and fine) sit at grid
    // positions with at most that radius, so when every grid position's window fits the image at
    // that radius the passes need no padding and their geometry is known here.
    bool grid_known = true;
    double fx_lo = 0, fx_hi = 0, fy_lo = 0, fy_hi = 0, prior_max = 0;
    int64_t r0x = 0, r0y = 0;
    for (int64_t k = 0; k < N && grid_known; k++) {
        const double x = grid->x[k], y = grid->y[k];
        const double ax = std::fabs(grid->dx_prior[k]), ay = std::fabs(grid->dy_prior[k]);
        if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(ax) || !std::isfinite(ay) ||
            std::fabs(x) > 1e15 || std::fabs(y) > 1e15 || ax > 1e15 || ay > 1e15) {
            grid_known = false;
            break;
        }
        const double fx = std::floor(x), fy = std::floor(y);
        if (k == 0) { fx_lo = fx_hi = fx; fy_lo = fy_hi = fy; }
        fx_lo = std::min(fx_lo, fx); fx_hi = std::max(fx_hi, fx);
        fy_lo = std::min(fy_lo, fy); fy_hi = std::max(fy_hi, fy);
        prior_max = std::max(prior_max, std::max(ax, ay));
        if (grid->radius_x[k] > 0 && grid->radius_y[k] > 0) {
            r0x = std::max(r0x, grid->radius_x[k]);
            r0y = std::max(r0y, grid->radius_y[k]);
        }
    }

    ar_status st;
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
        wanted_kernel<<<gblocks(N), GK, 0, s>>>(rcs, N, pend);
        level_points_kernel<<<gblocks(N), GK, 0, s>>>(grx, gry, rcs, N, csx, csy,
                                                      p.min_search_radius, lrx, lry, lcx, lcy, flags);
        if ((st = count_flags2(ctx, pend, flags, N, cnt, c2))) return st;
        if (c2[0] == 0) break;
        if (c2[1] == 0) continue;
        unsigned long long c = 0;
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
            coarse_points_kernel<<<gblocks(M), GK, 0, s>>>(gx, gy, gpdx, gpdy, lrx, lry, nr, nc, stride,
                                                           nrc, ncc, csx, csy, cxp, cyp, cpdx, cpdy,
                                                           crx, cry, ccx, ccy, cs8);
            // (no coarse point searchable: the pass measures nothing and the count below ends
            // the level)
            MCHK(hipGetLastError());
            ar_points cp = {cxp, cyp, crx, cry, cpdx, cpdy, ccx, ccy};
            if ((st = ar_track_pass(ctx, cp, M, 1, measure, cdx, cdy, ccr, cs8, hp))) return st;
            measured_kernel<<<gblocks(M), GK, 0, s>>>(cdx, M, cmeas);
            if ((st = count_flags(ctx, cmeas, M, cnt, &c))) return st;
            if (c == 0) continue;
            uint8_t* keep = nullptr;
            int64_t iters = p.gardner_iterations - 1 > 1 ? p.gardner_iterations - 1 : 1;
            if ((st = reject(ctx, cdx, cdy, crx, cry, cmeas, nrc, ncc, iters, cnx, cny, cka, ckb,
                             &keep)))
                return st;
            and_kernel<<<gblocks(M), GK, 0, s>>>(keep, cmeas, M);
            MCHK(hipMemsetAsync(cnt, 0, 16, s));
            fraction_kernel<<<gblocks(M), GK, 0, s>>>(crx, cmeas, keep, M, cnt);
            unsigned long long fr[2];
            MCHK(hipGetLastError());
            MCHK(hipMemcpyAsync(fr, cnt, 16, hipMemcpyDeviceToHost, s));
            MCHK(hipStreamSynchronize(s));
            if (fr[0] == 0) continue;
            if ((double)fr[1] / (double)fr[0] < p.min_coarse_valid_fraction) continue;
            double r2 = (double)p.coarse_buffer * (double)p.coarse_buffer;
            edt_col_kernel<<<gblocks(ncc), GK, 0, s>>>(keep, nrc, ncc, d2);
            edt_row_dilate_kernel<<<gblocks(nrc), GK, 0, s>>>(d2, nrc, ncc, r2, ev, ez, edd, grown);
            nearest_gate_kernel<<<gblocks(N), GK, 0, s>>>(grown, nrc, ncc, nr, nc, lrx, lry, flags);
            MCHK(hipGetLastError());
            if ((st = count_flags(ctx, flags, N, cnt, &c))) return st;
            if (c == 0) continue;
        }

        // ---- fine pass, rejection, hole filling, merge
        ar_points lp = {gx, gy, lrx, lry, gpdx, gpdy, lcx, lcy};
        if ((st = ar_track_pass(ctx, lp, N, p.upsampling, measure, fdx, fdy, fc, fs, hp)))
            return st;
        measured_kernel<<<gblocks(N), GK, 0, s>>>(fdx, N, meas);
        uint8_t* keep = nullptr;
        if ((st = reject(ctx, fdx, fdy, lrx, lry, meas, nr, nc, p.gardner_iterations, nxb, nyb, ka,
                         kb, &keep)))
            return st;
        reject_apply_kernel<<<gblocks(N), GK, 0, s>>>(keep, fdx, fdy, fc, N);
        MCHK(hipMemsetAsync(filled, 0, N, s));
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
    MCHK(hipMemcpyAsync(dx, rdx, 4 * N, hipMemcpyDeviceToHost, s));
    MCHK(hipMemcpyAsync(dy, rdy, 4 * N, hipMemcpyDeviceToHost, s));
    MCHK(hipMemcpyAsync(corr, rc, 4 * N, hipMemcpyDeviceToHost, s));
    MCHK(hipMemcpyAsync(chip_size, rcs, 2 * N, hipMemcpyDeviceToHost, s));
    MCHK(hipMemcpyAsync(interpolated, rit, N, hipMemcpyDeviceToHost, s));
    MCHK(hipStreamSynchronize(s));
    return AR_OK;
}
