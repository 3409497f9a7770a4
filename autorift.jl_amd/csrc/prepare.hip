// Image preparation on the device: the mask-aware high-pass pre-filter with mask erosion
// (src/preprocess.jl:275-312, 371-399, 447-460), the NaN-aware banded running-sum box mean it is
// built on (src/window.jl:289-439), and the UInt8 quantiser (src/preprocess.jl:546-592).
//
// The box mean reproduces the reference's summation order exactly: 16-row bands, a Float64
// running sum down each column of a band (one thread per band column), then a Float64 running
// sum along each row (one thread per row). Erosion is an exact boolean reduction. The quantiser's
// global mean and standard deviation are Float64 sums in a fixed (deterministic) parallel order
// rather than the reference's left-to-right order; they agree to ~1e-16 relative, which can move
// a pixel only if (v - lo) * scale lies within ~1e-12 of a rounding boundary.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/autorift_hip.h"
#include "ar_internal.h"
#include "seqsum.h"

namespace ar {

template <typename T>
__device__ inline double as_f64(const T* a, int64_t k) { return (double)a[k]; }
template <typename T>
__device__ inline float as_f32(const T* a, int64_t k) { return (float)a[k]; }

// Value the box mean sees at k. MASKED: the reference's `masked` Float32 copy (NaN where the mask
// is false); otherwise Float64 of the raw element.
template <typename T, bool MASKED>
__device__ inline double boxval(const T* img, const uint8_t* mask, int64_t k) {
    if (MASKED) return mask[k] ? (double)as_f32(img, k) : (double)NAN;
    return as_f64(img, k);
}

// ---- exact-sum certificate -------------------------------------------------------------------
// The reference's box mean is a chain of Float64 running sums (down each column of a band, then
// along each row). When every value that enters them is a multiple of 2^kmin and below 2^(emax+1)
// in magnitude, every partial sum of at most (wx + 1)(wy + 1) of them is a multiple of 2^kmin
// below 2^(lg + emax + 1), lg = ceil(log2((wx + 1)(wy + 1))); if that is at most 2^(kmin + 53),
// each partial sum is exactly representable, no addition rounds, and any summation order gives
// the reference's values. stats = {kmin, emax, non-finite seen} over the finite nonzero summed
// values; with no such value every sum is 0.
struct Cert {
    const int* stats;   // nullptr: no certificate (the sequential kernels always run)
    int wx, wy;
};

__device__ inline bool cert_holds(const Cert& c) {
    if (!c.stats) return false;
    if (c.stats[2]) return false;
    const int kmin = c.stats[0], emax = c.stats[1];
    if (kmin == INT_MAX) return true;
    const int64_t cnt = (int64_t)(c.wx + 1) * (int64_t)(c.wy + 1);
    int lg = 0;
    while (((int64_t)1 << lg) < cnt) lg++;
    return lg + emax + 1 <= kmin + 53;
}

__global__ void stats_init_kernel(int* stats) {
    if (threadIdx.x == 0) {
        stats[0] = INT_MAX;
        stats[1] = INT_MIN;
        stats[2] = 0;
        stats[3] = 0;   // some pixel's mask is false
    }
}

// kmin (exponent of the lowest set bit) and emax (floor(log2|v|)) of the values the masked box
// mean sums: Float32(img) where the mask is set; NaN is skipped by the sums, infinities fail the
// certificate.
template <typename T>
__global__ __launch_bounds__(256) void exact_stats_kernel(const T* img, const uint8_t* mask,
                                                          int64_t n, int* stats) {
    int kmin = INT_MAX, emax = INT_MIN, bad = 0, gap = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x) {
        if (!mask[k]) { gap = 1; continue; }
        const double v = (double)as_f32(img, k);
        if (isnan(v) || v == 0.0) continue;
        if (isinf(v)) { bad = 1; continue; }
        const unsigned long long b = (unsigned long long)__double_as_longlong(v);
        const int E = (int)((b >> 52) & 0x7ff);
        const unsigned long long S = (b & 0xfffffffffffffull) | (1ull << 52);
        const int e = E - 1023;
        const int lo = e - 52 + __builtin_ctzll(S);
        kmin = lo < kmin ? lo : kmin;
        emax = e > emax ? e : emax;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int a = __shfl_xor(kmin, o), b2 = __shfl_xor(emax, o), c = __shfl_xor(bad, o);
        const int g = __shfl_xor(gap, o);
        kmin = a < kmin ? a : kmin;
        emax = b2 > emax ? b2 : emax;
        bad |= c;
        gap |= g;
    }
    // one set of atomics per workgroup
    __shared__ int red[4][4];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wv] = kmin;
        red[1][wv] = emax;
        red[2][wv] = bad;
        red[3][wv] = gap;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
            kmin = red[0][q] < kmin ? red[0][q] : kmin;
            emax = red[1][q] > emax ? red[1][q] : emax;
            bad |= red[2][q];
            gap |= red[3][q];
        }
        if (kmin != INT_MAX) atomicMin(stats + 0, kmin);
        if (emax != INT_MIN) atomicMax(stats + 1, emax);
        if (bad) atomicOr(stats + 2, 1);
        if (gap) atomicOr(stats + 3, 1);
    }
}

// Certified box mean, column sums: thread per pixel, the NaN-aware sum and count of the valid
// Float32 values in its column window (the masked form of src/window.jl:299-325; under the
// certificate equal to the banded running sums).
template <typename T>
__global__ __launch_bounds__(256) void box_col_direct_kernel(const T* img, const uint8_t* mask,
                                                             int64_t nr, int64_t nc, int ly,
                                                             int ry, Cert cert, double* colsum,
                                                             int32_t* colcnt) {
    if (!cert_holds(cert)) return;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr * nc) return;
    const int64_t j = k / nr, i = k - j * nr;
    const int64_t a0 = i - ly > 0 ? i - ly : 0, a1 = i + ry < nr - 1 ? i + ry : nr - 1;
    double s = 0.0;
    int32_t c = 0;
    for (int64_t ii = a0; ii <= a1; ii++) {
        const double v = boxval<T, true>(img, mask, ii + j * nr);
        if (!isnan(v)) { s += v; c += 1; }
    }
    colsum[k] = s;
    colcnt[k] = c;
}

// Certified box mean, row sums fused with highpass (src/preprocess.jl:304-312): thread per pixel.
template <typename T>
__global__ __launch_bounds__(256) void box_row_direct_highpass_kernel(
    const T* img, const uint8_t* mask, int64_t nr, int64_t nc, int lx, int rx, Cert cert,
    const double* colsum, const int32_t* colcnt, float* out) {
    if (!cert_holds(cert)) return;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr * nc) return;
    const int64_t j = k / nr, i = k - j * nr;
    const int64_t b0 = j - lx > 0 ? j - lx : 0, b1 = j + rx < nc - 1 ? j + rx : nc - 1;
    double s = 0.0;
    int32_t c = 0;
    for (int64_t jj = b0; jj <= b1; jj++) {
        s += colsum[i + jj * nr];
        c += colcnt[i + jj * nr];
    }
    const float mean = c > 0 ? (float)(s / (double)c) : NAN;
    out[k] = (mask[k] && isfinite(mean)) ? as_f32(img, k) - mean : NAN;
}

// Column pass of one 16-row band (src/window.jl:299-325 masked, 397-409 dense).
// Thread per (band, column). colsum/colcnt hold, for every row i, the band's running state at i.
template <typename T, bool MASKED, bool NANAWARE>
__global__ void box_colpass_kernel(const T* img, const uint8_t* mask, int64_t nr, int64_t nc,
                                   int ly, int ry, double* colsum, int32_t* colcnt,
                                   Cert cert = Cert{nullptr, 0, 0}) {
    if (cert_holds(cert)) return;   // the certified direct kernels ran instead
    int64_t nb = nr < 16 ? nr : 16;
    int64_t nbands = (nr + nb - 1) / nb;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbands * nc) return;
    int64_t j = t / nbands, b = t - j * nbands;   // adjacent threads: adjacent bands, same column
    int64_t i0 = 1 + b * nb;
    int64_t i1 = i0 + nb - 1 < nr ? i0 + nb - 1 : nr;
    const int64_t cj = (j)*nr;
    double s = 0.0;
    int32_t c = 0;
    int64_t a0 = i0 - ly > 1 ? i0 - ly : 1, a1 = i0 + ry < nr ? i0 + ry : nr;
    for (int64_t i = a0; i <= a1; i++) {
        double v = boxval<T, MASKED>(img, mask, (i - 1) + cj);
        if (NANAWARE) {
            if (!isnan(v)) { s += v; c += 1; }
        } else {
            s += v;
        }
    }
    for (int64_t i = i0; i <= i1; i++) {
        colsum[(i - 1) + cj] = s;
        if (NANAWARE) colcnt[(i - 1) + cj] = c;
        int64_t oi = i - ly;
        if (1 <= oi && oi <= nr) {
            double v = boxval<T, MASKED>(img, mask, (oi - 1) + cj);
            if (NANAWARE) {
                if (!isnan(v)) { s -= v; c -= 1; }
            } else {
                s -= v;
            }
        }
        int64_t ii = i + ry + 1;
        if (ii <= nr) {
            double v = boxval<T, MASKED>(img, mask, (ii - 1) + cj);
            if (NANAWARE) {
                if (!isnan(v)) { s += v; c += 1; }
            } else {
                s += v;
            }
        }
    }
}

// Row pass (src/window.jl:327-361 masked, 411-435 dense) fused with highpass
// (src/preprocess.jl:304-312): out = mask & isfinite(mean) ? Float32(img) - mean : NaN.
// Thread per row; adjacent threads read adjacent rows of one column (column-major), coalesced.
// The running sum is sequential along the row, as in the reference; the loads it consumes do not
// depend on it, so they are fetched RU columns ahead.
#define RU 8
template <typename T, bool NANAWARE>
__global__ void box_rowpass_highpass_kernel(const T* img, const uint8_t* mask, int64_t nr,
                                            int64_t nc, int lx, int rx, int ly, int ry,
                                            const double* colsum, const int32_t* colcnt,
                                            float* out, Cert cert = Cert{nullptr, 0, 0}) {
    if (cert_holds(cert)) return;   // the certified direct kernels ran instead
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // 0-based row
    if (i >= nr) return;
    int64_t I = i + 1;
    int64_t ci = (I + ry < nr ? I + ry : nr) - (I - ly > 1 ? I - ly : 1) + 1;
    double s = 0.0;
    int32_t c = 0;
    int64_t b1 = rx + 1 < nc ? rx + 1 : nc;
    for (int64_t j = 1; j <= b1; j++) {
        s += colsum[i + (j - 1) * nr];
        if (NANAWARE) c += colcnt[i + (j - 1) * nr];
    }
    for (int64_t j0 = 1; j0 <= nc; j0 += RU) {
        double xo[RU], xi[RU];
        int32_t co[RU], cn[RU];
        uint8_t mk[RU];
        float iv[RU];
#pragma unroll
        for (int u = 0; u < RU; u++) {
            int64_t j = j0 + u;
            int64_t oj = j - lx, jj = j + rx + 1;
            bool in = j <= nc;
            bool vo = in && 1 <= oj && oj <= nc, vi = in && jj <= nc;
            xo[u] = vo ? colsum[i + (oj - 1) * nr] : 0.0;
            xi[u] = vi ? colsum[i + (jj - 1) * nr] : 0.0;
            if (NANAWARE) {
                co[u] = vo ? colcnt[i + (oj - 1) * nr] : 0;
                cn[u] = vi ? colcnt[i + (jj - 1) * nr] : 0;
            }
            int64_t k = i + (in ? j - 1 : 0) * nr;
            mk[u] = mask[k];
            iv[u] = as_f32(img, k);
        }
#pragma unroll
        for (int u = 0; u < RU; u++) {
            int64_t j = j0 + u;
            if (j > nc) break;
            float mean;
            if (NANAWARE) {
                mean = c > 0 ? (float)(s / (double)c) : NAN;
            } else {
                int64_t cj = (j + rx < nc ? j + rx : nc) - (j - lx > 1 ? j - lx : 1) + 1;
                mean = (float)(s / (double)(ci * cj));
            }
            out[i + (j - 1) * nr] = (mk[u] && isfinite(mean)) ? iv[u] - mean : NAN;
            int64_t oj = j - lx, jj = j + rx + 1;
            if (1 <= oj && oj <= nc) {
                s -= xo[u];
                if (NANAWARE) c -= co[u];
            }
            if (jj <= nc) {
                s += xi[u];
                if (NANAWARE) c += cn[u];
            }
        }
    }
}

// Row pass producing Wallis's two intermediate planes (src/preprocess.jl:333-351, 416-438): the
// masked box mean (as _masked_boxmean returns it) and the Float32 squared deviation about it,
// dev2 = mask & isfinite(mean) ? (Float32(img) - mean)^2 : NaN. Same summation as the highpass
// row pass.
template <typename T, bool NANAWARE>
__global__ void box_rowpass_wallis_dev_kernel(const T* img, const uint8_t* mask, int64_t nr,
                                              int64_t nc, int lx, int rx, int ly, int ry,
                                              const double* colsum, const int32_t* colcnt,
                                              float* mean_out, float* dev2) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nr) return;
    int64_t I = i + 1;
    int64_t ci = (I + ry < nr ? I + ry : nr) - (I - ly > 1 ? I - ly : 1) + 1;
    double s = 0.0;
    int32_t c = 0;
    int64_t b1 = rx + 1 < nc ? rx + 1 : nc;
    for (int64_t j = 1; j <= b1; j++) {
        s += colsum[i + (j - 1) * nr];
        if (NANAWARE) c += colcnt[i + (j - 1) * nr];
    }
    for (int64_t j = 1; j <= nc; j++) {
        const int64_t k = i + (j - 1) * nr;
        float mean;
        if (NANAWARE) {
            mean = c > 0 ? (float)(s / (double)c) : NAN;
        } else {
            int64_t cj = (j + rx < nc ? j + rx : nc) - (j - lx > 1 ? j - lx : 1) + 1;
            mean = (float)(s / (double)(ci * cj));
        }
        mean_out[k] = mean;
        if (mask[k] && isfinite(mean)) {
            const float d = as_f32(img, k) - mean;
            dev2[k] = d * d;
        } else {
            dev2[k] = NAN;
        }
        int64_t oj = j - lx, jj = j + rx + 1;
        if (1 <= oj && oj <= nc) {
            s -= colsum[i + (oj - 1) * nr];
            if (NANAWARE) c -= colcnt[i + (oj - 1) * nr];
        }
        if (jj <= nc) {
            s += colsum[i + (jj - 1) * nr];
            if (NANAWARE) c += colcnt[i + (jj - 1) * nr];
        }
    }
}

// Final Wallis row pass: the NaN-aware box mean of dev2 (windowmean(dev2; hasnan = gaps): with no
// NaN in dev2 its dense path adds the same values in the same order and divides by the same
// count, so the NaN-aware form gives the same result either way), sd = sqrt(max(msd, 0)) with
// NaN kept, then out = (Float32(img) - mean) / max(sd, floor_sd), NaN where the mask is false,
// mean or sd is not finite, or the divisor is 0. The Float32 square root and quotient are formed
// in Float64 and rounded once, which gives the correctly rounded Float32 result.
template <typename T>
__global__ void box_rowpass_wallis_out_kernel(const T* img, const uint8_t* mask, int64_t nr,
                                              int64_t nc, int lx, int rx, const double* colsum,
                                              const int32_t* colcnt, const float* mean_in,
                                              float floor_sd, float* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nr) return;
    double s = 0.0;
    int32_t c = 0;
    int64_t b1 = rx + 1 < nc ? rx + 1 : nc;
    for (int64_t j = 1; j <= b1; j++) {
        s += colsum[i + (j - 1) * nr];
        c += colcnt[i + (j - 1) * nr];
    }
    for (int64_t j = 1; j <= nc; j++) {
        const int64_t k = i + (j - 1) * nr;
        const float msd = c > 0 ? (float)(s / (double)c) : NAN;
        const float sd = isnan(msd) ? NAN : (float)sqrt((double)(msd < 0.0f ? 0.0f : msd));
        const float mean = mean_in[k];
        float o = NAN;
        if (mask[k] && isfinite(mean) && isfinite(sd)) {
            const float sdiv = sd < floor_sd ? floor_sd : sd;
            if (sdiv > 0.0f) o = (float)((double)(as_f32(img, k) - mean) / (double)sdiv);
        }
        out[k] = o;
        int64_t oj = j - lx, jj = j + rx + 1;
        if (1 <= oj && oj <= nc) {
            s -= colsum[i + (oj - 1) * nr];
            c -= colcnt[i + (oj - 1) * nr];
        }
        if (jj <= nc) {
            s += colsum[i + (jj - 1) * nr];
            c += colcnt[i + (jj - 1) * nr];
        }
    }
}

// _erode_mask (src/preprocess.jl:447-460) and _filtered (275-288): a pixel stays valid only if
// every in-bounds pixel of its (clamped) w-window is valid; non-finite filtered values become 0
// and invalid.
__global__ void erode_finalize_kernel(float* out, const uint8_t* mask, uint8_t* vmask, int64_t nr,
                                      int64_t nc, int lx, int rx, int ly, int ry, int allmask,
                                      const int* stats = nullptr) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr * nc) return;
    if (stats) allmask = stats[3] == 0;   // every mask bit set (the device's own statistics)
    int64_t j = k / nr, i = k - j * nr;
    uint8_t v = mask[k];
    if (!allmask) {
        v = 1;
        int64_t j0 = j - lx > 0 ? j - lx : 0, j1 = j + rx < nc - 1 ? j + rx : nc - 1;
        int64_t i0 = i - ly > 0 ? i - ly : 0, i1 = i + ry < nr - 1 ? i + ry : nr - 1;
        for (int64_t jj = j0; jj <= j1 && v; jj++)
            for (int64_t ii = i0; ii <= i1; ii++)
                if (!mask[ii + jj * nr]) { v = 0; break; }
    }
    float o = out[k];
    if (!isfinite(o)) {
        out[k] = 0.0f;
        v = 0;
    }
    vmask[k] = v;
}

// Mask statistics: every pixel valid? any NaN in the image?
template <typename T>
__global__ void mask_stats_kernel(const T* img, const uint8_t* mask, int64_t n, int* flags) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int bad = 0, nan = 0;
    for (; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        bad |= !mask[k];
        nan |= isnan(as_f64(img, k)) ? 1 : 0;
    }
    if (__syncthreads_or(bad) && threadIdx.x == 0) atomicAnd(&flags[0], 0);
    if (__syncthreads_or(nan) && threadIdx.x == 0) atomicOr(&flags[1], 1);
}

template <typename T>
__global__ void finite_mask_t_kernel(const T* img, uint8_t* out, int64_t n) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; k < n; k += (int64_t)gridDim.x * blockDim.x) out[k] = isfinite(as_f64(img, k)) ? 1 : 0;
}

// The quantiser's parameters, formed on the device (no host round trip): count, mean, Bessel
// standard deviation, lo = mean - 3 sd, scale = 255 / (6 sd), and the mode: 0 nothing valid, 1
// sd == 0, 2 map (src/preprocess.jl:546-592).
struct QParams {
    double mean, lo, scale, sd;
    long long n;
    int mode;
    long long walk_items[2], walk_in_order[2];   // per sum: items walked, elements summed in order
};

// The summands of the quantiser's two sequential sums (src/preprocess.jl:551-570): pass 0 the
// valid finite values, pass 1 their squared deviations from the mean; 0 elsewhere (adding +0.0
// leaves a running sum unchanged).
template <typename T>
struct QTerm {
    const T* img;
    const uint8_t* mask;
    const QParams* qp;
    int pass;
    __device__ inline double operator()(int64_t k, bool& ok) const {
        const double v = as_f64(img, k);
        ok = mask[k] && isfinite(v);
        if (!ok) return 0.0;
        if (pass == 0) return v;
        const double d = v - qp->mean;
        return d * d;
    }
};

// The elements of a plain array, every one counted (ar_sequential_sum).
struct ArrayTerm {
    const double* x;
    __device__ inline double operator()(int64_t k, bool& ok) const {
        ok = true;
        return x[k];
    }
};

struct SumOut {
    double* out;
    long long* stats;
    __device__ inline void operator()(double sum, long long, const SqWalkStats& st) const {
        *out = sum;
        stats[0] = st.items;
        stats[1] = st.in_order;
    }
};

// x[0] + x[1] + ... left to right by one wave's whole-wave steps (non-negative finite terms).
__global__ __launch_bounds__(64) void wave_sum_kernel(const double* x, int64_t n, double* out) {
    const int lane = threadIdx.x;
    double s = 0.0;
    for (int64_t k0 = 0; k0 < n; k0 += 64)
        s = seq_sum64_ties(s, k0 + lane < n ? x[k0 + lane] : 0.0, lane);
    if (lane == 0) *out = s;
}

// After the walk of pass 0: n and mean = s / n; after pass 1: sd = sqrt(ss / (n - 1)) and the map.
struct QFinish {
    QParams* qp;
    int pass;
    __device__ inline void operator()(double sum, long long count, const SqWalkStats& st) const {
        qp->walk_items[pass] = st.items;
        qp->walk_in_order[pass] = st.in_order;
        if (pass == 0) {
            qp->n = count;
            qp->mean = sum / (double)count;
            return;
        }
        const long long cnt = qp->n;
        const double sd = cnt > 1 ? sqrt(sum / (double)(cnt - 1)) : 0.0;
        qp->sd = sd;
        qp->mode = cnt == 0 ? 0 : (sd == 0 ? 1 : 2);
        qp->lo = qp->mean - 3 * sd;
        qp->scale = 255.0 / (6 * sd);
    }
};

// round(UInt8, clamp((v - lo) * scale, 0, 255)), ties to even; invalid -> 0; flat -> 0x80.
template <typename T>
__global__ void q_map_kernel(const T* img, const uint8_t* mask, int64_t n, const QParams* qp,
                             uint8_t* out) {
    const int mode = qp->mode;
    const double lo = qp->lo, scale = qp->scale;
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        if (mode == 0) { out[k] = 0; continue; }                       // nothing valid
        if (mode == 1) { out[k] = mask[k] ? 0x80 : 0x00; continue; }   // sd == 0
        double v = as_f64(img, k);
        if (!mask[k] || !isfinite(v)) { out[k] = 0; continue; }
        double t = (v - lo) * scale;
        t = t < 0.0 ? 0.0 : (t > 255.0 ? 255.0 : t);
        out[k] = (uint8_t)rint(t);
    }
}

// quantize(::NoQuantize) (src/preprocess.jl:492-507): non-finite floats -> 0; integers unchanged.
template <typename T>
__global__ void noquant_kernel(const T* img, int64_t n, float* out) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        float v = as_f32(img, k);
        out[k] = isfinite(v) ? v : 0.0f;
    }
}

__global__ void copy_u8_kernel(const uint8_t* a, uint8_t* b, int64_t n) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; k < n; k += (int64_t)gridDim.x * blockDim.x) b[k] = a[k];
}

}  // namespace ar

using namespace ar;

static ar_status pfail(ar_context* ctx, ar_status st, const std::string& m) {
    ctx->err = m;
    return st;
}

#define PCHK(expr)                                                                            \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            return pfail(ctx, _e == hipErrorOutOfMemory ? AR_ERR_OUT_OF_MEMORY : AR_ERR_DEVICE, \
                         std::string(#expr) + ": " + hipGetErrorString(_e));                  \
    } while (0)

static int gsz(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : (b > 65535 ? 65535 : b));
}

// The stream and scratch one image's preparation runs on: the context's own, or, for the
// secondary of a pair prepared side by side, the second stream with its own scratch.
struct PrepLane {
    hipStream_t s;
    DevBuf *raw, *rawmask, *work_f32, *work_f64, *flags, *qsum;
};

static PrepLane prep_lane(ar_context* ctx, bool side) {
    if (!side)
        return {ctx->stream, &ctx->raw, &ctx->rawmask, &ctx->work_f32, &ctx->work_f64,
                &ctx->flags, &ctx->qsum};
    return {ctx->stream2, &ctx->raw2, &ctx->rawmask2, &ctx->work_f32_2, &ctx->work_f64_2,
            &ctx->flags2, &ctx->qsum2};
}

// Quantise a float (or raw) image on the device into out (u8); the parameters stay in the
// context's slot `slot` (0 reference, 1 secondary) for inspection.
template <typename T>
static ar_status quantize_u8(ar_context* ctx, const T* img, const uint8_t* mask, int64_t N,
                             uint8_t* out, int slot, const PrepLane& L) {
    ar_status st;
    const int64_t nch = (N + SQ_CHUNK - 1) / SQ_CHUNK;
    if (nch >= ((int64_t)1 << 31))
        return pfail(ctx, AR_ERR_UNSUPPORTED, "image too large for the UInt8 quantiser");
    const size_t recb = (size_t)nch * SQ_SLOTS * sizeof(SqRec);
    const size_t per = SQ_WG * 8 + 8 + 8 + 8 + 4 + 4 + 4;   // tsum, csum, coff, ccnt, clow, cg, cnt
    if ((st = L.qsum->reserve(ctx, 2 * recb + (size_t)nch * per + 256))) return st;
    if ((st = ctx->qstats.reserve(ctx, 2 * sizeof(QParams)))) return st;
    char* b = (char*)L.qsum->ptr;
    SqRec* rec = (SqRec*)b;
    SqRec* items = (SqRec*)(b + recb);
    double* tsum = (double*)(b + 2 * recb);
    double* csum = tsum + (size_t)nch * SQ_WG;
    double* coff = csum + nch;
    long long* ccnt = (long long*)(coff + nch);
    int* clow = (int*)(ccnt + nch);
    int* cg = clow + nch;
    int* cnt = cg + nch;
    QParams* qp = (QParams*)ctx->qstats.ptr + slot;
    const bool inline_scan = nch <= SQ_SCAN_INLINE;
    for (int pass = 0; pass < 2; pass++) {
        const QTerm<T> term{img, mask, qp, pass};
        sq_chunk_kernel<<<(unsigned)nch, SQ_WG, 0, L.s>>>(term, N, tsum, csum, clow, ccnt);
        if (!inline_scan)
            sq_scan_kernel<<<1, 1024, 0, L.s>>>(csum, clow, ccnt, nch, coff, cg);
        sq_segment_kernel<<<(unsigned)nch, SQ_WG, 0, L.s>>>(
            term, N, tsum, csum, clow, inline_scan ? nullptr : coff, cg, cnt, rec);
        sq_walk_kernel<<<1, SQ_WG, 0, L.s>>>(term, N, nch, ccnt, cnt, rec, items,
                                                     QFinish{qp, pass});
    }
    q_map_kernel<T><<<gsz(N), 256, 0, L.s>>>(img, mask, N, qp, out);
    PCHK(hipGetLastError());
    return AR_OK;
}

// preprocess + quantize of one raw image already on the device (raw, mask) into the prepared
// slot (dst image, dst mask).
template <typename T>
static ar_status prepare_one(ar_context* ctx, const T* raw, const uint8_t* mask, int64_t ny,
                             int64_t nx, void* dst, uint8_t* dmask, int slot, const PrepLane& L) {
    const ar_params& p = ctx->params;
    int64_t N = ny * nx;
    ar_status st;
    const uint8_t* curmask = mask;
    float* hp = nullptr;
    if (p.preprocess == AR_PRE_HIGHPASS || p.preprocess == AR_PRE_WALLIS) {
        const bool wallis = p.preprocess == AR_PRE_WALLIS;
        int w = p.filter_width;
        if (w < 1) return pfail(ctx, AR_ERR_INVALID_ARGUMENT, "filter width must be >= 1");
        if (wallis && !(p.min_std >= 0.0))
            return pfail(ctx, AR_ERR_INVALID_ARGUMENT, "Wallis `min_std` must be >= 0");
        if ((st = L.flags->reserve(ctx, 64))) return st;
        // Highpass of a non-Float64 image needs no host decision: Float32(x) == x for every such
        // element, so the masked NaN-aware box mean equals the reference's dense and unmasked
        // forms wherever those apply, and the erosion of an all-valid mask is the mask itself.
        // Float64 images (the masked copy rounds them) and Wallis read the mask statistics back.
        const bool certified = !wallis && !std::is_same<T, double>::value;
        bool allmask = false, anynan = true;
        if (!certified) {
            int f0[2] = {1, 0};
            PCHK(hipMemcpyAsync(L.flags->ptr, f0, 8, hipMemcpyHostToDevice, L.s));
            mask_stats_kernel<T><<<gsz(N) > 4096 ? 4096 : gsz(N), 256, 0, L.s>>>(
                raw, mask, N, (int*)L.flags->ptr);
            int f[2];
            PCHK(hipGetLastError());
            PCHK(hipMemcpyAsync(f, L.flags->ptr, 8, hipMemcpyDeviceToHost, L.s));
            PCHK(hipStreamSynchronize(L.s));
            allmask = f[0] != 0;
            anynan = f[1] != 0;
        }
        int wx = (int)(w < nx ? w : nx), wy = (int)(w < ny ? w : ny);
        int lx = wx / 2, ly = wy / 2, rx = wx - 1 - lx, ry = wy - 1 - ly;
        if ((st = L.work_f64->reserve(ctx, N * 8 + N * 4))) return st;
        if ((st = L.work_f32->reserve(ctx, N * 4 * (wallis ? 3 : 1)))) return st;
        double* colsum = (double*)L.work_f64->ptr;
        int32_t* colcnt = (int32_t*)(colsum + N);
        hp = (float*)L.work_f32->ptr;
        int64_t nb = ny < 16 ? ny : 16;
        int64_t nthr = ((ny + nb - 1) / nb) * nx;
        int gb = (int)((nthr + 255) / 256);
        int gr = (int)((ny + 63) / 64);
        if (wallis) {
            // mean and squared deviations, then the box mean of the deviations and the output
            float* mean = hp + N;
            float* dev2 = hp + 2 * N;
            if (allmask && !anynan) {
                box_colpass_kernel<T, false, false><<<gb, 256, 0, L.s>>>(raw, mask, ny, nx, ly, ry, colsum, colcnt);
                box_rowpass_wallis_dev_kernel<T, false><<<gr, 64, 0, L.s>>>(raw, mask, ny, nx, lx, rx, ly, ry, colsum, colcnt, mean, dev2);
            } else if (allmask) {
                box_colpass_kernel<T, false, true><<<gb, 256, 0, L.s>>>(raw, mask, ny, nx, ly, ry, colsum, colcnt);
                box_rowpass_wallis_dev_kernel<T, true><<<gr, 64, 0, L.s>>>(raw, mask, ny, nx, lx, rx, ly, ry, colsum, colcnt, mean, dev2);
            } else {
                box_colpass_kernel<T, true, true><<<gb, 256, 0, L.s>>>(raw, mask, ny, nx, ly, ry, colsum, colcnt);
                box_rowpass_wallis_dev_kernel<T, true><<<gr, 64, 0, L.s>>>(raw, mask, ny, nx, lx, rx, ly, ry, colsum, colcnt, mean, dev2);
            }
            box_colpass_kernel<float, false, true><<<gb, 256, 0, L.s>>>(dev2, mask, ny, nx, ly, ry, colsum, colcnt);
            box_rowpass_wallis_out_kernel<T><<<gr, 64, 0, L.s>>>(raw, mask, ny, nx, lx, rx, colsum, colcnt, mean, (float)p.min_std, hp);
        } else if (certified) {
            // exact-sum certificate on the device; the direct kernels run when it holds, the
            // reference's sequential running sums otherwise (each kernel checks it and exits)
            int* stats = (int*)L.flags->ptr;
            stats_init_kernel<<<1, 64, 0, L.s>>>(stats);
            // about 16 values per thread, at most 1024 workgroups
            const int64_t sb = (N + 4095) / 4096;
            exact_stats_kernel<T><<<(unsigned)(sb < 1 ? 1 : (sb > 1024 ? 1024 : sb)), 256, 0,
                                    L.s>>>(raw, mask, N, stats);
            const Cert cert{stats, wx, wy};
            const unsigned gp = (unsigned)((N + 255) / 256);
            box_col_direct_kernel<T><<<gp, 256, 0, L.s>>>(raw, mask, ny, nx, ly, ry, cert, colsum, colcnt);
            box_colpass_kernel<T, true, true><<<gb, 256, 0, L.s>>>(raw, mask, ny, nx, ly, ry, colsum, colcnt, cert);
            box_row_direct_highpass_kernel<T><<<gp, 256, 0, L.s>>>(raw, mask, ny, nx, lx, rx, cert, colsum, colcnt, hp);
            box_rowpass_highpass_kernel<T, true><<<gr, 64, 0, L.s>>>(raw, mask, ny, nx, lx, rx, ly, ry, colsum, colcnt, hp, cert);
        } else if (allmask && !anynan) {
            box_colpass_kernel<T, false, false><<<gb, 256, 0, L.s>>>(raw, mask, ny, nx, ly, ry, colsum, colcnt);
            box_rowpass_highpass_kernel<T, false><<<gr, 64, 0, L.s>>>(raw, mask, ny, nx, lx, rx, ly, ry, colsum, colcnt, hp);
        } else if (allmask) {
            box_colpass_kernel<T, false, true><<<gb, 256, 0, L.s>>>(raw, mask, ny, nx, ly, ry, colsum, colcnt);
            box_rowpass_highpass_kernel<T, true><<<gr, 64, 0, L.s>>>(raw, mask, ny, nx, lx, rx, ly, ry, colsum, colcnt, hp);
        } else {
            box_colpass_kernel<T, true, true><<<gb, 256, 0, L.s>>>(raw, mask, ny, nx, ly, ry, colsum, colcnt);
            box_rowpass_highpass_kernel<T, true><<<gr, 64, 0, L.s>>>(raw, mask, ny, nx, lx, rx, ly, ry, colsum, colcnt, hp);
        }
        PCHK(hipGetLastError());
        // erosion uses the full filter width w (clamped to the image) like windowmin(f, w)
        erode_finalize_kernel<<<(unsigned)((N + 255) / 256), 256, 0, L.s>>>(
            hp, mask, dmask, ny, nx, lx, rx, ly, ry, allmask ? 1 : 0,
            certified ? (const int*)L.flags->ptr : nullptr);
        PCHK(hipGetLastError());
        curmask = dmask;
    } else {
        copy_u8_kernel<<<gsz(N), 256, 0, L.s>>>(mask, dmask, N);
        PCHK(hipGetLastError());
        curmask = dmask;
    }
    if (p.quantize == AR_QUANT_UINT8) {
        if (hp) return quantize_u8<float>(ctx, hp, curmask, N, (uint8_t*)dst, slot, L);
        return quantize_u8<T>(ctx, raw, curmask, N, (uint8_t*)dst, slot, L);
    }
    // NoQuantize: Float32 output
    if (hp) {
        PCHK(hipMemcpyAsync(dst, hp, N * 4, hipMemcpyDeviceToDevice, L.s));
    } else {
        noquant_kernel<T><<<gsz(N), 256, 0, L.s>>>(raw, N, (float*)dst);
        PCHK(hipGetLastError());
    }
    return AR_OK;
}

static size_t dt_size(int32_t dt) {
    switch (dt) {
        case AR_U8: return 1;
        case AR_I16: return 2;
        case AR_F32: return 4;
        case AR_F64: return 8;
        case AR_C64: return 8;
    }
    return 0;
}

template <typename T>
static ar_status prepare_slot(ar_context* ctx, const void* host_img, const uint8_t* host_mask,
                              int64_t ny, int64_t nx, void* dst, uint8_t* dmask, int slot,
                              const PrepLane& L) {
    int64_t N = ny * nx;
    ar_status st;
    if ((st = L.raw->reserve(ctx, N * sizeof(T))) || (st = L.rawmask->reserve(ctx, N)))
        return st;
    // the caller's image may be host or device memory (unified addressing picks the direction)
    PCHK(hipMemcpyAsync(L.raw->ptr, host_img, N * sizeof(T), hipMemcpyDefault, L.s));
    if (host_mask) {
        PCHK(hipMemcpyAsync(L.rawmask->ptr, host_mask, N, hipMemcpyDefault, L.s));
    } else {
        finite_mask_t_kernel<T><<<gsz(N), 256, 0, L.s>>>((const T*)L.raw->ptr,
                                                                 (uint8_t*)L.rawmask->ptr, N);
        PCHK(hipGetLastError());
    }
    return prepare_one<T>(ctx, (const T*)L.raw->ptr, (const uint8_t*)L.rawmask->ptr, ny, nx,
                          dst, dmask, slot, L);
}

extern "C" ar_status ar_set_images(ar_context* ctx, const void* ref, const void* sec,
                                   const uint8_t* ref_valid, const uint8_t* sec_valid, int64_t ny,
                                   int64_t nx, int32_t dtype) {
    if (!ctx) return AR_ERR_INVALID_ARGUMENT;
    if (!ref && !sec) return AR_OK;
    if (ny <= 0 || nx <= 0) return pfail(ctx, AR_ERR_INVALID_ARGUMENT, "image size must be positive");
    if (dt_size(dtype) == 0) return pfail(ctx, AR_ERR_UNSUPPORTED, "unsupported element type");
    if ((!ref || !sec) && (!ctx->has_pair || ctx->ny != ny || ctx->nx != nx))
        return pfail(ctx, AR_ERR_DIMENSION_MISMATCH,
                     "new images must match the size the context was built for; build a new "
                     "context (init) for a different image size");
    const ar_params& p = ctx->params;
    if (p.preprocess != AR_PRE_NONE && p.preprocess != AR_PRE_HIGHPASS &&
        p.preprocess != AR_PRE_WALLIS && p.preprocess != AR_PRE_DERAMP)
        return pfail(ctx, AR_ERR_UNSUPPORTED, "pre-filter not available on the device path");
    // complex input takes Deramp or nothing and NoQuantize; real input takes no Deramp
    // (src/preprocess.jl:247-260, 473-531)
    if (dtype == AR_C64 && (p.preprocess == AR_PRE_HIGHPASS || p.preprocess == AR_PRE_WALLIS))
        return pfail(ctx, AR_ERR_INVALID_ARGUMENT,
                     std::string(p.preprocess == AR_PRE_WALLIS ? "Wallis" : "Highpass") +
                     " is an amplitude filter and rejects complex input: it would "
                     "produce something plausible-looking and meaningless. Use Deramp or no "
                     "pre-filter for SLC data");
    if (dtype == AR_C64 && p.quantize != AR_QUANT_NONE)
        return pfail(ctx, AR_ERR_INVALID_ARGUMENT,
                     "UInt8 quantization rejects complex input; use quantize = none for SLC data");
    if (dtype != AR_C64 && p.preprocess == AR_PRE_DERAMP)
        return pfail(ctx, AR_ERR_INVALID_ARGUMENT,
                     "Deramp needs complex input: a real image has no phase ramp to remove");
    if (dtype == AR_C64 && (p.deramp_axis < 0 || p.deramp_axis > 2))
        return pfail(ctx, AR_ERR_INVALID_ARGUMENT, "Deramp axis must be both, x or y");
    PCHK(hipSetDevice(ctx->device));
    int32_t pdt = p.quantize == AR_QUANT_UINT8 ? AR_U8 : AR_F32;
    if (p.preprocess == AR_PRE_NONE && p.quantize == AR_QUANT_NONE && dtype == AR_U8) pdt = AR_U8;
    if (dtype == AR_C64) pdt = AR_C64;
    if (ctx->has_pair && (!ref || !sec) && pdt != ctx->pdtype)
        return pfail(ctx, AR_ERR_INVALID_ARGUMENT, "prepared element type changed");
    int64_t N = ny * nx;
    size_t es = pdt == AR_U8 ? 1 : (pdt == AR_C64 ? 8 : 4);
    ar_status st;
    if (!ctx->has_pair || ctx->ny != ny || ctx->nx != nx || ctx->pdtype != pdt) {
        if ((st = ctx->pref.reserve(ctx, N * es)) || (st = ctx->psec.reserve(ctx, N * es)) ||
            (st = ctx->rmask.reserve(ctx, N)) || (st = ctx->smask.reserve(ctx, N)))
            return st;
    }
    // Both images of a pair of at most kSidePrepMax elements are prepared side by side: the
    // secondary on the second stream, which first waits for everything already queued on the
    // context's stream (a pass may still read the prepared images), and which the context's
    // stream then waits for.
    const bool side = ctx->side_prep && ref && sec && N <= kSidePrepMax && dtype != AR_C64 &&
                      !(pdt == AR_U8 && p.preprocess == AR_PRE_NONE &&
                        p.quantize == AR_QUANT_NONE);
    if (side) {
        if ((st = ar_stream2(ctx))) return st;
        PCHK(hipEventRecord(ctx->evf[0], ctx->stream));
        PCHK(hipStreamWaitEvent(ctx->stream2, ctx->evf[0], 0));
    }
    for (int slot = 0; slot < 2; slot++) {
        const void* img = slot == 0 ? ref : sec;
        const uint8_t* m = slot == 0 ? ref_valid : sec_valid;
        if (!img) continue;
        void* dst = slot == 0 ? ctx->pref.ptr : ctx->psec.ptr;
        ctx->ref_gen++;   // the int8 copies of the pair are stale
        uint8_t* dm = (uint8_t*)(slot == 0 ? ctx->rmask.ptr : ctx->smask.ptr);
        if (pdt == AR_U8 && p.preprocess == AR_PRE_NONE && p.quantize == AR_QUANT_NONE) {
            // an integer image passes through unchanged
            PCHK(hipMemcpyAsync(dst, img, N, hipMemcpyDefault, ctx->stream));
            if (m) PCHK(hipMemcpyAsync(dm, m, N, hipMemcpyDefault, ctx->stream));
            else {
                if ((st = ctx->rawmask.reserve(ctx, N))) return st;
                std::vector<uint8_t> ones(N, 1);
                PCHK(hipMemcpyAsync(dm, ones.data(), N, hipMemcpyHostToDevice, ctx->stream));
                PCHK(hipStreamSynchronize(ctx->stream));
            }
            continue;
        }
        const PrepLane L = prep_lane(ctx, side && slot == 1);
        switch (dtype) {
            case AR_C64: st = ar_prepare_complex(ctx, img, m, ny, nx, dst, dm); break;
            case AR_U8: st = prepare_slot<uint8_t>(ctx, img, m, ny, nx, dst, dm, slot, L); break;
            case AR_I16: st = prepare_slot<int16_t>(ctx, img, m, ny, nx, dst, dm, slot, L); break;
            case AR_F32: st = prepare_slot<float>(ctx, img, m, ny, nx, dst, dm, slot, L); break;
            default: st = prepare_slot<double>(ctx, img, m, ny, nx, dst, dm, slot, L); break;
        }
        if (st) {
            if (side) hipStreamSynchronize(ctx->stream2);
            return st;
        }
        // (otherwise no wait: the next slot's upload into the same raw buffer is ordered on the
        // stream)
    }
    if (side) {
        PCHK(hipEventRecord(ctx->evf[1], ctx->stream2));
        PCHK(hipStreamWaitEvent(ctx->stream, ctx->evf[1], 0));
    }
    ctx->ny = ny;
    ctx->nx = nx;
    ctx->pdtype = pdt;
    ctx->has_pair = true;
    // masks are already in rmask/smask: combine without re-uploading
    return ar_upload_masks(ctx, nullptr, nullptr, N);
}

extern "C" ar_status ar_quantizer_stats(ar_context* ctx, double* stats) {
    if (!ctx || !stats) return AR_ERR_INVALID_ARGUMENT;
    if (!ctx->has_pair || ctx->params.quantize != AR_QUANT_UINT8 || !ctx->qstats.ptr)
        return pfail(ctx, AR_ERR_INVALID_ARGUMENT,
                     "no UInt8-quantised pair: the quantiser has no parameters to report");
    PCHK(hipSetDevice(ctx->device));
    QParams q[2];
    PCHK(hipMemcpyAsync(q, ctx->qstats.ptr, sizeof(q), hipMemcpyDeviceToHost, ctx->stream));
    PCHK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < 2; i++) {
        double* o = stats + 8 * i;
        o[0] = (double)q[i].n;
        o[1] = q[i].mean;
        o[2] = q[i].sd;
        o[3] = (double)q[i].mode;
        for (int k = 0; k < 2; k++) {
            o[4 + k] = (double)q[i].walk_items[k];
            o[6 + k] = (double)q[i].walk_in_order[k];
        }
    }
    return AR_OK;
}

extern "C" ar_status ar_sequential_sum(ar_context* ctx, const double* x, int64_t n, int32_t method,
                                       double* out, int64_t* walk) {
    if (!ctx || !x || !out || n <= 0) return AR_ERR_INVALID_ARGUMENT;
    if (method != 0 && method != 1)
        return pfail(ctx, AR_ERR_INVALID_ARGUMENT, "method must be 0 (chunk walk) or 1 (one wave)");
    if (method == 1)
        for (int64_t k = 0; k < n; k++)
            if (!(x[k] >= 0.0) || !std::isfinite(x[k]))
                return pfail(ctx, AR_ERR_INVALID_ARGUMENT,
                             "the one-wave sum takes non-negative finite terms");
    const int64_t nch = (n + SQ_CHUNK - 1) / SQ_CHUNK;
    if (nch >= ((int64_t)1 << 31)) return pfail(ctx, AR_ERR_UNSUPPORTED, "array too long");
    PCHK(hipSetDevice(ctx->device));
    ar_status st;
    const size_t recb = (size_t)nch * SQ_SLOTS * sizeof(SqRec);
    const size_t per = SQ_WG * 8 + 8 + 8 + 8 + 4 + 4 + 4;
    if ((st = ctx->work_f64.reserve(ctx, (size_t)n * 8 + 64))) return st;
    if ((st = ctx->qsum.reserve(ctx, 2 * recb + (size_t)nch * per + 256))) return st;
    double* dx = (double*)ctx->work_f64.ptr;
    double* dout = dx + n;
    long long* dstats = (long long*)(dout + 1);
    PCHK(hipMemcpyAsync(dx, x, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream));
    if (method == 1) {
        wave_sum_kernel<<<1, 64, 0, ctx->stream>>>(dx, n, dout);
    } else {
        char* b = (char*)ctx->qsum.ptr;
        SqRec* rec = (SqRec*)b;
        SqRec* items = (SqRec*)(b + recb);
        double* tsum = (double*)(b + 2 * recb);
        double* csum = tsum + (size_t)nch * SQ_WG;
        double* coff = csum + nch;
        long long* ccnt = (long long*)(coff + nch);
        int* clow = (int*)(ccnt + nch);
        int* cg = clow + nch;
        int* cnt = cg + nch;
        const bool inline_scan = nch <= SQ_SCAN_INLINE;
        const ArrayTerm term{dx};
        sq_chunk_kernel<<<(unsigned)nch, SQ_WG, 0, ctx->stream>>>(term, n, tsum, csum, clow, ccnt);
        if (!inline_scan)
            sq_scan_kernel<<<1, 1024, 0, ctx->stream>>>(csum, clow, ccnt, nch, coff, cg);
        sq_segment_kernel<<<(unsigned)nch, SQ_WG, 0, ctx->stream>>>(
            term, n, tsum, csum, clow, inline_scan ? nullptr : coff, cg, cnt, rec);
        sq_walk_kernel<<<1, SQ_WG, 0, ctx->stream>>>(term, n, nch, ccnt, cnt, rec, items,
                                                     SumOut{dout, dstats});
    }
    PCHK(hipGetLastError());
    long long ws[2] = {0, 0};
    PCHK(hipMemcpyAsync(out, dout, 8, hipMemcpyDeviceToHost, ctx->stream));
    if (method == 0) PCHK(hipMemcpyAsync(ws, dstats, 16, hipMemcpyDeviceToHost, ctx->stream));
    PCHK(hipStreamSynchronize(ctx->stream));
    if (walk) {
        walk[0] = ws[0];
        walk[1] = ws[1];
    }
    return AR_OK;
}
