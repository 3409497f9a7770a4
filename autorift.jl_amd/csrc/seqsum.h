// The reference's left-to-right Float64 sum (`s = 0.0; for v in xs: s += v`, as in the UInt8
// quantiser's mean and squared deviations, src/preprocess.jl:546-592) computed on the device with
// the same result bit for bit, in parallel.
//
// Why it can be done. Let s be the running sum and U = 2^(e - 52) the spacing of Float64 values in
// s's binade [2^e, 2^(e+1)). While s and s + x stay strictly inside that binade (one spacing from
// either edge), fl(s + x) = s + RN_U(x): the addition rounds x to a multiple of U, and only a tie
// (x exactly halfway between two multiples) depends on s, through the parity of s / U (ties go to
// the even multiple). So a run of steps that stays inside one binade is an integer sum in units of
// U, and the effect of a run is a function of the parity of s / U on entry: its total, its least and
// largest partial sums and its parity on exit, for each entry parity. Likewise, while every partial
// sum stays below 2^(g + 53) in magnitude and every x and s are multiples of 2^g, no step rounds at
// all and the run is an exact integer sum in units of 2^g.
//
// How. The elements are cut into chunks of SQ_CHUNK (one workgroup: 256 threads, SQ_T consecutive
// elements each). An approximate prefix sum (any order) predicts, for every step, the binade of
// the running sum before and after it; steps are classified as exact-zone steps (key: the grid g,
// the least set bit of every element so far), in-binade steps (key: the binade e) or steps taken on
// their own (binade changes, leading zeros aside), and each maximal run of one key becomes a
// segment record. A single wave then walks the records in order with the true running sum: an
// exact or in-binade record is applied in O(1) once the true sum is checked to be in its binade
// (or on its grid) and the record's least and largest partial sums are checked to keep every step
// of the run strictly inside it; a single step is added as is; any record that fails its check,
// and any chunk with more than SQ_SLOTS records, is summed element by element in order. The
// prediction only decides how much work the walk does: every result is either proven equal to the
// sequential one or computed sequentially.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

namespace ar {

#define SQ_T 16                           // consecutive elements per thread
#define SQ_WG 256                         // threads per chunk
#define SQ_CHUNK (SQ_WG * SQ_T)           // elements per chunk
#define SQ_SLOTS 64                       // segment records per chunk
#define SQ_SEQ 0x7fff0003                 // walk item: a chunk summed in order
#define SQ_RUN_CHUNKS (1 << 18)           // merged runs stop every 2^18 chunks (len < 2^31)
#define SQ_SCAN_INLINE 1024               // up to this many chunks the segment kernel forms the
                                          // chunk prefixes itself (no scan kernel)
#define SQ_DEP 0x7fff0001                 // a step taken on its own
#define SQ_ZERO 0x7fff0002                // no nonzero element so far: nothing to add
#define SQ_EXACT 0x40000000               // + (g + 4096): exact-zone run on the grid 2^g
#define SQ_GNONE INT_MAX                  // no finite nonzero element so far
#define SQ_SENT ((int64_t)1 << 61)        // partial-sum sentinels (|any real partial sum| < 2^60)
#define SQ_MAGIC 6755399441055744.0       // 1.5 * 2^52: y + SQ_MAGIC rounds y to an integer

// One segment record (64 bytes): elements [start, start + len) under one key. For a run, t/mn/mx
// are the total and the least / largest inclusive partial sums in grid units for entry parity 0
// and 1; for a single step t[0] holds the bits of the element.
struct SqRec {
    int64_t start;
    int32_t len;
    int32_t key;
    int64_t t[2];
    int64_t mn[2];
    int64_t mx[2];
};

// The effect of a run for entry parity 0 and 1 (p: parity on exit).
struct SqT {
    int64_t t0, t1, mn0, mn1, mx0, mx1;
    int p0, p1;
};

__device__ inline SqT sq_identity() { return SqT{0, 0, SQ_SENT, SQ_SENT, -SQ_SENT, -SQ_SENT, 0, 1}; }

// a then b
__device__ inline SqT sq_compose(const SqT& a, const SqT& b) {
    SqT r;
    const int q0 = a.p0, q1 = a.p1;
    const int64_t bt0 = q0 ? b.t1 : b.t0, bmn0 = q0 ? b.mn1 : b.mn0, bmx0 = q0 ? b.mx1 : b.mx0;
    const int64_t bt1 = q1 ? b.t1 : b.t0, bmn1 = q1 ? b.mn1 : b.mn0, bmx1 = q1 ? b.mx1 : b.mx0;
    r.t0 = a.t0 + bt0;
    r.t1 = a.t1 + bt1;
    r.mn0 = min(a.mn0, a.t0 + bmn0);
    r.mn1 = min(a.mn1, a.t1 + bmn1);
    r.mx0 = max(a.mx0, a.t0 + bmx0);
    r.mx1 = max(a.mx1, a.t1 + bmx1);
    r.p0 = q0 ? b.p1 : b.p0;
    r.p1 = q1 ? b.p1 : b.p0;
    return r;
}

// Append one step to a run: d0 is the even candidate of a tie (or the step), d1 the odd one
// (equal to d0 off ties); the running parity picks the candidate that makes the sum even.
__device__ inline void sq_push(SqT& a, int64_t d0, int64_t d1) {
    const int64_t e0 = a.p0 ? d1 : d0, e1 = a.p1 ? d1 : d0;
    a.t0 += e0;
    a.t1 += e1;
    a.p0 ^= (int)(e0 & 1);
    a.p1 ^= (int)(e1 & 1);
    a.mn0 = min(a.mn0, a.t0);
    a.mn1 = min(a.mn1, a.t1);
    a.mx0 = max(a.mx0, a.t0);
    a.mx1 = max(a.mx1, a.t1);
}

// exponent of the least significant set bit of a finite nonzero x
__device__ inline int sq_lowbit(double x) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const uint64_t m = b & 0xfffffffffffffull;
    const int ex = (int)((b >> 52) & 0x7ff);
    if (ex == 0) return -1074 + __builtin_ctzll(m);
    return ex - 1075 + __builtin_ctzll(m | 0x10000000000000ull);
}

// binade exponent of a nonzero finite x
__device__ inline int sq_binade(double x) { return ilogb(x); }

// nearest integer of y (ties to even), |y| < 2^51, as int64
__device__ inline int64_t sq_rne(double y, double& r) {
    const double t = y + SQ_MAGIC;
    r = t - SQ_MAGIC;
    return (int64_t)__double_as_longlong(t) - (int64_t)__double_as_longlong(SQ_MAGIC);
}

// Key of the step x taken from the predicted running sum ab (g: least set bit of every element
// so far; lim = 2^(g + 52), 0 when there is no exact zone); d0 / d1: what the step adds in grid
// units when the running sum's parity is even / odd (they differ only on a tie).
__device__ inline int sq_key(double ab, double x, int g, double lim, int64_t& d0, int64_t& d1) {
    d0 = d1 = 0;
    if (g == SQ_GNONE) return x == 0.0 ? SQ_ZERO : SQ_DEP;   // no finite nonzero element so far
    const double aa = ab + x;
    if (!isfinite(x) || !isfinite(aa)) return SQ_DEP;
    if (fabs(ab) < lim && fabs(aa) < lim) {
        const double y = ldexp(x, -g);   // an integer: x is a multiple of 2^g
        if (fabs(y) < 0x1p48) {
            double r;
            d0 = d1 = sq_rne(y, r);
            return SQ_EXACT + g + 4096;
        }
    }
    if (ab != 0.0 && aa != 0.0) {
        const int eb = sq_binade(ab);
        if (eb == sq_binade(aa) && eb >= -1022) {
            const double y = ldexp(x, 52 - eb);
            if (fabs(y) < 0x1p48) {
                double r;
                d0 = d1 = sq_rne(y, r);
                const double dt = r - y;   // exact
                if (dt == 0.5) d1 = d0 - 1;          // y = d0 - 1/2: the other candidate is below
                else if (dt == -0.5) d1 = d0 + 1;
                return eb;
            }
        }
    }
    return SQ_DEP;
}

__device__ inline double sq_shfl_up_d(double v, int d) {
    return __longlong_as_double(__shfl_up((long long)__double_as_longlong(v), d, 64));
}
__device__ inline int64_t sq_shfl_up_l(int64_t v, int d) {
    return (int64_t)__shfl_up((long long)v, d, 64);
}

// Called by one whole workgroup (any multiple of 64 threads): predicted running sum at each chunk
// start (exclusive scan of the chunk sums, any fixed order), the least set bit of every element up
// to and including each chunk (inclusive prefix minimum).
__device__ inline void sq_scan_block(const double* csum, const int* clow, const long long* ccnt,
                                     int64_t nch, double* coff, int* cg) {
    __shared__ double ws[16];
    __shared__ int wl[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, nw = blockDim.x >> 6;
    double carry = 0.0;
    int gcarry = SQ_GNONE;
    for (int64_t base = 0; base < nch; base += blockDim.x) {
        const int64_t c = base + t;
        const double v = c < nch ? csum[c] : 0.0;
        const int l = c < nch ? clow[c] : SQ_GNONE;
        double sv = v;
        int sl = l;
        for (int d = 1; d < 64; d <<= 1) {
            const double o = sq_shfl_up_d(sv, d);
            const int ol = __shfl_up(sl, d, 64);
            if (lane >= d) {
                sv += o;
                sl = min(sl, ol);
            }
        }
        if (lane == 63) {
            ws[w] = sv;
            wl[w] = sl;
        }
        __syncthreads();
        double pre = carry;
        int gpre = gcarry;
        for (int i = 0; i < w; i++) {
            pre += ws[i];
            gpre = min(gpre, wl[i]);
        }
        if (c < nch) {
            coff[c] = pre + (sv - v);
            cg[c] = min(gpre, sl);
        }
        for (int i = 0; i < nw; i++) {
            carry += ws[i];
            gcarry = min(gcarry, wl[i]);
        }
        __syncthreads();
    }
}

// Per chunk (one workgroup, SQ_T consecutive elements a thread): each thread's sum in order,
// the chunk's sum (a fixed tree over the thread sums), the least set bit of its finite nonzero
// elements and how many are valid; the last workgroup then scans the chunks (sq_scan_block).
// TERM: int64_t -> (double, bool valid).
template <class TERM>
__global__ __launch_bounds__(SQ_WG) void sq_chunk_kernel(TERM term, int64_t n, double* tsum,
                                                        double* csum, int* clow,
                                                        long long* ccnt) {
    __shared__ double ws[SQ_WG / 64];
    __shared__ int wl[SQ_WG / 64];
    __shared__ long long wc[SQ_WG / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t c = blockIdx.x;
    const int64_t k0 = c * SQ_CHUNK + (int64_t)t * SQ_T;
    const int64_t k1 = min(k0 + SQ_T, n);
    double s = 0.0;
    int low = SQ_GNONE;
    long long cnt = 0;
    for (int64_t k = k0; k < k1; k++) {
        bool ok;
        const double x = term(k, ok);
        s += x;
        cnt += ok;
        if (x != 0.0 && isfinite(x)) low = min(low, sq_lowbit(x));
    }
    tsum[c * SQ_WG + t] = s;
    for (int o = 32; o > 0; o >>= 1) {
        s += __longlong_as_double(__shfl_xor((long long)__double_as_longlong(s), o, 64));
        low = min(low, __shfl_xor(low, o, 64));
        cnt += __shfl_xor(cnt, o, 64);
    }
    if (lane == 0) {
        ws[w] = s;
        wl[w] = low;
        wc[w] = cnt;
    }
    __syncthreads();
    if (t == 0) {
        double a = ws[0];
        int l = wl[0];
        long long m = wc[0];
        for (int i = 1; i < SQ_WG / 64; i++) {
            a += ws[i];
            l = min(l, wl[i]);
            m += wc[i];
        }
        csum[c] = a;
        clow[c] = l;
        ccnt[c] = m;
    }
}

// One workgroup: sq_scan_block over every chunk (images of more than SQ_SCAN_INLINE chunks; the
// segment kernel forms the prefixes of smaller ones itself).
static __global__ __launch_bounds__(1024) void sq_scan_kernel(const double* csum, const int* clow,
                                                       const long long* ccnt, int64_t nch,
                                                       double* coff, int* cg) {
    sq_scan_block(csum, clow, ccnt, nch, coff, cg);
}

// A run's piece while the records are formed: effect, first element, key, element of a single
// step.
struct SqPiece {
    SqT tr;
    int64_t start;
    int key;
    double x;
};

__device__ inline SqPiece sq_shfl_up_p(const SqPiece& a, int d) {
    SqPiece r;
    r.tr = SqT{sq_shfl_up_l(a.tr.t0, d), sq_shfl_up_l(a.tr.t1, d), sq_shfl_up_l(a.tr.mn0, d),
               sq_shfl_up_l(a.tr.mn1, d), sq_shfl_up_l(a.tr.mx0, d), sq_shfl_up_l(a.tr.mx1, d),
               __shfl_up(a.tr.p0, d, 64), __shfl_up(a.tr.p1, d, 64)};
    r.start = sq_shfl_up_l(a.start, d);
    r.key = __shfl_up(a.key, d, 64);
    r.x = sq_shfl_up_d(a.x, d);
    return r;
}

// the run `a` continued by the piece `b` (which holds no start)
__device__ inline SqPiece sq_extend(const SqPiece& a, const SqT& b) {
    SqPiece r = a;
    r.tr = sq_compose(a.tr, b);
    return r;
}

__device__ inline void sq_write(SqRec* r, const SqPiece& p, int64_t end) {
    SqRec q;
    q.start = p.start;
    q.len = (int32_t)(end - p.start);
    q.key = p.key;
    if (p.key == SQ_DEP) {
        q.t[0] = (int64_t)__double_as_longlong(p.x);
        q.t[1] = 0;
        q.mn[0] = q.mn[1] = q.mx[0] = q.mx[1] = 0;
    } else {
        q.t[0] = p.tr.t0;
        q.t[1] = p.tr.t1;
        q.mn[0] = p.tr.mn0;
        q.mn[1] = p.tr.mn1;
        q.mx[0] = p.tr.mx0;
        q.mx[1] = p.tr.mx1;
    }
    *r = q;
}

// Elements [a, b) added to s in order, by one wave (uniform s). Lane L loads elements
// 16 L .. 16 L + 15 of each group of 1024 (the next group is in flight while one is summed);
// elements past b are +0.0, which leaves s unchanged.
#define SQ_IO 16
template <class TERM>
__device__ inline double sq_in_order(const TERM& term, int64_t a, int64_t b, double s) {
    const int lane = threadIdx.x & 63;
    auto load = [&](int64_t k0, double (&v)[SQ_IO]) {
#pragma unroll
        for (int i = 0; i < SQ_IO; i++) {
            const int64_t k = k0 + (int64_t)lane * SQ_IO + i;
            bool ok;
            v[i] = k < b ? term(k, ok) : 0.0;
        }
    };
    double x[SQ_IO];
    load(a, x);
    for (int64_t k0 = a; k0 < b; k0 += 64 * SQ_IO) {
        double y[SQ_IO];
        const bool more = k0 + 64 * SQ_IO < b;
        if (more) load(k0 + 64 * SQ_IO, y);
        const int lanes = (int)min((int64_t)64, (b - k0 + SQ_IO - 1) / SQ_IO);
        for (int L = 0; L < lanes; L++) {
#pragma unroll
            for (int i = 0; i < SQ_IO; i++) {
                const long long xb = __double_as_longlong(x[i]);
                const uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)(xb & 0xffffffff), L);
                const uint32_t h = (uint32_t)__builtin_amdgcn_readlane((int)(xb >> 32), L);
                s += __longlong_as_double((long long)(((uint64_t)h << 32) | l));
            }
        }
        if (more) {
#pragma unroll
            for (int i = 0; i < SQ_IO; i++) x[i] = y[i];
        }
    }
    return s;
}

// Apply one record to the running sum s; false when its check fails (the caller then sums its
// elements in order).
__device__ inline bool sq_apply(const SqRec& q, double& s) {
    if (q.key == SQ_ZERO) return true;   // only zeros: s + 0 == s
    if (q.key == SQ_DEP) {
        s += __longlong_as_double((long long)q.t[0]);
        return true;
    }
    if (q.key >= SQ_EXACT) {
        const int g = q.key - SQ_EXACT - 4096;
        const double qs = ldexp(s, -g);
        if (!(fabs(qs) < 0x1p50)) return false;
        double rq;
        const int64_t S = sq_rne(qs, rq);
        if (rq != qs) return false;   // s is not on the grid 2^g
        const int64_t lim = (int64_t)1 << 53;
        if (S + q.mn[0] <= -lim || S + q.mx[0] >= lim) return false;
        s = ldexp((double)(S + q.t[0]), g);
        return true;
    }
    // in-binade run: s = +-M 2^(e - 52) with M in [2^52, 2^53) read from its bits
    const int e = q.key;
    const uint64_t bits = (uint64_t)__double_as_longlong(s);
    if ((int)((bits >> 52) & 0x7ff) != e + 1023) return false;   // also rejects 0, inf, NaN
    const int64_t M = (int64_t)((bits & 0xfffffffffffffull) | 0x10000000000000ull);
    const bool neg = (bits >> 63) != 0;
    const int p = (int)(M & 1);   // the parity of s / U
    const int64_t mn = p ? q.mn[1] : q.mn[0], mx = p ? q.mx[1] : q.mx[0], t = p ? q.t[1] : q.t[0];
    const int64_t lo = ((int64_t)1 << 52) + 1, hi = ((int64_t)1 << 53) - 1;
    // every partial sum strictly inside the binade: M +- (partial sum) in [lo, hi]
    const int64_t a = neg ? M - mx : M + mn, b = neg ? M - mn : M + mx;
    if (a < lo || b > hi) return false;
    const int64_t M2 = neg ? M - t : M + t;
    s = __longlong_as_double((long long)((bits & 0xfff0000000000000ull) |
                                         ((uint64_t)M2 & 0xfffffffffffffull)));
    return true;
}

__device__ inline int64_t sq_rl64(int64_t v, int j) {
    const uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)(v & 0xffffffff), j);
    const uint32_t h = (uint32_t)__builtin_amdgcn_readlane((int)(v >> 32), j);
    return (int64_t)(((uint64_t)h << 32) | l);
}


// Called by one whole workgroup of SQ_WG threads after every chunk's records are written: the walk
// items in order. Chunks with one record whose key is shared with the neighbouring such chunks
// are merged into one item per run (composed effect, first element, total length); every other
// chunk contributes its records as they are, or one SQ_SEQ item. Returns the item count.
__device__ inline int64_t sq_merge_block(const int* cnt, const SqRec* rec, SqRec* items,
                                         int64_t nch, int64_t n) {
    __shared__ int mk[SQ_WG];          // key of a mergeable chunk, SQ_DEP otherwise
    __shared__ int hd[SQ_WG + 1];      // chunk starts a new item
    __shared__ int wcount[SQ_WG / 64];
    __shared__ SqPiece wv[SQ_WG / 64];
    __shared__ int wflag[SQ_WG / 64];
    __shared__ SqPiece carry;          // the run still open at the end of the previous tile
    __shared__ int carry_open, carry_key;
    __shared__ long long pos_s;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) {
        carry_open = 0;
        carry_key = SQ_DEP;
        pos_s = 0;
    }
    __syncthreads();
    for (int64_t base = 0; base < nch; base += SQ_WG) {
        const int64_t c = base + t;
        const int m = c < nch ? cnt[c] : 0;
        SqRec r0{};
        if (c < nch && m >= 1) r0 = rec[c * SQ_SLOTS];
        const bool mergeable = c < nch && m == 1 && r0.key != SQ_DEP;
        mk[t] = mergeable ? r0.key : SQ_DEP;
        __syncthreads();
        const int pkey = t > 0 ? mk[t - 1] : (carry_open ? carry_key : SQ_DEP);
        const bool head = !mergeable || pkey == SQ_DEP || pkey != r0.key || (c % SQ_RUN_CHUNKS) == 0;
        hd[t] = head || !mergeable;
        if (t == 0) hd[SQ_WG] = 1;   // the next tile decides for itself
        __syncthreads();
        // the run open from the previous tile ends before this tile's first chunk
        const bool emit_carry = t == 0 && carry_open && head;
        const bool last_in_tile = t == SQ_WG - 1 || c == nch - 1;
        const bool run_end = mergeable && ((c == nch - 1) || (!last_in_tile && hd[t + 1]));
        int nout = 0;
        if (c < nch) nout = mergeable ? (run_end ? 1 : 0) : (m < 0 ? 1 : m);
        nout += emit_carry ? 1 : 0;
        // segmented scan of the mergeable chunks' effects (heads start a new run)
        SqPiece V{mergeable ? SqT{r0.t[0], r0.t[1], r0.mn[0], r0.mn[1], r0.mx[0], r0.mx[1],
                                  0, 1}
                            : sq_identity(),
                  r0.start, r0.key, 0.0};
        if (mergeable) {
            // exit parity of a run record: for an in-binade run the parity flips with each odd
            // grid step; the record holds the totals, whose parity gives it
            V.tr.p0 = (int)(r0.t[0] & 1);
            V.tr.p1 = (int)((1 + r0.t[1]) & 1);
        }
        bool flag = head;
        if (t == 0 && !head && carry_open) {
            V = SqPiece{sq_compose(carry.tr, V.tr), carry.start, carry.key, 0.0};
            flag = true;
        }
        for (int d = 1; d < 64; d <<= 1) {
            const SqPiece o = sq_shfl_up_p(V, d);
            const bool of = __shfl_up((int)flag, d, 64) != 0;
            if (lane >= d && !flag) V = SqPiece{sq_compose(o.tr, V.tr), o.start, o.key, o.x};
            if (lane >= d) flag = flag || of;
        }
        // exclusive scan of the item counts
        int inc = nout;
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        if (lane == 63) {
            wv[w] = V;
            wflag[w] = flag;
            wcount[w] = inc;
        }
        __syncthreads();
        if (w > 0 && !flag) {
            SqPiece cv = wv[0];
            for (int i = 1; i < w; i++)
                if (wflag[i]) cv = wv[i];
                else cv.tr = sq_compose(cv.tr, wv[i].tr);
            V = SqPiece{sq_compose(cv.tr, V.tr), cv.start, cv.key, cv.x};
        }
        long long pos = pos_s + inc - nout;
        for (int i = 0; i < w; i++) pos += wcount[i];
        if (emit_carry) {   // the open run ends where this tile's first chunk starts
            sq_write(items + pos, carry, base * SQ_CHUNK);
            pos++;
        }
        if (c < nch) {
            if (!mergeable) {
                if (m < 0) {
                    SqRec q{};
                    q.start = c * SQ_CHUNK;
                    q.len = (int32_t)(min(c * SQ_CHUNK + SQ_CHUNK, n) - c * SQ_CHUNK);
                    q.key = SQ_SEQ;
                    items[pos] = q;
                } else {
                    for (int i = 0; i < m; i++) items[pos + i] = rec[c * SQ_SLOTS + i];
                }
            } else if (run_end) {
                sq_write(items + pos, V, min(c * SQ_CHUNK + SQ_CHUNK, n));
            }
        }
        __syncthreads();
        if (t == SQ_WG - 1) {
            // a run that reaches the tile's last chunk and is not ended stays open
            const bool open = c < nch && mergeable && !run_end;
            carry_open = open;
            if (open) {
                carry = V;
                carry_key = r0.key;
            }
            long long tot = 0;
            for (int i = 0; i < SQ_WG / 64; i++) tot += wcount[i];
            pos_s += tot;
        }
        __syncthreads();
    }
    return pos_s;
}

// One wave: walk the items in order from s = 0; the sequential sum. Lane j holds item b + j; the
// next 64 items are loaded while this batch is walked. A merged run whose check fails is retried
// chunk by chunk, and any record or chunk whose check fails is summed in order.
template <class TERM>
__device__ inline double sq_walk_items(const TERM& term, int64_t n, const int* cnt,
                                       const SqRec* rec, const SqRec* items, int64_t nit,
                                       long long& in_order) {
    const int lane = threadIdx.x & 63;
    SqRec qc{};
    if (lane < nit) qc = items[lane];
    double s = 0.0;
    for (int64_t b = 0; b < nit; b += 64) {
        SqRec qn{};
        if (b + 64 + lane < nit) qn = items[b + 64 + lane];
        const int jn = (int)min((int64_t)64, nit - b);
        for (int j = 0; j < jn; j++) {
            SqRec q;
            q.key = __builtin_amdgcn_readlane(qc.key, j);
            if (q.key == SQ_DEP) {   // a single step: only its element is needed
                s += __longlong_as_double((long long)sq_rl64(qc.t[0], j));
                continue;
            }
            if (q.key == SQ_ZERO) continue;
            q.start = sq_rl64(qc.start, j);
            q.len = __builtin_amdgcn_readlane(qc.len, j);
            for (int u = 0; u < 2; u++) {
                q.t[u] = sq_rl64(qc.t[u], j);
                q.mn[u] = sq_rl64(qc.mn[u], j);
                q.mx[u] = sq_rl64(qc.mx[u], j);
            }
            if (q.key == SQ_SEQ) {
                s = sq_in_order(term, q.start, q.start + q.len, s);
                in_order += q.len;
                continue;
            }
            if (sq_apply(q, s)) continue;
            const int64_t c1 = q.start / SQ_CHUNK, c2 = (q.start + q.len - 1) / SQ_CHUNK;
            if (c1 == c2) {
                s = sq_in_order(term, q.start, q.start + q.len, s);
                in_order += q.len;
                continue;
            }
            for (int64_t c = c1; c <= c2; c++) {   // each chunk of a merged run holds one record
                const SqRec r = rec[c * SQ_SLOTS];
                if (!sq_apply(r, s)) {
                    s = sq_in_order(term, r.start, r.start + r.len, s);
                    in_order += r.len;
                }
            }
        }
        qc = qn;
    }
    return s;
}

// Per chunk (one workgroup): the segment records of its elements, or cnt = -1 when they would take
// more than SQ_SLOTS records. A record starts at the chunk's first element, at every step taken on
// its own and after it, and wherever the key changes.
template <class TERM>
__global__ __launch_bounds__(SQ_WG) void sq_segment_kernel(TERM term, int64_t n,
                                                          const double* tsum, const double* csum,
                                                          const int* clow, const double* coff,
                                                          const int* cg, int* cnt, SqRec* rec) {
    __shared__ int keys[SQ_T][SQ_WG];
    __shared__ int64_t dd[SQ_T][SQ_WG];
    __shared__ int8_t dflag[SQ_T][SQ_WG];
    __shared__ double wsum[SQ_WG / 64];
    __shared__ int wlast[SQ_WG / 64], wst[SQ_WG / 64];
    __shared__ SqPiece wv[SQ_WG / 64];
    __shared__ int wflag[SQ_WG / 64];
    __shared__ int total_s;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t c = blockIdx.x;
    const int64_t k0 = c * SQ_CHUNK + (int64_t)t * SQ_T;
    const int64_t k1 = min(k0 + SQ_T, n);
    const int64_t cend = min(c * SQ_CHUNK + SQ_CHUNK, n);
    // the chunk's predicted start and grid: from the scan kernel, or formed here for few chunks
    __shared__ double cbase;
    __shared__ int cgrid;
    if (coff) {
        if (t == 0) {
            cbase = coff[c];
            cgrid = cg[c];
        }
    } else {
        double ps = 0.0;
        int pl = SQ_GNONE;
        for (int64_t i = t; i <= c; i += SQ_WG) {
            if (i < c) ps += csum[i];
            pl = min(pl, clow[i]);
        }
        for (int o = 32; o > 0; o >>= 1) {
            ps += __longlong_as_double(__shfl_xor((long long)__double_as_longlong(ps), o, 64));
            pl = min(pl, __shfl_xor(pl, o, 64));
        }
        __shared__ double rps[SQ_WG / 64];
        __shared__ int rpl[SQ_WG / 64];
        if (lane == 0) {
            rps[w] = ps;
            rpl[w] = pl;
        }
        __syncthreads();
        if (t == 0) {
            double a = rps[0];
            int l = rpl[0];
            for (int i = 1; i < SQ_WG / 64; i++) {
                a += rps[i];
                l = min(l, rpl[i]);
            }
            cbase = a;
            cgrid = l;
        }
    }
    __syncthreads();
    const int g = cgrid;
    const double lim = g != SQ_GNONE && g + 52 < 1023 ? ldexp(1.0, g + 52) : 0.0;
    // predicted running sum at this thread's first element: exclusive scan of the thread sums
    const double ts = tsum[c * SQ_WG + t];
    double inc = ts;
    for (int d = 1; d < 64; d <<= 1) {
        const double o = sq_shfl_up_d(inc, d);
        if (lane >= d) inc += o;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    double base = cbase + (inc - ts);
    for (int i = 0; i < w; i++) base += wsum[i];
    // keys and steps of this thread's elements; starts after the first element; first / last key
    int first = SQ_ZERO, prev = SQ_ZERO, nin = 0;
    {
        double a = base;
        for (int i = 0; i < SQ_T; i++) {
            const int64_t k = k0 + i;
            int key = SQ_ZERO;
            int64_t d0 = 0, d1 = 0;
            double x = 0.0;
            if (k < k1) {
                bool ok;
                x = term(k, ok);
                key = sq_key(a, x, g, lim, d0, d1);
                if (i == 0) first = key;
                else if (key == SQ_DEP || prev == SQ_DEP || key != prev) nin++;
                prev = key;
                a += x;
            }
            keys[i][t] = key;
            dd[i][t] = key == SQ_DEP ? (int64_t)__double_as_longlong(x) : d0;
            dflag[i][t] = (int8_t)(d1 - d0);
        }
    }
    const bool nonempty = k0 < k1;
    // the previous thread's last key
    int plast = __shfl_up(prev, 1, 64);
    if (lane == 63) wlast[w] = prev;
    __syncthreads();
    if (lane == 0 && w > 0) plast = wlast[w - 1];
    const bool s0 = nonempty && (t == 0 || first == SQ_DEP || plast == SQ_DEP || first != plast);
    const int nst = nin + (s0 ? 1 : 0);
    int idx = nst;   // inclusive scan of the starts
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(idx, d, 64);
        if (lane >= d) idx += o;
    }
    if (lane == 63) wst[w] = idx;
    __syncthreads();
    for (int i = 0; i < w; i++) idx += wst[i];
    if (t == SQ_WG - 1) total_s = idx;
    __syncthreads();
    const int total = total_s;
    if (total > SQ_SLOTS) {
        if (t == 0) cnt[c] = -1;
    } else {
    SqRec* r = rec + c * SQ_SLOTS;
    int seg = idx - nst;   // index of the first record that starts in this thread
    // pieces: F (elements before the thread's first start), P (the current piece)
    SqPiece P{sq_identity(), k0, SQ_ZERO, 0.0};
    SqT F = sq_identity();
    bool started = false;
    int64_t fstart = k1;
    {
        int pk = plast;
        for (int i = 0; i < SQ_T; i++) {
            const int64_t k = k0 + i;
            if (k >= k1) break;
            const int key = keys[i][t];
            const int64_t d0 = dd[i][t];
            const bool st = i == 0 ? s0 : (key == SQ_DEP || pk == SQ_DEP || key != pk);
            if (st) {
                if (started) sq_write(r + seg++, P, k);   // a record wholly inside the thread
                else {
                    F = P.tr;
                    fstart = k;
                }
                started = true;
                P = SqPiece{sq_identity(), k, key,
                            key == SQ_DEP ? __longlong_as_double((long long)d0) : 0.0};
            }
            if (key != SQ_DEP) sq_push(P.tr, d0, d0 + dflag[i][t]);
            pk = key;
        }
    }
    // segmented inclusive scan of the threads' trailing pieces: the run that reaches each
    // thread's last element (a thread without a start continues the run before it)
    SqPiece V = P;
    bool flag = started;
    for (int d = 1; d < 64; d <<= 1) {
        const SqPiece o = sq_shfl_up_p(V, d);
        const bool of = __shfl_up((int)flag, d, 64) != 0;
        if (lane >= d && !flag) V = SqPiece{sq_compose(o.tr, V.tr), o.start, o.key, o.x};
        if (lane >= d) flag = flag || of;
    }
    if (lane == 63) {
        wv[w] = V;
        wflag[w] = flag;
    }
    __syncthreads();
    if (w > 0 && !flag) {
        // the run reaching the end of the previous waves
        SqPiece cv = wv[0];
        for (int i = 1; i < w; i++)
            if (wflag[i]) cv = wv[i];
            else cv.tr = sq_compose(cv.tr, wv[i].tr);
        V = SqPiece{sq_compose(cv.tr, V.tr), cv.start, cv.key, cv.x};
    }
    __syncthreads();
    if (lane == 63) wv[w] = V;   // final: the run reaching each wave's last thread
    __syncthreads();
    SqPiece before = sq_shfl_up_p(V, 1);
    if (lane == 0 && w > 0) before = wv[w - 1];
    // the run that ends at this thread's first start began in an earlier thread
    if (t > 0 && started) sq_write(r + (idx - nst) - 1, sq_extend(before, F), fstart);
    if (t == SQ_WG - 1) {
        sq_write(r + total - 1, V, cend);
        cnt[c] = total;
    }
    }
}

// Walk statistics: items walked, elements summed one by one.
struct SqWalkStats {
    long long items, in_order;
};

// One workgroup of SQ_WG threads: merge the chunks' records into walk items, walk them with one
// wave, count the valid elements, and hand (sum, count, statistics) to fin.
template <class TERM, class FIN>
__global__ __launch_bounds__(SQ_WG) void sq_walk_kernel(TERM term, int64_t n, int64_t nch,
                                                       const long long* ccnt, const int* cnt,
                                                       const SqRec* rec, SqRec* items, FIN fin) {
    __shared__ long long rc[SQ_WG / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    long long m = 0;
    for (int64_t c = t; c < nch; c += SQ_WG) m += ccnt[c];
    for (int o = 32; o > 0; o >>= 1) m += __shfl_xor(m, o, 64);
    if (lane == 0) rc[w] = m;
    const int64_t nit = sq_merge_block(cnt, rec, items, nch, n);
    if (t < 64) {
        SqWalkStats st{nit, 0};
        const double sum = sq_walk_items(term, n, cnt, rec, items, nit, st.in_order);
        if (t == 0) {
            long long count = 0;
            for (int i = 0; i < SQ_WG / 64; i++) count += rc[i];
            fin(sum, count, st);
        }
    }
}

// ---- the reference's left-to-right Float64 sum of non-negative terms, one wave at a time

__device__ inline double wave_lane_f64(double v, int l) {
    const uint64_t u = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// s + t_0 + t_1 + ... + t_63 added left to right in Float64 (t_l >= 0 finite, this lane's t) for a
// wave-uniform s >= 0: the bits of the sequential loop. While the running sum stays strictly
// inside one binade [2^E, 2^(E+1)) every addition rounds its term to the grid u = 2^(E - 52) that
// s lies on: s_k = s + u (r_0 + ... + r_k) with r_l = t_l / u rounded to the nearest integer,
// where a tie (t_l / u halfway between two integers) takes the neighbour that makes s_k / u even.
// A tie's choice depends on the parity of the running sum, so each lane's step is a map of the
// entry parity (what it adds and the parity it leaves), and an inclusive scan of those maps over
// the lanes gives every partial sum for the actual parity of s / u. The first lane whose partial
// sum would leave the binade is added on its own, the reference's way, and the scan resumes after
// it. Every step is exact: the r_l and their prefix sums are integers below 2^53.
__device__ inline double seq_sum64_ties(double s, double t, int lane) {
    int start = 0;
    while (start < 64) {
        if (!(s >= 0x1p-900)) {   // zero or tiny: one term at a time
            s = s + wave_lane_f64(t, start);
            start++;
            continue;
        }
        const int E = (int)((__double_as_longlong(s) >> 52) & 0x7ff) - 1023;
        const double u = __longlong_as_double((long long)(E - 52 + 1023) << 52);
        const double inv = __longlong_as_double((long long)(52 - E + 1023) << 52);
        const double q = lane >= start ? t * inv : 0.0;   // exact: scaling by a power of two
        const double fl = floor(q), fr = q - fl;
        const bool tie = fr == 0.5;
        const bool fl_even = fmod(fl, 2.0) == 0.0;
        // what the lane adds when the running sum is even (a) / odd (b), and the parity it leaves
        const double a = tie ? (fl_even ? fl : fl + 1.0) : (fr > 0.5 ? fl + 1.0 : fl);
        const double b = tie ? (fl_even ? fl + 1.0 : fl) : a;
        double T0 = a, T1 = b;
        int P0 = (int)(fmod(a, 2.0) != 0.0), P1 = (int)(fmod(1.0 + b, 2.0) != 0.0);
        for (int d = 1; d < 64; d <<= 1) {   // inclusive scan of the maps (earlier lanes first)
            const double o0 = __shfl_up(T0, d, 64), o1 = __shfl_up(T1, d, 64);
            const int q0 = __shfl_up(P0, d, 64), q1 = __shfl_up(P1, d, 64);
            if (lane >= d) {
                const double n0 = o0 + (q0 ? T1 : T0), n1 = o1 + (q1 ? T1 : T0);
                const int m0 = q0 ? P1 : P0, m1 = q1 ? P1 : P0;
                T0 = n0;
                T1 = n1;
                P0 = m0;
                P1 = m1;
            }
        }
        const int par = (int)((__double_as_longlong(s) >> 0) & 1);   // parity of s / u
        const double P = par ? T1 : T0;
        const double lim = __longlong_as_double((long long)(E + 1 + 1023) << 52) - u;
        const bool bad = lane >= start && !(s + P * u <= lim);
        const uint64_t bm = __ballot(bad);
        const int f = bm ? __builtin_ctzll(bm) : 64;
        if (f > start) s = s + wave_lane_f64(P, f - 1) * u;   // lanes start .. f-1: exact
        if (f < 64) {
            s = s + wave_lane_f64(t, f);
            start = f + 1;
        } else {
            start = 64;
        }
    }
    return s;
}

}  // namespace ar
