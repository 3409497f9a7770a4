// Serving loop: image pairs of different sizes arrive at any time; each is prepared and run
// through the multi-chip search (src/api.jl:165-170 -> src/multichip.jl:106-137) on one of a
// fixed set of slots. A slot is a worker thread with its own context (device buffers and HIP
// stream), so pairs admitted to different slots overlap on the device; a slot keeps its device
// buffers between pairs and reuses them.
//
// Admission (FIFO): the head of the queue goes to the first idle slot once the device memory it
// needs fits the budget, counting what every other slot holds. A slot's need for a pair is the
// larger of what it already holds and the pair's footprint estimate; when the head does not fit,
// idle slots give their buffers back first. A pair whose footprint alone exceeds the budget is
// refused at submission (AR_ERR_OUT_OF_MEMORY).
//
// The submitting thread's buffers are copied at submission (the library keeps no host pointer)
// into a page-locked staging buffer from a pool the server keeps (16 GiB at most; beyond it,
// pageable memory); results wait in the server until ar_server_wait collects them.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ar_internal.h"
#include "layout.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t ticket = 0;
    int64_t ny = 0, nx = 0, nr = 0, nc = 0;
    int32_t dtype = 0;
    // images and masks, packed [ref | sec | ref_valid | sec_valid] in one staging buffer:
    // page-locked (from the server's pool) when possible, so the upload runs at DMA speed
    uint8_t* stage = nullptr;
    size_t stage_bytes = 0;
    int pinned = -1;                  // pool index, or -1: `spill` holds the bytes
    std::vector<uint8_t> spill;
    bool has_rmask = false, has_smask = false;
    std::vector<double> x, y, pdx, pdy;
    std::vector<int64_t> rx, ry;
    size_t footprint = 0;
    // results
    std::vector<float> dx, dy, corr;
    std::vector<uint16_t> cs;
    std::vector<uint8_t> interp;
    ar_status status = AR_OK;
    std::string err;
    bool done = false;
    int slot = -1;
    Clock::time_point t_submit, t_admit, t_done;
};

size_t elem_size(int32_t dtype) {
    switch (dtype) {
        case AR_U8: return 1;
        case AR_I16: return 2;
        case AR_F32: return 4;
        case AR_F64: return 8;
        case AR_C64: return 8;
        default: return 0;
    }
}

}  // namespace

struct PinnedBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    bool used = false;
};

struct ar_server {
    ar_params params;
    std::vector<PinnedBuf> pinned;    // page-locked staging buffers, reused across requests
    size_t pinned_total = 0;
    size_t pinned_limit = (size_t)16 << 30;
    int device = 0;
    int64_t budget = 0;
    std::vector<ar_context*> ctx;
    std::vector<size_t> held;     // device bytes each slot holds (or has been granted)
    std::vector<bool> busy;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<int64_t, Request*> reqs;
    int64_t next_ticket = 1;
    bool stop = false;
    // counters
    int64_t completed = 0, peak_running = 0, running = 0;
    size_t peak_held = 0;
    std::string err;
};

// Device bytes a context needs for one pair: raw upload and preparation work images, the
// prepared pair and masks, box images (the Float64 planes of the 16-bit and transform paths),
// the four byte-shifted int8 copies of each prepared image, per-point buffers and the
// multi-chip grid planes, the tracker scratch of the
// largest chip-size level the pair can need (the direct Float64 tracker's staging for Float32
// pairs with exact arithmetic or chip areas that are not a power of two, the generic tracker's
// per-workgroup regions for a rotation search, the Coherence tracker's for complex pairs), plus a
// fixed margin.
static size_t pair_footprint(const ar_params& p, int64_t ny, int64_t nx, int32_t dtype,
                             int64_t npts) {
    const size_t N = (size_t)ny * (size_t)nx;
    const size_t es = elem_size(dtype);
    const size_t raw = 2 * N * es + 2 * N;          // raw pair and masks
    // Float64 column sums and counts, Float32 planes (Wallis: output, mean, squared deviations)
    const size_t work = N * (8 + 4) + N * 4 * (p.preprocess == AR_PRE_WALLIS ? 3 : 1);
    const size_t pe = dtype == AR_C64 ? 8 : 4;      // prepared element: ComplexF32 or Float32
    const size_t prep = 2 * N * pe + 3 * N;         // prepared pair (at most pe bytes), masks
    // box images: the 16-bit and transform paths' Float64 column sums and box sums and Float32
    // window variances; otherwise the UInt8 path's int32 planes (bounded by 2 N 8)
    const bool fbox = p.quantize == AR_QUANT_NONE &&
                      (p.precision == AR_PREC_BF16 || p.precision == AR_PREC_FP16 ||
                       p.precision == AR_PREC_FFT);
    const size_t box = fbox ? 4 * N * 8 + N * 4 : 2 * N * 8;
    const size_t copies = 8 * N + 4096;             // int8 copies of the UInt8 pair
    const size_t pts = (size_t)npts * 512;          // point fields, outputs, hand-off, grid planes
    // the largest level: chip_size_max (its rows by the aspect), the largest radius
    const int cx = (int)(p.chip_size_max < 512 ? p.chip_size_max : 512);
    const int cy = 2 * (int)std::llround((double)cx * p.chip_aspect / 2.0);
    const int r = (int)(p.search_radius_x > p.search_radius_y ? p.search_radius_x
                                                              : p.search_radius_y);
    const int rr = r < 256 ? r : 256;
    const bool pow2 = cx > 0 && cy > 0 && ((int64_t)cx * cy & ((int64_t)cx * cy - 1)) == 0;
    size_t tracker = 0;
    if (dtype == AR_C64) {
        tracker = (size_t)2048 * ar_coherence_scratch(cy, cx, rr, rr);
    } else if (p.rotation_count > 0) {
        tracker = (size_t)2048 * ar::gen_scratch(cy, cx, rr, rr).total;
    } else if ((p.quantize == AR_QUANT_NONE && p.precision == AR_PREC_EXACT) || !pow2) {
        const size_t slot = ar::f64_slot(cy, cx, rr, rr).bytes;
        const size_t half = (size_t)1 << 30;        // half of the tracker's staging
        const size_t want = (size_t)npts * slot;
        tracker = want <= half ? want : 2 * half;
    }
    // the quantiser's sequential-sum records (under 3 bytes an element); slot contexts prepare
    // their two images one after the other, so one set of preparation scratch
    const size_t qs = p.quantize == AR_QUANT_UINT8 ? 3 * N + ((size_t)1 << 20) : 0;
    return raw + work + prep + box + copies + pts + tracker + qs + ((size_t)64 << 20);
}

// Host copy split over up to 8 threads for large buffers: one thread's memcpy reaches only a
// fraction of the host's memory bandwidth, and the submitting thread waits for the copy.
static void par_copy(uint8_t* dst, const uint8_t* src, size_t n) {
    const size_t chunk = (size_t)8 << 20;
    int k = (int)((n + chunk - 1) / chunk);
    if (k > 8) k = 8;
    if (k <= 1) {
        std::memcpy(dst, src, n);
        return;
    }
    const size_t per = (n + k - 1) / k;
    std::vector<std::thread> ts;
    for (int t = 1; t < k; t++) {
        const size_t a = (size_t)t * per, b = a + per < n ? a + per : n;
        if (a < b) ts.emplace_back([=] { std::memcpy(dst + a, src + a, b - a); });
    }
    std::memcpy(dst, src, per < n ? per : n);
    for (std::thread& t : ts) t.join();
}

static void worker(ar_server* S, int s) {
    ar_context* ctx = S->ctx[s];
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        Request* r = nullptr;
        for (;;) {
            if (S->stop) return;
            if (!S->queue.empty()) {
                Request* h = S->queue.front();
                size_t others = 0;
                for (size_t t = 0; t < S->held.size(); t++)
                    if ((int)t != s) others += S->held[t];
                size_t need = h->footprint > S->held[s] ? h->footprint : S->held[s];
                if (others + need > (size_t)S->budget) {
                    // idle slots give their buffers back, then check again
                    for (size_t t = 0; t < S->held.size(); t++)
                        if ((int)t != s && !S->busy[t] && S->held[t] > 0) {
                            ar_context_release(S->ctx[t]);
                            S->held[t] = 0;
                        }
                    others = 0;
                    for (size_t t = 0; t < S->held.size(); t++)
                        if ((int)t != s) others += S->held[t];
                    // this slot's own buffers from a larger earlier pair may be what is in the
                    // way: it gives them back too and asks for the pair's footprint alone
                    if (others + need > (size_t)S->budget && S->held[s] > h->footprint) {
                        ar_context_release(ctx);
                        S->held[s] = 0;
                        need = h->footprint;
                    }
                }
                if (others + need <= (size_t)S->budget) {
                    r = h;
                    S->queue.pop_front();
                    S->held[s] = need;
                    break;
                }
            }
            S->cv_work.wait(lk);
        }
        S->busy[s] = true;
        S->running++;
        if (S->running > S->peak_running) S->peak_running = S->running;
        size_t tot = 0;
        for (size_t h : S->held) tot += h;
        if (tot > S->peak_held) S->peak_held = tot;
        r->slot = s;
        r->t_admit = Clock::now();
        lk.unlock();

        // the pair: prepare on the device, then the multi-chip search
        const int64_t n = r->nr * r->nc;
        r->dx.resize(n);
        r->dy.resize(n);
        r->corr.resize(n);
        r->cs.resize(n);
        r->interp.resize(n);
        const size_t N = (size_t)r->ny * r->nx, ib = N * elem_size(r->dtype);
        const uint8_t* base = r->stage;
        ar_status st = ar_set_images(ctx, base, base + ib, r->has_rmask ? base + 2 * ib : nullptr,
                                     r->has_smask ? base + 2 * ib + N : nullptr, r->ny, r->nx,
                                     r->dtype);
        if (st == AR_OK) {
            ar_points g;
            g.x = r->x.data();
            g.y = r->y.data();
            g.dx_prior = r->pdx.data();
            g.dy_prior = r->pdy.data();
            g.radius_x = r->rx.data();
            g.radius_y = r->ry.data();
            g.chip_size_x = nullptr;
            g.chip_size_y = nullptr;
            st = ar_run_multichip(ctx, &g, r->nr, r->nc, r->dx.data(), r->dy.data(),
                                  r->corr.data(), r->cs.data(), r->interp.data());
        }
        const size_t actual = ar_context_bytes(ctx);

        lk.lock();
        r->status = st;
        if (st != AR_OK) r->err = ar_last_error(ctx);
        // the inputs are no longer needed: the staging buffer goes back to the pool
        if (r->pinned >= 0) S->pinned[r->pinned].used = false;
        r->pinned = -1;
        r->stage = nullptr;
        std::vector<uint8_t>().swap(r->spill);
        r->t_done = Clock::now();
        r->done = true;
        S->held[s] = actual;
        S->busy[s] = false;
        S->running--;
        S->completed++;
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
}

extern "C" {

ar_status ar_server_create(const ar_params* p, int32_t device, int32_t slots, int64_t hbm_budget,
                           ar_server** out) {
    if (!out) return AR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (slots < 1 || slots > 64) return AR_ERR_INVALID_ARGUMENT;
    ar_server* S = new ar_server();
    if (p) S->params = *p;
    else ar_params_default(&S->params);
    S->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete S;
        return AR_ERR_NO_DEVICE;
    }
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) {
        delete S;
        return AR_ERR_DEVICE;
    }
    // default budget: 90 % of the memory free now
    S->budget = hbm_budget > 0 ? hbm_budget : (int64_t)(freeb / 10 * 9);
    for (int s = 0; s < slots; s++) {
        ar_context* c = nullptr;
        ar_status st = ar_context_create(&S->params, device, &c);
        if (st != AR_OK) {
            for (ar_context* k : S->ctx) ar_context_destroy(k);
            delete S;
            return st;
        }
        c->side_prep = false;
        S->ctx.push_back(c);
    }
    S->held.assign(slots, 0);
    S->busy.assign(slots, false);
    for (int s = 0; s < slots; s++) S->workers.emplace_back(worker, S, s);
    *out = S;
    return AR_OK;
}

void ar_server_destroy(ar_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (std::thread& t : S->workers) t.join();
    for (ar_context* c : S->ctx) ar_context_destroy(c);
    for (auto& kv : S->reqs) delete kv.second;   // queued and finished requests alike
    for (PinnedBuf& b : S->pinned) hipHostFree(b.ptr);
    delete S;
}

ar_status ar_server_submit(ar_server* S, const void* ref, const void* sec, const uint8_t* ref_valid,
                           const uint8_t* sec_valid, int64_t ny, int64_t nx, int32_t dtype,
                           const ar_points* grid, int64_t nr, int64_t nc, int64_t* ticket) {
    if (!S || !ref || !sec || !grid || !ticket) return AR_ERR_INVALID_ARGUMENT;
    if (ny < 1 || nx < 1 || nr < 0 || nc < 0) return AR_ERR_INVALID_ARGUMENT;
    const size_t es = elem_size(dtype);
    if (es == 0) return AR_ERR_UNSUPPORTED;
    if (!grid->x || !grid->y || !grid->radius_x || !grid->radius_y || !grid->dx_prior ||
        !grid->dy_prior)
        return AR_ERR_INVALID_ARGUMENT;
    Request* r = new Request();
    r->ny = ny;
    r->nx = nx;
    r->nr = nr;
    r->nc = nc;
    r->dtype = dtype;
    const size_t N = (size_t)ny * nx;
    const int64_t n = nr * nc;
    r->footprint = pair_footprint(S->params, ny, nx, dtype, n);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (r->footprint > (size_t)S->budget) {
            S->err = "pair footprint exceeds the server's device-memory budget";
            delete r;
            return AR_ERR_OUT_OF_MEMORY;
        }
    }
    r->has_rmask = ref_valid != nullptr;
    r->has_smask = sec_valid != nullptr;
    r->stage_bytes = 2 * N * es + 2 * N;
    {
        // smallest free page-locked buffer that fits, else a new one within the pool limit
        std::unique_lock<std::mutex> lk(S->mu);
        int best = -1;
        for (size_t k = 0; k < S->pinned.size(); k++)
            if (!S->pinned[k].used && S->pinned[k].cap >= r->stage_bytes &&
                (best < 0 || S->pinned[k].cap < S->pinned[best].cap))
                best = (int)k;
        if (best < 0 && S->pinned_total + r->stage_bytes <= S->pinned_limit) {
            S->pinned_total += r->stage_bytes;
            lk.unlock();
            void* p = nullptr;
            const bool ok = hipHostMalloc(&p, r->stage_bytes, hipHostMallocDefault) == hipSuccess;
            lk.lock();
            if (ok) {
                PinnedBuf b;
                b.ptr = p;
                b.cap = r->stage_bytes;
                S->pinned.push_back(b);
                best = (int)S->pinned.size() - 1;
            } else {
                S->pinned_total -= r->stage_bytes;
            }
        }
        if (best >= 0) {
            S->pinned[best].used = true;
            r->pinned = best;
            r->stage = (uint8_t*)S->pinned[best].ptr;
        }
    }
    if (!r->stage) {
        r->spill.resize(r->stage_bytes);
        r->stage = r->spill.data();
    }
    par_copy(r->stage, (const uint8_t*)ref, N * es);
    par_copy(r->stage + N * es, (const uint8_t*)sec, N * es);
    if (ref_valid) std::memcpy(r->stage + 2 * N * es, ref_valid, N);
    if (sec_valid) std::memcpy(r->stage + 2 * N * es + N, sec_valid, N);
    r->x.assign(grid->x, grid->x + n);
    r->y.assign(grid->y, grid->y + n);
    r->pdx.assign(grid->dx_prior, grid->dx_prior + n);
    r->pdy.assign(grid->dy_prior, grid->dy_prior + n);
    r->rx.assign(grid->radius_x, grid->radius_x + n);
    r->ry.assign(grid->radius_y, grid->radius_y + n);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        r->ticket = S->next_ticket++;
        S->reqs[r->ticket] = r;
        S->queue.push_back(r);
        *ticket = r->ticket;
    }
    S->cv_work.notify_all();
    return AR_OK;
}

ar_status ar_server_wait(ar_server* S, int64_t ticket, float* dx, float* dy, float* corr,
                         uint16_t* chip_size, uint8_t* interpolated, double* timing) {
    if (!S) return AR_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) {
        S->err = "unknown ticket";
        return AR_ERR_INVALID_ARGUMENT;
    }
    Request* r = it->second;
    S->cv_done.wait(lk, [&] { return r->done; });
    S->reqs.erase(it);
    lk.unlock();
    const size_t n = r->dx.size();
    if (r->status == AR_OK) {
        if (dx) std::memcpy(dx, r->dx.data(), n * 4);
        if (dy) std::memcpy(dy, r->dy.data(), n * 4);
        if (corr) std::memcpy(corr, r->corr.data(), n * 4);
        if (chip_size) std::memcpy(chip_size, r->cs.data(), n * 2);
        if (interpolated) std::memcpy(interpolated, r->interp.data(), n);
    }
    if (timing) {
        using ms = std::chrono::duration<double, std::milli>;
        timing[0] = ms(r->t_done - r->t_submit).count();   // latency
        timing[1] = ms(r->t_admit - r->t_submit).count();  // queueing
        timing[2] = ms(r->t_done - r->t_admit).count();    // service
        timing[3] = (double)r->slot;
    }
    const ar_status st = r->status;
    if (st != AR_OK) {
        std::lock_guard<std::mutex> g(S->mu);
        S->err = r->err;
    }
    delete r;
    return st;
}

ar_status ar_server_stats(ar_server* S, int64_t* stats) {
    if (!S || !stats) return AR_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(S->mu);
    size_t tot = 0;
    for (size_t h : S->held) tot += h;
    stats[0] = S->completed;
    stats[1] = (int64_t)S->queue.size();
    stats[2] = S->running;
    stats[3] = S->peak_running;
    stats[4] = (int64_t)tot;
    stats[5] = (int64_t)S->peak_held;
    stats[6] = S->budget;
    stats[7] = (int64_t)S->ctx.size();
    return AR_OK;
}

const char* ar_server_last_error(ar_server* S) {
    if (!S) return "null server";
    // a copy for the calling thread: the server's own message may change under another thread
    thread_local std::string msg;
    std::lock_guard<std::mutex> lk(S->mu);
    msg = S->err;
    return msg.c_str();
}

}  // extern "C"
