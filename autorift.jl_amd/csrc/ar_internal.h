// Host-side state behind the opaque ar_context handle.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/autorift_hip.h"

struct ar_context;

// Dynamic LDS any one kernel of the library may take: a CU's 160 KiB less room for the kernels'
// static shared arrays. Every dispatcher sizes its layouts against this one figure.
#define AR_LDS_LIMIT ((size_t)160 * 1024 - 2048)

// A growable device allocation owned by a context.
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    ar_status reserve(ar_context* ctx, size_t n);
    void release();
};

struct ar_context {
    ar_params params;
    int device = 0;
    int ncu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evm = nullptr, evb = nullptr;
    // the direct Float64 tracker's second stream (preparation of the next chunk of points) and
    // the events that order it with `stream`; made on first use
    hipStream_t stream2 = nullptr;
    hipEvent_t evf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    std::string err;

    // prepared pair (what the tracker correlates), column-major ny x nx
    bool has_pair = false;
    int64_t ny = 0, nx = 0;
    int32_t pdtype = AR_U8;
    DevBuf pref, psec;          // prepared images
    DevBuf rmask, smask;        // prepared validity masks
    DevBuf okmask;              // rmask & smask
    bool all_valid = true;

    // raw images of the current pair (preparation inputs)
    DevBuf raw, rawmask, work_f32, work_f64;
    // the secondary's own preparation scratch when the two images are prepared side by side
    // (pairs of at most kSidePrepMax elements, on `stream2`). A server's slot contexts clear
    // it: they prepare one image after the other and stage a grid on one host thread, since
    // the slots already run side by side (more streams than the process's hardware queues would
    // share queues, and more host threads than cores would contend).
    bool side_prep = true;
    DevBuf raw2, rawmask2, work_f32_2, work_f64_2, flags2, qsum2;

    // scratch
    DevBuf flags, geom, pts, out, scratch, grid_a, grid_b;
    DevBuf refine;              // per-point refinement hand-off: 5x5 patches and origins
    DevBuf box;                 // box images of W and W^2 for a pass with one chip size
    // what the integer box images in `box` were made for (chip, padding, pair version, buffer):
    // a pass with the same key reuses them (a multi-chip level's coarse and fine passes)
    struct {
        int ch = 0, cw = 0;
        int64_t pady = -1, padx = -1;
        uint64_t gen = ~0ULL;
        void* ptr = nullptr;
    } box_key;
    DevBuf refx;                // int8 copies (value ^ 0x80) of the UInt8 pair, for the matrix cores
    DevBuf prec;                // per-point geometry records of the current pass (u8 path)
    DevBuf mc;                  // multi-chip driver planes (grid, level, coarse and result planes)
    DevBuf qsum;                // the quantiser's sequential-sum chunks and segment records
    DevBuf qstats;              // the quantiser's parameters of the two prepared images
    // page-locked host staging for the multi-chip driver's uploads, results and readbacks
    void* pin = nullptr;
    size_t pin_cap = 0;
    uint64_t ref_gen = 0, refx_gen = ~0ULL;   // versions of the pair and of the copies made from it
    float absmax = 0.0f;        // largest |value| of the prepared Float32 pair (FP16 operand scale)
    uint64_t absmax_gen = ~0ULL;

    // workgroups per CU by (kernel, threads per workgroup, dynamic LDS bytes)
    std::map<std::tuple<const void*, int, size_t>, int> occupancy;

    double last_ms = 0, last_ms_box = 0, last_ms_track = 0, last_ms_refine = 0;
    // a pass that returned without waiting (multi-chip levels): its times are read from the
    // events when asked for
    bool timing_pending = false, pending_box = false, pending_refine = false;
    int32_t last_path = -1;

    std::vector<DevBuf*> all_bufs() {
        return {&pref, &psec, &rmask, &smask, &okmask, &raw, &rawmask, &work_f32, &work_f64,
                &flags, &geom, &pts, &out, &scratch, &grid_a, &grid_b, &refine, &box, &refx, &prec,
                &mc, &qsum, &qstats, &raw2, &rawmask2, &work_f32_2, &work_f64_2, &flags2,
                &qsum2};
    }
};

// Pairs of at most this many elements per image prepare their two images side by side (the
// secondary on the context's second stream, with its own scratch): at such sizes each image's
// preparation is a chain of latency-bound launches.
constexpr int64_t kSidePrepMax = (int64_t)1 << 24;

// The context's second stream and its ordering events (evf), made on first use.
ar_status ar_stream2(ar_context* ctx);

// internal entry points shared between translation units
// Complex pairs: preparation of one image (isfinite / given mask, Deramp, NoQuantize) and the
// Coherence tracker's per-workgroup scratch bytes.
ar_status ar_prepare_complex(ar_context* ctx, const void* img, const uint8_t* mask, int64_t ny,
                             int64_t nx, void* dst, uint8_t* dmask);
size_t ar_coherence_scratch(int ch, int cw, int rx, int ry);
size_t ar_coherence_lds(int ch, int cw, int rx, int ry);
ar_status ar_upload_masks(ar_context* ctx, const uint8_t* ref_valid, const uint8_t* sec_valid,
                          int64_t N);
// Pass geometry a caller already knows (an upper bound is enough: the kernels' layouts only need
// to cover every point, and a wider virtual padding adds invalid zeros that change no result):
// largest chip extents, radii and padding reach, whether every window lies inside the image.
struct ar_pass_hint {
    int64_t cx, cy, rx, ry, px, py;
    bool fits;
};
// With a hint the pass also returns without waiting for the stream (the caller's next readback
// orders it); its kernel times are resolved by ar_resolve_timing.
void ar_resolve_timing(ar_context* ctx);
// searchable_max: an upper bound on the pass's searchable points when the caller knows one
// (-1: every point may be searched); it only informs the choice of box images.
// reuse_box: the pass may read box images an earlier pass made from the same pair, chip and
// padding (a multi-chip level's coarse and fine passes); otherwise it builds its own.
ar_status ar_track_pass(ar_context* ctx, const ar_points& dp, int64_t n, int64_t up,
                        int32_t measure, float* dx, float* dy, float* corr, uint8_t* srch,
                        const ar_pass_hint* hint = nullptr, int64_t searchable_max = -1,
                        bool reuse_box = false);
ar_status ar_box_images(ar_context* ctx, int ch, int cw, int64_t pady, int64_t padx, bool reuse,
                        int32_t** b1, int32_t** b2, int64_t* ld);
// Page-locked host staging of at least n bytes (grow-only; the stream is idle when it grows).
ar_status ar_pinned(ar_context* ctx, size_t n, void** out);
// device bytes a context holds, and giving them all back (serving: idle slots make room)
size_t ar_context_bytes(ar_context* ctx);
void ar_context_release(ar_context* ctx);
namespace ar { struct TrackArgs; }
// mixed-precision pass set-up (track_mp.hip): operand scales, Float64 box images, kernel choice
ar_status ar_mp_setup(ar_context* ctx, ar::TrackArgs& a, void (**kern)(ar::TrackArgs),
                      size_t* lds);
// Float64 box images of the pass's chip size (mixed-precision and FFT paths)
ar_status ar_f64_box_images(ar_context* ctx, ar::TrackArgs& a);
// FFT path set-up (track_fft.hip): transform sizes, box images, kernel choice; AR_ERR_UNSUPPORTED
// when the transform does not fit the kernel's LDS
ar_status ar_fft_setup(ar_context* ctx, ar::TrackArgs& a, void (**kern)(ar::TrackArgs),
                       size_t* lds);
// threads per workgroup of the FFT tracker
#define AR_FFT_WG 512
// next_fft_size (src/correlate.jl:357-403)
int64_t ar_next_fft_size(int64_t n);
ar_status ar_upload_points(ar_context* ctx, const ar_points* pts, int64_t n, ar_points* dp,
                           int64_t csx_fill, int64_t csy_fill);
