/*
 * autorift_hip -- C ABI of the MI355X (gfx950) dense ZNCC tracker.
 *
 * Drop-in boundary for AutoRIFT.jl's hot path. A Julia host keeps `autorift`, `init`,
 * `reinit!`, `autorift!`, the `Cache` lifecycle and `PointSet` (src/api.jl:30-295,
 * src/points.jl:50-74) and calls these entry points through `ccall`, the way src/plans.jl
 * calls FFTW (opaque handle, caller-owned buffers, status returns).
 *
 * Conventions
 *   - Images and gridded point fields are column-major (Julia layout); pixel coordinates are
 *     1-based Float64 as in src/points.jl.
 *   - The caller owns every host buffer for the duration of a call; the library never keeps a
 *     host pointer. The context owns its device memory.
 *   - No exception crosses the ABI: every entry point returns an ar_status and
 *     ar_last_error() gives the message.
 *   - A context is not thread-safe (one per task, like Cache, src/api.jl:27-28); distinct
 *     contexts may run concurrently, each on its own HIP stream.
 */
#ifndef AUTORIFT_HIP_H
#define AUTORIFT_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    AR_OK = 0,
    AR_ERR_INVALID_ARGUMENT = 1,   /* ArgumentError in the reference */
    AR_ERR_DIMENSION_MISMATCH = 2, /* DimensionMismatch, e.g. src/api.jl:120-123 */
    AR_ERR_OUT_OF_MEMORY = 3,      /* device allocation failed */
    AR_ERR_DEVICE = 4,             /* HIP runtime error */
    AR_ERR_UNSUPPORTED = 5,        /* element type or geometry the device path does not take */
    AR_ERR_NO_DEVICE = 6           /* no gfx950 device visible */
} ar_status;

/* AR_C64: ComplexF32 (single-look-complex radar), interleaved (re, im) float pairs. */
typedef enum { AR_U8 = 0, AR_F32 = 1, AR_F64 = 2, AR_I16 = 3, AR_C64 = 4 } ar_dtype;
/* AR_COHERENCE: complex coherence magnitude (src/types.jl:154-200), complex pairs only; ZNCC and
 * NCC need real pairs. */
typedef enum { AR_ZNCC = 0, AR_NCC = 1, AR_COHERENCE = 2 } ar_measure;
/* AR_PRE_DERAMP: Deramp(axis) (src/types.jl:383-423), complex pairs only; Highpass and Wallis
 * (src/types.jl:299-322, src/preprocess.jl:333-351) real only. */
typedef enum { AR_PRE_NONE = 0, AR_PRE_HIGHPASS = 1, AR_PRE_DERAMP = 2, AR_PRE_WALLIS = 3 } ar_preprocess;
typedef enum { AR_QUANT_NONE = 0, AR_QUANT_UINT8 = 1 } ar_quantize;
/* Arithmetic of the correlation contraction for Float32 prepared images (quantize = :none).
 * EXACT follows the reference's Float64 direct numerator (src/correlate.jl:799-814); BF16 and
 * FP16 run it on the matrix cores with 16-bit operands and FP32 accumulation, keeping Float64
 * for the box sums, the normalisation and the integer-peak / subpixel-patch values. FFT is the
 * reference's own dispatch (src/correlate.jl:785-797): above DIRECT_THRESHOLD = 20 000
 * multiply-adds per point the numerator is a single-precision transform
 * (_numerators_fft!, 816-848), below it the exact direct form; transforms larger than 128 per
 * side take the exact direct form too. UInt8 prepared images always take the exact integer
 * path. */
typedef enum { AR_PREC_EXACT = 0, AR_PREC_BF16 = 1, AR_PREC_FP16 = 2, AR_PREC_FFT = 3 } ar_precision;

/* Flattened Params (src/types.jl:781-829) with the defaults of params() (src/params.jl:306-392).
 * Built by the host from its resolved parameter set; the library never parses strings. */
typedef struct {
    int32_t preprocess;           /* ar_preprocess, default AR_PRE_HIGHPASS */
    int32_t filter_width;         /* Highpass / Wallis width, default 5 */
    int32_t quantize;             /* ar_quantize, default AR_QUANT_UINT8 */
    int32_t measure;              /* ar_measure, default AR_ZNCC */
    int64_t upsampling;           /* PyramidRefine(64); 1 means NoRefine */
    int32_t outliers;             /* 1 GardnerFilter, 0 NoOutlierFilter */
    int32_t _pad0;
    int64_t gardner_window;       /* 5 */
    int64_t gardner_iterations;   /* 3 */
    double min_agree_fraction;    /* 8/25 */
    double agree_tolerance;       /* 0.2 */
    double mad_scale;             /* 4.0 */
    int64_t chip_size_base;       /* 32 */
    int64_t chip_size_min;        /* = base */
    int64_t chip_size_max;        /* = 4 base */
    double chip_aspect;           /* 1.0 */
    int64_t grid_spacing;         /* = base */
    int64_t search_radius_x;      /* 25 */
    int64_t search_radius_y;      /* 25 */
    int64_t min_search_radius;    /* 6 */
    int64_t coarse_stride;        /* 4 */
    int64_t coarse_buffer;        /* 8 */
    double min_coarse_valid_fraction; /* 0.01 */
    double dx_prior;              /* 0.0 */
    double dy_prior;              /* 0.0 */
    int64_t fill_window;          /* 3 */
    int32_t precision;            /* ar_precision, default AR_PREC_EXACT (not in the reference) */
    int32_t rotation_count;       /* RotationSearch (src/types.jl:426-577): 0 = NoRotationSearch,
                                     else the number of chip rotations (at most 8) */
    double rotation_angles[8];    /* the chip rotations in degrees, `about` already subtracted
                                     (angles(m), src/types.jl:577) */
    int32_t deramp_axis;          /* Deramp axis: 0 both, 1 x only, 2 y only */
    int32_t level_measure_count;  /* the similarity tuple (src/types.jl:770-870): a measure per
                                     chip-size level, the last repeating; 0 = `measure` at every
                                     level (at most 8) */
    int32_t level_measures[8];
    double min_std;               /* Wallis divisor floor (min_std), default 0.0 */
} ar_params;

/* PointSet fields (src/points.jl:50-74), structure of arrays of equal length. */
typedef struct {
    const double* x;
    const double* y;
    const int64_t* radius_x;
    const int64_t* radius_y;
    const double* dx_prior;
    const double* dy_prior;
    const int64_t* chip_size_x;
    const int64_t* chip_size_y;
} ar_points;

typedef struct ar_context ar_context;

/* Fill `p` with the defaults of params() (src/params.jl:306-392). */
void ar_params_default(ar_params* p);

/* init (src/api.jl:66-76): opaque per-Cache state on `device` (HIP ordinal). */
ar_status ar_context_create(const ar_params* p, int32_t device, ar_context** out);
void ar_context_destroy(ar_context* ctx);

/* init / reinit! image preparation (src/api.jl:104-134, 305-311): uploads the raw pair, runs
 * preprocess (src/preprocess.jl:126-159, 304-312, 447-460) and quantize
 * (src/preprocess.jl:473-531, 546-592) on the device and keeps the prepared pair resident.
 * A NULL mask is derived from isfinite (src/preprocess.jl:74-79). A NULL image keeps the
 * current one of that slot; sizes must match the context's (DimensionMismatch otherwise). */
ar_status ar_set_images(ar_context* ctx, const void* ref, const void* sec,
                        const uint8_t* ref_valid, const uint8_t* sec_valid, int64_t ny,
                        int64_t nx, int32_t dtype);

/* An already prepared ImagePair (src/preprocess.jl:41-79), used as is: what track! sees.
 * dtype AR_U8 or AR_F32. */
ar_status ar_set_pair(ar_context* ctx, const void* ref, const void* sec,
                      const uint8_t* ref_valid, const uint8_t* sec_valid, int64_t ny, int64_t nx,
                      int32_t dtype);

/* Copies the prepared pair back (for inspection and tests). Element type is the prepared one:
 * u8 after AR_QUANT_UINT8, c64 for a complex pair, f32 otherwise. Any pointer may be NULL. */
ar_status ar_get_pair(ar_context* ctx, void* ref, void* sec, uint8_t* ref_valid,
                      uint8_t* sec_valid, int32_t* dtype);

/* The UInt8 quantiser's statistics of the prepared pair (src/preprocess.jl:546-592), for
 * inspection and tests: stats[8 i + 0..3] = valid count n, mean, Bessel standard deviation and
 * mode (0 nothing valid, 1 sd == 0, 2 mapped) of image i (0 reference, 1 secondary); stats[8 i +
 * 4..7] = for the sum of values and the sum of squared deviations, the number of runs the device
 * walked and the number of elements it added one by one. The mean and the sum of squared
 * deviations are the reference's left-to-right Float64 sums, bit for bit.
 * AR_ERR_INVALID_ARGUMENT unless the context holds a pair prepared with AR_QUANT_UINT8. */
ar_status ar_quantizer_stats(ar_context* ctx, double* stats);

/* x[0] + x[1] + ... + x[n-1] added left to right in Float64 -- the reference's sequential sums
 * (e.g. src/preprocess.jl:551-570, src/correlate.jl:435-447) -- computed on the context's device
 * by the library's parallel emulations, for inspection and tests: method 0 the chunk walk the
 * quantiser uses (any finite or non-finite terms; walk[0..1] = runs walked, elements added one
 * by one; walk may be NULL), method 1 the one-wave replay the trackers use for chip energies
 * (non-negative finite terms). No counterpart in the reference. */
ar_status ar_sequential_sum(ar_context* ctx, const double* x, int64_t n, int32_t method,
                            double* out, int64_t* walk);

/* track (src/track.jl:103-160) over n points with host buffers. upsampling 1 = integer peak
 * (NoRefine), else PyramidRefine(upsampling), a power of two <= 64. Writes dx, dy,
 * correlation (NaN where there is no measurement) and searched (0/1). */
ar_status ar_track(ar_context* ctx, const ar_points* pts, int64_t n, int64_t upsampling,
                   int32_t measure, float* dx, float* dy, float* corr, uint8_t* searched);

/* As ar_track, with every pointer (point fields and outputs) in device memory of the context's
 * device. Runs on the context's stream and returns after the stream has drained. */
ar_status ar_track_device(ar_context* ctx, const ar_points* dev_pts, int64_t n,
                          int64_t upsampling, int32_t measure, float* dev_dx, float* dev_dy,
                          float* dev_corr, uint8_t* dev_searched);

/* correlate_multichip (src/multichip.jl:106-137) over a gridded point set (nr x nc,
 * column-major fields; chip sizes are set per level, so chip_size_x/y may be NULL).
 * interpolated is returned as bytes (the host packs a BitMatrix). */
ar_status ar_run_multichip(ar_context* ctx, const ar_points* grid, int64_t nr, int64_t nc,
                           float* dx, float* dy, float* corr, uint16_t* chip_size,
                           uint8_t* interpolated);

/* Debug path: one chip against one window, as correlate! (src/correlate.jl:528-557) followed by
 * peak_offset / subpixel_peak (src/peak.jl:102-106, 456-522). search is
 * (ch + 2ry - 1) x (cw + 2rx - 1) and chip ch x cw, both column-major; ch and cw even.
 * dx, dy are NaN for a degenerate (constant) chip. Uses the context's device only. */
ar_status ar_correlate_point(ar_context* ctx, const void* search, const void* chip,
                             int32_t dtype, int64_t ch, int64_t cw, int64_t rx, int64_t ry,
                             int64_t upsampling, int32_t measure, double* dx, double* dy,
                             float* corr);

/* As ar_correlate_point, allocating no device memory: no counterpart in the reference. For callers
 * that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The window, chip, staging and outputs live in a library-owned 1 MiB device array
 * (one per GPU, kept until the process ends); a chip and window too large for it are refused
 * with AR_ERR_UNSUPPORTED. Every call on one GPU uses that GPU's array: calls that overlap on the
 * same GPU (from different threads or contexts) race and must be serialized by the caller. Uses
 * the context's device and stream only; the context's image pair is left untouched. */
ar_status ar_correlate_point_noalloc(ar_context* ctx, const void* search, const void* chip,
                                     int32_t dtype, int64_t ch, int64_t cw, int64_t rx,
                                     int64_t ry, int64_t upsampling, int32_t measure, double* dx,
                                     double* dy, float* corr);

/* Device timing of the last ar_track / ar_track_device pass, in milliseconds (HIP events on
 * the context's stream: correlation kernel plus refinement kernel), and the correlation path it
 * took: 0 u8 matrix-core path, 1 generic Float64, 2 BF16 / FP16 matrix-core path. */
ar_status ar_last_kernel(const ar_context* ctx, double* ms, int32_t* path);

/* The same pass split by kernel, in milliseconds: the box images of W and W^2 (0 when the pass
 * computes its box sums per point), the correlation kernel (integer peak, exact 5x5 patch) and
 * the refinement kernel (pyrUp cascade; 0 when upsampling is 1). No counterpart in the
 * reference, whose track loop is not timed per stage. */
ar_status ar_last_timing(const ar_context* ctx, double* box_ms, double* track_ms,
                         double* refine_ms);

/* Message of the last failed call on ctx (or of the last failed ar_context_create when ctx is
 * NULL). Valid until the next call on the same context. */
const char* ar_last_error(const ar_context* ctx);

/* Library version string. */
const char* ar_version(void);

/* ---- Batched passes (no counterpart in the reference, which tracks one pair per call).
 * A batch keeps many prepared UInt8 pairs resident (default pipeline: the batch's params must
 * quantise to UInt8) and runs one tracking pass over points of all of them, each point naming
 * its pair: track (src/track.jl:103-160) over every point, with uneven chip sizes and radii.
 * Power-of-two chip areas take the integer matrix-core tracker and equal a pass over each pair
 * alone, bit for bit; other chip areas take the same tracker with the unrounded mean-removed chip,
 * or, after ar_batch_set_exact(b, 1), the direct Float64 tracker (the reference's
 * Float32(x - mean) chip and Float64 sums, src/correlate.jl:423-453, 799-814), bit-equal too. */
typedef struct ar_batch ar_batch;
#define AR_BATCH_GRAPH 1      /* capture the pass in a HIP graph, replay it for the same buffers */
#define AR_BATCH_SPLIT_SET 2  /* bits 8..15 (R1) and 16..23 (R2) give the radius classes
                                 instead of the tuner's choice: one launch per class, max radius
                                 <= R1, <= R2, the rest (R1 = 0: one launch; R2 = 0: two) */
ar_status ar_batch_create(const ar_params* p, int32_t device, ar_batch** out);
void ar_batch_destroy(ar_batch* b);
/* Prepares a raw pair as ar_set_images does and keeps it in the batch; *index gets its pair
 * index. The images may be host or device memory. */
ar_status ar_batch_add_pair(ar_batch* b, const void* ref, const void* sec,
                            const uint8_t* ref_valid, const uint8_t* sec_valid, int64_t ny,
                            int64_t nx, int32_t dtype, int32_t* index);
/* One pass over n points (host buffers; pair[i] is point i's pair index). The radius classes
 * (one tracker launch each, LDS carved per class) are chosen by timing the candidates once per
 * pass shape. */
ar_status ar_batch_track(ar_batch* b, const ar_points* pts, const int32_t* pair, int64_t n,
                         int64_t upsampling, int32_t measure, float* dx, float* dy, float* corr,
                         uint8_t* searched);
/* As ar_batch_track with every pointer in device memory; flags: AR_BATCH_GRAPH,
 * AR_BATCH_SPLIT_SET | R1 << 8 | R2 << 16. A captured pass is replayed while the buffers and
 * the points' radii stay the same (the class sizes are fixed at capture). */
ar_status ar_batch_track_device(ar_batch* b, const ar_points* dev_pts, const int32_t* dev_pair,
                                int64_t n, int64_t upsampling, int32_t measure, float* dev_dx,
                                float* dev_dy, float* dev_corr, uint8_t* dev_searched,
                                int32_t flags);
/* Last pass: device time (ms); info = pairs, radius classes (R1 | R2 << 8), tracker launches,
 * replayed (0/1). */
ar_status ar_batch_info(const ar_batch* b, double* ms, int32_t* info);
const char* ar_batch_last_error(const ar_batch* b);
/* Chips whose area is not a power of two: exact = 1 sends them to the direct Float64 tracker
 * (every result bit-equal to a pass over its pair alone); exact = 0 (the default) keeps them on
 * the integer matrix-core tracker with the unrounded mean-removed chip, (n x - sum) / n, in place
 * of the reference's Float32(x - mean) chip: integer peaks agree with the exact pass to the
 * tolerance stated for the batched form, subpixel offsets to a small fraction of a pixel.
 * Power-of-two areas are bit-equal either way. */
ar_status ar_batch_set_exact(ar_batch* b, int32_t exact);

/* ---- Tuning cache: the batched pass's tuner records its choice of radius classes per pass
 * shape in a small per-device file, as the reference keeps its FFTW wisdom
 * (src/plans.jl:133-257). Directory: $AR_TUNE_CACHE_DIR, else $XDG_CACHE_HOME/autorift_hip,
 * else $HOME/.cache/autorift_hip, resolved on every call; AR_TUNE_CACHE=off disables it. The
 * file is named for `device` (e.g. "gfx950-256cu") and the library version. These calls never
 * fail their caller: a missing, unreadable or corrupt file reads as empty, and a write that
 * cannot be made (read-only directory, full disk) is dropped. Writes go to a temporary renamed
 * over the file. */
/* Path of the cache file for `device` (copied to buf, NUL-terminated, when len > 0); returns its
 * length, or 0 when caching is off or no directory is usable. */
int64_t ar_tune_cache_path(const char* device, char* buf, int64_t len);
/* 1 and *value when the file holds `key`, else 0. */
int32_t ar_tune_cache_lookup(const char* device, const char* key, int32_t* value);
/* Records key = value; 1 when the file was written, 0 when the write was dropped. */
int32_t ar_tune_cache_record(const char* device, const char* key, int32_t value);

/* ---- Serving (no counterpart in the reference, which correlates one pair per call).
 * A server owns `slots` contexts, each with its own worker thread and HIP stream. Pairs of any
 * size are submitted at any time and run as init + autorift! (src/api.jl:66-76, 165-170):
 * preparation, then the multi-chip search over the caller's grid. The head of the FIFO queue is
 * admitted to an idle slot once its device-memory footprint fits `hbm_budget` (bytes; <= 0:
 * 90 % of the memory free at creation) next to what the other slots hold; pairs on different
 * slots overlap on the device. Input buffers are copied at submission. */
typedef struct ar_server ar_server;
ar_status ar_server_create(const ar_params* p, int32_t device, int32_t slots, int64_t hbm_budget,
                           ar_server** out);
/* Stops the workers after the pairs they are running; pairs still queued are dropped. */
void ar_server_destroy(ar_server* s);
/* Queues one pair with its gridded point set (nr x nc, column-major fields as for
 * ar_run_multichip; chip sizes are set per level). AR_ERR_OUT_OF_MEMORY: the pair's footprint
 * alone exceeds the budget (admission refused). */
ar_status ar_server_submit(ar_server* s, const void* ref, const void* sec,
                           const uint8_t* ref_valid, const uint8_t* sec_valid, int64_t ny,
                           int64_t nx, int32_t dtype, const ar_points* grid, int64_t nr,
                           int64_t nc, int64_t* ticket);
/* Blocks until the pair of `ticket` is done and copies its MultichipResult planes out (any
 * pointer may be NULL). timing (4 doubles, may be NULL): latency from submission to completion,
 * time queued, time on its slot (ms), and the slot index. The ticket is consumed. */
ar_status ar_server_wait(ar_server* s, int64_t ticket, float* dx, float* dy, float* corr,
                         uint16_t* chip_size, uint8_t* interpolated, double* timing);
/* 8 counters: completed, queued, running, peak running, device bytes held by the slots, peak
 * bytes held, budget, slots. */
ar_status ar_server_stats(ar_server* s, int64_t* stats);
/* The server's last error message; the pointer stays valid in the calling thread until its next
 * call of this function. */
const char* ar_server_last_error(ar_server* s);

#ifdef __cplusplus
}
#endif
#endif
