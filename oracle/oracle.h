/*
 * CPU oracle for the dense ZNCC tracker: a plain-C restatement of the AutoRIFT.jl hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or calls this code;
 * only the test suite, the smoke check and the CPU baseline leg of the benchmark do, and there
 * only as the checker. Every function cites the reference file:line it restates
 * (paths relative to the AutoRIFT.jl checkout).
 *
 * Conventions: images and grids are column-major (element (i, j), 0-based, at i + j*ld), as in
 * Julia. Point coordinates are 1-based Float64, as in the reference. Arithmetic follows the
 * reference's operation order and is compiled without FP contraction.
 */
#ifndef AUTORIFT_ORACLE_H
#define AUTORIFT_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ORC_C64: ComplexF32, interleaved (re, im) pairs of float. */
enum { ORC_U8 = 0, ORC_F32 = 1, ORC_F64 = 2, ORC_I16 = 3, ORC_C64 = 4 };
enum { ORC_ZNCC = 0, ORC_NCC = 1, ORC_COHERENCE = 2 };
enum { ORC_NUM_DIRECT = 0, ORC_NUM_FFT = 1 };

int64_t orc_next_fft_size(int64_t n);

/* correlate! + prepare_chip!; returns 1 if the chip carried signal, 0 if degenerate. */
int orc_correlate(const void* search, int64_t ld_search, const void* chip, int64_t ld_chip,
                  int dtype, int64_t ch, int64_t cw, int64_t rx, int64_t ry, int measure,
                  int num_mode, float* surface /* (2ry) x (2rx), column-major */);

void orc_integral_both(const void* a, int64_t ld, int dtype, int64_t m, int64_t n,
                       double* s, double* s2);

void orc_peak_index(const float* surface, int64_t nr, int64_t nc, int64_t* row, int64_t* col);

void orc_pyrup(float* dst, const float* src, int64_t sh, int64_t sw);

void orc_subpixel_peak(const float* surface, int64_t nr, int64_t nc, int64_t rx, int64_t ry,
                       int64_t up, double* dx, double* dy, float* val);

/* track! over a flat point set. Writes dx, dy, corr (NaN when unmeasured) and searched. */
typedef struct {
    const double* x;
    const double* y;
    const int64_t* radius_x;
    const int64_t* radius_y;
    const double* dx_prior;
    const double* dy_prior;
    const int64_t* chip_size_x;
    const int64_t* chip_size_y;
} orc_points;

/* nrot > 0: RotationSearch over angles[0..nrot) (degrees, `about` already subtracted). */
int orc_track(const void* ref, const void* sec, const uint8_t* ok_mask, int dtype, int64_t ny,
              int64_t nx, const orc_points* pts, int64_t n, int64_t up, int measure,
              int num_mode, int nthreads, int nrot, const double* angles, float* dx, float* dy,
              float* corr, uint8_t* searched);
/* orc_track returns 0, or ORC_ERR_MEASURE when the measure does not suit the element type
 * (Coherence needs complex images, ZNCC and NCC real ones), or ORC_ERR_ROTATION for a rotation
 * search over complex images. */
enum { ORC_ERR_MEASURE = -2, ORC_ERR_ROTATION = -3 };

/* Complex (SLC) path: ramp_phase and deramp (src/preprocess.jl:184-245) over an nr x nc
 * ComplexF32 image (interleaved); axis 0 both, 1 x only, 2 y only. */
void orc_ramp_phase(const float* z, const uint8_t* mask, int64_t nr, int64_t nc, double* phase_x,
                    double* phase_y);
void orc_deramp(const float* z, const uint8_t* mask, int64_t nr, int64_t nc, int axis,
                float* out);

/* _rotate_chip: ch x cw chip (column-major, leading dimension ldc) rotated by deg degrees. */
void orc_rotate_chip(const void* chip, int64_t ldc, int dtype, int64_t ch, int64_t cw,
                     double deg, float* dst);

/* Pre-filter and quantizer. */
void orc_windowmean(const float* a, int64_t nr, int64_t nc, int64_t wx, int64_t wy, int hasnan,
                    float* out);
void orc_window_extremum(const float* a, int64_t nr, int64_t nc, int64_t wx, int64_t wy,
                         int is_max, float* out);
void orc_highpass_filtered(const void* img, const uint8_t* mask, int dtype, int64_t nr,
                           int64_t nc, int64_t w, float* out, uint8_t* out_mask);
void orc_wallis_filtered(const void* img, const uint8_t* mask, int dtype, int64_t nr, int64_t nc,
                         int64_t w, double min_std, float* out, uint8_t* vmask);
void orc_to_uint8(const float* img, const uint8_t* mask, int64_t nr, int64_t nc, uint8_t* out);
void orc_quant_stats(const double* v, const uint8_t* mask, int64_t N, double* out);
void orc_noquantize(const float* img, int64_t nr, int64_t nc, float* out);

/* Grid-level reductions and the outlier filter. */
void orc_count_agreeing(const float* a, int64_t nr, int64_t nc, int64_t w, float tol, float* out);
void orc_windowmedmad(const float* a, int64_t nr, int64_t nc, int64_t w, float* med, float* mad);
typedef struct {
    int64_t window;
    int64_t iterations;
    double min_agree_fraction;
    double agree_tolerance;
    double mad_scale;
} orc_gardner;
void orc_reject_gardner(const float* dx, const float* dy, const int64_t* rx, const int64_t* ry,
                        const uint8_t* valid, int64_t nr, int64_t nc, int64_t up,
                        const orc_gardner* f, uint8_t* keep);
void orc_sqdist_transform(const uint8_t* mask, int64_t nr, int64_t nc, double* d2);
void orc_dilate_within(const uint8_t* mask, int64_t nr, int64_t nc, double radius, uint8_t* out);
void orc_resample_nearest(const float* a, int64_t sr, int64_t sc, int64_t dr, int64_t dc,
                          float* out);
int64_t orc_fill_holes(float* dx, float* dy, int64_t nr, int64_t nc, int64_t w,
                       int64_t* filled /* capacity nr*nc */);

/* correlate_multichip over a gridded point set (nr x nc, column-major fields). */
typedef struct {
    int64_t chip_size_base, chip_size_min, chip_size_max;
    double chip_aspect;
    int64_t min_search_radius, coarse_stride, coarse_buffer;
    double min_coarse_valid_fraction;
    int64_t fill_window;
    int64_t upsampling;
    int measure;
    int outliers_on;
    orc_gardner gardner;
    int nrot;               /* RotationSearch angles (0: NoRotationSearch) */
    double angles[8];
    int nmeas;              /* a measure per chip-size level, the last repeating (measure_at,
                               src/types.jl:842-870); 0: `measure` at every level */
    int measures[8];
} orc_params;
/* Steps of the multi-chip loop, exported so the tests can pin each one on its own. */
void orc_cell_max_radius(const int64_t* radius, int64_t nr, int64_t nc, int64_t stride,
                         int64_t* out);
int orc_coarse_fraction_ok(const uint8_t* keep, const uint8_t* measured, const int64_t* radius_x,
                           int64_t n, double min_fraction);
int orc_coarse_pass(const void* ref, const void* sec, const uint8_t* ok, int dt, int64_t ny,
                    int64_t nx, const orc_points* lv, int64_t nr, int64_t nc, int64_t csx,
                    int64_t csy, const orc_params* p, int num_mode, int nthreads,
                    uint8_t* mask);
void orc_merge_level(float* rdx, float* rdy, float* rcorr, uint16_t* rcs, uint8_t* rinterp,
                     const float* fdx, const float* fdy, const float* fcorr,
                     const int64_t* filled, int64_t nfilled, int64_t n, int64_t chip_size,
                     uint8_t* wasfilled);
int orc_multichip(const void* ref, const void* sec, const uint8_t* ok_mask, int dtype,
                  int64_t ny, int64_t nx, const orc_points* grid, int64_t nr, int64_t nc,
                  const orc_params* p, int num_mode, int nthreads, float* dx, float* dy,
                  float* corr, uint16_t* chip_size, uint8_t* interpolated);

#ifdef __cplusplus
}
#endif
#endif
