/*
 * autorift -- dense normalized cross-correlation feature tracking on the MI355X library.
 *
 * The raw-plane command line of the reference's standalone binary (app/src/AutoRIFTApp.jl:38-170):
 *
 *   autorift <reference> <secondary> <ny> <nx> <output> [chip_size] [search_radius]
 *
 * Inputs are headerless raw Float32, column-major, ny * nx * 4 bytes each. The output is three
 * raw Float32 planes, dx, dy and correlation, each column-major over the output grid (the grid
 * of params(): chip sizes chip_size .. 4 chip_size, spacing chip_size, the margin of the largest
 * chip plus the search radius, src/api.jl:339-347 and src/points.jl:231-258). The pipeline is the
 * app's parameter set (app/src/AutoRIFTApp.jl:110-122): ZNCC, high-pass, UInt8 quantisation,
 * 64x pyramid refinement, Gardner outlier filter, hole filling. Argument errors print a message
 * and exit with status 2, as the reference's does; a failed run exits with status 1.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/autorift_hip.h"

static const char* USAGE =
    "autorift - dense normalized cross-correlation feature tracking\n"
    "\n"
    "  autorift <reference> <secondary> <ny> <nx> <output> [chip_size] [search_radius]\n"
    "\n"
    "  <reference>, <secondary>  headerless raw Float32, column-major, ny*nx*4 bytes\n"
    "  <ny> <nx>                 image shape in pixels\n"
    "  <output>                  written as three raw Float32 planes: dx, dy, correlation\n"
    "  [chip_size]               finest chip size, multiple of 4 (default 32)\n"
    "  [search_radius]           search half-extent in pixels (default 25)\n";

/* Whole-string positive integer, or -1. */
static long long parse_int(const char* s) {
    char* end = NULL;
    long long v = strtoll(s, &end, 10);
    if (end == s || *end != '\0') return -1;
    return v;
}

static float* read_image(const char* path, long long ny, long long nx) {
    const size_t n = (size_t)ny * (size_t)nx;
    float* a = (float*)malloc(n * sizeof(float));
    FILE* f = fopen(path, "rb");
    if (!a || !f) {
        fprintf(stderr, "error: cannot read %s\n", path);
        if (f) fclose(f);
        free(a);
        return NULL;
    }
    const size_t got = fread(a, sizeof(float), n, f);
    fclose(f);
    if (got != n) {
        fprintf(stderr, "error: %s holds %zu of the %zu Float32 values ny*nx needs\n", path, got, n);
        free(a);
        return NULL;
    }
    return a;
}

int main(int argc, char** argv) {
    if (argc < 6 || argc > 8) {
        fputs(USAGE, stdout);
        return 2;
    }
    const long long ny = parse_int(argv[3]), nx = parse_int(argv[4]);
    if (ny < 1 || nx < 1) {
        fputs("error: <ny> and <nx> must be positive integers\n", stdout);
        return 2;
    }
    long long chip = 32, radius = 25;
    if (argc >= 7) {
        chip = parse_int(argv[6]);
        if (chip < 4 || chip % 4 != 0) {
            fputs("error: [chip_size] must be a positive multiple of 4\n", stdout);
            return 2;
        }
    }
    if (argc >= 8) {
        radius = parse_int(argv[7]);
        if (radius < 1) {
            fputs("error: [search_radius] must be a positive integer\n", stdout);
            return 2;
        }
    }

    /* params() for the app's keywords */
    ar_params p;
    ar_params_default(&p);
    p.chip_size_base = chip;
    p.chip_size_min = chip;
    p.chip_size_max = 4 * chip;
    p.grid_spacing = chip;
    p.search_radius_x = radius;
    p.search_radius_y = radius;

    /* the grid of _build_grid: centres (m + 1):spacing:(n - m), m = cld(cs_max, 2) + r + 1 */
    const long long cmax = p.chip_size_max;
    const long long m = (cmax + 1) / 2 + radius + 1;
    long long nc = 0, nr = 0;
    for (long long x = m + 1; x <= nx - m; x += chip) nc++;
    for (long long y = m + 1; y <= ny - m; y += chip) nr++;
    if (nc == 0 || nr == 0) {
        fprintf(stdout,
                "error: no search center fits in a %lldx%lld image with chip %lldx%lld and search "
                "radius %lldx%lld: a margin of %lldx%lld pixels is needed on each side\n",
                ny, nx, cmax, cmax, radius, radius, m, m);
        return 1;
    }
    float* ref = read_image(argv[1], ny, nx);
    float* sec = ref ? read_image(argv[2], ny, nx) : NULL;
    if (!ref || !sec) {
        free(ref);
        return 1;
    }
    const size_t n = (size_t)nr * (size_t)nc;
    double* X = (double*)malloc(n * sizeof(double));
    double* Y = (double*)malloc(n * sizeof(double));
    double* zero = (double*)calloc(n, sizeof(double));
    int64_t* R = (int64_t*)malloc(n * sizeof(int64_t));
    float* out = (float*)malloc(3 * n * sizeof(float));
    uint16_t* cs = (uint16_t*)malloc(n * sizeof(uint16_t));
    uint8_t* interp = (uint8_t*)malloc(n);
    if (!X || !Y || !zero || !R || !out || !cs || !interp) {
        fputs("error: out of host memory\n", stderr);
        return 1;
    }
    for (long long j = 0; j < nc; j++)
        for (long long i = 0; i < nr; i++) {
            const size_t k = (size_t)i + (size_t)j * (size_t)nr;   /* column-major */
            X[k] = (double)(m + 1 + j * chip);
            Y[k] = (double)(m + 1 + i * chip);
            R[k] = radius;
        }
    ar_points g = {X, Y, R, R, zero, zero, NULL, NULL};

    ar_context* ctx = NULL;
    ar_status st = ar_context_create(&p, 0, &ctx);
    if (st == AR_OK) st = ar_set_images(ctx, ref, sec, NULL, NULL, ny, nx, AR_F32);
    if (st == AR_OK)
        st = ar_run_multichip(ctx, &g, nr, nc, out, out + n, out + 2 * n, cs, interp);
    if (st != AR_OK) {
        fprintf(stderr, "error: %s (status %d)\n", ctx ? ar_last_error(ctx) : ar_last_error(NULL),
                (int)st);
        ar_context_destroy(ctx);
        return 1;
    }
    ar_context_destroy(ctx);
    FILE* f = fopen(argv[5], "wb");
    if (!f || fwrite(out, sizeof(float), 3 * n, f) != 3 * n) {
        fprintf(stderr, "error: cannot write %s\n", argv[5]);
        if (f) fclose(f);
        return 1;
    }
    fclose(f);
    free(ref); free(sec); free(X); free(Y); free(zero); free(R); free(out); free(cs); free(interp);
    return 0;
}
