/*
 * san_check.c -- drives the oracle's functions (smip_oracle.c) on small inputs,
 * for a build with AddressSanitizer and UndefinedBehaviorSanitizer.  Test
 * infrastructure (built and run by tests/test_sanitize.py), not product code.
 *
 * It covers the pipeline of smip-cli/main.cpp:242-376 with both bispectrum
 * types, power-of-two and other frame extents, every bispectrum lookup of a
 * small descriptor including the out-of-range ones (include/bispectrum.h:
 * 162-174,324-349,482-589), the transforms, the cross-correlation with its
 * displacement (include/crosscorrel.h:96-179), Array2::shifted beyond the frame
 * (include/array2.h:603-622), the window (include/window_function.h:69-84) and
 * the recovery spiral with a short output buffer (src/phasereco.cpp:17-36).
 * Exit status 0 when every call returned what is expected of it; the
 * sanitizers end the process on the first memory error or undefined operation.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "smip_oracle.h"

static int fails;
#define EXPECT(cond)                                                          \
    do {                                                                      \
        if (!(cond)) {                                                        \
            fprintf(stderr, "%s:%d: expected %s\n", __FILE__, __LINE__, #cond); \
            ++fails;                                                          \
        }                                                                     \
    } while (0)

/* nframes x nrows x ncols doubles: a star that moves by a pixel or two between
 * frames, over 8-bit-like noise (a linear congruential generator). */
static double* make_frames(int64_t n, int64_t nc, int64_t nr, uint32_t seed)
{
    double* f = malloc(sizeof(double) * (size_t)(n * nc * nr));
    uint32_t s = seed * 2654435761u + 1u;
    for (int64_t k = 0; k < n; ++k)
        for (int64_t y = 0; y < nr; ++y)
            for (int64_t x = 0; x < nc; ++x) {
                s = s * 1664525u + 1013904223u;
                const double dx = (double)(x - nc / 2 - k % 3 + 1);
                const double dy = (double)(y - nr / 2 + k % 2);
                f[(k * nr + y) * nc + x] = 200.0 * exp(-(dx * dx + dy * dy) / 4.0) + (double)(s >> 24) / 16.0;
            }
    return f;
}

static int all_finite(const double* a, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(a[i]))
            return 0;
    return 1;
}

static void pipeline_case(int64_t nc, int64_t nr, int64_t n, uint64_t b, double p, int f32)
{
    double* fr = make_frames(n, nc, nr, (uint32_t)(nc * 31 + nr));
    so_desc d;
    so_desc_init(&d, (uint64_t)nc, (uint64_t)nr, b, b);
    const size_t px = (size_t)(nc * nr);
    so_pipeline_out o;
    memset(&o, 0, sizeof o);
    o.bispectrum = calloc(d.base_size, f32 ? 2 * sizeof(float) : 2 * sizeof(double));
    o.powerspec = malloc(px * sizeof(double));
    o.sum_image = malloc(px * sizeof(double));
    o.shifts = malloc((size_t)n * 2 * sizeof(int32_t));
    o.phases = malloc(px * 2 * sizeof(double));
    o.flags = malloc(px);
    o.consistency = malloc(px * sizeof(double));
    o.image = malloc(px * 2 * sizeof(double));
    const int st = so_pipeline(fr, n, nc, nr, b, p, f32, &o);
    EXPECT(st == SO_OK);
    if (st == SO_OK) {
        EXPECT(o.stats.visits > 0 && o.stats.lookups > 0);
        EXPECT(all_finite(o.image, px * 2));
        EXPECT(all_finite(o.sum_image, px));
        EXPECT(all_finite(o.powerspec, px));
    }
    /* the same with every output left out */
    so_pipeline_out none;
    memset(&none, 0, sizeof none);
    EXPECT(so_pipeline(fr, n, nc, nr, b, p, f32, &none) == SO_OK);
    EXPECT(so_pipeline(fr, 0, nc, nr, b, p, f32, &none) == SO_INVALID_ARGUMENT);
    free(o.bispectrum);
    free(o.powerspec);
    free(o.sum_image);
    free(o.shifts);
    free(o.phases);
    free(o.flags);
    free(o.consistency);
    free(o.image);
    free(fr);
}

/* Every index 4-tuple within two of the descriptor's range: lookups either
 * resolve to a cell inside the array or report out_of_range. */
static void lookup_case(uint64_t nc, uint64_t nr, uint64_t b0, uint64_t b1)
{
    so_desc d;
    so_desc_init(&d, nc, nr, b0, b1);
    double* data = calloc(d.base_size, 2 * sizeof(double));
    float* data32 = calloc(d.base_size, 2 * sizeof(float));
    for (uint64_t a = 0; a < d.base_size; ++a) {
        data[2 * a] = (double)a;
        data[2 * a + 1] = -(double)a;
        data32[2 * a] = (float)a;
    }
    long ok = 0, oor = 0;
    int32_t idx[4];
    for (idx[0] = d.min_idx[0] - 2; idx[0] <= d.max_idx[0] + 2; ++idx[0])
        for (idx[1] = d.min_idx[1] - 2; idx[1] <= d.max_idx[1] + 2; ++idx[1])
            for (idx[2] = d.min_idx[2] - 2; idx[2] <= d.max_idx[2] + 2; ++idx[2])
                for (idx[3] = d.min_idx[3] - 2; idx[3] <= d.max_idx[3] + 2; ++idx[3]) {
                    uint64_t addr = 0;
                    int conj = 0, aliased = 0;
                    const int st = so_lookup(&d, idx, &addr, &conj, &aliased);
                    double v[2], w[2];
                    const int sg = so_get_element(&d, data, 0, idx, v);
                    const int sg32 = so_get_element(&d, data32, 1, idx, w);
                    EXPECT(sg == st && sg32 == st);
                    if (st == SO_OK) {
                        ++ok;
                        EXPECT(addr < d.base_size);
                        EXPECT(v[0] == (double)addr);
                    } else {
                        ++oor;
                        EXPECT(st == SO_OUT_OF_RANGE);
                    }
                    uint64_t off = so_calc_offset(&d, idx);
                    int32_t back[4];
                    so_calc_indices(&d, off % (d.totalsize ? d.totalsize : 1), back);
                }
    EXPECT(ok > 0 && oor > 0);
    /* put_element writes at the raw offset (with its mod-2^64 wrap) and only
     * when that offset is inside the array */
    const int32_t zero[4] = {0, 0, 0, 0};
    const double one[2] = {1.0, 0.0};
    EXPECT(so_put_element(&d, data, 0, zero, one) == SO_OK);
    const int32_t far[3][4] = {{d.max_idx[0] + 5, 0, 0, 0}, {0, d.max_idx[1] + 3, 0, 0}, {0, 0, 0, d.min_idx[3] - 4}};
    for (int k = 0; k < 3; ++k) {
        const int put_ok = so_put_element(&d, data32, 1, far[k], one) == SO_OK;
        EXPECT(put_ok == (so_calc_offset(&d, far[k]) < d.base_size));
    }
    free(data);
    free(data32);
}

static void transform_case(int64_t nr, int64_t nc)
{
    const size_t px = (size_t)(nr * nc);
    double* a = malloc(px * 2 * sizeof(double));
    double* a0 = malloc(px * 2 * sizeof(double));
    for (size_t i = 0; i < 2 * px; ++i)
        a0[i] = a[i] = sin(0.37 * (double)i) + 0.01 * (double)(i % 7);
    so_fft2_c2c(a, nr, nc, -1);
    so_fft2_c2c(a, nr, nc, +1);
    double e_raw = 0.0, e_div = 0.0;
    for (size_t i = 0; i < 2 * px; ++i) {
        e_raw = fmax(e_raw, fabs(a[i] - a0[i]));
        e_div = fmax(e_div, fabs(a[i] / (double)px - a0[i]));
    }
    EXPECT(fmin(e_raw, e_div) < 1e-9);
    double* re = malloc(px * sizeof(double));
    double* spec = malloc(px * 2 * sizeof(double));
    double* back = malloc(px * sizeof(double));
    for (size_t i = 0; i < px; ++i)
        re[i] = a0[2 * i];
    so_fft2_r2c(re, spec, nr, nc);
    so_fft2_c2r(spec, back, nr, nc);
    EXPECT(all_finite(back, px));
    free(a);
    free(a0);
    free(re);
    free(spec);
    free(back);
}

static void xcorr_case(int64_t nc, int64_t nr)
{
    const size_t px = (size_t)(nc * nr);
    double* fr = make_frames(1, nc, nr, 7u);
    double* moved = malloc(px * sizeof(double));
    double* corr = malloc(px * sizeof(double));
    EXPECT(so_shifted(fr, nc, nr, 3, -2, moved) == SO_OK);
    EXPECT(so_xcorr(fr, moved, nc, nr, corr) == SO_OK);
    int32_t sh[2];
    so_displacement(corr, nc, nr, sh);
    EXPECT(sh[0] >= -nc && sh[0] <= nc && sh[1] >= -nr && sh[1] <= nr);
    /* shifts of a frame extent or more leave nothing of the frame */
    const int sb = so_shifted(fr, nc, nr, (int32_t)nc + 3, -(int32_t)nr - 1, moved);
    EXPECT(sb >= SO_OK && sb <= SO_DOMAIN_ERROR);
    free(fr);
    free(moved);
    free(corr);
}

static void window_case(void)
{
    double* w = malloc(24 * 20 * sizeof(double));
    EXPECT(so_general_hamming(24, 20, 10.0, 0.5, w) == SO_OK);
    EXPECT(all_finite(w, 24 * 20));
    const int sn = so_general_hamming(24, 20, -1.0, 0.5, w);
    EXPECT(sn >= SO_OK && sn <= SO_DOMAIN_ERROR);
    free(w);
}

static void spiral_case(void)
{
    const int64_t need = so_spiral_visits(20.0, NULL, 0);
    EXPECT(need > 0);
    int32_t* ij = malloc((size_t)need * 2 * sizeof(int32_t));
    EXPECT(so_spiral_visits(20.0, ij, need / 2) == need); /* a short buffer is not overrun */
    EXPECT(so_spiral_visits(20.0, ij, need) == need);
    free(ij);
}

int main(void)
{
    pipeline_case(32, 32, 6, 8, 12.0, 1);
    pipeline_case(32, 32, 6, 8, 12.0, 0);
    pipeline_case(24, 20, 4, 6, 9.0, 0);
    pipeline_case(15, 17, 3, 5, 7.0, 1);
    pipeline_case(16, 16, 2, 15, 8.0, 0); /* bispectrum depth up to the frame extent */
    lookup_case(8, 8, 3, 3);
    lookup_case(9, 6, 4, 2);
    transform_case(8, 16);
    transform_case(12, 10);
    transform_case(7, 5);
    xcorr_case(16, 16);
    xcorr_case(20, 12);
    window_case();
    spiral_case();
    if (fails) {
        fprintf(stderr, "san_check: %d expectation(s) failed\n", fails);
        return 1;
    }
    printf("san_check: ok\n");
    return 0;
}
