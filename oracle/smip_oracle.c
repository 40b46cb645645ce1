/*
 * smip_oracle.c -- FP64 CPU restatement of the SMIP speckle-masking hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke check
 * and bench.py's cpu_baseline leg load it; the product library never does.
 *
 * Every function cites the reference (hangeza/speckle-stacker 0.3.2) file:line it
 * restates.  The restatement keeps the reference's arithmetic order wherever the
 * result depends on it (complex products, the calc_phase mean, the spiral) and is
 * compiled with -ffp-contract=off so no FMA contraction changes a rounding.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   descriptor / calc_offset / calc_indices / get_element / put_element:
 *       the reference's own tests (tests/bispectrum_test.cpp:27-79,132-149),
 *       a brute-force symmetry check of every canonical lookup against direct
 *       triple products of a known spectrum (aliased lookups, which the
 *       reference's index arithmetic produces, are counted and skipped);
 *   accumulate: delta frame -> B == 1 on every written cell, translation
 *       invariance, brute force over an explicit (u,v) enumeration on 8x8 and
 *       6x5 grids, two-delta closed form;
 *   FFT: brute-force DFT and numpy.fft on random inputs, Parseval;
 *   NextRecoIndex / spiral: first visits and visit counts per radius;
 *   reconstruct_phases: delta object (all flags in the disc, unit phases),
 *       noise-free single frame of a point pair (true phases up to tilt),
 *       lookup and aliasing counts;
 *   xcorr / displacement / shifted: brute-force circular correlation, exact
 *       recovery of integer shifts, tests/array2_test.cpp:329-345;
 *   window: closed form; pipeline: composition of the above.
 *
 * Parity unpinned (no reference output can be produced here; each is fixed
 * only by restating the reference's arithmetic):
 *   get_element on aliased (T3/T9) lookups: the values follow calc_offset's
 *       verbatim restatement, with no independent check;
 *   NextRecoIndex: the visit order beyond its first visits and per-radius
 *       counts depends on libm cos/sin;
 *   displacement: maxima that tie to the last bit of the FFT rounding (FFTW is
 *       absent); the synthetic frames avoid near-ties.
 */
#include "smip_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SO_2PI (3.14159265358979323846 * 2.0) /* constants.h:16, pi<double>*2 */
#define SO_EPS 2.220446049250313e-16          /* constants.h:20, DBL_EPSILON */

static inline int64_t so_imin(int64_t a, int64_t b) { return a < b ? a : b; }
static inline int64_t so_imax(int64_t a, int64_t b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------- */
/* Descriptor: include/bispectrum.h:143-160 (compute_descriptor),            */
/* :203-211 (sizes), :213-220 (base_sizes)                                   */
/* ------------------------------------------------------------------------- */
void so_desc_init(so_desc* d, uint64_t ncols, uint64_t nrows, uint64_t b0, uint64_t b1)
{
    const uint64_t dims[4] = { ncols, nrows, b0, b1 };
    memset(d, 0, sizeof(*d));
    uint64_t base = 1, total = 1;
    for (int k = 0; k < 4; ++k) {
        d->dims[k] = dims[k];
        d->sizes[k] = (dims[k] / 2) * 2 + 1;
        /* sizes * {1,0,1,0} / 2 is subtracted from the x dimensions only */
        d->base_sizes[k] = (k == 0 || k == 2) ? d->sizes[k] - d->sizes[k] / 2 : d->sizes[k];
        base *= d->base_sizes[k];
        total *= d->sizes[k];
        /* -static_cast<int>(x) / 2 truncates toward zero */
        d->min_idx[k] = (int32_t)(-(int32_t)d->sizes[k] / 2);
        d->max_idx[k] = (int32_t)d->sizes[k] + d->min_idx[k] - 1;
    }
    d->base_size = base;
    d->totalsize = total;
}

/* Empty (default-constructed) bispectrum: every descriptor field is zero. */
static void so_desc_empty(so_desc* d) { memset(d, 0, sizeof(*d)); }

/* calc_offset, include/bispectrum.h:324-349.  Indices are negated in the x
 * dimensions, negative values get the full size added (int result), and the
 * address is built in size_t arithmetic, i.e. modulo 2^64, with no range check. */
uint64_t so_calc_offset(const so_desc* d, const int32_t idx[4])
{
    int32_t a[4];
    const int sign[4] = { -1, 1, -1, 1 };
    for (int k = 0; k < 4; ++k) {
        int32_t v = idx[k] * sign[k];
        a[k] = v < 0 ? (int32_t)((uint64_t)(int64_t)v + d->sizes[k]) : v;
    }
    uint64_t addr = (uint64_t)(int64_t)a[0];
    addr *= d->base_sizes[1];
    addr += (uint64_t)(int64_t)a[1];
    addr *= d->base_sizes[2];
    addr += (uint64_t)(int64_t)a[2];
    addr *= d->base_sizes[3];
    addr += (uint64_t)(int64_t)a[3];
    return addr;
}

/* calc_indices, include/bispectrum.h:280-316 */
void so_calc_indices(const so_desc* d, uint64_t addr, int32_t idx[4])
{
    uint64_t temp = d->base_size / d->base_sizes[0];
    uint64_t rest = addr;
    uint64_t a0 = rest / temp;
    rest -= temp * a0;
    temp = d->base_sizes[2] * d->base_sizes[3];
    uint64_t a1 = rest / temp;
    rest -= temp * a1;
    temp = d->base_sizes[3];
    uint64_t a2 = rest / temp;
    rest -= temp * a2;
    int32_t v[4] = { (int32_t)a0, (int32_t)a1, (int32_t)a2, (int32_t)rest };
    const int sign[4] = { -1, 1, -1, 1 };
    for (int k = 0; k < 4; ++k) {
        int32_t half = (int32_t)d->sizes[k] / 2;
        if (v[k] > half)
            v[k] = (int32_t)((uint64_t)(int64_t)v[k] - d->sizes[k]);
        idx[k] = v[k] * sign[k];
    }
}

/* classify_indices + canonicalize_indices, include/bispectrum.h:162-174,553-589 */
static void so_canonicalize(const int32_t in[4], int32_t out[4], int* conj)
{
    const int32_t ux = in[0], uy = in[1], vx = in[2], vy = in[3];
    *conj = 0;
    if (ux <= 0 && vx <= 0) { /* T1 */
        out[0] = ux; out[1] = uy; out[2] = vx; out[3] = vy;
    } else if (ux > 0 && vx > 0) { /* T7 */
        *conj = 1;
        out[0] = -ux; out[1] = -uy; out[2] = -vx; out[3] = -vy;
    } else if (ux > 0 && vx <= 0) {
        if (ux + vx > 0) { /* T6 */
            out[0] = -ux - vx; out[1] = -uy - vy; out[2] = vx; out[3] = vy;
        } else { /* T9 */
            *conj = 1;
            out[0] = -ux; out[1] = -uy; out[2] = vx + ux; out[3] = vy + uy;
        }
    } else { /* ux <= 0 && vx > 0 */
        if (ux + vx > 0) { /* T3 */
            out[0] = ux; out[1] = uy; out[2] = -vx - ux; out[3] = -vy - uy;
        } else { /* T12 */
            *conj = 1;
            out[0] = ux + vx; out[1] = uy + vy; out[2] = -vx; out[3] = -vy;
        }
    }
}

static inline int32_t so_iabs(int32_t x) { return x < 0 ? -x : x; }

/* get_element's address resolution, include/bispectrum.h:482-507 */
int so_lookup(const so_desc* d, const int32_t idx_in[4], uint64_t* addr, int* conj, int* aliased)
{
    int32_t idx[4] = { idx_in[0], idx_in[1], idx_in[2], idx_in[3] };
    const int32_t* mx = d->max_idx;
    if (so_iabs(idx[2]) > mx[2] || so_iabs(idx[3]) > mx[3]) {
        int32_t t = idx[0]; idx[0] = idx[2]; idx[2] = t;
        t = idx[1]; idx[1] = idx[3]; idx[3] = t;
    }
    for (int k = 0; k < 4; ++k) /* Range::contains is inclusive, range.h:112-126 */
        if (idx[k] < d->min_idx[k] || idx[k] > d->max_idx[k])
            return SO_OUT_OF_RANGE;
    int32_t uv[4];
    so_canonicalize(idx, uv, conj);
    if (so_iabs(uv[2]) > mx[2] || so_iabs(uv[3]) > mx[3]) {
        int32_t t = uv[0]; uv[0] = uv[2]; uv[2] = t;
        t = uv[1]; uv[1] = uv[3]; uv[3] = t;
    }
    int alias = 0;
    for (int k = 0; k < 4; ++k)
        if (uv[k] < d->min_idx[k] || uv[k] > d->max_idx[k] || ((k == 0 || k == 2) && uv[k] > 0))
            alias = 1;
    if (aliased)
        *aliased = alias;
    *addr = so_calc_offset(d, uv);
    if (*addr >= d->base_size)
        return SO_OUT_OF_RANGE;
    return SO_OK;
}

int so_get_element(const so_desc* d, const void* data, int is_f32, const int32_t idx[4], double out[2])
{
    uint64_t addr;
    int conj, alias;
    int rc = so_lookup(d, idx, &addr, &conj, &alias);
    if (rc)
        return rc;
    double re, im;
    if (is_f32) {
        re = ((const float*)data)[2 * addr];
        im = ((const float*)data)[2 * addr + 1];
    } else {
        re = ((const double*)data)[2 * addr];
        im = ((const double*)data)[2 * addr + 1];
    }
    out[0] = re;
    out[1] = conj ? -im : im;
    return SO_OK;
}

/* put_element, include/bispectrum.h:525-535: no canonicalisation. */
int so_put_element(const so_desc* d, void* data, int is_f32, const int32_t idx[4], const double val[2])
{
    uint64_t addr = so_calc_offset(d, idx);
    if (addr >= d->base_size)
        return SO_OUT_OF_RANGE;
    if (is_f32) {
        ((float*)data)[2 * addr] = (float)val[0];
        ((float*)data)[2 * addr + 1] = (float)val[1];
    } else {
        ((double*)data)[2 * addr] = val[0];
        ((double*)data)[2 * addr + 1] = val[1];
    }
    return SO_OK;
}

/* ------------------------------------------------------------------------- */
/* accumulate_from_fft, include/bispectrum.h:356-409                         */
/* ------------------------------------------------------------------------- */

/* Array2 signed index range, include/array2.h:675-688 */
static void so_array2_range(int64_t ncols, int64_t nrows, int32_t mn[2], int32_t mx[2])
{
    mn[0] = -(int32_t)(ncols >> 1);
    mn[1] = -(int32_t)(nrows >> 1);
    mx[0] = -(int32_t)ncols / 2 + (int32_t)ncols - 1;
    mx[1] = -(int32_t)nrows / 2 + (int32_t)nrows - 1;
}

typedef struct so_loop {
    int32_t min1, min2, min3, min4, max1, max2, max4;
} so_loop;

static void so_loop_bounds(const so_desc* d, int64_t ncols, int64_t nrows, so_loop* L)
{
    int32_t fmn[2], fmx[2];
    so_array2_range(ncols, nrows, fmn, fmx);
    L->min1 = (int32_t)so_imax(fmn[0], d->min_idx[0]);
    L->min2 = (int32_t)so_imax(fmn[1], d->min_idx[1]);
    L->min3 = (int32_t)so_imax(fmn[0], d->min_idx[2]);
    L->min4 = (int32_t)so_imax(fmn[1], d->min_idx[3]);
    L->max1 = (int32_t)so_imin(fmx[0], d->max_idx[0]);
    L->max2 = (int32_t)so_imin(fmx[1], d->max_idx[1]);
    L->max4 = (int32_t)so_imin(fmx[1], d->max_idx[3]);
}

/* Array2::at with negative-index wrap, include/array2.h:310-325 */
static inline uint64_t so_at(int64_t ncols, int64_t nrows, int32_t x, int32_t y)
{
    int64_t xx = x < 0 ? x + ncols : x;
    int64_t yy = y < 0 ? y + nrows : y;
    return (uint64_t)(xx + yy * ncols);
}

uint64_t so_count_triples(const so_desc* d, uint64_t ncols, uint64_t nrows)
{
    so_loop L;
    so_loop_bounds(d, (int64_t)ncols, (int64_t)nrows, &L);
    uint64_t n = 0;
    for (int32_t i = L.min1; i <= 0; ++i)
        for (int32_t j = L.min2; j <= L.max2; ++j)
            for (int32_t k = L.min3; k <= 0; ++k) {
                if (i + k < L.min1 || i + k > L.max1)
                    continue;
                int32_t lo = (int32_t)so_imax(L.min4, L.min2 - j);
                int32_t hi = (int32_t)so_imin(L.max4, L.max2 - j);
                if (hi >= lo)
                    n += (uint64_t)(hi - lo + 1);
            }
    return n;
}

int so_accumulate_f64(const so_desc* d, double* B, const double* F, uint64_t ncols, uint64_t nrows)
{
    so_loop L;
    so_loop_bounds(d, (int64_t)ncols, (int64_t)nrows, &L);
    const int64_t X = (int64_t)ncols, Y = (int64_t)nrows;
    for (int32_t i = L.min1; i <= 0; ++i)
        for (int32_t j = L.min2; j <= L.max2; ++j)
            for (int32_t k = L.min3; k <= 0; ++k)
                for (int32_t l = L.min4; l <= L.max4; ++l) {
                    if ((i + k) >= L.min1 && (j + l) >= L.min2 && (i + k) <= L.max1 && (j + l) <= L.max2) {
                        const double* a = F + 2 * so_at(X, Y, i, j);
                        const double* b = F + 2 * so_at(X, Y, k, l);
                        const double* c = F + 2 * so_at(X, Y, i + k, j + l);
                        /* t = a; t *= b; t *= conj(c) */
                        double tr = a[0] * b[0] - a[1] * b[1];
                        double ti = a[0] * b[1] + a[1] * b[0];
                        double cr = c[0], ci = -c[1];
                        double ur = tr * cr - ti * ci;
                        double ui = tr * ci + ti * cr;
                        int32_t idx[4] = { i, j, k, l };
                        uint64_t addr = so_calc_offset(d, idx);
                        if (addr >= d->base_size)
                            return SO_OUT_OF_RANGE;
                        B[2 * addr] += ur;
                        B[2 * addr + 1] += ui;
                    }
                }
    return SO_OK;
}

int so_accumulate_f32(const so_desc* d, float* B, const double* F, uint64_t ncols, uint64_t nrows)
{
    so_loop L;
    so_loop_bounds(d, (int64_t)ncols, (int64_t)nrows, &L);
    const int64_t X = (int64_t)ncols, Y = (int64_t)nrows;
    for (int32_t i = L.min1; i <= 0; ++i)
        for (int32_t j = L.min2; j <= L.max2; ++j)
            for (int32_t k = L.min3; k <= 0; ++k)
                for (int32_t l = L.min4; l <= L.max4; ++l) {
                    if ((i + k) >= L.min1 && (j + l) >= L.min2 && (i + k) <= L.max1 && (j + l) <= L.max2) {
                        const double* a = F + 2 * so_at(X, Y, i, j);
                        const double* b = F + 2 * so_at(X, Y, k, l);
                        const double* c = F + 2 * so_at(X, Y, i + k, j + l);
                        /* complex<float> t{F(u)}; t *= F(v); t *= conj(F(u+v)) with every
                         * factor narrowed to float before the multiply */
                        float ar = (float)a[0], ai = (float)a[1];
                        float br = (float)b[0], bi = (float)b[1];
                        float cr = (float)c[0], ci = (float)(-c[1]);
                        float tr = ar * br - ai * bi;
                        float ti = ar * bi + ai * br;
                        float ur = tr * cr - ti * ci;
                        float ui = tr * ci + ti * cr;
                        int32_t idx[4] = { i, j, k, l };
                        uint64_t addr = so_calc_offset(d, idx);
                        if (addr >= d->base_size)
                            return SO_OUT_OF_RANGE;
                        B[2 * addr] += ur;
                        B[2 * addr + 1] += ui;
                    }
                }
    return SO_OK;
}

/* ------------------------------------------------------------------------- */
/* FFT with FFTW conventions (FFTW is the reference's FFT,                   */
/* smip-cli/main.cpp:256-259,340-346; include/crosscorrel.h:112-141).        */
/* Radix-2 for powers of two, direct O(n^2) DFT otherwise.                   */
/* ------------------------------------------------------------------------- */
static int so_is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

/* In-place 1-D transform of n complex values spaced by `stride` complex elements. */
static void so_fft1(double* x, int64_t n, int64_t stride, int sign, double* work)
{
    if (n <= 1)
        return;
    for (int64_t k = 0; k < n; ++k) {
        work[2 * k] = x[2 * k * stride];
        work[2 * k + 1] = x[2 * k * stride + 1];
    }
    if (so_is_pow2(n)) {
        /* bit reversal */
        for (int64_t i = 1, j = 0; i < n; ++i) {
            int64_t bit = n >> 1;
            for (; j & bit; bit >>= 1)
                j ^= bit;
            j ^= bit;
            if (i < j) {
                double t0 = work[2 * i], t1 = work[2 * i + 1];
                work[2 * i] = work[2 * j];
                work[2 * i + 1] = work[2 * j + 1];
                work[2 * j] = t0;
                work[2 * j + 1] = t1;
            }
        }
        for (int64_t len = 2; len <= n; len <<= 1) {
            int64_t half = len >> 1;
            for (int64_t m = 0; m < half; ++m) {
                /* twiddle exp(sign * 2 pi i m / len), angle reduced exactly */
                double ang = SO_2PI * (double)m / (double)len;
                double wr = cos(ang), wi = sign * sin(ang);
                for (int64_t s = 0; s < n; s += len) {
                    double* p = work + 2 * (s + m);
                    double* q = work + 2 * (s + m + half);
                    double qr = q[0] * wr - q[1] * wi;
                    double qi = q[0] * wi + q[1] * wr;
                    q[0] = p[0] - qr;
                    q[1] = p[1] - qi;
                    p[0] += qr;
                    p[1] += qi;
                }
            }
        }
        for (int64_t k = 0; k < n; ++k) {
            x[2 * k * stride] = work[2 * k];
            x[2 * k * stride + 1] = work[2 * k + 1];
        }
    } else {
        for (int64_t k = 0; k < n; ++k) {
            double sr = 0, si = 0;
            for (int64_t t = 0; t < n; ++t) {
                double ang = SO_2PI * (double)((k * t) % n) / (double)n;
                double wr = cos(ang), wi = sign * sin(ang);
                sr += work[2 * t] * wr - work[2 * t + 1] * wi;
                si += work[2 * t] * wi + work[2 * t + 1] * wr;
            }
            x[2 * k * stride] = sr;
            x[2 * k * stride + 1] = si;
        }
    }
}

void so_fft2_c2c(double* data, int64_t nrows, int64_t ncols, int sign)
{
    int64_t m = nrows > ncols ? nrows : ncols;
    double* work = (double*)malloc(sizeof(double) * 2 * (size_t)m);
    for (int64_t r = 0; r < nrows; ++r)
        so_fft1(data + 2 * r * ncols, ncols, 1, sign, work);
    for (int64_t c = 0; c < ncols; ++c)
        so_fft1(data + 2 * c, nrows, ncols, sign, work);
    free(work);
}

void so_fft2_r2c(const double* in, double* out, int64_t nrows, int64_t ncols)
{
    const int64_t hx = ncols / 2 + 1;
    double* full = (double*)malloc(sizeof(double) * 2 * (size_t)(nrows * ncols));
    for (int64_t i = 0; i < nrows * ncols; ++i) {
        full[2 * i] = in[i];
        full[2 * i + 1] = 0.0;
    }
    so_fft2_c2c(full, nrows, ncols, -1);
    for (int64_t r = 0; r < nrows; ++r)
        for (int64_t c = 0; c < hx; ++c) {
            out[2 * (r * hx + c)] = full[2 * (r * ncols + c)];
            out[2 * (r * hx + c) + 1] = full[2 * (r * ncols + c) + 1];
        }
    free(full);
}

void so_fft2_c2r(const double* in, double* out, int64_t nrows, int64_t ncols)
{
    const int64_t hx = ncols / 2 + 1;
    double* full = (double*)malloc(sizeof(double) * 2 * (size_t)(nrows * ncols));
    for (int64_t r = 0; r < nrows; ++r)
        for (int64_t c = 0; c < ncols; ++c) {
            double re, im;
            if (c < hx) {
                re = in[2 * (r * hx + c)];
                im = in[2 * (r * hx + c) + 1];
            } else { /* Hermitian completion */
                int64_t rr = (nrows - r) % nrows, cc = ncols - c;
                re = in[2 * (rr * hx + cc)];
                im = -in[2 * (rr * hx + cc) + 1];
            }
            full[2 * (r * ncols + c)] = re;
            full[2 * (r * ncols + c) + 1] = im;
        }
    so_fft2_c2c(full, nrows, ncols, +1);
    for (int64_t i = 0; i < nrows * ncols; ++i)
        out[i] = full[2 * i];
    free(full);
}

/* ------------------------------------------------------------------------- */
/* Phase reconstruction                                                      */
/* ------------------------------------------------------------------------- */

/* NextRecoIndex, src/phasereco.cpp:17-36 */
void so_next_reco_index(double* r, double* phi, int32_t* i, int32_t* j)
{
    const double dr = 1.0;
    const double dphi = 1.0 / (SO_2PI * *r);
    int32_t ii = *i, jj = *j;
    while (*i == ii && *j == jj) {
        if ((*phi += dphi) > SO_2PI) {
            *phi = 0;
            *r += dr;
        }
        ii = (int32_t)(*r * cos(*phi));
        jj = (int32_t)(*r * sin(*phi));
    }
    *i = ii;
    *j = jj;
}

int64_t so_spiral_visits(double reco_radius, int32_t* ij, int64_t cap)
{
    double r = 0.0, phi = 0.0;
    int32_t i = 0, j = 0;
    int64_t n = 0;
    while (r <= reco_radius) { /* include/phasereco.h:60-68 */
        so_next_reco_index(&r, &phi, &i, &j);
        if (ij && n < cap) {
            ij[2 * n] = i;
            ij[2 * n + 1] = j;
        }
        ++n;
    }
    return n;
}

typedef struct so_pm {
    int64_t xs, ys;
    int32_t mn[2], mx[2];
    double* ph;      /* complex */
    uint8_t* flag;
    double* cons;
} so_pm;

static inline int so_pm_contains(const so_pm* pm, int32_t x, int32_t y)
{
    return x >= pm->mn[0] && x <= pm->mx[0] && y >= pm->mn[1] && y <= pm->mx[1];
}

static inline uint64_t so_pm_at(const so_pm* pm, int32_t x, int32_t y)
{
    return so_at(pm->xs, pm->ys, x, y);
}

/* calc_phase, include/phasereco.h:75-123.  alias_mode 0 is the reference;
 * alias_mode 1 (a sensitivity probe, not the reference) leaves out every lookup
 * whose address calc_offset forms from out-of-range indices, as if its |beta|
 * were below DBL_EPSILON. */
static int so_calc_phase(const so_desc* d, const void* B, int is_f32, so_pm* pm,
    int32_t wx, int32_t wy, so_reco_stats* st, double* sbuf, int64_t sbuf_cap, int alias_mode)
{
    /* bispec_u_range / bispec_v_range from the descriptor, :81-82 */
    if (abs(wx) + abs(wy) <= 1 || wx < d->min_idx[0] || wx > d->max_idx[0] || wy < d->min_idx[1]
        || wy > d->max_idx[1])
        return SO_OK;
    int64_t n = 0;
    /* for (auto u : pm.range()) iterates x fastest (range.h:161-192) */
    for (int32_t uy = pm->mn[1]; uy <= pm->mx[1]; ++uy) {
        const int32_t vy = wy - uy;
        if (vy < d->min_idx[3] || vy > d->max_idx[3])
            continue;
        for (int32_t ux = pm->mn[0]; ux <= pm->mx[0]; ++ux) {
            const int32_t vx = wx - ux;
            if (vx < d->min_idx[2] || vx > d->max_idx[2])
                continue;
            /* pm.at(v) wraps like Array2::at; v lies inside pm for |v| <= b/2 */
            if (!pm->flag[so_pm_at(pm, ux, uy)] || !pm->flag[so_pm_at(pm, vx, vy)])
                continue;
            int32_t idx[4] = { ux, uy, vx, vy };
            double beta[2];
            uint64_t addr;
            int conj, alias;
            int rc = so_lookup(d, idx, &addr, &conj, &alias);
            if (st) {
                st->lookups++;
                st->aliased += alias;
            }
            if (rc)
                return rc; /* get_element throws std::out_of_range */
            if (alias && alias_mode == 1)
                continue;
            if (is_f32) {
                beta[0] = ((const float*)B)[2 * addr];
                beta[1] = ((const float*)B)[2 * addr + 1];
            } else {
                beta[0] = ((const double*)B)[2 * addr];
                beta[1] = ((const double*)B)[2 * addr + 1];
            }
            if (conj)
                beta[1] = -beta[1];
            const double* pu = pm->ph + 2 * so_pm_at(pm, ux, uy);
            const double* pv = pm->ph + 2 * so_pm_at(pm, vx, vy);
            /* T ph{phases.at(u)}; ph *= phases.at(v) */
            double phr = pu[0] * pv[0] - pu[1] * pv[1];
            double phi = pu[0] * pv[1] + pu[1] * pv[0];
            double ab = hypot(beta[0], beta[1]);
            if (ab > SO_EPS) {
                /* temp /= abs(temp); temp = conj(temp); ph *= temp */
                double tr = beta[0] / ab, ti = beta[1] / ab;
                ti = -ti;
                double qr = phr * tr - phi * ti;
                double qi = phr * ti + phi * tr;
                double aq = hypot(qr, qi);
                if (n >= sbuf_cap)
                    return SO_RUNTIME_ERROR;
                sbuf[2 * n] = qr / aq;
                sbuf[2 * n + 1] = qi / aq;
                ++n;
            }
        }
    }
    /* std::accumulate in list order, then the mean */
    double mr = 0.0, mi = 0.0;
    for (int64_t k = 0; k < n; ++k) {
        mr += sbuf[2 * k];
        mi += sbuf[2 * k + 1];
    }
    if (n > 0) {
        const uint64_t w = so_pm_at(pm, wx, wy);
        pm->flag[w] = 1;
        mr /= (double)n;
        mi /= (double)n;
        const double ap = hypot(mr, mi);
        pm->cons[w] = ap;
        if (ap > SO_EPS) {
            pm->ph[2 * w] = mr / ap;
            pm->ph[2 * w + 1] = mi / ap;
        } else {
            pm->ph[2 * w] = 0.0;
            pm->ph[2 * w + 1] = 0.0;
        }
        if (st)
            st->flagged++;
    }
    return SO_OK;
}

/* reconstruct_phases, include/phasereco.h:33-73 */
int so_reconstruct_phases(const so_desc* d, const void* B, int is_f32,
    uint64_t xsize, uint64_t ysize, double reco_radius,
    double* phases, uint8_t* flags, double* consistency, so_reco_stats* stats)
{
    return so_reconstruct_phases_alias(d, B, is_f32, xsize, ysize, reco_radius, phases, flags, consistency, stats, 0);
}

int so_reconstruct_phases_alias(const so_desc* d, const void* B, int is_f32,
    uint64_t xsize, uint64_t ysize, double reco_radius,
    double* phases, uint8_t* flags, double* consistency, so_reco_stats* stats, int alias_mode)
{
    so_pm pm;
    pm.xs = (int64_t)xsize;
    pm.ys = (int64_t)ysize;
    so_array2_range(pm.xs, pm.ys, pm.mn, pm.mx);
    pm.ph = phases;
    pm.flag = flags;
    pm.cons = consistency;
    memset(phases, 0, sizeof(double) * 2 * xsize * ysize);
    memset(flags, 0, xsize * ysize);
    memset(consistency, 0, sizeof(double) * xsize * ysize);
    so_reco_stats st;
    memset(&st, 0, sizeof(st));

    /* seeds, :42-54: init_phase = 1+0i; std::conj(init_phase) = 1-0i at (-1,0), (0,-1) */
    const int32_t seeds[5][2] = { { 0, 0 }, { 1, 0 }, { 0, 1 }, { -1, 0 }, { 0, -1 } };
    for (int s = 0; s < 5; ++s) {
        uint64_t a = so_pm_at(&pm, seeds[s][0], seeds[s][1]);
        phases[2 * a] = 1.0;
        phases[2 * a + 1] = s >= 3 ? -0.0 : 0.0;
        flags[a] = 1;
        consistency[a] = 1.0;
    }
    const int64_t cap = (int64_t)(2 * so_imax(d->max_idx[2], 0) + 1) * (2 * so_imax(d->max_idx[3], 0) + 1) + 16;
    double* sbuf = (double*)malloc(sizeof(double) * 2 * (size_t)cap);
    double r = 0.0, phi = 0.0;
    int32_t i = 0, j = 0;
    int rc = SO_OK;
    while (r <= reco_radius) {
        so_next_reco_index(&r, &phi, &i, &j);
        st.visits++;
        if (so_pm_contains(&pm, i, j) && !flags[so_pm_at(&pm, i, j)]) {
            st.calls++;
            rc = so_calc_phase(d, B, is_f32, &pm, i, j, &st, sbuf, cap, alias_mode);
            if (rc)
                break;
        }
    }
    free(sbuf);
    if (stats)
        *stats = st;
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Cross-correlation and shift-and-add                                       */
/* ------------------------------------------------------------------------- */
int so_xcorr(const double* ref, const double* frame, int64_t ncols, int64_t nrows, double* corr)
{
    const int64_t hx = ncols / 2 + 1;
    double* f1 = (double*)malloc(sizeof(double) * 2 * (size_t)(nrows * hx));
    double* f2 = (double*)malloc(sizeof(double) * 2 * (size_t)(nrows * hx));
    so_fft2_r2c(ref, f1, nrows, ncols);
    so_fft2_r2c(frame, f2, nrows, ncols);
    for (int64_t k = 0; k < nrows * hx; ++k) { /* conj(fft1) * fft2, crosscorrel.h:127-130 */
        double ar = f1[2 * k], ai = -f1[2 * k + 1];
        double br = f2[2 * k], bi = f2[2 * k + 1];
        f1[2 * k] = ar * br - ai * bi;
        f1[2 * k + 1] = ar * bi + ai * br;
    }
    so_fft2_c2r(f1, corr, nrows, ncols);
    free(f1);
    free(f2);
    return SO_OK;
}

void so_displacement(const double* corr, int64_t ncols, int64_t nrows, int32_t shift[2])
{
    int64_t best = 0; /* std::max_element: first maximum */
    for (int64_t k = 1; k < ncols * nrows; ++k)
        if (corr[best] < corr[k])
            best = k;
    int32_t dist = (int32_t)best, stride = (int32_t)ncols;
    shift[0] = dist % stride;
    shift[1] = dist / stride;
    if (!(shift[0] < (int32_t)(ncols + 1) / 2))
        shift[0] -= (int32_t)ncols;
    if (!(shift[1] < (int32_t)(nrows + 1) / 2))
        shift[1] -= (int32_t)nrows;
}

static int32_t so_clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

int so_shifted(const double* in, int64_t ncols, int64_t nrows, int32_t dx, int32_t dy, double* out)
{
    memset(out, 0, sizeof(double) * (size_t)(ncols * nrows));
    const int32_t X = (int32_t)ncols, Y = (int32_t)nrows;
    const int32_t sx = so_clampi(dx, 0, X - 1), sy = so_clampi(dy, 0, Y - 1);
    const int32_t ex = so_clampi(dx + X, 1, X), ey = so_clampi(dy + Y, 1, Y);
    for (int32_t row = sy; row < ey; ++row)
        for (int32_t col = sx; col < ex; ++col) {
            int32_t c = col - dx, r = row - dy;
            if (c < 0 || c >= X || r < 0 || r >= Y) /* operator() bounds check, array2.h:300-308 */
                return SO_OUT_OF_RANGE;
            out[row * X + col] = in[r * X + c];
        }
    return SO_OK;
}

int so_general_hamming(int64_t xsize, int64_t ysize, double aperture, double alpha, double* out)
{
    if (!(aperture > 0.))
        return SO_DOMAIN_ERROR;
    int32_t mn[2], mx[2];
    so_array2_range(xsize, ysize, mn, mx);
    const double trigarg = SO_2PI / aperture;
    const double c1 = alpha, c2 = 1. - alpha;
    for (int32_t i = mn[0]; i <= mx[0]; ++i)
        for (int32_t j = mn[1]; j <= mx[1]; ++j) {
            double t1 = c1 + c2 * cos(trigarg * (double)i);
            double t2 = c1 + c2 * cos(trigarg * (double)j);
            out[so_at(xsize, ysize, i, j)] = t1 * t2;
        }
    return SO_OK;
}

/* ------------------------------------------------------------------------- */
/* smip-cli pipeline, smip-cli/main.cpp:242-376                              */
/* ------------------------------------------------------------------------- */
int so_power_and_bispectrum(const double* frames, int64_t nframes, int64_t X, int64_t Y,
    uint64_t bdepth, int bispec_f32, void* bispec, double* powerspec)
{
    if (nframes <= 0)
        return SO_INVALID_ARGUMENT;
    so_desc d;
    so_desc_init(&d, (uint64_t)X, (uint64_t)Y, bdepth, bdepth);
    const size_t np = (size_t)(X * Y);
    double* F = (double*)malloc(sizeof(double) * 2 * np);
    if (powerspec)
        memset(powerspec, 0, sizeof(double) * np);
    int rc = SO_OK;
    for (int64_t f = 0; f < nframes && rc == SO_OK; ++f) {
        const double* fr = frames + (size_t)f * np;
        for (size_t k = 0; k < np; ++k) {
            F[2 * k] = fr[k];
            F[2 * k + 1] = 0.0;
        }
        so_fft2_c2c(F, Y, X, -1);
        if (bispec)
            rc = bispec_f32 ? so_accumulate_f32(&d, (float*)bispec, F, (uint64_t)X, (uint64_t)Y)
                            : so_accumulate_f64(&d, (double*)bispec, F, (uint64_t)X, (uint64_t)Y);
        if (powerspec) /* P += complex{norm(F), 0} */
            for (size_t k = 0; k < np; ++k)
                powerspec[k] += F[2 * k] * F[2 * k] + F[2 * k + 1] * F[2 * k + 1];
    }
    free(F);
    if (rc)
        return rc;
    if (powerspec) /* powerspec /= nframes * powerspec.size() (size_t -> double) */
        for (size_t k = 0; k < np; ++k)
            powerspec[k] /= (double)((uint64_t)nframes * (uint64_t)np);
    if (bispec) { /* bispectrum /= complex(nframes, 0) */
        if (bispec_f32) {
            float* b = (float*)bispec;
            const float n = (float)nframes;
            for (uint64_t k = 0; k < d.base_size; ++k) {
                b[2 * k] /= n;
                b[2 * k + 1] /= n;
            }
        } else {
            double* b = (double*)bispec;
            const double n = (double)nframes;
            for (uint64_t k = 0; k < d.base_size; ++k) {
                b[2 * k] /= n;
                b[2 * k + 1] /= n;
            }
        }
    }
    return SO_OK;
}

int so_saa(const double* frames, int64_t nframes, int64_t X, int64_t Y, int32_t* shifts, double* sum)
{
    const size_t np = (size_t)(X * Y);
    double* ref = (double*)malloc(sizeof(double) * np);
    double* absf = (double*)malloc(sizeof(double) * np);
    double* corr = (double*)malloc(sizeof(double) * np);
    double* sh = (double*)malloc(sizeof(double) * np);
    int rc = SO_OK;
    /* sumarray = Re(frame0); CrossCorrelation(sumarray), main.cpp:263-268 */
    memcpy(ref, frames, sizeof(double) * np);
    if (sum)
        memcpy(sum, frames, sizeof(double) * np);
    if (shifts) {
        shifts[0] = 0;
        shifts[1] = 0;
    }
    for (int64_t f = 1; f < nframes; ++f) {
        const double* fr = frames + (size_t)f * np;
        for (size_t k = 0; k < np; ++k)
            absf[k] = fabs(fr[k]); /* complex_abs of a value with zero imaginary part */
        so_xcorr(ref, absf, X, Y, corr);
        int32_t s[2];
        so_displacement(corr, X, Y, s);
        if (shifts) {
            shifts[2 * f] = s[0];
            shifts[2 * f + 1] = s[1];
        }
        if (sum) {
            rc = so_shifted(absf, X, Y, -s[0], -s[1], sh);
            if (rc)
                break;
            for (size_t k = 0; k < np; ++k)
                sum[k] += sh[k];
        }
    }
    if (sum && rc == SO_OK) /* sumarray /= nframes (size_t -> double) */
        for (size_t k = 0; k < np; ++k)
            sum[k] /= (double)nframes;
    free(ref);
    free(absf);
    free(corr);
    free(sh);
    return rc;
}

int so_synthesize(const double* powerspec, const double* phases, int64_t X, int64_t Y,
    double reco_radius, double* image)
{
    const size_t np = (size_t)(X * Y);
    double* win = (double*)malloc(sizeof(double) * np);
    int rc = so_general_hamming(X, Y, reco_radius * 2, 0.5, win); /* Hann(…, 2p), main.cpp:329 */
    if (rc) {
        free(win);
        return rc;
    }
    for (size_t k = 0; k < np; ++k) {
        /* phases *= window (complex (w,0)); result = (sqrt(Re P), 0) * phases */
        double pr = phases[2 * k] * win[k] - phases[2 * k + 1] * 0.0;
        double pi = phases[2 * k] * 0.0 + phases[2 * k + 1] * win[k];
        double s = sqrt(powerspec[k]);
        image[2 * k] = s * pr - 0.0 * pi;
        image[2 * k + 1] = s * pi + 0.0 * pr;
    }
    free(win);
    so_fft2_c2c(image, Y, X, +1); /* FFTW_BACKWARD, unnormalised, main.cpp:340-346 */
    double mx = 0.0;
    for (size_t k = 0; k < np; ++k) {
        double a = hypot(image[2 * k], image[2 * k + 1]);
        if (a > mx)
            mx = a;
    }
    for (size_t k = 0; k < np; ++k) { /* result /= complex(normfact, 0) */
        image[2 * k] /= mx;
        image[2 * k + 1] /= mx;
    }
    return SO_OK;
}

int so_pipeline(const double* frames, int64_t nframes, int64_t X, int64_t Y,
    uint64_t bdepth, double reco_radius, int bispec_f32, so_pipeline_out* out)
{
    if (nframes <= 0 || X <= 0 || Y <= 0)
        return SO_INVALID_ARGUMENT;
    so_desc d;
    so_desc_init(&d, (uint64_t)X, (uint64_t)Y, bdepth, bdepth);
    const size_t np = (size_t)(X * Y);
    const size_t esz = bispec_f32 ? 2 * sizeof(float) : 2 * sizeof(double);
    void* B = out->bispectrum ? out->bispectrum : calloc(d.base_size, esz);
    if (out->bispectrum)
        memset(B, 0, d.base_size * esz);
    double* P = out->powerspec ? out->powerspec : (double*)malloc(sizeof(double) * np);
    int rc = so_power_and_bispectrum(frames, nframes, X, Y, bdepth, bispec_f32, B, P);
    if (rc == SO_OK && (out->sum_image || out->shifts))
        rc = so_saa(frames, nframes, X, Y, out->shifts, out->sum_image);
    double* ph = out->phases ? out->phases : (double*)malloc(sizeof(double) * 2 * np);
    uint8_t* fl = out->flags ? out->flags : (uint8_t*)malloc(np);
    double* co = out->consistency ? out->consistency : (double*)malloc(sizeof(double) * np);
    if (rc == SO_OK)
        rc = so_reconstruct_phases(&d, B, bispec_f32, (uint64_t)X, (uint64_t)Y, reco_radius, ph, fl, co, &out->stats);
    if (rc == SO_OK && out->image)
        rc = so_synthesize(P, ph, X, Y, reco_radius, out->image);
    if (!out->bispectrum)
        free(B);
    if (!out->powerspec)
        free(P);
    if (!out->phases)
        free(ph);
    if (!out->flags)
        free(fl);
    if (!out->consistency)
        free(co);
    return rc;
}

/* Keeps so_desc_empty referenced for the zero-size checks made from Python. */
void so_desc_init_empty(so_desc* d) { so_desc_empty(d); }
