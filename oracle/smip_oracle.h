/*
 * smip_oracle.h -- CPU restatement (test infrastructure, NOT product code) of the
 * speckle-masking hot path of hangeza/speckle-stacker 0.3.2 (SMIP).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker.  The product (libsmip_mi355x.so) never
 * links or calls it.
 *
 * Conventions
 *   - complex arrays are interleaved doubles (re, im);
 *   - 2-D arrays are row-major with x (column) fastest: a[row * ncols + col],
 *     the layout of smip::Array2 (include/array2.h:290-308);
 *   - the bispectrum is stored in the reference layout
 *     [-ux][uy wrapped][-vx][vy wrapped] (include/bispectrum.h:324-349).
 *
 * Status codes mirror the exception classes of the reference:
 *   0 ok, 1 invalid_argument, 2 out_of_range, 3 runtime_error,
 *   4 bad_function_call, 5 logic_error, 6 domain_error.
 */
#ifndef SMIP_ORACLE_H
#define SMIP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SO_OK = 0,
    SO_INVALID_ARGUMENT = 1,
    SO_OUT_OF_RANGE = 2,
    SO_RUNTIME_ERROR = 3,
    SO_BAD_FUNCTION_CALL = 4,
    SO_LOGIC_ERROR = 5,
    SO_DOMAIN_ERROR = 6
};

/* Bispectrum array descriptor, include/bispectrum.h:104-112,143-160,203-220 */
typedef struct so_desc {
    uint64_t dims[4];       /* constructor extents {ncols, nrows, b0, b1} */
    uint64_t sizes[4];      /* 2*(d/2)+1 */
    uint64_t base_sizes[4]; /* sizes - sizes*{1,0,1,0}/2 */
    uint64_t base_size;
    uint64_t totalsize;
    int32_t min_idx[4];
    int32_t max_idx[4];
} so_desc;

void so_desc_init(so_desc* d, uint64_t ncols, uint64_t nrows, uint64_t b0, uint64_t b1);
uint64_t so_calc_offset(const so_desc* d, const int32_t idx[4]);
void so_calc_indices(const so_desc* d, uint64_t addr, int32_t idx[4]);
/* Resolves a get_element lookup to (address, conjugate flag).  Returns SO_OK or
 * SO_OUT_OF_RANGE; *aliased is set when the final canonical (u,v) still lies
 * outside the stored v range, i.e. the address is produced by calc_offset's
 * modular arithmetic from out-of-range indices. */
int so_lookup(const so_desc* d, const int32_t idx[4], uint64_t* addr, int* conj, int* aliased);
int so_get_element(const so_desc* d, const void* data, int is_f32, const int32_t idx[4], double out[2]);
int so_put_element(const so_desc* d, void* data, int is_f32, const int32_t idx[4], const double val[2]);

/* accumulate_from_fft: fft is an nrows x ncols complex<double> array.
 * f64: Bispectrum<complex<double>>; f32: Bispectrum<complex<float>> (each factor
 * narrowed to float, product and sum in float). */
int so_accumulate_f64(const so_desc* d, double* bispec, const double* fft, uint64_t ncols, uint64_t nrows);
int so_accumulate_f32(const so_desc* d, float* bispec, const double* fft, uint64_t ncols, uint64_t nrows);
/* Number of (i,j,k,l) index tuples the accumulate loop adds per frame. */
uint64_t so_count_triples(const so_desc* d, uint64_t ncols, uint64_t nrows);

/* FFTs with the FFTW conventions: sign -1 forward, +1 backward, unnormalised. */
void so_fft2_c2c(double* data, int64_t nrows, int64_t ncols, int sign);
void so_fft2_r2c(const double* in, double* out, int64_t nrows, int64_t ncols);
void so_fft2_c2r(const double* in, double* out, int64_t nrows, int64_t ncols);

/* src/phasereco.cpp:17-36 */
void so_next_reco_index(double* r, double* phi, int32_t* i, int32_t* j);
/* Visits of reconstruct_phases' spiral while r <= reco_radius
 * (include/phasereco.h:56-68).  Writes up to cap (i,j) pairs into ij and returns
 * the total number of visits. */
int64_t so_spiral_visits(double reco_radius, int32_t* ij, int64_t cap);

typedef struct so_reco_stats {
    int64_t visits;        /* NextRecoIndex calls */
    int64_t calls;         /* calc_phase invocations */
    int64_t lookups;       /* get_element calls */
    int64_t aliased;       /* lookups resolved through out-of-range indices */
    int64_t flagged;       /* points flagged by calc_phase */
} so_reco_stats;

/* reconstruct_phases<complex<double>, U>, include/phasereco.h:33-123.
 * bispec holds U = complex<float> (is_f32) or complex<double>.
 * phases: xsize*ysize complex; flags: xsize*ysize bytes; consistency: doubles. */
int so_reconstruct_phases(const so_desc* d, const void* bispec, int is_f32,
    uint64_t xsize, uint64_t ysize, double reco_radius,
    double* phases, uint8_t* flags, double* consistency, so_reco_stats* stats);
/* The same with alias_mode 1: lookups that resolve through out-of-range indices
 * (the T3/T9 aliasing of calc_offset's wrap) are left out of the mean -- a
 * sensitivity probe of how much the reference's outputs depend on them, not a
 * restatement of the reference.  alias_mode 0 is so_reconstruct_phases. */
int so_reconstruct_phases_alias(const so_desc* d, const void* bispec, int is_f32,
    uint64_t xsize, uint64_t ysize, double reco_radius,
    double* phases, uint8_t* flags, double* consistency, so_reco_stats* stats, int alias_mode);

/* CrossCorrelation<double>::correlate, include/crosscorrel.h:96-143 */
int so_xcorr(const double* ref, const double* frame, int64_t ncols, int64_t nrows, double* corr);
/* calculate_displacement, include/crosscorrel.h:166-179 */
void so_displacement(const double* corr, int64_t ncols, int64_t nrows, int32_t shift[2]);
/* Array2::shifted, include/array2.h:603-622 */
int so_shifted(const double* in, int64_t ncols, int64_t nrows, int32_t dx, int32_t dy, double* out);
/* GeneralHamming / Hann, include/window_function.h:69-84 (output real) */
int so_general_hamming(int64_t xsize, int64_t ysize, double aperture, double alpha, double* out);

/* The smip-cli pipeline, smip-cli/main.cpp:242-376, on frames already converted
 * to doubles (nframes x nrows x ncols).  bispec_f32 selects the CLI's
 * complex<float> bispectrum (1) or the FP64 oracle (0).  Any output pointer may
 * be NULL. */
typedef struct so_pipeline_out {
    void* bispectrum;        /* base_size complex (float or double), normalised by n */
    double* powerspec;       /* nrows*ncols, Re of the normalised power spectrum */
    double* sum_image;       /* nrows*ncols, SAA sum / n (before the /255 for PNG) */
    int32_t* shifts;         /* nframes x 2, shift of frame f wrt frame 0 (0,0 for f=0) */
    double* phases;          /* nrows*ncols complex, reconstructed (before windowing) */
    uint8_t* flags;          /* nrows*ncols */
    double* consistency;     /* nrows*ncols */
    double* image;           /* nrows*ncols complex, result / max|result| */
    so_reco_stats stats;
} so_pipeline_out;

int so_pipeline(const double* frames, int64_t nframes, int64_t ncols, int64_t nrows,
    uint64_t bdepth, double reco_radius, int bispec_f32, so_pipeline_out* out);

/* Stage pieces of the pipeline (used to check GPU stages one by one). */
int so_power_and_bispectrum(const double* frames, int64_t nframes, int64_t ncols, int64_t nrows,
    uint64_t bdepth, int bispec_f32, void* bispectrum, double* powerspec);
int so_saa(const double* frames, int64_t nframes, int64_t ncols, int64_t nrows,
    int32_t* shifts, double* sum_image);
int so_synthesize(const double* powerspec, const double* phases, int64_t ncols, int64_t nrows,
    double reco_radius, double* image);

#ifdef __cplusplus
}
#endif
#endif
