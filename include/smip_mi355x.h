/*
 * smip_mi355x.h -- C ABI of the MI355X speckle-masking engine (libsmip_mi355x.so).
 *
 * This is the drop-in boundary for the hot path of hangeza/speckle-stacker 0.3.2
 * (SMIP).  The reference exposes that path as header-only C++ templates; each
 * entry point below replaces one of its members (file:line cited per
 * declaration).  The C++ mirror of the reference surface (namespace smip,
 * speckle-stacker_amd/smip/ headers) is written on top of these calls.
 *
 * Conventions
 *  - every function returns a status code (SMIP_OK = 0); the message of the last
 *    failure on the calling thread is returned by smip_last_error();
 *  - host arrays are plain pointers: complex values are interleaved (re, im),
 *    2-D arrays are row-major with x (column) fastest, the layout of
 *    smip::Array2 (include/array2.h:290-308);
 *  - the bispectrum is stored on the device in the reference layout
 *    [-ux][uy wrapped][-vx][vy wrapped] (include/bispectrum.h:324-349) so that
 *    exported buffers and .dat files are interchangeable with the reference's;
 *  - one HIP stream per job; calls on one object are not thread-safe, distinct
 *    objects may be used from distinct threads.
 */
#ifndef SMIP_MI355X_H
#define SMIP_MI355X_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__)
#define SMIP_API __attribute__((visibility("default")))
#else
#define SMIP_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: one per exception class thrown by the reference path. */
#define SMIP_OK 0
#define SMIP_ERR_INVALID_ARGUMENT 1  /* std::invalid_argument */
#define SMIP_ERR_OUT_OF_RANGE 2      /* std::out_of_range */
#define SMIP_ERR_RUNTIME 3           /* std::runtime_error */
#define SMIP_ERR_BAD_FUNCTION_CALL 4 /* std::bad_function_call */
#define SMIP_ERR_LOGIC 5             /* std::logic_error */
#define SMIP_ERR_DOMAIN 6            /* std::domain_error */
#define SMIP_ERR_HIP 7               /* HIP runtime failure */
#define SMIP_ERR_NO_MEMORY 8         /* device memory exhausted or admission refused */

/* Sample and element types. */
#define SMIP_DTYPE_U8 1
#define SMIP_DTYPE_U16 2
#define SMIP_DTYPE_F32 3
#define SMIP_DTYPE_F64 4
#define SMIP_DTYPE_C64 5  /* std::complex<float>, bispec_complex_t (include/types.h:33) */
#define SMIP_DTYPE_C128 6 /* std::complex<double>, complex_t (include/types.h:32) */

/* Arithmetic operators of Array_base / Bispectrum. */
#define SMIP_OP_ADD 0
#define SMIP_OP_SUB 1
#define SMIP_OP_MUL 2
#define SMIP_OP_DIV 3

/* Accumulation precision of an engine job. */
#define SMIP_PRECISION_FP64 0  /* plain drop-in: FP64 FFT, per-frame FP64 triple products */
#define SMIP_PRECISION_MIXED 1 /* batched: FP32 FFT + triple products per batch, FP64 fold */

/* Thread-local message of the last failing call (empty string if none). */
SMIP_API const char* smip_last_error(void);
/* Library version; the reference's is config/config.h:38-50 (0.3.2). */
SMIP_API int smip_version(int32_t* major, int32_t* minor, int32_t* patch);

/* ------------------------------------------------------------------------ */
/* Context: one GPU                                                         */
/* ------------------------------------------------------------------------ */
typedef struct smip_ctx smip_ctx;
SMIP_API int smip_ctx_create(int32_t device, smip_ctx** out);
/* Fails with SMIP_ERR_LOGIC (and destroys nothing) while objects made on the
 * context -- bispectra, frame stacks, cross-correlations, jobs, servers -- are
 * alive: each of them uses the context until it is destroyed. */
SMIP_API int smip_ctx_destroy(smip_ctx* ctx);
SMIP_API int smip_ctx_synchronize(smip_ctx* ctx);
SMIP_API int smip_ctx_mem_info(smip_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);

/* ------------------------------------------------------------------------ */
/* smip::Bispectrum<T> (include/bispectrum.h:28-125)                        */
/* ------------------------------------------------------------------------ */
typedef struct smip_bispec smip_bispec;

typedef struct smip_bispec_info {
    uint64_t dims[4];       /* constructor extents {ncols, nrows, b0, b1} */
    uint64_t sizes[4];      /* sizes(), :85-86,203-211 */
    uint64_t base_sizes[4]; /* base_sizes(), :87,213-220 */
    uint64_t base_size;     /* base_size() == size(), :85,88 */
    uint64_t totalsize;     /* totalsize(), :89 */
    int32_t min_indices[4]; /* min_indices(), :90,151-153 */
    int32_t max_indices[4]; /* max_indices(), :91,154-158 */
    int32_t dtype;          /* SMIP_DTYPE_C64 or SMIP_DTYPE_C128 */
} smip_bispec_info;

/* Bispectrum(const extents&) zero-filled, include/bispectrum.h:182-192; all-zero
 * dims give the default-constructed empty bispectrum (:176-180). */
SMIP_API int smip_bispec_create(smip_ctx* ctx, const uint64_t dims[4], int32_t dtype, smip_bispec** out);
SMIP_API int smip_bispec_destroy(smip_bispec* b);
/* sizes(), base_sizes(), min_indices() ..., include/bispectrum.h:85-91 */
SMIP_API int smip_bispec_info_get(const smip_bispec* b, smip_bispec_info* out);
/* accumulate_from_fft(const Array2<U>&), include/bispectrum.h:356-409.  fft is the
 * nrows x ncols complex<double> Array2 in FFTW output order. */
SMIP_API int smip_bispec_accumulate_from_fft(smip_bispec* b, const double* fft, uint64_t ncols, uint64_t nrows);
/* get_elements(idx_list, results), include/bispectrum.h:514-523 (get_element
 * :482-512 for n == 1): indices are n x 4 ints, values n complex doubles. */
SMIP_API int smip_bispec_get_elements(const smip_bispec* b, const int32_t* indices, uint64_t n, double* values);
/* smip_bispec_get_elements without any device or pinned allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate (code
 * between allocation-free sections of a serving loop, harnesses that forbid
 * hipMalloc).  The lookups, values and statuses are staged in a fixed scratch
 * array the library holds on each GPU for the life of the process (about 2.3 MB,
 * its address resolved when the first context is made on that GPU).  At most
 * 65536 lookups per call; more -> SMIP_ERR_INVALID_ARGUMENT.  Calls on one GPU
 * share that array, so the library runs them one at a time.  Same results and
 * errors as smip_bispec_get_elements. */
SMIP_API int smip_bispec_get_elements_noalloc(const smip_bispec* b, const int32_t* indices, uint64_t n,
    double* values);
/* put_element, include/bispectrum.h:525-535 */
SMIP_API int smip_bispec_put_element(smip_bispec* b, const int32_t idx[4], const double value[2]);
/* calc_offset (no range check, modulo 2^64), include/bispectrum.h:318-349 */
SMIP_API uint64_t smip_bispec_calc_offset(const smip_bispec* b, const int32_t idx[4]);
/* calc_indices, include/bispectrum.h:280-316 */
SMIP_API int smip_bispec_calc_indices(const smip_bispec* b, uint64_t addr, int32_t idx[4]);
/* Array_base<T>::data() contents (base_size elements of the bispectrum's dtype). */
SMIP_API int smip_bispec_download(const smip_bispec* b, void* host);
SMIP_API int smip_bispec_upload(smip_bispec* b, const void* host);
/* write_to_file / read_from_file, include/bispectrum.h:411-462 (same byte format). */
SMIP_API int smip_bispec_write_file(const smip_bispec* b, const char* path);
SMIP_API int smip_bispec_read_file(smip_bispec* b, const char* path);
/* Array_base operator op=(value), include/array_base.h:300-349 (/= 0 throws). */
SMIP_API int smip_bispec_scalar_op(smip_bispec* b, int32_t op, double re, double im);
/* Bispectrum operator op=(const Bispectrum&), include/bispectrum.h:232-278. */
SMIP_API int smip_bispec_array_op(smip_bispec* b, const smip_bispec* other, int32_t op);
/* operator=(const Bispectrum&), include/bispectrum.h:222-230 (dst keeps its dtype). */
SMIP_API int smip_bispec_assign(smip_bispec* dst, const smip_bispec* src);

/* ------------------------------------------------------------------------ */
/* Phase reconstruction (include/phasereco.h:15-27, src/phasereco.cpp)      */
/* ------------------------------------------------------------------------ */
/* reconstruct_phases<complex<double>, U>(bispec, xsize, ysize, reco_radius,
 * &phasemap), include/phasereco.h:33-73.  phases: xsize*ysize complex doubles;
 * flags / consistency: the PhaseMapElement fields (include/phasemap.h:10-20).
 * Any output pointer may be NULL. */
SMIP_API int smip_reconstruct_phases(const smip_bispec* b, uint64_t xsize, uint64_t ysize, double reco_radius,
    double* phases, uint8_t* flags, double* consistency);
/* calc_phase<complex<double>, U>(bispec, phases, phasemap, w),
 * include/phasereco.h:23-27,75-123: one recursion step on caller-held state.
 * phases / flags / consistency are read and updated in place (xsize*ysize). */
SMIP_API int smip_calc_phase(const smip_bispec* b, uint64_t xsize, uint64_t ysize, double* phases, uint8_t* flags,
    double* consistency, int32_t wx, int32_t wy);
/* NextRecoIndex, src/phasereco.cpp:17-36 (host function). */
SMIP_API void smip_next_reco_index(double* r, double* phi, int32_t* i, int32_t* j);

/* ------------------------------------------------------------------------ */
/* smip::CrossCorrelation<double> (include/crosscorrel.h:43-189)            */
/* ------------------------------------------------------------------------ */
typedef struct smip_xcorr smip_xcorr;
/* CrossCorrelation(const Array2<T>& ref), :89-94 */
SMIP_API int smip_xcorr_create(smip_ctx* ctx, const double* ref, uint64_t ncols, uint64_t nrows, smip_xcorr** out);
SMIP_API int smip_xcorr_destroy(smip_xcorr* x);
/* correlate(frame), :96-143 (size mismatch -> SMIP_ERR_INVALID_ARGUMENT) */
SMIP_API int smip_xcorr_correlate(smip_xcorr* x, const double* frame, uint64_t ncols, uint64_t nrows);
/* get_correlation_array(), :145-152 (before correlate -> SMIP_ERR_BAD_FUNCTION_CALL) */
SMIP_API int smip_xcorr_get_correlation(smip_xcorr* x, double* out);
/* get_displacement(), :154-164 */
SMIP_API int smip_xcorr_get_displacement(smip_xcorr* x, int32_t shift[2]);
/* operator()(frame) = correlate + displacement, :181-189 */
SMIP_API int smip_xcorr_displacement(smip_xcorr* x, const double* frame, uint64_t ncols, uint64_t nrows,
    int32_t shift[2]);

/* ------------------------------------------------------------------------ */
/* Window functions (include/window_function.h:11-84), host                 */
/* ------------------------------------------------------------------------ */
/* GeneralHamming(xsize, ysize, aperture, alpha); Hann is alpha = 0.5,
 * Hamming alpha = 0.54.  aperture <= 0 -> SMIP_ERR_DOMAIN.  out: real values. */
SMIP_API int smip_general_hamming(uint64_t xsize, uint64_t ysize, double aperture, double alpha, double* out);

/* ------------------------------------------------------------------------ */
/* Device-resident frame stacks                                             */
/* ------------------------------------------------------------------------ */
typedef struct smip_frames smip_frames;
SMIP_API int smip_frames_create(smip_ctx* ctx, int32_t dtype, uint64_t ncols, uint64_t nrows, uint64_t count,
    smip_frames** out);
SMIP_API int smip_frames_destroy(smip_frames* f);
SMIP_API int smip_frames_upload(smip_frames* f, uint64_t first, uint64_t count, const void* host);

/* ------------------------------------------------------------------------ */
/* Engine jobs: the smip-cli per-frame loop and post-processing            */
/* (smip-cli/main.cpp:242-376) on device-resident state                    */
/* ------------------------------------------------------------------------ */
typedef struct smip_job smip_job;

typedef struct smip_job_params {
    uint64_t ncols, nrows;     /* frame extents, any value in [2, 2048] */
    uint64_t bdepth;           /* -b, bispectrum extent of dims 3 and 4 */
    double reco_radius;        /* -p */
    int32_t frame_dtype;       /* SMIP_DTYPE_U8 / U16 / F32 */
    int32_t precision;         /* SMIP_PRECISION_FP64 or SMIP_PRECISION_MIXED */
    int32_t saa;               /* 1: cross-correlation shift-and-add against frame 0 */
    int32_t frames_per_batch;  /* FP32 frames per FP64 fold (MIXED); 0 = default (512) */
    int32_t vy_tile;           /* vy values per register tile (MIXED); 0 = default */
} smip_job_params;

typedef struct smip_job_stats {
    uint64_t frames;        /* frames accumulated */
    double accumulate_ms;   /* device time of FFT + |F|^2 + triple products */
    double saa_ms;          /* device time of the shift-and-add stage */
    double reco_ms;         /* device time of phase recovery */
    double synth_ms;        /* device time of image synthesis */
    uint64_t footprint;     /* device bytes held by the job */
    uint64_t reco_calls;    /* calc_phase invocations */
    uint64_t cells_per_frame;   /* accumulator cells the tile kernel computes per frame */
    uint64_t triples_per_frame; /* (u,v) tuples the reference loop adds per frame */
    double tile_ms;             /* device time of the triple-product kernel launches */
    uint64_t tile_launches;     /* number of those launches (graph replays not included) */
    uint64_t graph_launches;    /* per-batch HIP graph replays (full batches of pushed frames) */
} smip_job_stats;

/* Device bytes a job of these parameters holds (admission control). */
SMIP_API int smip_job_footprint(const smip_job_params* p, uint64_t* bytes);
SMIP_API int smip_job_create(smip_ctx* ctx, const smip_job_params* p, smip_job** out);
SMIP_API int smip_job_destroy(smip_job* j);
/* Enqueue frames [first, first+count) of a device stack (asynchronous). */
SMIP_API int smip_job_add_frames(smip_job* j, const smip_frames* f, uint64_t first, uint64_t count);
/* smip_job_add_frames for n jobs at once (no counterpart in the reference, which
 * processes one stack per process): frames [first[k], first[k] + count[k]) of
 * stacks[k] into jobs[k].  The jobs advance in lockstep rounds, each taking the
 * next batch of every job that has frames left -- the batches
 * smip_job_add_frames forms -- and each round runs every stage once for all of
 * them: the forward transform passes, |F|^2, the triple products (one launch per
 * vy tile width) and the shift-and-add passes, each block of a launch working for
 * one job.  Each job's results equal smip_job_add_frames's, bit for bit.  Jobs
 * in FP64 precision or with a frame side that is not a power of two >= 8 take
 * their own launches instead.  Jobs must share a device and appear once; the
 * call returns when the grouped work is done. */
SMIP_API int smip_job_add_frames_many(smip_job* const* jobs, const smip_frames* const* stacks, const uint64_t* first,
    const uint64_t* count, uint64_t n);
/* Upload host frames and enqueue them. */
SMIP_API int smip_job_add_frames_host(smip_job* j, const void* frames, uint64_t count);
/* Streaming form of the frame loop (smip-cli/main.cpp:281-306): appends count
 * host frames to the job's current batch and processes a batch whenever it is
 * full; smip_job_finalize processes the last, partial batch.  Batch boundaries
 * depend only on the frame sequence, not on how it is split into calls, so the
 * result is the same however the frames arrive.  The host buffer may be reused
 * on return.  Not to be mixed with smip_job_add_frames(_host) while pushed
 * frames are pending. */
SMIP_API int smip_job_push_frames_host(smip_job* j, const void* frames, uint64_t count);
/* The parameters a job of p is created with: vy_tile and frames_per_batch 0
 * replaced by the autotuner's cached choice for (ncols, nrows, bdepth) when
 * there is one, else by the built-in defaults. */
SMIP_API int smip_job_resolve_params(const smip_job_params* p, smip_job_params* out);
/* Normalise by n (main.cpp:307-312), recover phases (:319), synthesise (:328-376).
 * The division by n happens once per job: if phase recovery fails, the job stays
 * unfinalized with its bispectrum normalised, a later call retries the recovery
 * on that same bispectrum, and no more frames can be added. */
SMIP_API int smip_job_finalize(smip_job* j);
/* smip_job_finalize of n jobs at once (no counterpart in the reference): their
 * phase recoveries run as one group, one launch per stage and window index
 * for all of them, so the recursions of different jobs proceed side by side.
 * Each job's results equal those of smip_job_finalize.  Jobs must share a
 * device; none may be finalized or appear twice.  status (may be NULL)
 * receives each job's code; a job whose recovery fails (SMIP_ERR_OUT_OF_RANGE,
 * where get_element throws) is left unfinalized and the others complete; the
 * return value is the first failure. */
SMIP_API int smip_job_finalize_many(smip_job* const* jobs, uint64_t n, int32_t* status);
SMIP_API int smip_job_synchronize(smip_job* j);
SMIP_API int smip_job_get_stats(smip_job* j, smip_job_stats* out);
/* The job's bispectrum as a Bispectrum object (borrowed; valid until the job is
 * destroyed).  After finalize it holds the normalised values. */
SMIP_API int smip_job_bispectrum(smip_job* j, smip_bispec** out);
/* Outputs after finalize, each nrows x ncols: */
SMIP_API int smip_job_get_powerspec(smip_job* j, double* out);        /* Re of powerspec / (n N^2) */
SMIP_API int smip_job_get_sum_image(smip_job* j, double* out);        /* SAA sum / n */
SMIP_API int smip_job_get_shifts(smip_job* j, int32_t* out, uint64_t count); /* count x 2 */
SMIP_API int smip_job_get_phases(smip_job* j, double* phases, uint8_t* flags, double* consistency);
SMIP_API int smip_job_get_image(smip_job* j, double* out);            /* complex, /max|.| */

/* ------------------------------------------------------------------------ */
/* Autotuner of the batched triple-product kernel (no counterpart in the     */
/* reference)                                                                */
/* ------------------------------------------------------------------------ */
typedef struct smip_tune_result {
    int32_t vy_tile;          /* vy values per register tile */
    int32_t frames_per_batch; /* FP32 frames per FP64 fold */
    double frames_per_s;      /* triple-product kernel rate at that choice */
    double tile_tflops;       /* the same in algorithmic TFLOP/s (14 per triple) */
    int32_t candidates;       /* timed configurations */
} smip_tune_result;
/* Times the kernel on synthetic spectra for every vy tile, then for batch
 * sizes 32..1024 that fit max_bytes (0 = a quarter of free memory) beside the
 * bispectrum; picks the smallest batch within 3 % of the best frames/s and
 * caches the choice for (ncols, nrows, bdepth) in this process. */
SMIP_API int smip_autotune(smip_ctx* ctx, uint64_t ncols, uint64_t nrows, uint64_t bdepth, uint64_t max_bytes,
    smip_tune_result* out);
/* Clears the cache (jobs then use the built-in defaults again). */
SMIP_API int smip_autotune_forget(void);

/* ------------------------------------------------------------------------ */
/* Serving: a continuous-batching request loop (no counterpart in the        */
/* reference, which processes one stack per process, smip-cli/main.cpp)      */
/* ------------------------------------------------------------------------ */
/* Jobs are submitted, stream frames in chunks and are closed; they are
 * admitted first-fit in arrival order when a slot is free and their
 * smip_job_footprint fits the HBM budget beside the admitted jobs, keep their
 * state on the device while frames arrive, and overlap on the GPU (one stream
 * per job, worker threads drive the chunks).  A job's result is the same alone
 * and in any mix: batch boundaries follow its own frame sequence only.  Device
 * memory returns to the budget when the job is released. */
typedef struct smip_server smip_server;

typedef struct smip_server_params {
    uint64_t hbm_budget; /* device bytes admitted jobs may hold; 0 = 90% of free memory */
    int32_t max_active;  /* jobs admitted at once; 0 = 16 */
    int32_t workers;     /* host threads driving admitted jobs; 0 = 8 */
} smip_server_params;

#define SMIP_JOB_WAITING 0 /* submitted, not admitted */
#define SMIP_JOB_ACTIVE 1  /* admitted, frames streaming or being processed */
#define SMIP_JOB_DONE 2    /* finalised; results readable through smip_server_job */
#define SMIP_JOB_FAILED 3

typedef struct smip_server_job_info {
    int32_t state;
    uint64_t footprint;               /* device bytes */
    uint64_t frames_in, frames_done;  /* pushed / processed */
    double queue_ms;                  /* submit -> admission (or now) */
    double latency_ms;                /* submit -> completion (or now) */
    double tail_ms;                   /* close -> completion (or now) */
    char error[256];
} smip_server_job_info;

typedef struct smip_server_stats {
    uint64_t submitted, admitted, completed, failed;
    uint64_t waiting, active;       /* now */
    uint64_t bytes_in_use, budget;  /* admitted footprints / budget */
    uint64_t frames;                /* frames processed */
    uint64_t finalize_groups;       /* smip_job_finalize_many calls: jobs ready together finalise together */
    uint64_t finalized_in_groups;   /* jobs those calls held */
} smip_server_stats;

SMIP_API int smip_server_create(smip_ctx* ctx, const smip_server_params* p, smip_server** out);
/* Stops the workers and frees every job, finished or not. */
SMIP_API int smip_server_destroy(smip_server* s);
/* Footprint above the budget -> SMIP_ERR_NO_MEMORY; bad parameters -> SMIP_ERR_INVALID_ARGUMENT. */
SMIP_API int smip_server_submit(smip_server* s, const smip_job_params* p, uint64_t* id);
/* Copies count frames (the job's dtype and extents); the buffer may be reused on return. */
SMIP_API int smip_server_push(smip_server* s, uint64_t id, const void* frames, uint64_t count);
/* No more frames: the job is finalised once its frames are processed. */
SMIP_API int smip_server_close(smip_server* s, uint64_t id);
/* Blocks until the job is done or failed (timeout_s < 0: no limit); state may be NULL.
 * A failed job -> SMIP_ERR_RUNTIME with its error text. */
SMIP_API int smip_server_wait(smip_server* s, uint64_t id, double timeout_s, int32_t* state);
/* The finished job (borrowed, valid until release): read results with smip_job_get_*. */
SMIP_API int smip_server_job(smip_server* s, uint64_t id, smip_job** out);
SMIP_API int smip_server_get_job_info(smip_server* s, uint64_t id, smip_server_job_info* out);
/* Frees the job's device state and returns its footprint to the budget. */
SMIP_API int smip_server_release(smip_server* s, uint64_t id);
SMIP_API int smip_server_get_stats(smip_server* s, smip_server_stats* out);

#ifdef __cplusplus
}
#endif
#endif
