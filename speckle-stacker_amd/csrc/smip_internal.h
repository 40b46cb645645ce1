// smip_internal.h -- shared definitions of the MI355X speckle-masking engine
// (host launchers, device helpers, error plumbing).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "smip_mi355x.h"

namespace smip_amd {

// ---------------------------------------------------------------------------
// errors: C++ exceptions inside the library, mapped to status codes at the ABI
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m)
        : std::runtime_error(m)
        , code(c)
    {
    }
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

// HIP refuses calls that act on the legacy stream or the whole device
// (synchronous copies; allocation and free may synchronise) while another thread
// captures a stream into a graph.  Graph capture and every such call in this
// library take this process-wide lock, and copies run on explicit non-blocking
// streams (copy_sync), never on the legacy stream.
std::recursive_mutex& capture_guard();
void* dev_malloc(size_t bytes);                 // throws SMIP_ERR_NO_MEMORY
void dev_free(void* p);                         // null allowed
void* host_malloc(size_t bytes);                // pinned host memory
void host_free(void* p);
// While an IdleRelease lives on a thread, that thread's dev_free / host_free keep
// the block for reuse instead of handing it back to HIP (whose free waits for the
// whole device).  Only for blocks no queued work can still touch.
struct IdleRelease {
    IdleRelease();
    ~IdleRelease();
    IdleRelease(const IdleRelease&) = delete;
    IdleRelease& operator=(const IdleRelease&) = delete;
};
void cache_trim(int device);          // return the device's cached blocks to HIP
size_t cache_device_bytes(int device); // bytes cached for reuse on the device
hipStream_t make_stream();                      // non-blocking, default priority
hipStream_t make_queue_stream();                // the same, with a hardware queue of its own
hipError_t mem_get_info(size_t* free_b, size_t* total_b);
void destroy_stream(hipStream_t s);
// synchronous copy through stream s (which is synchronised)
void copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s);

// the calling thread's last error text (smip_last_error)
void set_last_error(const char* msg);
void clear_last_error();
const char* last_error_text();

// runs fn, mapping exceptions to status codes and the last-error text
template <typename F>
int guarded(F&& fn)
{
    try {
        fn();
        clear_last_error();
        return SMIP_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return SMIP_ERR_NO_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return SMIP_ERR_RUNTIME;
    } catch (...) {
        set_last_error("unknown error");
        return SMIP_ERR_RUNTIME;
    }
}

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        fail(SMIP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define SMIP_HIP(x) ::smip_amd::hip_check((x), #x)

// ---------------------------------------------------------------------------
// Bispectrum descriptor (include/bispectrum.h:104-112,143-160,203-220)
// ---------------------------------------------------------------------------
struct Desc {
    uint64_t dims[4];
    uint64_t sizes[4];
    uint64_t base_sizes[4];
    uint64_t base_size;
    uint64_t totalsize;
    int32_t min_idx[4];
    int32_t max_idx[4];
};

Desc make_desc(uint64_t ncols, uint64_t nrows, uint64_t b0, uint64_t b1);

// Reference-layout address arithmetic, shared by host and device.
// calc_offset keeps the size_t (mod 2^64) arithmetic and the absence of a range
// check of include/bispectrum.h:324-349.
__host__ __device__ inline uint64_t calc_offset(const Desc& d, int32_t i0, int32_t i1, int32_t i2, int32_t i3)
{
    int32_t a[4] = { -i0, i1, -i2, i3 };
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (a[k] < 0)
            a[k] = (int32_t)((uint64_t)(int64_t)a[k] + d.sizes[k]);
    uint64_t addr = (uint64_t)(int64_t)a[0];
    addr = addr * d.base_sizes[1] + (uint64_t)(int64_t)a[1];
    addr = addr * d.base_sizes[2] + (uint64_t)(int64_t)a[2];
    addr = addr * d.base_sizes[3] + (uint64_t)(int64_t)a[3];
    return addr;
}

__host__ __device__ inline int32_t iabs32(int32_t x) { return x < 0 ? -x : x; }

// get_element's address resolution: pre-swap, range check, symmetry
// canonicalisation (T1/T7/T6/T9/T3/T12), post-swap, calc_offset, address check
// (include/bispectrum.h:162-174,482-512,553-589).  Returns 0 or SMIP_ERR_OUT_OF_RANGE.
__host__ __device__ inline int resolve_element(const Desc& d, int32_t ux, int32_t uy, int32_t vx, int32_t vy,
    uint64_t* addr, bool* conj)
{
    if (iabs32(vx) > d.max_idx[2] || iabs32(vy) > d.max_idx[3]) {
        int32_t t = ux;
        ux = vx;
        vx = t;
        t = uy;
        uy = vy;
        vy = t;
    }
    if (ux < d.min_idx[0] || ux > d.max_idx[0] || uy < d.min_idx[1] || uy > d.max_idx[1] || vx < d.min_idx[2]
        || vx > d.max_idx[2] || vy < d.min_idx[3] || vy > d.max_idx[3])
        return SMIP_ERR_OUT_OF_RANGE;
    int32_t a0, a1, a2, a3;
    bool c = false;
    if (ux <= 0 && vx <= 0) { // T1
        a0 = ux, a1 = uy, a2 = vx, a3 = vy;
    } else if (ux > 0 && vx > 0) { // T7
        c = true;
        a0 = -ux, a1 = -uy, a2 = -vx, a3 = -vy;
    } else if (ux > 0) { // vx <= 0
        if (ux + vx > 0) { // T6
            a0 = -ux - vx, a1 = -uy - vy, a2 = vx, a3 = vy;
        } else { // T9
            c = true;
            a0 = -ux, a1 = -uy, a2 = vx + ux, a3 = vy + uy;
        }
    } else { // ux <= 0, vx > 0
        if (ux + vx > 0) { // T3
            a0 = ux, a1 = uy, a2 = -vx - ux, a3 = -vy - uy;
        } else { // T12
            c = true;
            a0 = ux + vx, a1 = uy + vy, a2 = -vx, a3 = -vy;
        }
    }
    if (iabs32(a2) > d.max_idx[2] || iabs32(a3) > d.max_idx[3]) {
        int32_t t = a0;
        a0 = a2;
        a2 = t;
        t = a1;
        a1 = a3;
        a3 = t;
    }
    uint64_t off = calc_offset(d, a0, a1, a2, a3);
    if (off >= d.base_size)
        return SMIP_ERR_OUT_OF_RANGE;
    *addr = off;
    *conj = c;
    return 0;
}

// Loop bounds of accumulate_from_fft (include/bispectrum.h:383-391) for an FFT
// array of ncols x nrows.
struct LoopBounds {
    int32_t min1, min2, min3, min4, max1, max2, max4;
};
LoopBounds loop_bounds(const Desc& d, int64_t ncols, int64_t nrows);

// ---------------------------------------------------------------------------
// Device spectrum layout ("half plane", x <= 0):
//   H[frame][c][PAD + r], c = ux + (ncols>>1) in [0, ncols>>1],
//   r = uy + (nrows>>1) in [0, nrows), zero rows outside [PAD, PAD + nrows).
// Row stride RS complex elements, frame stride HX * RS.
// ---------------------------------------------------------------------------
struct HalfPlane {
    int32_t ncols, nrows; // frame extents
    int32_t hx;           // number of stored columns = (ncols >> 1) + 1
    int32_t pad;          // zero rows before row 0
    int32_t rs;           // row stride (complex elements)
    int64_t fs;           // frame stride = hx * rs
};
HalfPlane make_half_plane(int32_t ncols, int32_t nrows, int32_t pad, int32_t tail);

// ---------------------------------------------------------------------------
// launchers (defined in the .hip files)
// ---------------------------------------------------------------------------
enum class FrameType : int { u8 = SMIP_DTYPE_U8, u16 = SMIP_DTYPE_U16, f32 = SMIP_DTYPE_F32, f64 = SMIP_DTYPE_F64 };
size_t dtype_size(int dtype);

struct FftPlan; // twiddle tables for one frame shape
constexpr int32_t kMaxFftLength = 2048; // longest transform held in LDS
// any extents in [2, kMaxFftLength]
FftPlan* fft_plan_create(int32_t ncols, int32_t nrows, hipStream_t stream);
void fft_plan_destroy(FftPlan*);

// frames -> half-plane spectrum (float2 or double2), optional |x| of the input
// and optional FP64 power-spectrum accumulation over the nf frames.
// work must hold nf * hx * nrows complex values of the spectrum precision.
void launch_forward_f32(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval,
    const HalfPlane& hp, float2* out, float2* work, double* power_acc, hipStream_t s);
void launch_forward_f64(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval,
    const HalfPlane& hp, double2* out, double2* work, double* power_acc, hipStream_t s);
// natural r2c spectrum R[f][ky][kx] (kx in [0, ncols/2]) in FP64 for the
// cross-correlation; work as above.
void launch_rfft_f64(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval,
    double2* out, double2* work, hipStream_t s);
// c = c2r(conj(R_ref) * R_f) for nf frames; out nf * nrows * ncols doubles.
// Per-workgroup first maximum of a correlation (value, row-major index).
struct RowBest {
    double v;
    int64_t i;
};
// The cross-correlation of each of nf frames with the reference spectrum,
// reduced straight to its displacement: shifts[2f], shifts[2f+1].  work holds
// nf * nrows * (ncols/2 + 1) complex values; best nf * nrows / 2 entries.
void launch_shift_f64(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval,
    const double2* ref_spec, RowBest* best, double2* work, int32_t* shifts, hipStream_t s);
// RowBest entries launch_shift_f64 writes per frame
int fft_shift_best_entries(const FftPlan* p);
void launch_xcorr_f64(const FftPlan* p, const double2* ref_spec, const double2* spec, int64_t nf,
    double* corr, double2* work, hipStream_t s);
// unnormalised backward complex 2-D FFT, FP64, in place on nrows x ncols.
void launch_c2c_inverse_f64(const FftPlan* p, double2* data, double2* work, hipStream_t s);

// bispectrum accumulation
void launch_accum_plain_f64(const Desc& d, const LoopBounds& lb, const HalfPlane& hp, const double2* H,
    double2* B, hipStream_t s);
void launch_accum_plain_f32(const Desc& d, const LoopBounds& lb, const HalfPlane& hp, const double2* H,
    float2* B, hipStream_t s);

// Arguments of the batched tile kernel, shared by the items of one job.
struct TileArgs {
    int32_t X, Y, pad, rs;
    int64_t fs;
    int32_t half_x, half_y; // ncols >> 1, nrows >> 1
    int32_t min4, max4;     // valid vy range (loop bounds)
    int32_t s1, s3;         // sizes[1], sizes[3]
    int32_t bs1, bs2, bs3;  // base_sizes[1..3]
};

struct TilePlan; // work-item table of the batched accumulate
TilePlan* tile_plan_create(const Desc& d, const LoopBounds& lb, const HalfPlane& hp, int vy_tile, hipStream_t stream);
void tile_plan_destroy(TilePlan*);
int tile_plan_vy_tile(const TilePlan*);
int tile_default_vy(const LoopBounds& L, int X, int Y); // vy tile of the batched kernel when none is given
int tile_plan_pad_needed(const Desc& d, int vy_tile);
int tile_plan_tail_needed(const Desc& d, int nrows, int vy_tile);
void launch_accum_tile(const TilePlan* tp, const HalfPlane& hp, const float2* H, int64_t nf, double2* B64,
    hipStream_t s);
uint64_t tile_plan_cells(const TilePlan* tp); // cells computed per frame
const int4* tile_plan_items(const TilePlan* tp);  // device item table
int64_t tile_plan_nitems(const TilePlan* tp);     // a multiple of 8
const TileArgs& tile_plan_args(const TilePlan* tp);

// ---------------------------------------------------------------------------
// Grouped launches (smip_job_add_frames_many): one launch per stage for the
// current batch of many jobs.  Block k of a grouped launch belongs to the job j
// with off[j] <= k < off[j + 1] (off[0] = 0, off[n] = the launch's blocks); the
// block then runs the single-job kernel's body for its local index.
// ---------------------------------------------------------------------------
// A device pointer read from a table in device memory, as a pointer into global
// memory.  A generic pointer loaded from memory makes every access through it a
// flat instruction, which also counts against the LDS wait counter, so each
// wait then covers LDS and memory together, and the operands of a uniform load
// land in vector registers.  The field is read as an integer because a cast of
// the loaded generic pointer to global and back is folded away.
template <class T>
__device__ __forceinline__ T* global_ptr(T* const& field)
{
    const uint64_t v = *reinterpret_cast<const uint64_t*>(&field);
    return (T*)(__attribute__((address_space(1))) T*)v;
}

// Element i of a table that no kernel writes, as a scalar (constant address
// space) load when i is uniform.
template <class T>
__device__ __forceinline__ T load_const(const T* const& field, int64_t i)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const uint64_t v = *reinterpret_cast<const uint64_t*>(&field);
    return ((const __attribute__((address_space(4))) T*)v)[i];
#else
    return field[i]; // the host pass only parses device functions
#endif
}

__device__ inline int group_job(const uint32_t* __restrict__ off, int n, uint32_t g)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// one job's forward FP32 transform, |F|^2 and tile inputs for one batch
struct GroupFwd {
    const void* frames;
    int32_t dtype, X, Y, lx, ly, hx, ppw, cb;
    int64_t nf;
    const float2* twx;
    const float2* twy;
    float2* G;
    float2* H;
    HalfPlane hp;
    double* P;
};
struct GroupTile {
    const int4* items;
    TileArgs a;
    const float2* H;
    int64_t nf;
    double2* B;
};
// one job's shift-and-add of one chunk of frames
struct GroupSaa {
    const void* frames;
    int32_t dtype, X, Y, lx, ly, hx, ppw, cb, nblk, parts;
    int64_t nf;
    const double2* twx;
    const double2* twy;
    const double2* ref;
    double2* G;
    RowBest* best;
    int32_t* shifts;
    double* sum;
};
// Row and column tiling of the power-of-two transforms (the grouped launches use
// the single-job launches' shapes).
struct FftTiling {
    int lx, ly; // log2 of the lengths
    int ppw32, cb32, ppw64, cb64, nblk64;
    size_t lds_rows32, lds_cols32, lds_rows64, lds_cols64;
    const float2* twx32;
    const float2* twy32;
    const double2* twx64;
    const double2* twy64;
};
FftTiling fft_tiling(const FftPlan* p);
bool fft_plan_pow2(const FftPlan* p);
// frame slices of the integer shift-and-add (1 for float frames)
unsigned saa_add_parts(int dtype, int64_t nf, int32_t ncols, int32_t nrows);
void launch_group_fwd(const GroupFwd* tab, const uint32_t* off_rows, const uint32_t* off_cols,
    const uint32_t* off_pow, int n, uint32_t nb_rows, uint32_t nb_cols, uint32_t nb_pow, size_t lds_rows,
    size_t lds_cols, hipStream_t s);
void launch_group_tile(int vy_tile, const GroupTile* tab, const uint32_t* off, int n, uint32_t nblocks, hipStream_t s);
// the five shift-and-add stages of one chunk index; off holds five (n + 1) arrays
// (rows, columns, c2r, merge, add)
void launch_group_saa(const GroupSaa* tab, const uint32_t* off, int n, const uint32_t nb[5], size_t lds_rows,
    size_t lds_cols, hipStream_t s);
void launch_group_saa_add(const GroupSaa* tab, const uint32_t* off, int n, uint32_t nblocks, hipStream_t s);

void launch_scale_c64(double2* B, uint64_t n, double re, double im, int op, hipStream_t s);
void launch_scale_c32(float2* B, uint64_t n, float re, float im, int op, hipStream_t s);
void launch_convert_c64_to_c32(const double2* in, float2* out, uint64_t n, hipStream_t s);
void launch_convert_c32_to_c64(const float2* in, double2* out, uint64_t n, hipStream_t s);
void launch_get_elements(const Desc& d, const void* B, int is_f32, const int32_t* idx, int64_t n, double2* out,
    int32_t* status, hipStream_t s);

// shift-and-add
void launch_argmax_shift(const double* corr, int64_t nf, int32_t ncols, int32_t nrows, int32_t* shifts,
    hipStream_t s);
void launch_saa_add(const void* frames, int dtype, int64_t nf, int32_t ncols, int32_t nrows,
    const int32_t* shifts, double* sum, hipStream_t s);

// synthesis
void launch_synth_prepare(const double* power_half, const double2* phases, int32_t ncols, int32_t nrows,
    double reco_radius, double norm, double2* out, hipStream_t s);
void launch_abs_max_normalize(double2* data, int64_t n, double* scratch, hipStream_t s);
void launch_power_expand(const double* power_half, int32_t ncols, int32_t nrows, double norm, double* out,
    hipStream_t s);

// phase reconstruction (device)
// in-range spiral visits of reconstruct_phases as (x, y) pairs, host side, cached
std::shared_ptr<const std::vector<int32_t>> visit_list(double reco_radius, int64_t X, int64_t Y);
// the visit list on the current device, cached per (device, radius, X, Y); the
// upload runs on `stream` (synchronously) the first time a geometry is asked for
struct RecoGeom;
std::shared_ptr<const RecoGeom> reco_geom(double reco_radius, int32_t xsize, int32_t ysize, hipStream_t stream);
uint64_t reco_geom_bytes_estimate(double reco_radius, int32_t xsize, int32_t ysize);
// per-recovery scratch (one device allocation, nothing uploaded)
struct RecoWork;
RecoWork* reco_work_create(const Desc& d, std::shared_ptr<const RecoGeom> geom);
uint64_t reco_scratch_bytes(const Desc& d, int32_t xsize, int32_t ysize);
void reco_work_destroy(RecoWork*);
// n recoveries together (one launch per stage and window index); calls[k] = job k's
// calc_phase count, status[k] = its status.  push: the push chain where a window
// allows it (as a single recovery), else the general schedule throughout.
void run_reconstruct_many(RecoWork* const* w, int n, const void* const* B, const int* is_f32,
    double2* const* phases, uint8_t* const* flags, double* const* consistency, int32_t* calls, int32_t* status,
    hipStream_t s, bool push);
void run_calc_phase(const Desc& d, int32_t xsize, int32_t ysize, const void* B, int is_f32, double2* phases,
    uint8_t* flags, double* consistency, int32_t wx, int32_t wy, hipStream_t s);
// message for a nonzero recovery status code
const char* reco_error_text(int32_t code);
int run_reconstruct(RecoWork* w, const void* B, int is_f32, double2* phases, uint8_t* flags, double* consistency,
    hipStream_t s);

// batched-kernel autotuner cache (tune.hip); false when nothing is cached
bool tune_lookup(uint64_t ncols, uint64_t nrows, uint64_t bdepth, int* vy_tile, int* batch);
// (u, v) tuples the reference accumulate loop adds per frame
uint64_t count_triples(const LoopBounds& L);

} // namespace smip_amd

// the device context behind smip_ctx handles
struct smip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // objects made on this context and not yet destroyed (bispectra, frame
    // stacks, cross-correlations, jobs, servers); smip_ctx_destroy refuses while
    // any is alive, since each of them reads ctx when it is destroyed
    std::atomic<int64_t> live { 0 };
};

namespace smip_amd {
// Member of every object made on a context: counts it in ctx->live for its lifetime.
struct CtxRef {
    smip_ctx* ctx = nullptr;
    CtxRef() = default;
    CtxRef(const CtxRef&) = delete;
    CtxRef& operator=(const CtxRef&) = delete;
    void hold(smip_ctx* c)
    {
        if (ctx)
            ctx->live.fetch_sub(1);
        ctx = c;
        if (ctx)
            ctx->live.fetch_add(1);
    }
    ~CtxRef()
    {
        if (ctx)
            ctx->live.fetch_sub(1);
    }
};
} // namespace smip_amd

namespace smip_amd {
inline void ensure_device(smip_ctx* ctx) { SMIP_HIP(hipSetDevice(ctx->device)); }
// smip_job_finalize_many with the group's phase recovery on group_stream (nullptr:
// the first job's stream); the server passes each worker's own-queue stream
int job_finalize_many_on(smip_job* const* jobs, uint64_t n, int32_t* status, hipStream_t group_stream);
} // namespace smip_amd
