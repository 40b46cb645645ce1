// abi.cpp -- C ABI of libsmip_mi355x.so (declared in include/smip_mi355x.h).
//
// Host side of the engine: object lifetimes, the job pipeline (the smip-cli
// per-frame loop and post-processing of smip-cli/main.cpp:242-376 on
// device-resident state), the spiral visit list (src/phasereco.cpp:17-36), the
// window function (include/window_function.h:69-84) and the mapping of C++
// exceptions to status codes.  All arithmetic on frames, spectra and the
// bispectrum runs in the HIP kernels of fft.hip, bispec.hip, reco.hip, post.hip.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <tuple>
#include <condition_variable>
#include <mutex>
#include <map>
#include <string>
#include <unordered_map>
#include <thread>
#include <vector>

#include "smip_internal.h"

using namespace smip_amd;

namespace smip_amd {
// Trivially destructible on purpose: callers may reach the ABI from static
// destructors, which run after the calling thread's thread_local objects have
// been destroyed; a std::string here would then be written after its heap
// buffer was freed.
static thread_local char g_last_error[512];

void set_last_error(const char* m)
{
    std::strncpy(g_last_error, m, sizeof(g_last_error) - 1);
    g_last_error[sizeof(g_last_error) - 1] = 0;
}
void clear_last_error() { g_last_error[0] = 0; }
const char* last_error_text() { return g_last_error; }

std::recursive_mutex& capture_guard()
{
    static std::recursive_mutex m;
    return m;
}

// Block cache.  hipFree and hipHostFree wait until every stream of the device is
// idle, so a job released while other jobs run would stall them all (and every
// thread queued behind the allocation lock).  Blocks that their owner has shown
// to be idle (all streams that used them synchronised: IdleRelease scope) are
// kept here instead, per device, and handed out again for requests of the same
// size up to 1/8 smaller, up to half of the device's memory.  Cached blocks count
// as free memory; an allocation that fails returns the device's cached blocks to
// HIP and tries again.
namespace {
thread_local int g_idle_release = 0;

struct BlockCache {
    std::mutex mu;
    std::unordered_map<void*, std::pair<int, size_t>> dev_live; // every dev_malloc block: device, bytes
    std::multimap<std::pair<int, size_t>, void*> dev_free;     // idle blocks by (device, bytes)
    std::map<int, size_t> dev_cached, dev_cap;
    std::unordered_map<void*, size_t> host_live;
    std::multimap<size_t, void*> host_free;
    size_t host_cached = 0;
    static constexpr size_t kHostCap = size_t(4) << 30;

    static BlockCache& get()
    {
        static auto* c = new BlockCache; // never destroyed: nothing is freed at exit
        return *c;
    }
    template <typename M, typename K>
    static void* take(M& m, const K& lo, const K& hi)
    {
        auto it = m.lower_bound(lo);
        if (it == m.end() || hi < it->first)
            return nullptr;
        void* p = it->second;
        m.erase(it);
        return p;
    }
};

size_t device_cap(BlockCache& c, int dev)
{
    auto it = c.dev_cap.find(dev);
    if (it != c.dev_cap.end())
        return it->second;
    size_t total = 0;
    if (hipDeviceTotalMem(&total, dev) != hipSuccess) {
        (void)hipGetLastError();
        total = 0;
    }
    return c.dev_cap[dev] = total / 2;
}

int current_device()
{
    int dev = 0;
    SMIP_HIP(hipGetDevice(&dev));
    return dev;
}
} // namespace

IdleRelease::IdleRelease() { ++g_idle_release; }
IdleRelease::~IdleRelease() { --g_idle_release; }

void cache_trim(int dev)
{
    BlockCache& c = BlockCache::get();
    std::vector<void*> v;
    {
        std::lock_guard<std::mutex> lk(c.mu);
        for (auto it = c.dev_free.lower_bound({ dev, 0 }); it != c.dev_free.end() && it->first.first == dev;) {
            v.push_back(it->second);
            c.dev_live.erase(it->second);
            it = c.dev_free.erase(it);
        }
        c.dev_cached[dev] = 0;
    }
    std::lock_guard<std::recursive_mutex> lk(capture_guard());
    for (void* p : v)
        (void)hipFree(p);
}

void* dev_malloc(size_t bytes)
{
    BlockCache& c = BlockCache::get();
    const int dev = current_device();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        if (void* p = BlockCache::take(c.dev_free, std::make_pair(dev, bytes), std::make_pair(dev, bytes + bytes / 8))) {
            c.dev_cached[dev] -= c.dev_live[p].second;
            return p;
        }
    }
    void* p = nullptr;
    for (int attempt = 0; attempt < 2; ++attempt) {
        {
            std::lock_guard<std::recursive_mutex> lk(capture_guard());
            if (hipMalloc(&p, std::max<size_t>(bytes, 1)) == hipSuccess)
                break;
            (void)hipGetLastError();
            p = nullptr;
        }
        if (attempt == 0)
            cache_trim(dev);
    }
    if (!p)
        fail(SMIP_ERR_NO_MEMORY, "device allocation of " + std::to_string(bytes) + " bytes failed");
    std::lock_guard<std::mutex> lk(c.mu);
    c.dev_live[p] = { dev, std::max<size_t>(bytes, 1) };
    return p;
}

void dev_free(void* p)
{
    if (!p)
        return;
    BlockCache& c = BlockCache::get();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.dev_live.find(p);
        if (it != c.dev_live.end()) {
            const int dev = it->second.first;
            const size_t n = it->second.second;
            if (g_idle_release > 0 && c.dev_cached[dev] + n <= device_cap(c, dev)) {
                c.dev_free.emplace(std::make_pair(dev, n), p);
                c.dev_cached[dev] += n;
                return;
            }
            c.dev_live.erase(it);
        }
    }
    std::lock_guard<std::recursive_mutex> lk(capture_guard());
    (void)hipFree(p);
}

size_t cache_device_bytes(int dev)
{
    BlockCache& c = BlockCache::get();
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.dev_cached.find(dev);
    return it == c.dev_cached.end() ? 0 : it->second;
}

void* host_malloc(size_t bytes)
{
    BlockCache& c = BlockCache::get();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        if (void* p = BlockCache::take(c.host_free, bytes, bytes + bytes / 8)) {
            c.host_cached -= c.host_live[p];
            return p;
        }
    }
    std::lock_guard<std::recursive_mutex> lk(capture_guard());
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        fail(SMIP_ERR_NO_MEMORY, "pinned host allocation of " + std::to_string(bytes) + " bytes failed");
    }
    std::lock_guard<std::mutex> lk2(c.mu);
    c.host_live[p] = bytes;
    return p;
}

void host_free(void* p)
{
    if (!p)
        return;
    BlockCache& c = BlockCache::get();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.host_live.find(p);
        if (it != c.host_live.end()) {
            if (g_idle_release > 0 && c.host_cached + it->second <= BlockCache::kHostCap) {
                c.host_free.emplace(it->second, p);
                c.host_cached += it->second;
                return;
            }
            c.host_live.erase(it);
        }
    }
    std::lock_guard<std::recursive_mutex> lk(capture_guard());
    (void)hipHostFree(p);
}

// free memory as HIP reports it plus the blocks cached for reuse
hipError_t mem_get_info(size_t* free_b, size_t* total_b)
{
    hipError_t e;
    {
        std::lock_guard<std::recursive_mutex> lk(capture_guard());
        e = hipMemGetInfo(free_b, total_b);
    }
    if (e == hipSuccess) {
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess)
            *free_b += cache_device_bytes(dev);
    }
    return e;
}

// Stream pool.  hipStreamDestroy takes 0.6-2 ms, longer while the device is busy,
// and every job makes and releases up to four streams, so released ordinary streams
// (idle: synchronised, not capturing) are kept per device and handed out again.
// Streams with a CU mask (a hardware queue of their own) are destroyed.  Like the
// block cache, the pool is never destroyed: nothing touches HIP at process exit.
namespace {
struct StreamPool {
    std::mutex mu;
    std::unordered_map<hipStream_t, int> ordinary; // made by make_stream: its device
    std::map<int, std::vector<hipStream_t>> idle;  // per device
    static constexpr size_t kMaxIdle = 256;
    static StreamPool& get()
    {
        static auto* p = new StreamPool;
        return *p;
    }
};
} // namespace

hipStream_t make_stream()
{
    const int dev = current_device();
    StreamPool& P = StreamPool::get();
    {
        std::lock_guard<std::mutex> lk(P.mu);
        auto& v = P.idle[dev];
        if (!v.empty()) {
            hipStream_t s = v.back();
            v.pop_back();
            return s;
        }
    }
    hipStream_t s = nullptr;
    {
        std::lock_guard<std::recursive_mutex> lk(capture_guard());
        SMIP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    }
    std::lock_guard<std::mutex> lk(P.mu);
    P.ordinary[s] = dev;
    return s;
}

// A stream with a hardware queue of its own: HIP maps a process's ordinary streams
// onto a few shared queues (four by default), where a phase recovery's long chain
// of short launches holds up every other stream on its queue.  A stream created
// with a CU mask gets a queue of its own; the mask here names every CU.  Falls
// back to an ordinary stream where such a stream cannot be made.
hipStream_t make_queue_stream()
{
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess
        && prop.multiProcessorCount > 0) {
        const int ncu = prop.multiProcessorCount;
        std::vector<uint32_t> mask((size_t)(ncu + 31) / 32, 0xffffffffu);
        if (ncu % 32)
            mask.back() = (1u << (ncu % 32)) - 1u;
        std::lock_guard<std::recursive_mutex> lk(capture_guard());
        hipStream_t s = nullptr;
        if (hipExtStreamCreateWithCUMask(&s, (uint32_t)mask.size(), mask.data()) == hipSuccess)
            return s;
    }
    (void)hipGetLastError();
    return make_stream();
}

void destroy_stream(hipStream_t s)
{
    if (!s)
        return;
    const bool idle = hipStreamSynchronize(s) == hipSuccess;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    const bool plain = idle && hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone;
    StreamPool& P = StreamPool::get();
    {
        std::lock_guard<std::mutex> lk(P.mu);
        auto it = P.ordinary.find(s);
        if (it != P.ordinary.end()) {
            auto& v = P.idle[it->second];
            if (plain && v.size() < StreamPool::kMaxIdle) {
                v.push_back(s);
                return;
            }
            P.ordinary.erase(it);
        }
    }
    std::lock_guard<std::recursive_mutex> lk(capture_guard());
    (void)hipStreamDestroy(s);
}

void copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s)
{
    if (!bytes)
        return;
    {
        std::lock_guard<std::recursive_mutex> lk(capture_guard());
        SMIP_HIP(hipMemcpyAsync(dst, src, bytes, kind, s));
    }
    SMIP_HIP(hipStreamSynchronize(s));
}
} // namespace smip_amd

namespace {

void require(bool cond, int code, const char* msg)
{
    if (!cond)
        fail(code, msg);
}

template <typename T>
T* dev_alloc(size_t n)
{
    if (n == 0)
        return nullptr;
    return static_cast<T*>(dev_malloc(sizeof(T) * n));
}

struct DevFree {
    void operator()(void* p) const { dev_free(p); }
};
template <typename T>
using dptr = std::unique_ptr<T, DevFree>;

template <typename T>
dptr<T> dmake(size_t n) { return dptr<T>(dev_alloc<T>(n)); }

constexpr double k2pi = 3.14159265358979323846 * 2.0; // constants::c_2pi<double>

} // namespace

// ---------------------------------------------------------------------------
// objects
// ---------------------------------------------------------------------------
struct smip_bispec {
    CtxRef ctx_ref;
    smip_ctx* ctx = nullptr;
    Desc d {};
    int dtype = SMIP_DTYPE_C128;
    void* data = nullptr; // device, base_size elements
    bool borrowed = false; // reference-backed (a job's accumulator)
    size_t esz() const { return dtype == SMIP_DTYPE_C64 ? sizeof(float2) : sizeof(double2); }
    ~smip_bispec()
    {
        if (!borrowed && data)
            dev_free(data);
    }
};

struct smip_xcorr {
    CtxRef ctx_ref;
    smip_ctx* ctx = nullptr;
    int X = 0, Y = 0;
    FftPlan* plan = nullptr;
    dptr<double> ref, frame, corr;
    dptr<double2> ref_spec, spec, work;
    dptr<int32_t> shift_dev;
    int readiness = 0; // 0 none, 1 correl, 2 shift
    int32_t shift[2] = { 0, 0 };
    ~smip_xcorr() { fft_plan_destroy(plan); }
};

struct smip_frames {
    CtxRef ctx_ref;
    smip_ctx* ctx = nullptr;
    int dtype = 0;
    uint64_t X = 0, Y = 0, count = 0;
    void* data = nullptr;
    ~smip_frames()
    {
        if (data)
            dev_free(data);
    }
};

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    int kind = 0; // 0 accumulate, 1 saa, 2 reco, 3 synth, 4 triple-product kernel
};

struct smip_job {
    CtxRef ctx_ref;
    smip_ctx* ctx = nullptr;
    smip_job_params p {};
    hipStream_t stream = nullptr;
    Desc d {};
    LoopBounds L {};
    HalfPlane hp {};
    FftPlan* plan = nullptr;
    TilePlan* tile = nullptr;
    int batch = 0;
    dptr<double2> B64;
    dptr<double> P_half;
    dptr<float2> H32, work32;
    dptr<double2> H64, work64;
    // shift-and-add
    dptr<double2> ref_spec, swork;
    dptr<double> corr, sum;
    dptr<int32_t> shifts;
    uint64_t shifts_cap = 0;
    // frames from host memory: two device staging buffers of one batch each.  The
    // copy into one runs on `ustream` while the batch in the other is processed
    // on `stream`; ev_up[k] marks buffer k's copies done, ev_used[k] its batch done.
    dptr<uint8_t> upload[2];
    hipStream_t ustream = nullptr;
    hipEvent_t ev_up[2] = { nullptr, nullptr }, ev_used[2] = { nullptr, nullptr };
    int up_cur = 0;
    // host data reaches the device through two pinned slots; pin_done[k] marks
    // slot k's copy to the device done
    uint8_t* pin[2] = { nullptr, nullptr };
    hipEvent_t pin_done[2] = { nullptr, nullptr };
    size_t pin_bytes = 0;
    int pin_cur = 0;
    // buffered streaming (smip_job_push_frames_host): frames collect in
    // upload[up_cur] until a full batch
    uint64_t staged = 0;
    // per-batch launch sequences captured as HIP graphs (full batches from a
    // staging buffer, one graph each): accumulate, and shift-and-add into a
    // fixed shift scratch
    hipGraphExec_t g_acc[2] = { nullptr, nullptr }, g_saa[2] = { nullptr, nullptr };
    dptr<int32_t> bshift;
    uint64_t graph_launches = 0;
    // overlap: the forward FFT of batch k+1 runs on fstream into the other half of
    // a double-buffered spectrum (H32, H32b) while batch k's triple products run.
    // Device stacks overlap within a call; host batches of large jobs
    // (host_overlap) across calls.  hcur is the half the next FFT writes,
    // hused[b] whether ev_tile[b] marks a pending reader of half b.
    hipStream_t fstream = nullptr;
    dptr<float2> H32b;
    hipEvent_t ev_fft[2] = { nullptr, nullptr }, ev_tile[2] = { nullptr, nullptr }, ev_fork = nullptr;
    int hcur = 0;
    bool hused[2] = { false, false };
    bool host_overlap = false;
    // ... and the shift-and-add runs beside both on its own stream
    hipStream_t sstream = nullptr;
    hipEvent_t ev_sjoin = nullptr;
    // outputs
    dptr<double2> phases, image;
    dptr<uint8_t> flags;
    dptr<double> cons, scratch;
    uint64_t nframes = 0;
    bool finalized = false;
    // B64 has been divided by nframes: set by the first finalisation attempt, so a
    // retry after a failed phase recovery does not divide again, and no frame can
    // be added to a divided sum
    bool scaled = false;
    uint64_t reco_calls = 0;
    RecoWork* reco = nullptr; // phase-recovery scratch, made with the job
    std::unique_ptr<smip_bispec> view;
    std::vector<EventPair> events;
    double ms[5] = { 0, 0, 0, 0, 0 };
    uint64_t tile_launches = 0;
    uint64_t triples = 0;
    uint64_t footprint = 0;
    ~smip_job()
    {
        for (auto& e : events) {
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
        if (ustream != fstream)
            destroy_stream(ustream);
        destroy_stream(fstream);
        destroy_stream(sstream);
        for (hipEvent_t e : { ev_fft[0], ev_fft[1], ev_tile[0], ev_tile[1], ev_fork, ev_sjoin, ev_up[0], ev_up[1],
                 ev_used[0], ev_used[1], pin_done[0], pin_done[1] })
            if (e)
                (void)hipEventDestroy(e);
        for (int k = 0; k < 2; ++k) {
            if (g_acc[k])
                (void)hipGraphExecDestroy(g_acc[k]);
            if (g_saa[k])
                (void)hipGraphExecDestroy(g_saa[k]);
            host_free(pin[k]);
        }
        reco_work_destroy(reco);
        tile_plan_destroy(tile);
        fft_plan_destroy(plan);
        destroy_stream(stream);
    }
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
namespace smip_amd {
size_t dtype_size(int dtype)
{
    switch (dtype) {
    case SMIP_DTYPE_U8:
        return 1;
    case SMIP_DTYPE_U16:
        return 2;
    case SMIP_DTYPE_F32:
        return 4;
    case SMIP_DTYPE_F64:
        return 8;
    case SMIP_DTYPE_C64:
        return 8;
    case SMIP_DTYPE_C128:
        return 16;
    default:
        fail(SMIP_ERR_INVALID_ARGUMENT, "unknown dtype");
    }
}
// NextRecoIndex, src/phasereco.cpp:17-36
void next_reco_index(double& r, double& phi, int32_t& i, int32_t& j)
{
    constexpr double dr = 1.0;
    const double dphi = 1.0 / (k2pi * r);
    int32_t ii = i, jj = j;
    while (i == ii && j == jj) {
        if ((phi += dphi) > k2pi) {
            phi = 0;
            r += dr;
        }
        ii = static_cast<int32_t>(r * std::cos(phi));
        jj = static_cast<int32_t>(r * std::sin(phi));
    }
    i = ii;
    j = jj;
}

// Visits of reconstruct_phases' loop (include/phasereco.h:56-68) that fall in
// the PhaseMap range; the others are no-ops.
std::vector<int32_t> make_visit_list(double reco_radius, int64_t X, int64_t Y)
{
    const int32_t mn0 = -(int32_t)(X >> 1), mn1 = -(int32_t)(Y >> 1);
    const int32_t mx0 = -(int32_t)X / 2 + (int32_t)X - 1, mx1 = -(int32_t)Y / 2 + (int32_t)Y - 1;
    std::vector<int32_t> v;
    double r = 0.0, phi = 0.0;
    int32_t i = 0, j = 0;
    // r never decreases and |(i, j)| > r - 1.5, so once r exceeds the map's
    // half-diagonal by 2 no later visit is in range
    const double hd = std::hypot((double)std::max<int64_t>(-mn0, mx0), (double)std::max<int64_t>(-mn1, mx1));
    const double stop = std::min(reco_radius, hd + 2.0);
    while (r <= stop) {
        next_reco_index(r, phi, i, j);
        if (i >= mn0 && i <= mx0 && j >= mn1 && j <= mx1) {
            v.push_back(i);
            v.push_back(j);
        }
    }
    return v;
}

// The list depends only on (radius, extents) and costs tens of thousands of
// libm cos/sin steps: kept per process (at most 64 geometries).
std::shared_ptr<const std::vector<int32_t>> visit_list(double reco_radius, int64_t X, int64_t Y)
{
    static std::mutex mu;
    static std::map<std::tuple<double, int64_t, int64_t>, std::shared_ptr<const std::vector<int32_t>>> cache;
    const auto key = std::make_tuple(reco_radius, X, Y);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end())
            return it->second;
    }
    auto v = std::make_shared<const std::vector<int32_t>>(make_visit_list(reco_radius, X, Y));
    std::lock_guard<std::mutex> lk(mu);
    if (cache.size() >= 64)
        cache.clear();
    cache[key] = v;
    return v;
}

} // namespace smip_amd

namespace {

void check_reco_geometry(const Desc& d, uint64_t xsize, uint64_t ysize)
{
    require(xsize > 0 && ysize > 0, SMIP_ERR_INVALID_ARGUMENT, "phase map size must be positive");
    require(xsize * ysize <= (1u << 22), SMIP_ERR_INVALID_ARGUMENT, "phase map too large");
    // pm.at(v) must stay inside the PhaseMap for every v of the stored range
    require((uint64_t)std::max(d.max_idx[2], 0) < xsize && (uint64_t)std::max(d.max_idx[3], 0) < ysize,
        SMIP_ERR_INVALID_ARGUMENT, "bispectrum v range exceeds the phase map");
}

__global__ void k_scatter_half(const double2* __restrict__ full, int32_t X, int32_t Y, HalfPlane hp,
    double2* __restrict__ H)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)hp.hx * Y)
        return;
    const int32_t c = (int32_t)(i / Y), r = (int32_t)(i - (int64_t)c * Y);
    const int32_t ux = c - (X >> 1), uy = r - (Y >> 1);
    const int32_t col = ux < 0 ? ux + X : ux, row = uy < 0 ? uy + Y : uy;
    H[(int64_t)c * hp.rs + hp.pad + r] = full[(int64_t)row * X + col];
}

template <typename V>
__global__ void k_array_op(V* __restrict__ a, const V* __restrict__ b, uint64_t n, int op)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    V x = a[i];
    const V y = b[i];
    switch (op) {
    case SMIP_OP_ADD:
        x.x += y.x;
        x.y += y.y;
        break;
    case SMIP_OP_SUB:
        x.x -= y.x;
        x.y -= y.y;
        break;
    case SMIP_OP_MUL: {
        const auto re = x.x * y.x - x.y * y.y, im = x.x * y.y + x.y * y.x;
        x.x = re;
        x.y = im;
        break;
    }
    default: {
        const auto den = y.x * y.x + y.y * y.y;
        const auto re = (x.x * y.x + x.y * y.y) / den, im = (x.y * y.x - x.x * y.y) / den;
        x.x = re;
        x.y = im;
    }
    }
    a[i] = x;
}

__global__ void k_frame_to_f64(const void* __restrict__ in, int dtype, int64_t n, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    double v;
    switch (dtype) {
    case SMIP_DTYPE_U8:
        v = (double)static_cast<const uint8_t*>(in)[i];
        break;
    case SMIP_DTYPE_U16:
        v = (double)static_cast<const uint16_t*>(in)[i];
        break;
    case SMIP_DTYPE_F32:
        v = (double)static_cast<const float*>(in)[i];
        break;
    default:
        v = static_cast<const double*>(in)[i];
    }
    out[i] = v;
}


} // namespace

// ---------------------------------------------------------------------------
// Scratch of smip_bispec_get_elements_noalloc: one copy of this array per GPU,
// held until the process ends.  Laid out as the staged lookups (4 x int32 each),
// their values (double2), their statuses (int32) and the first failure
// (int32 index, int32 code).  Host code must not pass `scratch` itself to a
// kernel or a copy (in host code the name is a stand-in, not a device address):
// it uses the address hipGetSymbolAddress returned for the current device.
// ---------------------------------------------------------------------------
namespace {
constexpr uint64_t kNoAllocLookups = (uint64_t)1 << 16;
constexpr size_t kNaIdx = 0;
constexpr size_t kNaOut = kNaIdx + sizeof(int32_t) * 4 * kNoAllocLookups;
constexpr size_t kNaStatus = kNaOut + sizeof(double2) * kNoAllocLookups;
constexpr size_t kNaFirst = kNaStatus + sizeof(int32_t) * kNoAllocLookups;
constexpr size_t kNaBytes = kNaFirst + 256;
static_assert(kNaOut % 256 == 0 && kNaStatus % 256 == 0 && kNaFirst % 256 == 0, "scratch parts 256-byte aligned");
} // namespace

static __device__ __align__(256) uint8_t scratch[kNaBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch`, per device
std::mutex g_scratch_use[kMaxDevices];        // one get_elements_noalloc call per device at a time

// Resolves this device's copy of `scratch` (the device must be current); called
// by smip_ctx_create, once per device.
void noalloc_scratch_setup(int device)
{
    require(device >= 0 && device < kMaxDevices, SMIP_ERR_INVALID_ARGUMENT, "device index beyond the library's table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device])
        return;
    void* ptr = nullptr;
    SMIP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    require(ptr != nullptr, SMIP_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = static_cast<uint8_t*>(ptr);
}

uint8_t* noalloc_scratch(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    uint8_t* p = (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
    require(p != nullptr, SMIP_ERR_LOGIC, "no scratch address for this device (no context was created on it)");
    return p;
}

// first[0] = lowest index i < n with status[i] != 0 (n if none), first[1] = its
// status; one workgroup.
__global__ __launch_bounds__(1024) void k_first_failure(const int32_t* __restrict__ status, int64_t n,
    int32_t* __restrict__ first)
{
    __shared__ int32_t lo[16];
    int32_t m = (int32_t)n;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
        if (status[i] != 0) {
            m = (int32_t)i;
            break; // later entries of this thread's stride have larger indices
        }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        m = min(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0)
        lo[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k)
            m = min(m, lo[k]);
        first[0] = m;
        first[1] = m < n ? status[m] : 0;
    }
}
} // namespace

// ---------------------------------------------------------------------------
// misc
// ---------------------------------------------------------------------------
extern "C" SMIP_API const char* smip_last_error(void) { return last_error_text(); }

extern "C" SMIP_API int smip_version(int32_t* major, int32_t* minor, int32_t* patch)
{
    return guarded([&] {
        require(major && minor && patch, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        *major = 0;
        *minor = 3;
        *patch = 2;
    });
}

extern "C" SMIP_API int smip_ctx_create(int32_t device, smip_ctx** out)
{
    return guarded([&] {
        require(out != nullptr, SMIP_ERR_INVALID_ARGUMENT, "null output pointer");
        int n = 0;
        SMIP_HIP(hipGetDeviceCount(&n));
        require(device >= 0 && device < n, SMIP_ERR_INVALID_ARGUMENT, "no such HIP device");
        auto c = std::make_unique<smip_ctx>();
        c->device = device;
        SMIP_HIP(hipSetDevice(device));
        noalloc_scratch_setup(device);
        c->stream = make_stream();
        *out = c.release();
    });
}

extern "C" SMIP_API int smip_ctx_destroy(smip_ctx* ctx)
{
    return guarded([&] {
        if (!ctx)
            return;
        const int64_t live = ctx->live.load();
        if (live > 0)
            fail(SMIP_ERR_LOGIC, "context still has " + std::to_string(live)
                    + " live objects (bispectra, frame stacks, cross-correlations, jobs or servers); destroy them first");
        destroy_stream(ctx->stream);
        cache_trim(ctx->device);
        delete ctx;
    });
}

extern "C" SMIP_API int smip_ctx_synchronize(smip_ctx* ctx)
{
    return guarded([&] {
        require(ctx != nullptr, SMIP_ERR_INVALID_ARGUMENT, "null context");
        ensure_device(ctx);
        SMIP_HIP(hipDeviceSynchronize());
    });
}

extern "C" SMIP_API int smip_ctx_mem_info(smip_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes)
{
    return guarded([&] {
        require(ctx && free_bytes && total_bytes, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        ensure_device(ctx);
        size_t f = 0, t = 0;
        SMIP_HIP(mem_get_info(&f, &t));
        *free_bytes = f;
        *total_bytes = t;
    });
}

// ---------------------------------------------------------------------------
// Bispectrum
// ---------------------------------------------------------------------------
extern "C" SMIP_API int smip_bispec_create(smip_ctx* ctx, const uint64_t dims[4], int32_t dtype, smip_bispec** out)
{
    return guarded([&] {
        require(ctx && dims && out, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(dtype == SMIP_DTYPE_C64 || dtype == SMIP_DTYPE_C128, SMIP_ERR_INVALID_ARGUMENT,
            "bispectrum dtype must be complex64 or complex128");
        ensure_device(ctx);
        auto b = std::make_unique<smip_bispec>();
        b->ctx_ref.hold(ctx);
        b->ctx = ctx;
        b->dtype = dtype;
        if (dims[0] || dims[1] || dims[2] || dims[3])
            b->d = make_desc(dims[0], dims[1], dims[2], dims[3]);
        else
            b->d = Desc {}; // default-constructed: every descriptor field zero
        if (dims[0] == 0 && dims[1] == 0 && dims[2] == 0 && dims[3] == 0)
            b->d.base_size = 0;
        if (b->d.base_size) {
            b->data = dev_alloc<uint8_t>(b->d.base_size * b->esz());
            SMIP_HIP(hipMemsetAsync(b->data, 0, b->d.base_size * b->esz(), ctx->stream));
            SMIP_HIP(hipStreamSynchronize(ctx->stream));
        }
        *out = b.release();
    });
}

extern "C" SMIP_API int smip_bispec_destroy(smip_bispec* b)
{
    return guarded([&] {
        if (!b)
            return;
        require(!b->borrowed, SMIP_ERR_LOGIC, "a job's bispectrum is owned by the job");
        delete b;
    });
}

extern "C" SMIP_API int smip_bispec_info_get(const smip_bispec* b, smip_bispec_info* out)
{
    return guarded([&] {
        require(b && out, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        for (int k = 0; k < 4; ++k) {
            out->dims[k] = b->d.dims[k];
            out->sizes[k] = b->d.sizes[k];
            out->base_sizes[k] = b->d.base_sizes[k];
            out->min_indices[k] = b->d.min_idx[k];
            out->max_indices[k] = b->d.max_idx[k];
        }
        out->base_size = b->d.base_size;
        out->totalsize = b->d.totalsize;
        out->dtype = b->dtype;
    });
}

extern "C" SMIP_API int smip_bispec_accumulate_from_fft(smip_bispec* b, const double* fft, uint64_t ncols,
    uint64_t nrows)
{
    return guarded([&] {
        require(b && fft, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(ncols > 0 && nrows > 0 && ncols <= 65536 && nrows <= 65536, SMIP_ERR_INVALID_ARGUMENT,
            "fft extents out of range");
        if (b->d.base_size == 0)
            return; // the loops of an empty bispectrum are empty
        ensure_device(b->ctx);
        hipStream_t s = b->ctx->stream;
        const int32_t X = (int32_t)ncols, Y = (int32_t)nrows;
        const HalfPlane hp = make_half_plane(X, Y, 0, 0);
        auto full = dmake<double2>(ncols * nrows);
        auto H = dmake<double2>((size_t)hp.fs);
        SMIP_HIP(hipMemcpyAsync(full.get(), fft, sizeof(double2) * ncols * nrows, hipMemcpyHostToDevice, s));
        const int64_t n = (int64_t)hp.hx * Y;
        hipLaunchKernelGGL(k_scatter_half, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, full.get(), X, Y, hp,
            H.get());
        SMIP_HIP(hipGetLastError());
        const LoopBounds L = loop_bounds(b->d, X, Y);
        if (b->dtype == SMIP_DTYPE_C128)
            launch_accum_plain_f64(b->d, L, hp, H.get(), static_cast<double2*>(b->data), s);
        else
            launch_accum_plain_f32(b->d, L, hp, H.get(), static_cast<float2*>(b->data), s);
        SMIP_HIP(hipStreamSynchronize(s));
    });
}

extern "C" SMIP_API int smip_bispec_get_elements(const smip_bispec* b, const int32_t* indices, uint64_t n,
    double* values)
{
    return guarded([&] {
        require(b && (n == 0 || (indices && values)), SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        if (n == 0)
            return;
        ensure_device(b->ctx);
        hipStream_t s = b->ctx->stream;
        if (b->d.base_size == 0) {
            // every lookup of an empty bispectrum ends at addr >= base_size (0)
            fail(SMIP_ERR_OUT_OF_RANGE, "Bispectrum: element address out of bounds.");
        }
        auto idx = dmake<int32_t>(4 * n);
        auto out = dmake<double2>(n);
        auto st = dmake<int32_t>(n);
        SMIP_HIP(hipMemcpyAsync(idx.get(), indices, sizeof(int32_t) * 4 * n, hipMemcpyHostToDevice, s));
        launch_get_elements(b->d, b->data, b->dtype == SMIP_DTYPE_C64, idx.get(), (int64_t)n, out.get(), st.get(),
            s);
        std::vector<int32_t> hst(n);
        SMIP_HIP(hipMemcpyAsync(values, out.get(), sizeof(double2) * n, hipMemcpyDeviceToHost, s));
        SMIP_HIP(hipMemcpyAsync(hst.data(), st.get(), sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
        SMIP_HIP(hipStreamSynchronize(s));
        for (uint64_t i = 0; i < n; ++i)
            if (hst[i]) {
                const int32_t* q = indices + 4 * i;
                fail(hst[i], "Bispectrum: element access out of bounds: indices = [" + std::to_string(q[0]) + ", "
                        + std::to_string(q[1]) + ", " + std::to_string(q[2]) + ", " + std::to_string(q[3]) + "]");
            }
    });
}

extern "C" SMIP_API int smip_bispec_get_elements_noalloc(const smip_bispec* b, const int32_t* indices, uint64_t n,
    double* values)
{
    return guarded([&] {
        require(b && (n == 0 || (indices && values)), SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        if (n > kNoAllocLookups)
            fail(SMIP_ERR_INVALID_ARGUMENT, "smip_bispec_get_elements_noalloc takes at most "
                    + std::to_string(kNoAllocLookups) + " lookups per call, got " + std::to_string(n));
        if (n == 0)
            return;
        if (b->d.base_size == 0)
            fail(SMIP_ERR_OUT_OF_RANGE, "Bispectrum: element address out of bounds.");
        ensure_device(b->ctx);
        const int dev = b->ctx->device;
        uint8_t* base = noalloc_scratch(dev);
        int32_t* idx = reinterpret_cast<int32_t*>(base + kNaIdx);
        double2* out = reinterpret_cast<double2*>(base + kNaOut);
        int32_t* st = reinterpret_cast<int32_t*>(base + kNaStatus);
        int32_t* first = reinterpret_cast<int32_t*>(base + kNaFirst);
        hipStream_t s = b->ctx->stream;
        int32_t bad[2] = { 0, 0 };
        {
            // calls on one device share its scratch: one at a time
            std::lock_guard<std::mutex> use(g_scratch_use[dev]);
            SMIP_HIP(hipMemcpyAsync(idx, indices, sizeof(int32_t) * 4 * n, hipMemcpyHostToDevice, s));
            launch_get_elements(b->d, b->data, b->dtype == SMIP_DTYPE_C64, idx, (int64_t)n, out, st, s);
            hipLaunchKernelGGL(k_first_failure, dim3(1), dim3(1024), 0, s, st, (int64_t)n, first);
            SMIP_HIP(hipGetLastError());
            SMIP_HIP(hipMemcpyAsync(values, out, sizeof(double2) * n, hipMemcpyDeviceToHost, s));
            SMIP_HIP(hipMemcpyAsync(bad, first, sizeof(bad), hipMemcpyDeviceToHost, s));
            SMIP_HIP(hipStreamSynchronize(s));
        }
        if ((uint64_t)bad[0] < n) {
            const int32_t* q = indices + 4 * (uint64_t)bad[0];
            fail(bad[1], "Bispectrum: element access out of bounds: indices = [" + std::to_string(q[0]) + ", "
                    + std::to_string(q[1]) + ", " + std::to_string(q[2]) + ", " + std::to_string(q[3]) + "]");
        }
    });
}

extern "C" SMIP_API uint64_t smip_bispec_calc_offset(const smip_bispec* b, const int32_t idx[4])
{
    if (!b || !idx)
        return 0;
    return calc_offset(b->d, idx[0], idx[1], idx[2], idx[3]);
}

extern "C" SMIP_API int smip_bispec_calc_indices(const smip_bispec* b, uint64_t addr, int32_t idx[4])
{
    return guarded([&] {
        require(b && idx, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(b->d.base_size > 0 && addr < b->d.base_size, SMIP_ERR_OUT_OF_RANGE, "address out of range");
        // include/bispectrum.h:280-316
        const Desc& d = b->d;
        uint64_t temp = d.base_size / d.base_sizes[0], rest = addr;
        const uint64_t a0 = rest / temp;
        rest -= temp * a0;
        temp = d.base_sizes[2] * d.base_sizes[3];
        const uint64_t a1 = rest / temp;
        rest -= temp * a1;
        temp = d.base_sizes[3];
        const uint64_t a2 = rest / temp;
        rest -= temp * a2;
        int32_t v[4] = { (int32_t)a0, (int32_t)a1, (int32_t)a2, (int32_t)rest };
        const int sign[4] = { -1, 1, -1, 1 };
        for (int k = 0; k < 4; ++k) {
            if (v[k] > (int32_t)d.sizes[k] / 2)
                v[k] = (int32_t)((uint64_t)(int64_t)v[k] - d.sizes[k]);
            idx[k] = v[k] * sign[k];
        }
    });
}

extern "C" SMIP_API int smip_bispec_put_element(smip_bispec* b, const int32_t idx[4], const double value[2])
{
    return guarded([&] {
        require(b && idx && value, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        const uint64_t addr = calc_offset(b->d, idx[0], idx[1], idx[2], idx[3]);
        if (addr >= b->d.base_size)
            fail(SMIP_ERR_OUT_OF_RANGE, "Bispectrum: put_element address out of bounds.");
        ensure_device(b->ctx);
        if (b->dtype == SMIP_DTYPE_C128) {
            double2 v = make_double2(value[0], value[1]);
            copy_sync(static_cast<double2*>(b->data) + addr, &v, sizeof(v), hipMemcpyHostToDevice, b->ctx->stream);
        } else {
            float2 v = make_float2((float)value[0], (float)value[1]);
            copy_sync(static_cast<float2*>(b->data) + addr, &v, sizeof(v), hipMemcpyHostToDevice, b->ctx->stream);
        }
    });
}

extern "C" SMIP_API int smip_bispec_download(const smip_bispec* b, void* host)
{
    return guarded([&] {
        require(b && (host || b->d.base_size == 0), SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        if (!b->d.base_size)
            return;
        ensure_device(b->ctx);
        SMIP_HIP(hipStreamSynchronize(b->ctx->stream));
        copy_sync(host, b->data, b->d.base_size * b->esz(), hipMemcpyDeviceToHost, b->ctx->stream);
    });
}

extern "C" SMIP_API int smip_bispec_upload(smip_bispec* b, const void* host)
{
    return guarded([&] {
        require(b && (host || b->d.base_size == 0), SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        if (!b->d.base_size)
            return;
        ensure_device(b->ctx);
        SMIP_HIP(hipStreamSynchronize(b->ctx->stream));
        copy_sync(b->data, host, b->d.base_size * b->esz(), hipMemcpyHostToDevice, b->ctx->stream);
    });
}

extern "C" SMIP_API int smip_bispec_write_file(const smip_bispec* b, const char* path)
{
    return guarded([&] {
        require(b && path, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        std::vector<uint8_t> buf(b->d.base_size * b->esz());
        if (b->d.base_size) {
            ensure_device(b->ctx);
            SMIP_HIP(hipStreamSynchronize(b->ctx->stream));
            copy_sync(buf.data(), b->data, buf.size(), hipMemcpyDeviceToHost, b->ctx->stream);
        }
        // header: size_t base_size, 4 x size_t constructor dims; then T[base_size]
        FILE* f = std::fopen(path, "wb");
        if (!f)
            fail(SMIP_ERR_RUNTIME, std::string("unable to open file ") + path + " for writing");
        const uint64_t hdr[5] = { b->d.base_size, b->d.dims[0], b->d.dims[1], b->d.dims[2], b->d.dims[3] };
        std::fwrite(hdr, sizeof(uint64_t), 5, f);
        if (!buf.empty())
            std::fwrite(buf.data(), 1, buf.size(), f);
        std::fclose(f);
    });
}

extern "C" SMIP_API int smip_bispec_read_file(smip_bispec* b, const char* path)
{
    return guarded([&] {
        require(b && path, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        FILE* f = std::fopen(path, "rb");
        if (!f)
            fail(SMIP_ERR_RUNTIME, std::string("unable to open file ") + path);
        uint64_t hdr[5];
        if (std::fread(hdr, sizeof(uint64_t), 5, f) != 5) {
            std::fclose(f);
            fail(SMIP_ERR_RUNTIME, std::string("error reading bispectrum metadata from file ") + path);
        }
        if (b->borrowed) {
            std::fclose(f);
            fail(SMIP_ERR_LOGIC, "Cannot resize a reference-backed Array_base.");
        }
        const uint64_t size = hdr[0];
        std::vector<uint8_t> buf(size * b->esz());
        const size_t got = buf.empty() ? 0 : std::fread(buf.data(), b->esz(), size, f);
        std::fclose(f);
        ensure_device(b->ctx);
        // Array_base::resize(size), then the descriptor of the stored dims
        if (size != b->d.base_size) {
            if (b->data)
                dev_free(b->data);
            b->data = size ? dev_alloc<uint8_t>(size * b->esz()) : nullptr;
        }
        const bool zero = !(hdr[1] || hdr[2] || hdr[3] || hdr[4]);
        b->d = zero ? Desc {} : make_desc(hdr[1], hdr[2], hdr[3], hdr[4]);
        b->d.base_size = size;
        if (got != size)
            fail(SMIP_ERR_RUNTIME, std::string("error reading data block from file ") + path);
        if (size)
            copy_sync(b->data, buf.data(), buf.size(), hipMemcpyHostToDevice, b->ctx->stream);
    });
}

extern "C" SMIP_API int smip_bispec_scalar_op(smip_bispec* b, int32_t op, double re, double im)
{
    return guarded([&] {
        require(b != nullptr, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(op >= SMIP_OP_ADD && op <= SMIP_OP_DIV, SMIP_ERR_INVALID_ARGUMENT, "unknown operator");
        const bool f32 = b->dtype == SMIP_DTYPE_C64;
        const bool zero = f32 ? ((float)re == 0.f && (float)im == 0.f) : (re == 0.0 && im == 0.0);
        if (op == SMIP_OP_DIV && zero)
            fail(SMIP_ERR_RUNTIME, "Array_base<T>::operator/=(T) : division by zero");
        ensure_device(b->ctx);
        if (f32)
            launch_scale_c32(static_cast<float2*>(b->data), b->d.base_size, (float)re, (float)im, op,
                b->ctx->stream);
        else
            launch_scale_c64(static_cast<double2*>(b->data), b->d.base_size, re, im, op, b->ctx->stream);
        SMIP_HIP(hipStreamSynchronize(b->ctx->stream));
    });
}

extern "C" SMIP_API int smip_bispec_array_op(smip_bispec* b, const smip_bispec* o, int32_t op)
{
    return guarded([&] {
        require(b && o, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(op >= SMIP_OP_ADD && op <= SMIP_OP_DIV, SMIP_ERR_INVALID_ARGUMENT, "unknown operator");
        if (std::memcmp(b->d.dims, o->d.dims, sizeof(b->d.dims)) != 0)
            fail(SMIP_ERR_INVALID_ARGUMENT, "Bispectrum::operator+=(const Bispectrum) : operand dimension size mismatch");
        require(b->dtype == o->dtype, SMIP_ERR_INVALID_ARGUMENT, "bispectrum element types differ");
        if (!b->d.base_size)
            return;
        ensure_device(b->ctx);
        const uint64_t n = b->d.base_size;
        dim3 g((unsigned)((n + 255) / 256));
        if (b->dtype == SMIP_DTYPE_C128)
            hipLaunchKernelGGL(k_array_op<double2>, g, dim3(256), 0, b->ctx->stream, static_cast<double2*>(b->data),
                static_cast<const double2*>(o->data), n, op);
        else
            hipLaunchKernelGGL(k_array_op<float2>, g, dim3(256), 0, b->ctx->stream, static_cast<float2*>(b->data),
                static_cast<const float2*>(o->data), n, op);
        SMIP_HIP(hipGetLastError());
        SMIP_HIP(hipStreamSynchronize(b->ctx->stream));
    });
}

extern "C" SMIP_API int smip_bispec_assign(smip_bispec* dst, const smip_bispec* src)
{
    return guarded([&] {
        require(dst && src, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        if (dst == src)
            return;
        ensure_device(dst->ctx);
        if (dst->d.base_size != src->d.base_size) {
            require(!dst->borrowed, SMIP_ERR_LOGIC, "Cannot resize a reference-backed Array_base.");
            if (dst->data)
                dev_free(dst->data);
            dst->data = nullptr;
            dst->data = src->d.base_size ? dev_alloc<uint8_t>(src->d.base_size * dst->esz()) : nullptr;
        }
        dst->d = src->d;
        if (src->d.base_size) {
            hipStream_t s = dst->ctx->stream;
            SMIP_HIP(hipStreamSynchronize(src->ctx->stream));
            const uint64_t n = src->d.base_size;
            if (dst->dtype == src->dtype)
                SMIP_HIP(hipMemcpyAsync(dst->data, src->data, n * src->esz(), hipMemcpyDeviceToDevice, s));
            else if (dst->dtype == SMIP_DTYPE_C64) // elements converted, dst keeps its type
                launch_convert_c64_to_c32(static_cast<const double2*>(src->data), static_cast<float2*>(dst->data), n,
                    s);
            else
                launch_convert_c32_to_c64(static_cast<const float2*>(src->data), static_cast<double2*>(dst->data), n,
                    s);
            SMIP_HIP(hipStreamSynchronize(s));
        }
    });
}

// ---------------------------------------------------------------------------
// phase reconstruction
// ---------------------------------------------------------------------------
extern "C" SMIP_API void smip_next_reco_index(double* r, double* phi, int32_t* i, int32_t* j)
{
    if (r && phi && i && j)
        next_reco_index(*r, *phi, *i, *j);
}

static void reconstruct_into(smip_ctx* ctx, hipStream_t s, const Desc& d, const void* B, int is_f32, uint64_t xs,
    uint64_t ys, double reco_radius, double2* ph, uint8_t* fl, double* co, uint64_t* calls)
{
    check_reco_geometry(d, xs, ys);
    (void)ctx;
    std::unique_ptr<RecoWork, void (*)(RecoWork*)> w(
        reco_work_create(d, reco_geom(reco_radius, (int32_t)xs, (int32_t)ys, s)), reco_work_destroy);
    const int c = run_reconstruct(w.get(), B, is_f32, ph, fl, co, s);
    {
        IdleRelease idle; // run_reconstruct synchronised s
        w.reset();
    }
    if (calls)
        *calls = (uint64_t)c;
}

extern "C" SMIP_API int smip_reconstruct_phases(const smip_bispec* b, uint64_t xsize, uint64_t ysize,
    double reco_radius, double* phases, uint8_t* flags, double* consistency)
{
    return guarded([&] {
        require(b != nullptr, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        ensure_device(b->ctx);
        hipStream_t s = b->ctx->stream;
        const uint64_t n = xsize * ysize;
        require(b->d.base_size > 0 || reco_radius < 1.0, SMIP_ERR_OUT_OF_RANGE,
            "Bispectrum: element address out of bounds.");
        auto ph = dmake<double2>(n);
        auto fl = dmake<uint8_t>(n);
        auto co = dmake<double>(n);
        reconstruct_into(b->ctx, s, b->d, b->data, b->dtype == SMIP_DTYPE_C64, xsize, ysize, reco_radius, ph.get(),
            fl.get(), co.get(), nullptr);
        if (phases)
            SMIP_HIP(hipMemcpyAsync(phases, ph.get(), sizeof(double2) * n, hipMemcpyDeviceToHost, s));
        if (flags)
            SMIP_HIP(hipMemcpyAsync(flags, fl.get(), n, hipMemcpyDeviceToHost, s));
        if (consistency)
            SMIP_HIP(hipMemcpyAsync(consistency, co.get(), sizeof(double) * n, hipMemcpyDeviceToHost, s));
        SMIP_HIP(hipStreamSynchronize(s));
    });
}

extern "C" SMIP_API int smip_calc_phase(const smip_bispec* b, uint64_t xsize, uint64_t ysize, double* phases,
    uint8_t* flags, double* consistency, int32_t wx, int32_t wy)
{
    return guarded([&] {
        require(b && phases && flags && consistency, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        check_reco_geometry(b->d, xsize, ysize);
        ensure_device(b->ctx);
        hipStream_t s = b->ctx->stream;
        const uint64_t n = xsize * ysize;
        auto ph = dmake<double2>(n);
        auto fl = dmake<uint8_t>(n);
        auto co = dmake<double>(n);
        SMIP_HIP(hipMemcpyAsync(ph.get(), phases, sizeof(double2) * n, hipMemcpyHostToDevice, s));
        SMIP_HIP(hipMemcpyAsync(fl.get(), flags, n, hipMemcpyHostToDevice, s));
        SMIP_HIP(hipMemcpyAsync(co.get(), consistency, sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (b->d.base_size == 0)
            fail(SMIP_ERR_OUT_OF_RANGE, "Bispectrum: element address out of bounds.");
        run_calc_phase(b->d, (int32_t)xsize, (int32_t)ysize, b->data, b->dtype == SMIP_DTYPE_C64, ph.get(), fl.get(),
            co.get(), wx, wy, s);
        SMIP_HIP(hipMemcpyAsync(phases, ph.get(), sizeof(double2) * n, hipMemcpyDeviceToHost, s));
        SMIP_HIP(hipMemcpyAsync(flags, fl.get(), n, hipMemcpyDeviceToHost, s));
        SMIP_HIP(hipMemcpyAsync(consistency, co.get(), sizeof(double) * n, hipMemcpyDeviceToHost, s));
        SMIP_HIP(hipStreamSynchronize(s));
    });
}

// ---------------------------------------------------------------------------
// CrossCorrelation<double>
// ---------------------------------------------------------------------------
extern "C" SMIP_API int smip_xcorr_create(smip_ctx* ctx, const double* ref, uint64_t ncols, uint64_t nrows,
    smip_xcorr** out)
{
    return guarded([&] {
        require(ctx && ref && out, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        ensure_device(ctx);
        auto x = std::make_unique<smip_xcorr>();
        x->ctx_ref.hold(ctx);
        x->ctx = ctx;
        x->X = (int)ncols;
        x->Y = (int)nrows;
        x->plan = fft_plan_create((int32_t)ncols, (int32_t)nrows, ctx->stream);
        const uint64_t np = ncols * nrows, nh = (ncols / 2 + 1) * nrows;
        x->ref = dmake<double>(np);
        x->frame = dmake<double>(np);
        x->corr = dmake<double>(np);
        x->ref_spec = dmake<double2>(nh);
        x->spec = dmake<double2>(nh);
        x->work = dmake<double2>(nh);
        x->shift_dev = dmake<int32_t>(2);
        hipStream_t s = ctx->stream;
        SMIP_HIP(hipMemcpyAsync(x->ref.get(), ref, sizeof(double) * np, hipMemcpyHostToDevice, s));
        launch_rfft_f64(x->plan, x->ref.get(), SMIP_DTYPE_F64, 1, false, x->ref_spec.get(), x->work.get(), s);
        SMIP_HIP(hipStreamSynchronize(s));
        *out = x.release();
    });
}

extern "C" SMIP_API int smip_xcorr_destroy(smip_xcorr* x)
{
    return guarded([&] { delete x; });
}

extern "C" SMIP_API int smip_xcorr_correlate(smip_xcorr* x, const double* frame, uint64_t ncols, uint64_t nrows)
{
    return guarded([&] {
        require(x && frame, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        if ((int)ncols != x->X || (int)nrows != x->Y)
            fail(SMIP_ERR_INVALID_ARGUMENT, "Matrix dimensions must match for correlation");
        ensure_device(x->ctx);
        hipStream_t s = x->ctx->stream;
        SMIP_HIP(hipMemcpyAsync(x->frame.get(), frame, sizeof(double) * ncols * nrows, hipMemcpyHostToDevice, s));
        launch_rfft_f64(x->plan, x->frame.get(), SMIP_DTYPE_F64, 1, false, x->spec.get(), x->work.get(), s);
        launch_xcorr_f64(x->plan, x->ref_spec.get(), x->spec.get(), 1, x->corr.get(), x->work.get(), s);
        SMIP_HIP(hipStreamSynchronize(s));
        x->readiness = 1;
    });
}

extern "C" SMIP_API int smip_xcorr_get_correlation(smip_xcorr* x, double* out)
{
    return guarded([&] {
        require(x && out, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        if (x->readiness < 1)
            fail(SMIP_ERR_BAD_FUNCTION_CALL, "bad_function_call");
        ensure_device(x->ctx);
        copy_sync(out, x->corr.get(), sizeof(double) * x->X * x->Y, hipMemcpyDeviceToHost, x->ctx->stream);
    });
}

extern "C" SMIP_API int smip_xcorr_get_displacement(smip_xcorr* x, int32_t shift[2])
{
    return guarded([&] {
        require(x && shift, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        if (x->readiness < 1)
            fail(SMIP_ERR_BAD_FUNCTION_CALL, "bad_function_call");
        if (x->readiness == 1) {
            ensure_device(x->ctx);
            hipStream_t s = x->ctx->stream;
            launch_argmax_shift(x->corr.get(), 1, x->X, x->Y, x->shift_dev.get(), s);
            SMIP_HIP(hipMemcpyAsync(x->shift, x->shift_dev.get(), sizeof(int32_t) * 2, hipMemcpyDeviceToHost, s));
            SMIP_HIP(hipStreamSynchronize(s));
            x->readiness = 2;
        }
        shift[0] = x->shift[0];
        shift[1] = x->shift[1];
    });
}

extern "C" SMIP_API int smip_xcorr_displacement(smip_xcorr* x, const double* frame, uint64_t ncols, uint64_t nrows,
    int32_t shift[2])
{
    if (x)
        x->readiness = 0;
    int rc = smip_xcorr_correlate(x, frame, ncols, nrows);
    if (rc)
        return rc;
    return smip_xcorr_get_displacement(x, shift);
}

// ---------------------------------------------------------------------------
// window
// ---------------------------------------------------------------------------
extern "C" SMIP_API int smip_general_hamming(uint64_t xsize, uint64_t ysize, double aperture, double alpha,
    double* out)
{
    return guarded([&] {
        require(out || xsize * ysize == 0, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        if (!(aperture > 0.))
            fail(SMIP_ERR_DOMAIN, "WindowFunction aperture size <= 0");
        const int32_t X = (int32_t)xsize, Y = (int32_t)ysize;
        const double trigarg = k2pi / aperture, c1 = alpha, c2 = 1. - alpha;
        for (int32_t i = -(X >> 1); i <= -X / 2 + X - 1; ++i)
            for (int32_t j = -(Y >> 1); j <= -Y / 2 + Y - 1; ++j) {
                const double t1 = c1 + c2 * std::cos(trigarg * static_cast<double>(i));
                const double t2 = c1 + c2 * std::cos(trigarg * static_cast<double>(j));
                const int32_t col = i < 0 ? i + X : i, row = j < 0 ? j + Y : j;
                out[(int64_t)row * X + col] = t1 * t2;
            }
    });
}

// ---------------------------------------------------------------------------
// frames
// ---------------------------------------------------------------------------
extern "C" SMIP_API int smip_frames_create(smip_ctx* ctx, int32_t dtype, uint64_t ncols, uint64_t nrows,
    uint64_t count, smip_frames** out)
{
    return guarded([&] {
        require(ctx && out, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(dtype >= SMIP_DTYPE_U8 && dtype <= SMIP_DTYPE_F64, SMIP_ERR_INVALID_ARGUMENT, "bad frame dtype");
        ensure_device(ctx);
        auto f = std::make_unique<smip_frames>();
        f->ctx_ref.hold(ctx);
        f->ctx = ctx;
        f->dtype = dtype;
        f->X = ncols;
        f->Y = nrows;
        f->count = count;
        f->data = dev_alloc<uint8_t>(count * ncols * nrows * dtype_size(dtype));
        *out = f.release();
    });
}

extern "C" SMIP_API int smip_frames_destroy(smip_frames* f)
{
    return guarded([&] { delete f; });
}

extern "C" SMIP_API int smip_frames_upload(smip_frames* f, uint64_t first, uint64_t count, const void* host)
{
    return guarded([&] {
        require(f && host, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(first + count <= f->count, SMIP_ERR_OUT_OF_RANGE, "frame range outside the stack");
        ensure_device(f->ctx);
        const size_t fb = f->X * f->Y * dtype_size(f->dtype);
        copy_sync(static_cast<uint8_t*>(f->data) + first * fb, host, count * fb, hipMemcpyHostToDevice, f->ctx->stream);
    });
}

// ---------------------------------------------------------------------------
// jobs
// ---------------------------------------------------------------------------
namespace {

// Frames per shift-and-add pass, which bounds its scratch: 16 Mi pixels per pass,
// 64 to 256 frames, so that passes over small frames still fill the GPU (at 128x128,
// 256-frame passes take a third less time per frame than 64-frame ones).
int saa_chunk(uint64_t ncols, uint64_t nrows, int batch)
{
    const uint64_t c = std::clamp<uint64_t>(((uint64_t)1 << 24) / (ncols * nrows), 64, 256);
    return std::min(batch, (int)c);
}

// Bytes of per-workgroup correlation maxima the shift-and-add writes per frame:
// nrows / 2 entries on the power-of-two path (two rows per transform), at most
// nrows (one row per transform) on the other.
uint64_t saa_best_bytes(uint64_t ncols, uint64_t nrows)
{
    auto pow2 = [](uint64_t n) { return n >= 8 && (n & (n - 1)) == 0; };
    const uint64_t entries = pow2(ncols) && pow2(nrows) ? nrows / 2 : nrows;
    return entries * sizeof(RowBest);
}

struct JobSizes {
    Desc d;
    LoopBounds L;
    HalfPlane hp;
    int vy_tile;
    int batch;
    uint64_t bytes;
};

JobSizes job_sizes(const smip_job_params& p)
{
    // any extents FFTW would take, up to the longest transform the LDS holds
    require(p.ncols >= 2 && p.nrows >= 2 && p.ncols <= (uint64_t)kMaxFftLength && p.nrows <= (uint64_t)kMaxFftLength,
        SMIP_ERR_INVALID_ARGUMENT, "job frame extents must be in [2, 2048]");
    // the accumulate loops clip v to the frame's range (include/bispectrum.h:383-391), so
    // b may exceed N; phase recovery evaluates a visit's (2h+1)^2 candidates in one
    // workgroup, at most 65536 of them: sizes(b) = 2(b/2)+1 <= 255
    require(p.bdepth >= 1 && p.bdepth <= 255, SMIP_ERR_INVALID_ARGUMENT,
        "bispectrum depth must satisfy 1 <= b <= 255 (phase recovery takes (2(b/2)+1)^2 <= 65536 candidates per visit)");
    require(p.frame_dtype == SMIP_DTYPE_U8 || p.frame_dtype == SMIP_DTYPE_U16 || p.frame_dtype == SMIP_DTYPE_F32,
        SMIP_ERR_INVALID_ARGUMENT, "frame dtype must be u8, u16 or f32");
    require(p.precision == SMIP_PRECISION_FP64 || p.precision == SMIP_PRECISION_MIXED, SMIP_ERR_INVALID_ARGUMENT,
        "unknown precision mode");
    require(p.frames_per_batch >= 0 && p.frames_per_batch <= 4096, SMIP_ERR_INVALID_ARGUMENT,
        "frames_per_batch out of range");
    JobSizes z;
    z.d = make_desc(p.ncols, p.nrows, p.bdepth, p.bdepth);
    z.L = loop_bounds(z.d, (int64_t)p.ncols, (int64_t)p.nrows);
    const bool mixed = p.precision == SMIP_PRECISION_MIXED;
    z.vy_tile = mixed ? (p.vy_tile > 0 ? p.vy_tile : 0) : 0;
    int tuned_vy = 0, tuned_batch = 0;
    const bool tuned = mixed && tune_lookup(p.ncols, p.nrows, p.bdepth, &tuned_vy, &tuned_batch);
    if (mixed && z.vy_tile == 0 && tuned)
        z.vy_tile = tuned_vy;
    if (mixed && z.vy_tile == 0) {
        z.vy_tile = tile_default_vy(z.L, (int)p.ncols, (int)p.nrows); // as tile_plan_create
    }
    const int pad = mixed ? tile_plan_pad_needed(z.d, z.vy_tile) : 0;
    const int tail = mixed ? tile_plan_tail_needed(z.d, (int)p.nrows, z.vy_tile) : 0;
    z.hp = make_half_plane((int32_t)p.ncols, (int32_t)p.nrows, pad, tail);
    z.batch = p.frames_per_batch > 0 ? p.frames_per_batch : (tuned ? tuned_batch : (mixed ? 512 : 16));
    const uint64_t X = p.ncols, Y = p.nrows, hx = X / 2 + 1, np = X * Y;
    uint64_t b = z.d.base_size * sizeof(double2) + hx * Y * sizeof(double);
    if (mixed) // two spectra buffers (FFT / tile overlap on device stacks) + FFT work
        b += (uint64_t)z.batch * (2 * (uint64_t)z.hp.fs * sizeof(float2) + hx * Y * sizeof(float2));
    else
        b += (uint64_t)z.batch * ((uint64_t)z.hp.fs * sizeof(double2) + hx * Y * sizeof(double2));
    if (p.saa)
        b += hx * Y * sizeof(double2)
            + (uint64_t)saa_chunk(X, Y, z.batch) * (hx * Y * sizeof(double2) + saa_best_bytes(X, Y))
            + np * sizeof(double) + (uint64_t)z.batch * 2 * sizeof(int32_t);
    b += 2 * (uint64_t)z.batch * np * dtype_size(p.frame_dtype);         // two host staging buffers
    b += np * (2 * sizeof(double2) + sizeof(uint8_t) + sizeof(double)); // outputs
    // phase-recovery scratch, and the geometry's visit list (shared between jobs of
    // one geometry, counted for each)
    b += reco_scratch_bytes(z.d, (int32_t)X, (int32_t)Y) + reco_geom_bytes_estimate(p.reco_radius, (int32_t)X, (int32_t)Y);
    z.bytes = b;
    return z;
}

void record(smip_job* j, int kind, hipEvent_t a, hipEvent_t b) { j->events.push_back({ a, b, kind }); }


hipEvent_t mark_on(smip_job* j, hipStream_t s)
{
    (void)j;
    hipEvent_t e;
    SMIP_HIP(hipEventCreate(&e));
    SMIP_HIP(hipEventRecord(e, s));
    return e;
}

hipEvent_t mark(smip_job* j) { return mark_on(j, j->stream); }

void drain_events(smip_job* j)
{
    for (auto& e : j->events) {
        float ms = 0.f;
        SMIP_HIP(hipEventSynchronize(e.b));
        SMIP_HIP(hipEventElapsedTime(&ms, e.a, e.b));
        j->ms[e.kind] += ms;
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    j->events.clear();
}

// FFT + |F|^2 + triple products of nf frames at fr (no timing events: also
// captured into a graph)
void enqueue_accumulate(smip_job* j, const uint8_t* fr, int64_t nf)
{
    const smip_job_params& p = j->p;
    hipStream_t s = j->stream;
    if (p.precision == SMIP_PRECISION_MIXED) {
        launch_forward_f32(j->plan, fr, p.frame_dtype, nf, false, j->hp, j->H32.get(), j->work32.get(),
            j->P_half.get(), s);
        launch_accum_tile(j->tile, j->hp, j->H32.get(), nf, j->B64.get(), s);
    } else {
        launch_forward_f64(j->plan, fr, p.frame_dtype, nf, false, j->hp, j->H64.get(), j->work64.get(),
            j->P_half.get(), s);
        for (int64_t f = 0; f < nf; ++f) // one frame per launch: the plain drop-in
            launch_accum_plain_f64(j->d, j->L, j->hp, j->H64.get() + f * j->hp.fs, j->B64.get(), s);
    }
}

// shift-and-add of m frames at fm (none of them frame 0): shifts into sh; in
// chunks of at most saa_chunk frames, added in frame order
void enqueue_saa(smip_job* j, const uint8_t* fm, int64_t m, int32_t* sh, hipStream_t s)
{
    const smip_job_params& p = j->p;
    const int32_t X = (int32_t)p.ncols, Y = (int32_t)p.nrows;
    const size_t fb = p.ncols * p.nrows * dtype_size(p.frame_dtype);
    const int64_t chunk = saa_chunk(p.ncols, p.nrows, j->batch);
    for (int64_t o = 0; o < m; o += chunk) {
        const int64_t c = std::min<int64_t>(chunk, m - o);
        const uint8_t* f = fm + o * fb;
        // the correlation buffer (c N^2 doubles) holds the per-workgroup maxima
        launch_shift_f64(j->plan, f, p.frame_dtype, c, true, j->ref_spec.get(),
            reinterpret_cast<RowBest*>(j->corr.get()), j->swork.get(), sh + 2 * o, s);
        launch_saa_add(f, p.frame_dtype, c, X, Y, sh + 2 * o, j->sum.get(), s);
    }
}

// hipGraphLaunch under the library's capture lock, as capture_graph instantiates
void graph_launch(hipGraphExec_t g, hipStream_t s)
{
    std::lock_guard<std::recursive_mutex> lk(capture_guard());
    SMIP_HIP(hipGraphLaunch(g, s));
}

bool graphs_enabled()
{
    const char* e = std::getenv("SMIP_NO_GRAPH");
    return !(e && e[0] == '1');
}

template <typename F>
hipGraphExec_t capture_graph(hipStream_t s, F&& enqueue)
{
    hipGraph_t g = nullptr;
    {
        // no other thread of this library touches the legacy stream or allocates meanwhile
        std::lock_guard<std::recursive_mutex> lk(capture_guard());
        SMIP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            enqueue();
        } catch (...) {
            if (hipStreamEndCapture(s, &g) == hipSuccess && g)
                (void)hipGraphDestroy(g);
            throw;
        }
        SMIP_HIP(hipStreamEndCapture(s, &g));
    }
    hipGraphExec_t ex = nullptr;
    hipError_t e;
    {
        // graph instantiation and launches are serialised (graph_launch): a host
        // SIGSEGV inside hipGraphLaunch was once recorded right after another
        // thread's instantiation
        std::lock_guard<std::recursive_mutex> lk(capture_guard());
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
    }
    SMIP_HIP(e);
    return ex;
}

void enable_overlap(smip_job* j)
{
    if (j->fstream)
        return;
    if (!j->H32b) { // made with the job for mixed precision; padding rows zeroed there
        j->H32b = dmake<float2>((size_t)j->batch * j->hp.fs);
        SMIP_HIP(hipMemsetAsync(j->H32b.get(), 0, sizeof(float2) * j->batch * j->hp.fs, j->stream));
    }
    for (int k = 0; k < 2; ++k) {
        SMIP_HIP(hipEventCreateWithFlags(&j->ev_fft[k], hipEventDisableTiming));
        SMIP_HIP(hipEventCreateWithFlags(&j->ev_tile[k], hipEventDisableTiming));
    }
    SMIP_HIP(hipEventCreateWithFlags(&j->ev_fork, hipEventDisableTiming));
    j->fstream = make_stream();
    if (j->p.saa) {
        SMIP_HIP(hipEventCreateWithFlags(&j->ev_sjoin, hipEventDisableTiming));
        j->sstream = make_stream();
    }
    // fstream starts after everything so far on the job stream (padding rows,
    // earlier batches in H32)
    SMIP_HIP(hipEventRecord(j->ev_fork, j->stream));
    SMIP_HIP(hipStreamWaitEvent(j->fstream, j->ev_fork, 0));
}

// A batch on the job stream alone used half 0 (H32): a later overlapped FFT
// into it waits for that batch (before overlap is enabled, its initial fork
// orders it).
void note_h32_reader(smip_job* j)
{
    if (!j->fstream)
        return;
    SMIP_HIP(hipEventRecord(j->ev_tile[0], j->stream));
    j->hused[0] = true;
}

// which staging buffer `frames` is (-1: none, frames from a device stack)
int upload_index(const smip_job* j, const void* frames)
{
    for (int k = 0; k < 2; ++k)
        if (j->upload[k] && frames == j->upload[k].get())
            return k;
    return -1;
}

void job_process(smip_job* j, const void* frames, uint64_t count)
{
    const smip_job_params& p = j->p;
    const size_t fb = p.ncols * p.nrows * dtype_size(p.frame_dtype);
    const int32_t X = (int32_t)p.ncols, Y = (int32_t)p.nrows;
    hipStream_t s = j->stream;
    const bool mixed = p.precision == SMIP_PRECISION_MIXED;
    // FFT / tile overlap: device stacks of more than a batch, and host batches
    // (a staging buffer) of large jobs
    const int host_buf = upload_index(j, frames);
    const bool overlap = mixed && (host_buf < 0 ? count > (uint64_t)j->batch : j->host_overlap);
    hipStream_t ss = s; // stream of the shift-and-add
    if (overlap) {
        enable_overlap(j);
        // fork: sstream follows everything before this call; so does fstream for
        // a device stack, while for a host batch it needs only the batch's copy
        // (and, per half, the triple products that last read it), which lets
        // the FFT run beside the previous call's triple products
        SMIP_HIP(hipEventRecord(j->ev_fork, s));
        SMIP_HIP(hipStreamWaitEvent(j->fstream, host_buf < 0 ? j->ev_fork : j->ev_up[host_buf], 0));
        if (p.saa) {
            ss = j->sstream;
            SMIP_HIP(hipStreamWaitEvent(ss, j->ev_fork, 0));
        }
    }
    for (uint64_t off = 0; off < count;) {
        const int64_t nf = (int64_t)std::min<uint64_t>((uint64_t)j->batch, count - off);
        const uint8_t* fr = static_cast<const uint8_t*>(frames) + off * fb;
        // full batches from a staging buffer after frame 0 read and write the same
        // buffers every time: one graph launch each instead of a dozen launches
        const int ub = upload_index(j, fr);
        const bool graph = graphs_enabled() && !overlap && nf == j->batch && ub >= 0 && j->nframes > 0;
        hipEvent_t e0 = mark(j);
        if (graph) {
            if (!j->g_acc[ub])
                j->g_acc[ub] = capture_graph(s, [&] { enqueue_accumulate(j, fr, nf); });
            graph_launch(j->g_acc[ub], s);
            j->graph_launches++;
            note_h32_reader(j);
        } else if (overlap) {
            const int b = j->hcur;
            j->hcur ^= 1;
            float2* H = b ? j->H32b.get() : j->H32.get();
            if (j->hused[b]) // the triple products that last read this half have finished
                SMIP_HIP(hipStreamWaitEvent(j->fstream, j->ev_tile[b], 0));
            launch_forward_f32(j->plan, fr, p.frame_dtype, nf, false, j->hp, H, j->work32.get(), j->P_half.get(),
                j->fstream);
            SMIP_HIP(hipEventRecord(j->ev_fft[b], j->fstream));
            SMIP_HIP(hipStreamWaitEvent(s, j->ev_fft[b], 0));
            hipEvent_t t0 = mark(j);
            launch_accum_tile(j->tile, j->hp, H, nf, j->B64.get(), s);
            hipEvent_t t1 = mark(j);
            SMIP_HIP(hipEventRecord(j->ev_tile[b], s));
            j->hused[b] = true;
            record(j, 4, t0, t1);
            j->tile_launches++;
        } else if (mixed) {
            launch_forward_f32(j->plan, fr, p.frame_dtype, nf, false, j->hp, j->H32.get(), j->work32.get(),
                j->P_half.get(), s);
            hipEvent_t t0 = mark(j);
            launch_accum_tile(j->tile, j->hp, j->H32.get(), nf, j->B64.get(), s);
            hipEvent_t t1 = mark(j);
            note_h32_reader(j);
            record(j, 4, t0, t1);
            j->tile_launches++;
        } else {
            enqueue_accumulate(j, fr, nf);
        }
        hipEvent_t e1 = mark(j);
        record(j, 0, e0, e1);
        if (p.saa) {
            hipEvent_t e2 = mark_on(j, ss);
            if (graph) {
                if (!j->g_saa[ub])
                    j->g_saa[ub] = capture_graph(s, [&] { enqueue_saa(j, fr, nf, j->bshift.get(), s); });
                graph_launch(j->g_saa[ub], s);
                j->graph_launches++;
                SMIP_HIP(hipMemcpyAsync(j->shifts.get() + 2 * j->nframes, j->bshift.get(), sizeof(int32_t) * 2 * nf,
                    hipMemcpyDeviceToDevice, s));
            } else {
                int64_t first = 0;
                if (j->nframes == 0) {
                    // sumarray = Re(frame 0); the reference spectrum is frame 0's r2c
                    const int64_t np = (int64_t)X * Y;
                    hipLaunchKernelGGL(k_frame_to_f64, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ss, fr,
                        p.frame_dtype, np, j->sum.get());
                    launch_rfft_f64(j->plan, fr, p.frame_dtype, 1, false, j->ref_spec.get(), j->swork.get(), ss);
                    first = 1;
                }
                const int64_t m = nf - first;
                if (m > 0)
                    enqueue_saa(j, fr + first * fb, m, j->shifts.get() + 2 * (j->nframes + first), ss);
            }
            hipEvent_t e3 = mark_on(j, ss);
            record(j, 1, e2, e3);
        }
        j->nframes += (uint64_t)nf;
        off += (uint64_t)nf;
    }
    if (overlap) { // join: the power spectrum, spectra and sum image are complete for what follows on the main stream
        SMIP_HIP(hipEventRecord(j->ev_fork, j->fstream));
        SMIP_HIP(hipStreamWaitEvent(s, j->ev_fork, 0));
        if (p.saa) {
            SMIP_HIP(hipEventRecord(j->ev_sjoin, ss));
            SMIP_HIP(hipStreamWaitEvent(s, j->ev_sjoin, 0));
        }
    }
}

// B /= complex(n, 0) (smip-cli/main.cpp:312), once per job whatever happens after it
void scale_once(smip_job* j, hipStream_t s)
{
    if (j->scaled)
        return;
    launch_scale_c64(j->B64.get(), j->d.base_size, (double)j->nframes, 0.0, SMIP_OP_DIV, s);
    j->scaled = true;
}

// Failure injection for tests of the error paths: SMIP_INJECT_FAIL=<stage> makes
// that stage fail with SMIP_ERR_OUT_OF_RANGE (the code of a failing get_element)
// before it launches anything.  Read on every call, so a test can clear it.
void injected_failure(const char* stage)
{
    const char* e = std::getenv("SMIP_INJECT_FAIL");
    if (e && std::strcmp(e, stage) == 0)
        fail(SMIP_ERR_OUT_OF_RANGE, std::string("injected failure of stage ") + stage);
}

void grow_shifts(smip_job* j, uint64_t total)
{
    if (!j->p.saa || total <= j->shifts_cap)
        return;
    uint64_t cap = std::max<uint64_t>(total, 2 * j->shifts_cap);
    auto n = dmake<int32_t>(2 * cap);
    SMIP_HIP(hipMemsetAsync(n.get(), 0, sizeof(int32_t) * 2 * cap, j->stream));
    if (j->shifts_cap)
        SMIP_HIP(hipMemcpyAsync(n.get(), j->shifts.get(), sizeof(int32_t) * 2 * j->shifts_cap,
            hipMemcpyDeviceToDevice, j->stream));
    for (hipStream_t st : { j->sstream, j->fstream, j->stream })
        if (st)
            SMIP_HIP(hipStreamSynchronize(st));
    IdleRelease idle; // every stream of the job is synchronised: the old array is idle
    j->shifts = std::move(n);
    j->shifts_cap = cap;
}

} // namespace

extern "C" SMIP_API int smip_job_resolve_params(const smip_job_params* p, smip_job_params* out)
{
    return guarded([&] {
        require(p && out, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        const JobSizes z = job_sizes(*p);
        smip_job_params q = *p;
        q.frames_per_batch = z.batch;
        q.vy_tile = p->precision == SMIP_PRECISION_MIXED ? z.vy_tile : p->vy_tile;
        *out = q;
    });
}

extern "C" SMIP_API int smip_job_footprint(const smip_job_params* p, uint64_t* bytes)
{
    return guarded([&] {
        require(p && bytes, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        *bytes = job_sizes(*p).bytes;
    });
}

extern "C" SMIP_API int smip_job_create(smip_ctx* ctx, const smip_job_params* p, smip_job** out)
{
    return guarded([&] {
        require(ctx && p && out, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        ensure_device(ctx);
        const JobSizes z = job_sizes(*p);
        size_t fr = 0, tot = 0;
        SMIP_HIP(mem_get_info(&fr, &tot));
        if (z.bytes > fr)
            fail(SMIP_ERR_NO_MEMORY, "job footprint " + std::to_string(z.bytes) + " B exceeds free device memory");
        auto j = std::make_unique<smip_job>();
        j->ctx_ref.hold(ctx);
        j->ctx = ctx;
        j->p = *p;
        j->d = z.d;
        j->L = z.L;
        j->hp = z.hp;
        j->batch = z.batch;
        j->footprint = z.bytes;
        j->triples = count_triples(z.L);
        // host batches of at least ~1 ms of triple products overlap their FFT with
        // the previous batch; smaller ones replay a captured graph per batch
        j->host_overlap = p->precision == SMIP_PRECISION_MIXED && 14.0 * (double)j->triples * z.batch > 5e10;
        j->stream = make_stream();
        hipStream_t s = j->stream;
        const uint64_t X = p->ncols, Y = p->nrows, hx = X / 2 + 1, np = X * Y;
        j->plan = fft_plan_create((int32_t)X, (int32_t)Y, s);
        j->B64 = dmake<double2>(z.d.base_size);
        SMIP_HIP(hipMemsetAsync(j->B64.get(), 0, sizeof(double2) * z.d.base_size, s));
        j->P_half = dmake<double>(hx * Y);
        SMIP_HIP(hipMemsetAsync(j->P_half.get(), 0, sizeof(double) * hx * Y, s));
        if (p->precision == SMIP_PRECISION_MIXED) {
            j->tile = tile_plan_create(z.d, z.L, z.hp, z.vy_tile, s);
            j->H32 = dmake<float2>((size_t)j->batch * z.hp.fs);
            SMIP_HIP(hipMemsetAsync(j->H32.get(), 0, sizeof(float2) * j->batch * z.hp.fs, s)); // zero padding rows
            // the second half of the double-buffered spectrum (the footprint counts
            // both): made now rather than when the first overlapped batch needs it
            j->H32b = dmake<float2>((size_t)j->batch * z.hp.fs);
            SMIP_HIP(hipMemsetAsync(j->H32b.get(), 0, sizeof(float2) * j->batch * z.hp.fs, s));
            j->work32 = dmake<float2>((size_t)j->batch * hx * Y);
        } else {
            j->H64 = dmake<double2>((size_t)j->batch * z.hp.fs);
            SMIP_HIP(hipMemsetAsync(j->H64.get(), 0, sizeof(double2) * j->batch * z.hp.fs, s));
            j->work64 = dmake<double2>((size_t)j->batch * hx * Y);
        }
        if (p->saa) {
            j->ref_spec = dmake<double2>(hx * Y);
            const size_t sc = (size_t)saa_chunk(X, Y, j->batch);
            j->swork = dmake<double2>(sc * hx * Y);
            // per-workgroup correlation maxima (RowBest, 16 B each) of every frame of a pass
            require((uint64_t)fft_shift_best_entries(j->plan) * sizeof(RowBest) <= saa_best_bytes(X, Y), SMIP_ERR_LOGIC,
                "shift-and-add maxima exceed their buffer");
            j->corr = dmake<double>(sc * saa_best_bytes(X, Y) / sizeof(double));
            j->sum = dmake<double>(np);
            SMIP_HIP(hipMemsetAsync(j->sum.get(), 0, sizeof(double) * np, s));
            j->bshift = dmake<int32_t>(2 * (size_t)j->batch);
        }
        j->phases = dmake<double2>(np);
        j->image = dmake<double2>(np);
        j->flags = dmake<uint8_t>(np);
        j->cons = dmake<double>(np);
        j->scratch = dmake<double>(16);
        j->reco = reco_work_create(z.d, reco_geom(p->reco_radius, (int32_t)X, (int32_t)Y, s));
        SMIP_HIP(hipStreamSynchronize(s));
        auto v = std::make_unique<smip_bispec>();
        v->ctx = ctx;
        v->d = z.d;
        v->dtype = SMIP_DTYPE_C128;
        v->data = j->B64.get();
        v->borrowed = true;
        j->view = std::move(v);
        *out = j.release();
    });
}

extern "C" SMIP_API int smip_job_destroy(smip_job* j)
{
    return guarded([&] {
        if (!j)
            return;
        ensure_device(j->ctx);
        // copies of pushed frames still pending on the copy stream read the pinned
        // slots freed below
        for (hipStream_t st : { j->ustream, j->fstream, j->sstream, j->stream })
            if (st)
                (void)hipStreamSynchronize(st);
        IdleRelease idle; // nothing queued can touch the job's memory now
        delete j;
    });
}

extern "C" SMIP_API int smip_job_add_frames(smip_job* j, const smip_frames* f, uint64_t first, uint64_t count)
{
    return guarded([&] {
        require(j && f, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(!j->finalized && !j->scaled, SMIP_ERR_LOGIC, "job already finalized");
        require(j->staged == 0, SMIP_ERR_LOGIC, "frames pushed with smip_job_push_frames_host are pending");
        require(f->dtype == j->p.frame_dtype && f->X == j->p.ncols && f->Y == j->p.nrows, SMIP_ERR_INVALID_ARGUMENT,
            "frame stack does not match the job");
        require(first + count <= f->count, SMIP_ERR_OUT_OF_RANGE, "frame range outside the stack");
        ensure_device(j->ctx);
        grow_shifts(j, j->nframes + count);
        const size_t fb = f->X * f->Y * dtype_size(f->dtype);
        job_process(j, static_cast<const uint8_t*>(f->data) + first * fb, count);
    });
}

// ---------------------------------------------------------------------------
// Many jobs at once: one launch per stage for the current batch of every job
// ---------------------------------------------------------------------------
namespace {

// Host image of one round's grouped tables, uploaded in one copy.
struct TableImage {
    std::vector<uint8_t> bytes;
    template <typename T>
    size_t add(const T* p, size_t n)
    {
        const size_t at = (bytes.size() + 15) & ~(size_t)15;
        bytes.resize(at + sizeof(T) * n);
        if (n)
            std::memcpy(bytes.data() + at, p, sizeof(T) * n);
        return at;
    }
};

// Two pinned + device slots for the round tables: the copy of round r + 2 into a
// slot waits (host side) for round r's copy out of it; kernels reading a slot run
// before the next copy into it on the same stream.
struct TableRing {
    uint8_t* host[2] = { nullptr, nullptr };
    uint8_t* dev[2] = { nullptr, nullptr };
    hipEvent_t copied[2] = { nullptr, nullptr };
    size_t cap = 0;
    int cur = 0;
    explicit TableRing(size_t bytes)
        : cap(std::max<size_t>(bytes, 4096))
    {
        for (int k = 0; k < 2; ++k) {
            host[k] = static_cast<uint8_t*>(host_malloc(cap));
            dev[k] = static_cast<uint8_t*>(dev_malloc(cap));
            SMIP_HIP(hipEventCreateWithFlags(&copied[k], hipEventDisableTiming));
        }
    }
    ~TableRing()
    {
        for (int k = 0; k < 2; ++k) {
            if (copied[k])
                (void)hipEventDestroy(copied[k]);
            host_free(host[k]);
            dev_free(dev[k]);
        }
    }
    uint8_t* upload(const TableImage& t, hipStream_t s)
    {
        require(t.bytes.size() <= cap, SMIP_ERR_LOGIC, "grouped tables exceed their buffer");
        const int k = cur;
        cur ^= 1;
        SMIP_HIP(hipEventSynchronize(copied[k]));
        std::memcpy(host[k], t.bytes.data(), t.bytes.size());
        SMIP_HIP(hipMemcpyAsync(dev[k], host[k], t.bytes.size(), hipMemcpyHostToDevice, s));
        SMIP_HIP(hipEventRecord(copied[k], s));
        return dev[k];
    }
};

// Per device: the table ring and the three streams of smip_job_add_frames_many,
// and the lock that gives them to one call at a time.  The streams live as long
// as the ring, whose copy events they record (an event is not waited on after
// the stream it was recorded on is gone).  Never destroyed (no HIP calls at
// process exit).
struct RingSlot {
    std::mutex use;
    std::unique_ptr<TableRing> ring;
    hipStream_t gs = nullptr, fs = nullptr, ss = nullptr;
};
RingSlot& ring_slot(int dev)
{
    static std::mutex mu;
    static auto* slots = new std::map<int, std::unique_ptr<RingSlot>>;
    std::lock_guard<std::mutex> lk(mu);
    auto& p = (*slots)[dev];
    if (!p)
        p = std::make_unique<RingSlot>();
    return *p;
}
std::unique_lock<std::mutex> table_ring_lock(int dev) { return std::unique_lock<std::mutex>(ring_slot(dev).use); }
// the grouped streams of the device (caller holds its lock), made on first use
void group_streams(int dev, hipStream_t& gs, hipStream_t& fs, hipStream_t& ss)
{
    RingSlot& r = ring_slot(dev);
    for (hipStream_t* st : { &r.gs, &r.fs, &r.ss })
        if (!*st)
            *st = make_stream();
    gs = r.gs;
    fs = r.fs;
    ss = r.ss;
}
// the ring of the device (caller holds its lock), at least `cap` bytes per slot;
// a smaller ring is replaced once idle (its slots' copies have completed)
TableRing& table_ring(int dev, size_t cap)
{
    RingSlot& r = ring_slot(dev);
    if (!r.ring || r.ring->cap < cap) {
        if (r.ring) {
            for (hipEvent_t e : r.ring->copied)
                SMIP_HIP(hipEventSynchronize(e));
            IdleRelease idle; // its last copies have landed and its readers were synchronised
            r.ring.reset();
        }
        r.ring = std::make_unique<TableRing>(cap + cap / 2);
    }
    return *r.ring;
}

struct GroupMember {
    smip_job* j;
    const uint8_t* base; // frames of the job's stack, from its first frame
    uint64_t pos = 0, total = 0;
    FftTiling t {};
    size_t fb = 0;
};

// Offsets of one grouped launch: off[k] blocks before member k.
struct Offsets {
    std::vector<uint32_t> off { 0 };
    void add(uint64_t blocks)
    {
        const uint64_t next = (uint64_t)off.back() + blocks;
        require(next < (1ull << 31), SMIP_ERR_INVALID_ARGUMENT, "grouped launch too large");
        off.push_back((uint32_t)next);
    }
    uint32_t total() const { return off.back(); }
};

// The lockstep rounds of smip_job_add_frames_many for jobs the grouped kernels
// take (mixed precision, power-of-two frame sides).  Each round takes the next
// batch of every job that has frames left -- the batches smip_job_add_frames
// would form -- and runs each stage once for all of them: the forward transforms
// and |F|^2 on fs into one half of each job's double-buffered spectrum, the
// triple products on gs once those are done, and the shift-and-add on ss beside
// both, in chunks of saa_chunk frames as the single-job path runs it.  So the
// next round's transforms and this round's shift-and-add overlap this round's
// triple products, as they do across the streams of separately driven jobs.
// Synchronous.
void run_group(std::vector<GroupMember>& M)
{
    const size_t n = M.size();
    // table bytes of the largest round
    int max_chunks = 1;
    for (auto& m : M) {
        const smip_job_params& p = m.j->p;
        if (p.saa)
            max_chunks = std::max(max_chunks, (m.j->batch + saa_chunk(p.ncols, p.nrows, m.j->batch) - 1)
                    / saa_chunk(p.ncols, p.nrows, m.j->batch));
        enable_overlap(m.j); // the second spectrum buffer (padding rows zeroed on the job stream)
        SMIP_HIP(hipStreamSynchronize(m.j->stream));
    }
    // tables and offsets of the largest round; every class (shape, tile width,
    // chunk index and shape) adds at most one entry per member plus one offset
    // array terminator and alignment
    const size_t cap = n * (sizeof(GroupFwd) + sizeof(GroupTile)) + (size_t)max_chunks * n * sizeof(GroupSaa)
        + sizeof(uint32_t) * (4 * n + 5 * (size_t)max_chunks * n) + (n + 8) * (size_t)(max_chunks + 2) * 5 * 32;
    // the device's table ring, kept between calls (it grows to the largest round
    // seen); calls on one device take it in turn
    int dev = 0;
    SMIP_HIP(hipGetDevice(&dev));
    std::unique_lock<std::mutex> ring_lock = table_ring_lock(dev);
    hipStream_t gs = nullptr, fs = nullptr, ss = nullptr;
    group_streams(dev, gs, fs, ss);
    TableRing& ring = table_ring(dev, cap);
    struct RoundEvents {
        hipEvent_t f0, t0, t1, s0, s1;
        std::vector<size_t> members;
    };
    std::vector<RoundEvents> rounds;
    std::vector<hipEvent_t> sync_ev; // order-only events
    auto order_event = [&] {
        hipEvent_t e;
        SMIP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        sync_ev.push_back(e);
        return e;
    };
    auto cleanup = [&] {
        for (hipStream_t st : { fs, ss, gs })
            (void)hipStreamSynchronize(st);
        for (auto& r : rounds)
            for (hipEvent_t e : { r.f0, r.t0, r.t1, r.s0, r.s1 })
                (void)hipEventDestroy(e);
        for (hipEvent_t e : sync_ev)
            (void)hipEventDestroy(e);
    };
    try {
        hipEvent_t e_up = order_event(), e_fft = order_event();
        hipEvent_t e_tile[2] = { order_event(), order_event() };     // last reader of spectrum half h
        hipEvent_t e_fs_done[2] = { order_event(), order_event() };  // fs's last reads of table slot k
        hipEvent_t e_ss_done[2] = { order_event(), order_event() };  // ss's last reads of table slot k
        bool tile_used[2] = { false, false }, slot_used[2] = { false, false };
        auto timing_event = [&] {
            hipEvent_t e;
            SMIP_HIP(hipEventCreate(&e));
            return e;
        };
        for (int round = 0;; ++round) {
            std::vector<size_t> R;
            for (size_t k = 0; k < n; ++k)
                if (M[k].pos < M[k].total)
                    R.push_back(k);
            if (R.empty())
                break;
            const int half = round & 1;
            std::vector<int64_t> nf(n, 0);
            // forward transforms and |F|^2 per frame shape (a launch's dynamic LDS is
            // its largest member's), triple products per tile width
            struct FwdClass {
                std::vector<GroupFwd> tab;
                Offsets rows, cols, pow;
                size_t lds_rows = 0, lds_cols = 0;
            };
            std::map<std::pair<int, int>, FwdClass> fwdc;
            std::map<int, std::pair<std::vector<GroupTile>, Offsets>> tiles;
            for (size_t k : R) {
                GroupMember& m = M[k];
                smip_job* j = m.j;
                const smip_job_params& p = j->p;
                nf[k] = (int64_t)std::min<uint64_t>((uint64_t)j->batch, m.total - m.pos);
                float2* H = half ? j->H32b.get() : j->H32.get();
                GroupFwd g {};
                g.frames = m.base + m.pos * m.fb;
                g.dtype = p.frame_dtype;
                g.X = (int32_t)p.ncols;
                g.Y = (int32_t)p.nrows;
                g.lx = m.t.lx;
                g.ly = m.t.ly;
                g.hx = g.X / 2 + 1;
                g.ppw = m.t.ppw32;
                g.cb = m.t.cb32;
                g.nf = nf[k];
                g.twx = m.t.twx32;
                g.twy = m.t.twy32;
                g.G = j->work32.get();
                g.H = H;
                g.hp = j->hp;
                g.P = j->P_half.get();
                FwdClass& fc = fwdc[{ g.X, g.Y }];
                fc.tab.push_back(g);
                fc.rows.add((uint64_t)((g.Y / 2) / g.ppw) * g.nf);
                fc.cols.add((uint64_t)((g.hx + g.cb - 1) / g.cb) * g.nf);
                fc.pow.add(((uint64_t)g.hx * g.Y + 255) / 256);
                fc.lds_rows = std::max(fc.lds_rows, m.t.lds_rows32);
                fc.lds_cols = std::max(fc.lds_cols, m.t.lds_cols32);
                GroupTile t {};
                t.items = tile_plan_items(j->tile);
                t.a = tile_plan_args(j->tile);
                t.H = H;
                t.nf = nf[k];
                t.B = j->B64.get();
                auto& cls = tiles[tile_plan_vy_tile(j->tile)];
                cls.first.push_back(t);
                cls.second.add((uint64_t)tile_plan_nitems(j->tile));
            }
            // shift-and-add chunks: chunk c of every job that has one; frame 0 of a
            // job starts its sum and reference spectrum on its own
            struct SaaClass {
                std::vector<GroupSaa> tab;
                std::array<Offsets, 5> off;
                size_t lds_rows = 0, lds_cols = 0;
            };
            std::map<std::pair<int64_t, std::pair<int, int>>, SaaClass> saac; // (chunk index, shape)
            std::vector<size_t> first_frame;
            for (size_t k : R) {
                GroupMember& m = M[k];
                smip_job* j = m.j;
                const smip_job_params& p = j->p;
                if (!p.saa)
                    continue;
                const int64_t first = j->nframes == 0 ? 1 : 0;
                if (first)
                    first_frame.push_back(k);
                const int64_t mm = nf[k] - first;
                const int64_t chunk = saa_chunk(p.ncols, p.nrows, j->batch);
                for (int64_t c = 0; c * chunk < mm; ++c) {
                    SaaClass& sc = saac[{ c, { (int)p.ncols, (int)p.nrows } }];
                    sc.lds_rows = std::max(sc.lds_rows, m.t.lds_rows64);
                    sc.lds_cols = std::max(sc.lds_cols, m.t.lds_cols64);
                    GroupSaa g {};
                    const int64_t o = first + c * chunk;
                    g.frames = m.base + (m.pos + (uint64_t)o) * m.fb;
                    g.dtype = p.frame_dtype;
                    g.X = (int32_t)p.ncols;
                    g.Y = (int32_t)p.nrows;
                    g.lx = m.t.lx;
                    g.ly = m.t.ly;
                    g.hx = g.X / 2 + 1;
                    g.ppw = m.t.ppw64;
                    g.cb = m.t.cb64;
                    g.nblk = m.t.nblk64;
                    g.nf = std::min(chunk, mm - c * chunk);
                    g.parts = (int32_t)saa_add_parts(g.dtype, g.nf, g.X, g.Y);
                    g.twx = m.t.twx64;
                    g.twy = m.t.twy64;
                    g.ref = j->ref_spec.get();
                    g.G = j->swork.get();
                    g.best = reinterpret_cast<RowBest*>(j->corr.get());
                    g.shifts = j->shifts.get() + 2 * (j->nframes + (uint64_t)o);
                    g.sum = j->sum.get();
                    sc.tab.push_back(g);
                    auto& of = sc.off;
                    of[0].add((uint64_t)((g.Y / 2) / g.ppw) * g.nf);
                    of[1].add((uint64_t)((g.hx + g.cb - 1) / g.cb) * g.nf);
                    of[2].add((uint64_t)g.nblk * g.nf);
                    of[3].add((uint64_t)g.nf);
                    of[4].add(((uint64_t)g.X * g.Y + 255) / 256 * (uint64_t)g.parts);
                }
            }
            // one upload for the round, on gs, once fs and ss are done with the slot
            TableImage img;
            std::vector<std::array<size_t, 4>> at_fwd;
            for (auto& kv : fwdc)
                at_fwd.push_back({ img.add(kv.second.tab.data(), kv.second.tab.size()),
                    img.add(kv.second.rows.off.data(), kv.second.rows.off.size()),
                    img.add(kv.second.cols.off.data(), kv.second.cols.off.size()),
                    img.add(kv.second.pow.off.data(), kv.second.pow.off.size()) });
            std::vector<std::pair<size_t, size_t>> at_tile;
            for (auto& kv : tiles)
                at_tile.push_back({ img.add(kv.second.first.data(), kv.second.first.size()),
                    img.add(kv.second.second.off.data(), kv.second.second.off.size()) });
            std::vector<std::pair<size_t, size_t>> at_saa;
            for (auto& kv : saac) {
                std::vector<uint32_t> offs;
                for (int q = 0; q < 5; ++q)
                    offs.insert(offs.end(), kv.second.off[q].off.begin(), kv.second.off[q].off.end());
                at_saa.push_back({ img.add(kv.second.tab.data(), kv.second.tab.size()), img.add(offs.data(), offs.size()) });
            }
            const int slot = ring.cur;
            if (slot_used[slot]) {
                SMIP_HIP(hipStreamWaitEvent(gs, e_fs_done[slot], 0));
                SMIP_HIP(hipStreamWaitEvent(gs, e_ss_done[slot], 0));
            }
            uint8_t* dt = ring.upload(img, gs);
            SMIP_HIP(hipEventRecord(e_up, gs));
            SMIP_HIP(hipStreamWaitEvent(fs, e_up, 0));
            SMIP_HIP(hipStreamWaitEvent(ss, e_up, 0));
            RoundEvents ev { timing_event(), timing_event(), timing_event(), timing_event(), timing_event(), R };
            rounds.push_back(ev);
            // transforms into spectrum half `half`, once round - 2's triple products have read it
            if (tile_used[half])
                SMIP_HIP(hipStreamWaitEvent(fs, e_tile[half], 0));
            SMIP_HIP(hipEventRecord(ev.f0, fs));
            size_t fi = 0;
            for (auto& kv : fwdc) {
                const auto& a = at_fwd[fi++];
                const FwdClass& fc = kv.second;
                launch_group_fwd(reinterpret_cast<const GroupFwd*>(dt + a[0]), reinterpret_cast<const uint32_t*>(dt + a[1]),
                    reinterpret_cast<const uint32_t*>(dt + a[2]), reinterpret_cast<const uint32_t*>(dt + a[3]),
                    (int)fc.tab.size(), fc.rows.total(), fc.cols.total(), fc.pow.total(), fc.lds_rows, fc.lds_cols, fs);
            }
            SMIP_HIP(hipEventRecord(e_fft, fs));
            SMIP_HIP(hipEventRecord(e_fs_done[slot], fs));
            // triple products
            SMIP_HIP(hipStreamWaitEvent(gs, e_fft, 0));
            SMIP_HIP(hipEventRecord(ev.t0, gs));
            size_t ti = 0;
            for (auto& kv : tiles) {
                launch_group_tile(kv.first, reinterpret_cast<const GroupTile*>(dt + at_tile[ti].first),
                    reinterpret_cast<const uint32_t*>(dt + at_tile[ti].second), (int)kv.second.first.size(),
                    kv.second.second.total(), gs);
                ++ti;
            }
            SMIP_HIP(hipEventRecord(ev.t1, gs));
            SMIP_HIP(hipEventRecord(e_tile[half], gs));
            tile_used[half] = true;
            // shift-and-add
            SMIP_HIP(hipEventRecord(ev.s0, ss));
            for (size_t k : first_frame) { // sumarray = Re(frame 0); reference spectrum of frame 0
                GroupMember& m = M[k];
                smip_job* j = m.j;
                const int64_t np = (int64_t)j->p.ncols * j->p.nrows;
                const uint8_t* fr = m.base + m.pos * m.fb;
                hipLaunchKernelGGL(k_frame_to_f64, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ss, fr,
                    j->p.frame_dtype, np, j->sum.get());
                launch_rfft_f64(j->plan, fr, j->p.frame_dtype, 1, false, j->ref_spec.get(), j->swork.get(), ss);
            }
            size_t si = 0;
            for (auto& kv : saac) { // chunk index ascending: each job's chunks in frame order
                const SaaClass& sc = kv.second;
                uint32_t nb[5];
                for (int q = 0; q < 5; ++q)
                    nb[q] = sc.off[q].total();
                launch_group_saa(reinterpret_cast<const GroupSaa*>(dt + at_saa[si].first),
                    reinterpret_cast<const uint32_t*>(dt + at_saa[si].second), (int)sc.tab.size(), nb, sc.lds_rows,
                    sc.lds_cols, ss);
                ++si;
            }
            SMIP_HIP(hipEventRecord(ev.s1, ss));
            SMIP_HIP(hipEventRecord(e_ss_done[slot], ss));
            slot_used[slot] = true;
            for (size_t k : R) {
                M[k].j->nframes += (uint64_t)nf[k];
                M[k].pos += (uint64_t)nf[k];
                M[k].j->tile_launches++;
            }
        }
        for (hipStream_t st : { fs, ss, gs })
            SMIP_HIP(hipStreamSynchronize(st));
        // the round times, shared by the round's jobs
        for (auto& r : rounds) {
            float acc = 0.f, tl = 0.f, sa = 0.f;
            SMIP_HIP(hipEventElapsedTime(&acc, r.f0, r.t1));
            SMIP_HIP(hipEventElapsedTime(&tl, r.t0, r.t1));
            SMIP_HIP(hipEventElapsedTime(&sa, r.s0, r.s1));
            for (size_t k : r.members) {
                M[k].j->ms[0] += acc;
                M[k].j->ms[4] += tl;
                M[k].j->ms[1] += sa;
            }
        }
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

} // namespace

extern "C" SMIP_API int smip_job_add_frames_many(smip_job* const* jobs, const smip_frames* const* stacks,
    const uint64_t* first, const uint64_t* count, uint64_t n)
{
    return guarded([&] {
        require(n == 0 || (jobs && stacks && first && count), SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        if (n == 0)
            return;
        require(n <= 4096, SMIP_ERR_INVALID_ARGUMENT, "too many jobs in one group");
        for (uint64_t k = 0; k < n; ++k) {
            smip_job* j = jobs[k];
            const smip_frames* f = stacks[k];
            require(j && f, SMIP_ERR_INVALID_ARGUMENT, "null job or frame stack");
            require(!j->finalized && !j->scaled, SMIP_ERR_LOGIC, "job already finalized");
            require(j->staged == 0, SMIP_ERR_LOGIC, "frames pushed with smip_job_push_frames_host are pending");
            require(f->dtype == j->p.frame_dtype && f->X == j->p.ncols && f->Y == j->p.nrows, SMIP_ERR_INVALID_ARGUMENT,
                "frame stack does not match the job");
            require(first[k] + count[k] <= f->count, SMIP_ERR_OUT_OF_RANGE, "frame range outside the stack");
            require(j->ctx->device == jobs[0]->ctx->device && f->ctx->device == jobs[0]->ctx->device,
                SMIP_ERR_INVALID_ARGUMENT, "jobs of one group must share a device");
            for (uint64_t m = 0; m < k; ++m)
                require(jobs[m] != j, SMIP_ERR_INVALID_ARGUMENT, "a job appears twice in the group");
        }
        ensure_device(jobs[0]->ctx);
        // every job's earlier work is done: the grouped launches use its buffers from another stream
        for (uint64_t k = 0; k < n; ++k)
            for (hipStream_t st : { jobs[k]->ustream, jobs[k]->fstream, jobs[k]->sstream, jobs[k]->stream })
                if (st)
                    SMIP_HIP(hipStreamSynchronize(st));
        std::vector<GroupMember> M;
        for (uint64_t k = 0; k < n; ++k) {
            smip_job* j = jobs[k];
            if (count[k] == 0)
                continue;
            grow_shifts(j, j->nframes + count[k]);
            const size_t fb = j->p.ncols * j->p.nrows * dtype_size(j->p.frame_dtype);
            const uint8_t* base = static_cast<const uint8_t*>(stacks[k]->data) + first[k] * fb;
            if (j->p.precision == SMIP_PRECISION_MIXED && fft_plan_pow2(j->plan)) {
                GroupMember m;
                m.j = j;
                m.base = base;
                m.total = count[k];
                m.t = fft_tiling(j->plan);
                m.fb = fb;
                M.push_back(m);
            } else {
                job_process(j, base, count[k]); // its own launches, on its own streams
            }
        }
        if (!M.empty())
            run_group(M);
    });
}

// Copies from caller memory into pinned memory, split over a few helper threads
// so that this host-side step keeps ahead of the kernels it feeds (one thread
// moves pageable memory at roughly a fifth of the PCIe rate).  The pool is
// never destroyed: its helpers may still be waiting when the process exits.
class CopyPool {
public:
    static CopyPool& get()
    {
        static CopyPool* pool = new CopyPool();
        return *pool;
    }
    void copy(void* dst, const void* src, size_t bytes)
    {
        if (helpers_ == 0 || bytes < ((size_t)1 << 20)) {
            std::memcpy(dst, src, bytes);
            return;
        }
        std::lock_guard<std::mutex> one(use_); // one split copy at a time
        {
            std::lock_guard<std::mutex> lk(m_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            bytes_ = bytes;
            pending_ = helpers_;
            ++gen_;
        }
        work_.notify_all();
        part(0);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return pending_ == 0; });
    }

private:
    CopyPool()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        helpers_ = hw >= 4 ? (int)std::min<unsigned>(7, hw / 2 - 1) : 0;
        for (int t = 1; t <= helpers_; ++t)
            std::thread([this, t] { run(t); }).detach();
    }
    void part(int i) const
    {
        const size_t per = ((bytes_ / (size_t)(helpers_ + 1)) + 63) & ~(size_t)63;
        const size_t a = std::min(bytes_, per * (size_t)i);
        const size_t b = i == helpers_ ? bytes_ : std::min(bytes_, a + per);
        if (b > a)
            std::memcpy(dst_ + a, src_ + a, b - a);
    }
    void run(int i)
    {
        uint64_t seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> lk(m_);
            work_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            lk.unlock();
            part(i);
            lk.lock();
            if (--pending_ == 0)
                done_.notify_one();
        }
    }
    std::mutex use_, m_;
    std::condition_variable work_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0;
    int helpers_ = 0, pending_ = 0;
    uint64_t gen_ = 0;
};

// Host frames enter through the two staging buffers: ingest_copy moves host data
// through the pinned slots into buffer k on the copy stream, once k's previous
// batch is done with it; ingest_process runs the batch in k on the job stream
// once its copies landed.  The copies of one batch overlap the processing of
// the other, and the caller's memory is free as soon as ingest_copy returns.
constexpr size_t kPinSlotBytes = (size_t)8 << 20;

void ensure_ingest(smip_job* j)
{
    if (j->ustream)
        return;
    const size_t fb = j->p.ncols * j->p.nrows * dtype_size(j->p.frame_dtype);
    for (int k = 0; k < 2; ++k) {
        j->upload[k] = dmake<uint8_t>((size_t)j->batch * fb);
        SMIP_HIP(hipEventCreateWithFlags(&j->ev_up[k], hipEventDisableTiming));
        SMIP_HIP(hipEventCreateWithFlags(&j->ev_used[k], hipEventDisableTiming));
        SMIP_HIP(hipEventRecord(j->ev_used[k], j->stream));
    }
    j->pin_bytes = std::min(kPinSlotBytes, (size_t)j->batch * fb);
    for (int k = 0; k < 2; ++k) {
        j->pin[k] = static_cast<uint8_t*>(host_malloc(j->pin_bytes));
        SMIP_HIP(hipEventCreateWithFlags(&j->pin_done[k], hipEventDisableTiming));
        SMIP_HIP(hipEventRecord(j->pin_done[k], j->stream));
    }
    // overlapped host batches copy on the FFT stream, right ahead of their FFT
    // (one stream fewer to share the hardware queues); others on a copy stream
    if (j->host_overlap) {
        enable_overlap(j);
        j->ustream = j->fstream;
    } else {
        j->ustream = make_stream();
    }
}

void ingest_copy(smip_job* j, int k, uint64_t at, const uint8_t* src, uint64_t n, size_t fb)
{
    SMIP_HIP(hipStreamWaitEvent(j->ustream, j->ev_used[k], 0));
    const size_t bytes = n * fb;
    uint8_t* dst = j->upload[k].get() + at * fb;
    for (size_t o = 0; o < bytes; o += j->pin_bytes) {
        const size_t len = std::min(j->pin_bytes, bytes - o);
        const int sl = j->pin_cur;
        j->pin_cur ^= 1;
        SMIP_HIP(hipEventSynchronize(j->pin_done[sl])); // the slot's previous piece has left it
        CopyPool::get().copy(j->pin[sl], src + o, len);
        SMIP_HIP(hipMemcpyAsync(dst + o, j->pin[sl], len, hipMemcpyHostToDevice, j->ustream));
        SMIP_HIP(hipEventRecord(j->pin_done[sl], j->ustream));
    }
}

void ingest_process(smip_job* j, int k, uint64_t nf)
{
    SMIP_HIP(hipEventRecord(j->ev_up[k], j->ustream));
    SMIP_HIP(hipStreamWaitEvent(j->stream, j->ev_up[k], 0));
    job_process(j, j->upload[k].get(), nf);
    SMIP_HIP(hipEventRecord(j->ev_used[k], j->stream));
}

// the batch collected by smip_job_push_frames_host, if any
void flush_staged(smip_job* j)
{
    if (j->staged == 0)
        return;
    ingest_process(j, j->up_cur, j->staged);
    j->up_cur ^= 1;
    j->staged = 0;
}

extern "C" SMIP_API int smip_job_add_frames_host(smip_job* j, const void* frames, uint64_t count)
{
    return guarded([&] {
        require(j && (frames || count == 0), SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(!j->finalized && !j->scaled, SMIP_ERR_LOGIC, "job already finalized");
        require(j->staged == 0, SMIP_ERR_LOGIC, "frames pushed with smip_job_push_frames_host are pending");
        ensure_device(j->ctx);
        grow_shifts(j, j->nframes + count);
        ensure_ingest(j);
        const size_t fb = j->p.ncols * j->p.nrows * dtype_size(j->p.frame_dtype);
        const uint8_t* src = static_cast<const uint8_t*>(frames);
        for (uint64_t off = 0; off < count;) {
            const uint64_t nf = std::min<uint64_t>((uint64_t)j->batch, count - off);
            const int k = j->up_cur;
            j->up_cur ^= 1;
            ingest_copy(j, k, 0, src + off * fb, nf, fb);
            ingest_process(j, k, nf);
            off += nf;
        }
        SMIP_HIP(hipStreamSynchronize(j->stream));
    });
}

extern "C" SMIP_API int smip_job_push_frames_host(smip_job* j, const void* frames, uint64_t count)
{
    return guarded([&] {
        require(j && (frames || count == 0), SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(!j->finalized && !j->scaled, SMIP_ERR_LOGIC, "job already finalized");
        ensure_device(j->ctx);
        ensure_ingest(j);
        const size_t fb = j->p.ncols * j->p.nrows * dtype_size(j->p.frame_dtype);
        const uint64_t batch = (uint64_t)j->batch;
        grow_shifts(j, j->nframes + j->staged + count);
        const uint8_t* src = static_cast<const uint8_t*>(frames);
        while (count > 0) {
            const uint64_t n = std::min(count, batch - j->staged);
            ingest_copy(j, j->up_cur, j->staged, src, n, fb);
            j->staged += n;
            src += n * fb;
            count -= n;
            if (j->staged == batch)
                flush_staged(j);
        }
    });
}

extern "C" SMIP_API int smip_job_finalize(smip_job* j)
{
    return guarded([&] {
        require(j != nullptr, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(!j->finalized, SMIP_ERR_LOGIC, "job already finalized");
        ensure_device(j->ctx);
        flush_staged(j); // the last, partial batch of pushed frames
        require(j->nframes > 0, SMIP_ERR_RUNTIME, "Array_base<T>::operator/=(T) : division by zero");
        hipStream_t s = j->stream;
        const smip_job_params& p = j->p;
        check_reco_geometry(j->d, p.ncols, p.nrows);
        scale_once(j, s); // bispectrum /= complex(n, 0)
        hipEvent_t e0 = mark(j);
        injected_failure("reco");
        j->reco_calls = (uint64_t)run_reconstruct(j->reco, j->B64.get(), 0, j->phases.get(), j->flags.get(),
            j->cons.get(), s);
        hipEvent_t e1 = mark(j);
        record(j, 2, e0, e1);
        hipEvent_t e2 = mark(j);
        const double norm = (double)(j->nframes * p.ncols * p.nrows);
        launch_synth_prepare(j->P_half.get(), j->phases.get(), (int32_t)p.ncols, (int32_t)p.nrows, p.reco_radius,
            norm, j->image.get(), s);
        launch_c2c_inverse_f64(j->plan, j->image.get(), nullptr, s);
        launch_abs_max_normalize(j->image.get(), (int64_t)(p.ncols * p.nrows), j->scratch.get(), s);
        hipEvent_t e3 = mark(j);
        record(j, 3, e2, e3);
        SMIP_HIP(hipStreamSynchronize(s));
        j->finalized = true;
    });
}

extern "C" SMIP_API int smip_job_finalize_many(smip_job* const* jobs, uint64_t n, int32_t* status)
{
    return smip_amd::job_finalize_many_on(jobs, n, status, nullptr);
}

// smip_job_finalize_many with the group's phase recovery on `group_stream`
// (nullptr: the first job's stream)
int smip_amd::job_finalize_many_on(smip_job* const* jobs, uint64_t n, int32_t* status, hipStream_t group_stream)
{
    int first_err = SMIP_OK;
    std::string first_msg;
    const int rc = guarded([&] {
        require(jobs != nullptr || n == 0, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        if (n == 0)
            return;
        require(n <= 4096, SMIP_ERR_INVALID_ARGUMENT, "too many jobs in one group");
        for (uint64_t k = 0; k < n; ++k) {
            require(jobs[k] != nullptr, SMIP_ERR_INVALID_ARGUMENT, "null job");
            require(!jobs[k]->finalized, SMIP_ERR_LOGIC, "job already finalized");
            require(jobs[k]->ctx->device == jobs[0]->ctx->device, SMIP_ERR_INVALID_ARGUMENT,
                "jobs of one group must share a device");
            for (uint64_t m = 0; m < k; ++m)
                require(jobs[m] != jobs[k], SMIP_ERR_INVALID_ARGUMENT, "a job appears twice in the group");
        }
        for (uint64_t k = 0; k < n; ++k) { // geometry first: nothing is changed for an invalid group
            check_reco_geometry(jobs[k]->d, jobs[k]->p.ncols, jobs[k]->p.nrows);
            require(jobs[k]->nframes + jobs[k]->staged > 0, SMIP_ERR_RUNTIME,
                "Array_base<T>::operator/=(T) : division by zero");
        }
        auto note = [&](uint64_t k, int code, const char* msg) {
            if (status)
                status[k] = code;
            if (code != SMIP_OK && first_err == SMIP_OK) {
                first_err = code;
                first_msg = msg;
            }
        };
        const char* seq = std::getenv("SMIP_RECO_SEQUENTIAL");
        if (seq && seq[0] == '1') { // the one-visit-at-a-time cross-check runs job by job
            for (uint64_t k = 0; k < n; ++k) {
                const int r = smip_job_finalize(jobs[k]);
                note(k, r, last_error_text());
            }
            return;
        }
        ensure_device(jobs[0]->ctx);
        hipStream_t cs = group_stream ? group_stream : jobs[0]->stream; // the group's stream
        std::vector<hipEvent_t> ev;
        std::vector<RecoWork*> work;
        auto cleanup = [&] {
            for (hipEvent_t e : ev)
                (void)hipEventDestroy(e);
        };
        try {
            auto event = [&](bool timing) {
                hipEvent_t e;
                SMIP_HIP(timing ? hipEventCreate(&e) : hipEventCreateWithFlags(&e, hipEventDisableTiming));
                ev.push_back(e);
                return e;
            };
            for (uint64_t k = 0; k < n; ++k) {
                smip_job* j = jobs[k];
                flush_staged(j); // the last, partial batch of pushed frames
                scale_once(j, j->stream); // bispectrum /= complex(n, 0)
                hipEvent_t e = event(false);
                SMIP_HIP(hipEventRecord(e, j->stream));
                SMIP_HIP(hipStreamWaitEvent(cs, e, 0));
            }
            std::vector<const void*> B(n);
            std::vector<int> f32(n, 0);
            std::vector<double2*> ph(n);
            std::vector<uint8_t*> fl(n);
            std::vector<double*> co(n);
            std::vector<int32_t> calls(n, 0), st(n, 0);
            for (uint64_t k = 0; k < n; ++k) {
                smip_job* j = jobs[k];
                work.push_back(j->reco);
                B[k] = j->B64.get();
                ph[k] = j->phases.get();
                fl[k] = j->flags.get();
                co[k] = j->cons.get();
            }
            hipEvent_t t0 = event(true), t1 = event(true);
            injected_failure("reco");
            SMIP_HIP(hipEventRecord(t0, cs));
            // the server's recoveries (on its own-queue streams) keep the general
            // schedule: the one hang seen with the push chain was there
            run_reconstruct_many(work.data(), (int)n, B.data(), f32.data(), ph.data(), fl.data(), co.data(),
                calls.data(), st.data(), cs, group_stream == nullptr);
            SMIP_HIP(hipEventRecord(t1, cs));
            SMIP_HIP(hipEventSynchronize(t1));
            float reco_ms = 0.f;
            SMIP_HIP(hipEventElapsedTime(&reco_ms, t0, t1));
            for (uint64_t k = 0; k < n; ++k) {
                smip_job* j = jobs[k];
                const smip_job_params& p = j->p;
                j->reco_calls = (uint64_t)calls[k];
                j->ms[2] += reco_ms; // the group's recovery time, shared
                note(k, st[k], reco_error_text(st[k]));
                if (st[k] != SMIP_OK)
                    continue; // left unfinalized, as smip_job_finalize leaves a job that throws
                SMIP_HIP(hipStreamWaitEvent(j->stream, t1, 0));
                hipEvent_t e2 = mark(j);
                const double norm = (double)(j->nframes * p.ncols * p.nrows);
                launch_synth_prepare(j->P_half.get(), j->phases.get(), (int32_t)p.ncols, (int32_t)p.nrows,
                    p.reco_radius, norm, j->image.get(), j->stream);
                launch_c2c_inverse_f64(j->plan, j->image.get(), nullptr, j->stream);
                launch_abs_max_normalize(j->image.get(), (int64_t)(p.ncols * p.nrows), j->scratch.get(), j->stream);
                hipEvent_t e3 = mark(j);
                record(j, 3, e2, e3);
            }
            for (uint64_t k = 0; k < n; ++k) {
                SMIP_HIP(hipStreamSynchronize(jobs[k]->stream));
                if (st[k] == SMIP_OK)
                    jobs[k]->finalized = true;
            }
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
    if (rc != SMIP_OK)
        return rc;
    if (first_err != SMIP_OK)
        set_last_error(first_msg.c_str());
    return first_err;
}

extern "C" SMIP_API int smip_job_synchronize(smip_job* j)
{
    return guarded([&] {
        require(j != nullptr, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        ensure_device(j->ctx);
        SMIP_HIP(hipStreamSynchronize(j->stream));
    });
}

extern "C" SMIP_API int smip_job_get_stats(smip_job* j, smip_job_stats* out)
{
    return guarded([&] {
        require(j && out, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        ensure_device(j->ctx);
        SMIP_HIP(hipStreamSynchronize(j->stream));
        drain_events(j);
        out->frames = j->nframes;
        out->accumulate_ms = j->ms[0];
        out->saa_ms = j->ms[1];
        out->reco_ms = j->ms[2];
        out->synth_ms = j->ms[3];
        out->footprint = j->footprint;
        out->reco_calls = j->reco_calls;
        out->cells_per_frame = j->tile ? tile_plan_cells(j->tile) : 0;
        out->triples_per_frame = j->triples;
        out->tile_ms = j->ms[4];
        out->tile_launches = j->tile_launches;
        out->graph_launches = j->graph_launches;
    });
}

extern "C" SMIP_API int smip_job_bispectrum(smip_job* j, smip_bispec** out)
{
    return guarded([&] {
        require(j && out, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        ensure_device(j->ctx);
        SMIP_HIP(hipStreamSynchronize(j->stream));
        *out = j->view.get();
    });
}

extern "C" SMIP_API int smip_job_get_powerspec(smip_job* j, double* out)
{
    return guarded([&] {
        require(j && out, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(j->nframes > 0, SMIP_ERR_RUNTIME, "no frames accumulated");
        ensure_device(j->ctx);
        const uint64_t np = j->p.ncols * j->p.nrows;
        auto tmp = dmake<double>(np);
        launch_power_expand(j->P_half.get(), (int32_t)j->p.ncols, (int32_t)j->p.nrows, (double)(j->nframes * np),
            tmp.get(), j->stream);
        SMIP_HIP(hipMemcpyAsync(out, tmp.get(), sizeof(double) * np, hipMemcpyDeviceToHost, j->stream));
        SMIP_HIP(hipStreamSynchronize(j->stream));
        IdleRelease idle;
        tmp.reset();
    });
}

extern "C" SMIP_API int smip_job_get_sum_image(smip_job* j, double* out)
{
    return guarded([&] {
        require(j && out, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(j->p.saa, SMIP_ERR_LOGIC, "job was created without shift-and-add");
        require(j->nframes > 0, SMIP_ERR_RUNTIME, "no frames accumulated");
        ensure_device(j->ctx);
        const uint64_t np = j->p.ncols * j->p.nrows;
        SMIP_HIP(hipMemcpyAsync(out, j->sum.get(), sizeof(double) * np, hipMemcpyDeviceToHost, j->stream));
        SMIP_HIP(hipStreamSynchronize(j->stream));
        const double n = (double)j->nframes; // sumarray /= nframes
        for (uint64_t k = 0; k < np; ++k)
            out[k] /= n;
    });
}

extern "C" SMIP_API int smip_job_get_shifts(smip_job* j, int32_t* out, uint64_t count)
{
    return guarded([&] {
        require(j && (out || count == 0), SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(j->p.saa, SMIP_ERR_LOGIC, "job was created without shift-and-add");
        require(count <= j->nframes, SMIP_ERR_OUT_OF_RANGE, "more shifts requested than frames");
        ensure_device(j->ctx);
        if (count)
            SMIP_HIP(hipMemcpyAsync(out, j->shifts.get(), sizeof(int32_t) * 2 * count, hipMemcpyDeviceToHost,
                j->stream));
        SMIP_HIP(hipStreamSynchronize(j->stream));
    });
}

extern "C" SMIP_API int smip_job_get_phases(smip_job* j, double* phases, uint8_t* flags, double* consistency)
{
    return guarded([&] {
        require(j != nullptr, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(j->finalized, SMIP_ERR_BAD_FUNCTION_CALL, "job not finalized");
        ensure_device(j->ctx);
        const uint64_t np = j->p.ncols * j->p.nrows;
        hipStream_t s = j->stream;
        if (phases)
            SMIP_HIP(hipMemcpyAsync(phases, j->phases.get(), sizeof(double2) * np, hipMemcpyDeviceToHost, s));
        if (flags)
            SMIP_HIP(hipMemcpyAsync(flags, j->flags.get(), np, hipMemcpyDeviceToHost, s));
        if (consistency)
            SMIP_HIP(hipMemcpyAsync(consistency, j->cons.get(), sizeof(double) * np, hipMemcpyDeviceToHost, s));
        SMIP_HIP(hipStreamSynchronize(s));
    });
}

extern "C" SMIP_API int smip_job_get_image(smip_job* j, double* out)
{
    return guarded([&] {
        require(j && out, SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        require(j->finalized, SMIP_ERR_BAD_FUNCTION_CALL, "job not finalized");
        ensure_device(j->ctx);
        const uint64_t np = j->p.ncols * j->p.nrows;
        SMIP_HIP(hipMemcpyAsync(out, j->image.get(), sizeof(double2) * np, hipMemcpyDeviceToHost, j->stream));
        SMIP_HIP(hipStreamSynchronize(j->stream));
    });
}

namespace smip_amd {
// (u, v) tuples the accumulate loop of include/bispectrum.h:393-398 adds per frame
uint64_t count_triples(const LoopBounds& L)
{
    uint64_t n = 0;
    for (int32_t i = L.min1; i <= 0; ++i)
        for (int32_t k = L.min3; k <= 0; ++k) {
            if (i + k < L.min1 || i + k > L.max1)
                continue;
            for (int32_t j = L.min2; j <= L.max2; ++j) {
                const int32_t lo = std::max(L.min4, L.min2 - j), hi = std::min(L.max4, L.max2 - j);
                if (hi >= lo)
                    n += (uint64_t)(hi - lo + 1);
            }
        }
    return n;
}
} // namespace smip_amd
