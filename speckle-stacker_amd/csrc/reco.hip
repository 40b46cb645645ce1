// reco.hip -- closure-phase recovery on the device.
//
// Replaces reconstruct_phases / calc_phase (include/phasereco.h:33-123).  The
// spiral visit order depends on libm cos/sin and is generated on the host by a
// literal restatement of NextRecoIndex (src/phasereco.cpp:17-36); only visits
// inside the PhaseMap range are uploaded (the others do nothing).
//
// calc_phase(w) averages, over every u with v = w - u in the stored v range and
// both flags set *at the time of the call*, the unit phasors
// phi(u) phi(v) conj(beta / |beta|) (|beta| > DBL_EPSILON), then sets flag,
// consistency and phase of w.  The recursion is sequential, so the schedule
// splits the visit list into windows of K visits:
//
//  * k_reco_bulk (K workgroups, one per visit of the window) sums the pairs
//    whose u and v were both flagged before the window -- exact, since flags
//    are never cleared -- and writes a "pending" record for every pair in which
//    u or v is a point visited earlier in the same window (it may get flagged
//    before this visit): beta-hat, the pre-window flags and phases, and the
//    window slots of the in-window points.  At most 2(K-1) records per visit.
//  * k_reco_fixup (one workgroup) walks the window's visits in order on one
//    wavefront; each step adds the pending pairs whose points are flagged by
//    now and finalises the point exactly as calc_phase does.  Helper
//    wavefronts run a few visits ahead and pre-sum every pending pair whose
//    outcome is already decided, so the step itself handles only the pairs
//    with a point flagged in the last few visits (reco_fixup_body).
//
// Eligibility is evaluated exactly, so the flag mask is bit-exact; only the
// order of the FP64 summation of the mean differs from the reference's raster
// order.  k_reco_sequential keeps the one-visit-at-a-time form as a cross-check.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "smip_complex.h"
#include "smip_internal.h"

namespace smip_amd {

constexpr double kEps = 2.220446049250313e-16; // constants::c_epsilon<double>
constexpr int kWindow = 512;                   // visits per window (256: 2% slower at 256^2, b=48; 128: 9%)
constexpr int kBulkThreads = 1024;            // lanes per visit in the bulk pass (latency-bound loads)

struct RecoArgs {
    Desc d;
    int32_t X, Y;
    int32_t pmn0, pmn1, pmx0, pmx1; // PhaseMap signed range (Array2::range)
    int64_t nvisits;
    int32_t push; // 1: windows whose schedule is static take the push chain
    int32_t pad;
};

enum : int32_t { kPreU = 1, kPreV = 2, kValid = 4, kErr = 8 };

// Per-visit status word E of the bulk pass: an eligible pre-window lookup was
// out of range; the window cannot take the push chain (see reco_push_window).
enum : int32_t { kEOutOfRange = 1, kEIrregular = 2 };

// The push chain: a visit's terms whose younger in-window point was flagged at
// most kPushAges visits before it are folded in by the chain itself, linearly
// (one wavefront lane per pending visit); helpers prepare everything else from
// the phases as they stood kPushLag visits before the visit.
constexpr int kPushAges = 64;
constexpr int kPushLag = 96;
constexpr int kAgeNone = 1 << 30; // a point flagged before the window

// c = conj(beta/|beta|) times the pre-window phase of whichever of u, v was
// flagged before the window (records never have both); the fixup multiplies in
// the phase(s) of the in-window point(s).
struct __align__(16) PendingRec {
    int32_t ju, jv; // window slots of u / v (-1: not an in-window point)
    int32_t f;      // kPreU | kPreV | kValid | kErr
    int32_t pad;
    double2 c;
};

// Device copy of one geometry's visit list: the in-range spiral visits and, per
// visit, the window slot of the first visit of the same point in its window.
// Read-only once built, so jobs of one geometry share it.
struct RecoGeom {
    int32_t X = 0, Y = 0;
    int64_t nvisits = 0;
    void* block = nullptr; // visits, then first_pos
    const int2* visits = nullptr;
    const int32_t* first_pos = nullptr;
    uint64_t bytes = 0;
    ~RecoGeom() { dev_free(block); }
};

// Per-recovery scratch, carved from one allocation.
struct RecoWork {
    RecoArgs args;
    int64_t nwin = 0;
    int cap = 0; // records per visit
    std::shared_ptr<const RecoGeom> geom;
    void* block = nullptr;
    const int2* visits = nullptr;
    const int32_t* first_pos = nullptr;
    int64_t* map = nullptr; // point -> (window << 32) | slot of its first visit
    double2* S = nullptr;
    int32_t* N = nullptr; // -1: the visit does nothing
    int32_t* E = nullptr;
    int32_t* cnt = nullptr;
    PendingRec* rec = nullptr;
    int32_t* status = nullptr; // [0] error code, [1] calls, [2] flagged
};

__device__ inline int32_t pm_at(const RecoArgs& a, int32_t x, int32_t y)
{
    const int32_t xx = x < 0 ? x + a.X : x;
    const int32_t yy = y < 0 ? y + a.Y : y;
    return yy * a.X + xx;
}

__device__ inline bool calc_phase_returns_early(const RecoArgs& a, int2 w)
{
    // include/phasereco.h:84-86
    return abs(w.x) + abs(w.y) <= 1 || w.x < a.d.min_idx[0] || w.x > a.d.max_idx[0] || w.y < a.d.min_idx[1]
        || w.y > a.d.max_idx[1];
}

__device__ inline double2 load_beta(const void* B, int is_f32, uint64_t addr, bool cj)
{
    double2 beta;
    if (is_f32) {
        const float2 f = static_cast<const float2*>(B)[addr];
        beta = mk<double2>(f.x, f.y);
    } else {
        beta = static_cast<const double2*>(B)[addr];
    }
    if (cj)
        beta.y = -beta.y;
    return beta;
}

// phi(u) phi(v) conj(beta/|beta|), normalised -- include/phasereco.h:97-105.
// |q| as sqrt(re^2 + im^2): the operands are unit phasors, far from the range
// where hypot's scaling matters.
__device__ inline double2 normalised(double2 q)
{
    // For |q|^2 = 1 + d, 1/sqrt(1 + d) = 1 - d/2 + O(d^2): one Newton step from 1
    // is exact to rounding when |d| < 1e-8; anything else (including q = 0,
    // which gives NaN as the reference's 0/0 does) takes the exact path.
    const double m2 = q.x * q.x + q.y * q.y;
    const double inv = fabs(m2 - 1.0) < 1e-8 ? 1.5 - 0.5 * m2 : 1.0 / sqrt(m2);
    return mk<double2>(q.x * inv, q.y * inv);
}

__device__ inline double2 unit_phasor(double2 pu, double2 pv, double2 bhat)
{
    return normalised(cmul(cmul(pu, pv), bhat));
}

// Sum over the 64 lanes of a wavefront with DPP row operations (quad xor 1 and
// 2, half-row and row mirrors, then row broadcasts 15 and 31); the total ends in
// lane 63 and is broadcast with readlane.  A few cycles per level, against an
// LDS round trip per level for shuffle-based butterflies.
template <int CTRL, int ROWS = 0xf>
__device__ inline int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
// the same where every lane reads an active lane of its own row (quad
// permutes and mirrors): no old value to carry
template <int CTRL>
__device__ inline double dpp_row_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL, int ROWS = 0xf>
__device__ inline double dpp_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i<CTRL, ROWS>((int)(b & 0xffffffff));
    const int hi = dpp_i<CTRL, ROWS>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double wave_sum(double v)
{
    v += dpp_d<0xB1>(v);        // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);        // quad_perm [2,3,0,1]
    v += dpp_d<0x141>(v);       // row_half_mirror
    v += dpp_d<0x140>(v);       // row_mirror
    v += dpp_d<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
    v += dpp_d<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// count and complex sum of a wavefront at once: three independent DPP chains
// interleaved, so their latencies overlap
__device__ inline void wave_sum3(int& n, double& re, double& im)
{
#define SMIP_LEVEL(CTRL, ROWS)                                                                              \
    {                                                                                                       \
        const int tn = dpp_i<CTRL, ROWS>(n);                                                                \
        const double tr = dpp_d<CTRL, ROWS>(re), ti = dpp_d<CTRL, ROWS>(im);                                \
        n += tn;                                                                                            \
        re += tr;                                                                                           \
        im += ti;                                                                                           \
    }
    SMIP_LEVEL(0xB1, 0xf)
    SMIP_LEVEL(0x4E, 0xf)
    SMIP_LEVEL(0x141, 0xf)
    SMIP_LEVEL(0x140, 0xf)
    SMIP_LEVEL(0x142, 0xa)
    SMIP_LEVEL(0x143, 0xc)
#undef SMIP_LEVEL
    n = __builtin_amdgcn_readlane(n, 63);
    const long long br = __double_as_longlong(re), bi = __double_as_longlong(im);
    re = __longlong_as_double(((long long)__builtin_amdgcn_readlane((int)(br >> 32), 63) << 32)
        | (unsigned int)__builtin_amdgcn_readlane((int)(br & 0xffffffff), 63));
    im = __longlong_as_double(((long long)__builtin_amdgcn_readlane((int)(bi >> 32), 63) << 32)
        | (unsigned int)__builtin_amdgcn_readlane((int)(bi & 0xffffffff), 63));
}

__device__ inline int wave_sum(int v)
{
    v += dpp_i<0xB1>(v);
    v += dpp_i<0x4E>(v);
    v += dpp_i<0x141>(v);
    v += dpp_i<0x140>(v);
    v += dpp_i<0x142, 0xa>(v);
    v += dpp_i<0x143, 0xc>(v);
    return __builtin_amdgcn_readlane(v, 63);
}

// Orders LDS accesses of one wavefront for the compiler; the hardware executes a
// wavefront's LDS instructions in issue order, so no wait is needed.
__device__ inline void wave_lds_order()
{
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// 1/sqrt(x) and 1/x from the hardware approximations (about 2^-23 relative)
// and two Newton-Raphson steps, each squaring the relative error: 2^-46, then
// rounding level.  A fraction of the latency of the IEEE sqrt and divide
// sequences on the recursion's serial path; phases and consistencies carry a
// tolerance, and the flags do not depend on them.
__device__ inline double rsqrt_nr(double x)
{
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double e = fma(-(x * y), y, 1.0); // 1 - x y^2
        y = fma(0.5 * y, e, y);
    }
    return y;
}
// the same with one Newton-Raphson step: about 2^-44 relative; the push chain
// scales a phasor by it, which leaves its direction exact to rounding
__device__ inline double rsqrt_nr1(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-(x * y), y, 1.0);
    return fma(0.5 * y, e, y);
}
__device__ inline double rcp_nr(double x)
{
    double y = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int k = 0; k < 2; ++k)
        y = fma(y, fma(-x, y, 1.0), y);
    return y;
}

// conj(beta) / |beta| when |beta| > epsilon (the reference's test), else false.
// |beta|^2 is compared with epsilon^2 and 1/|beta| comes from rsqrt_nr: the same
// decision as hypot(beta) > epsilon except within an ulp of epsilon, and the
// same unit phasor to rounding, for about a fifth of the FP64 instructions of
// hypot and two divisions.  The bulk pass evaluates it for every record of every
// visit.
__device__ inline bool beta_hat(double2 beta, double2* bh)
{
    const double m2 = fma(beta.x, beta.x, beta.y * beta.y);
    if (!(m2 > kEps * kEps))
        return false;
    const double ib = rsqrt_nr(m2);
    *bh = mk<double2>(beta.x * ib, -(beta.y * ib));
    return true;
}

// phase sum/|sum| and consistency |sum|/n of a point (include/phasereco.h:107-121)
__device__ inline double2 point_phase(double sr, double si, int n, double* consistency)
{
    // mean = sum / n; consistency = |mean|; phase = mean / |mean| = sum / |sum|
    const double m2 = sr * sr + si * si;
    const double ia = m2 > 0.0 ? rsqrt_nr(m2) : 0.0;
    const double ap = m2 * ia * rcp_nr((double)n);
    *consistency = ap;
    return ap > kEps ? mk<double2>(sr * ia, si * ia) : mk<double2>(0.0, 0.0);
}

__device__ inline void finalize_point(double sr, double si, int n, double2* phases, uint8_t* flags, double* cons,
    int32_t pw, double2* out_phase)
{
    // mean = sum / n; consistency = |mean|; phase = mean / |mean| = sum / |sum|
    double ap;
    const double2 ph = point_phase(sr, si, n, &ap);
    flags[pw] = 1;
    cons[pw] = ap;
    phases[pw] = ph;
    if (out_phase)
        *out_phase = ph;
}

// ---------------------------------------------------------------------------
// windowed schedule
// ---------------------------------------------------------------------------
__device__ __forceinline__ void reco_map_body(int bx, const int2* __restrict__ visits,
    const int32_t* __restrict__ first_pos, const RecoArgs& a, int64_t s, int64_t e, int64_t win,
    int64_t* __restrict__ map)
{
    const int64_t k = s + (int64_t)bx * blockDim.x + threadIdx.x;
    if (k >= e)
        return;
    if (first_pos[k] == (int32_t)(k - s)) {
        const int2 w = visits[k];
        map[pm_at(a, w.x, w.y)] = (win << 32) | (int64_t)(k - s);
    }
}

// A pending record as the push chain classifies it, for visit slot j: whether
// it counts once the window's points are flagged at their first visits (every
// in-window point is one whose calc_phase can flag it), and the ages (visits
// between the flagging and visit j) of its two points, kAgeNone for a point
// flagged before the window.
__device__ inline bool push_counts(const RecoArgs& a, const PendingRec& r, int2 u, int2 v)
{
    return (r.f & kValid) && ((r.f & kPreU) || !calc_phase_returns_early(a, u))
        && ((r.f & kPreV) || !calc_phase_returns_early(a, v));
}
__device__ inline void push_ages(const PendingRec& r, int j, int& young, int& old)
{
    const int au = (r.f & kPreU) ? kAgeNone : j - r.ju, av = (r.f & kPreV) ? kAgeNone : j - r.jv;
    young = min(au, av);
    old = max(au, av);
}
// The record fits the push chain: its younger point is at most kPushAges old and
// the other one older than kPushLag (the chain adds it), or both are older than
// kPushLag (a helper adds it).
__device__ inline bool push_fits(int young, int old)
{
    return young <= kPushAges ? old > kPushLag : young > kPushLag;
}

template <bool PUSH>
__device__ __forceinline__ void reco_bulk_body(int bx, const RecoArgs& a, const int2* __restrict__ visits,
    const int32_t* __restrict__ first_pos, int64_t s, int64_t win, const int64_t* __restrict__ map,
    const uint8_t* __restrict__ flags, const double2* __restrict__ phases, const void* __restrict__ B, int is_f32,
    double2* __restrict__ S, int32_t* __restrict__ N, int32_t* __restrict__ E, int32_t* __restrict__ cnt,
    PendingRec* __restrict__ rec, int cap, const int32_t* __restrict__ status)
{
    __shared__ double red_re[kBulkThreads / 64], red_im[kBulkThreads / 64];
    __shared__ int red_n[kBulkThreads / 64], red_e[kBulkThreads / 64], red_p[kBulkThreads / 64];
    __shared__ int push_irregular, push_counted;
    const int j = bx;
    const int tid = threadIdx.x;
    if (status[0])
        return; // an earlier window raised
    const int2 w = visits[s + j];
    const int32_t pw = pm_at(a, w.x, w.y);
    if (flags[pw] || calc_phase_returns_early(a, w)) {
        if (tid == 0) {
            N[j] = -1;
            cnt[j] = 0;
            E[j] = 0;
        }
        return;
    }
    if (PUSH && tid == 0) {
        push_irregular = 0;
        push_counted = 0;
    }
    const Desc& d = a.d;
    const int vmn0 = d.min_idx[2], vmn1 = d.min_idx[3];
    const int nvx = d.max_idx[2] - vmn0 + 1, nvy = d.max_idx[3] - vmn1 + 1;
    const int ncand = nvx * nvy;
    double sr = 0.0, si = 0.0;
    int n = 0, err = 0;
    uint64_t pend = 0; // bit it: candidate tid + it * blockDim is pending
    int it = 0;
    for (int c = tid; c < ncand; c += blockDim.x, ++it) {
        const int vy = vmn1 + c / nvx, vx = vmn0 + c % nvx;
        const int ux = w.x - vx, uy = w.y - vy;
        if (ux < a.pmn0 || ux > a.pmx0 || uy < a.pmn1 || uy > a.pmx1)
            continue;
        const int32_t pu = pm_at(a, ux, uy), pv = pm_at(a, vx, vy);
        const bool fu = flags[pu] != 0, fv = flags[pv] != 0;
        if (fu && fv) {
            uint64_t addr;
            bool cj;
            if (resolve_element(d, ux, uy, vx, vy, &addr, &cj)) {
                err = 1;
                continue;
            }
            double2 bh;
            if (beta_hat(load_beta(B, is_f32, addr, cj), &bh)) {
                const double2 q = unit_phasor(phases[pu], phases[pv], bh);
                sr += q.x;
                si += q.y;
                ++n;
            }
        } else {
            const int64_t mu = map[pu], mv = map[pv];
            const bool iu = (mu >> 32) == win && (int32_t)(mu & 0xffffffff) < j;
            const bool iv = (mv >> 32) == win && (int32_t)(mv & 0xffffffff) < j;
            if ((fu || iu) && (fv || iv))
                pend |= 1ull << it;
        }
    }
    // block reductions in a fixed order
    int np = __popcll(pend);
    int incl = np; // inclusive scan of pending counts over the block
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o);
        si += __shfl_xor(si, o);
        n += __shfl_xor(n, o);
        err |= __shfl_xor(err, o);
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if ((tid & 63) >= o)
            incl += t;
    }
    if ((tid & 63) == 63)
        red_p[tid >> 6] = incl;
    if ((tid & 63) == 0) {
        red_re[tid >> 6] = sr;
        red_im[tid >> 6] = si;
        red_n[tid >> 6] = n;
        red_e[tid >> 6] = err;
    }
    __syncthreads();
    int base = incl - np;
    for (int wv = 0; wv < (tid >> 6); ++wv)
        base += red_p[wv];
    int tn = 0, te = 0;
    if (tid == 0) {
        double tr = 0.0, ti = 0.0;
        int tp = 0;
        for (int wv = 0; wv < (int)(blockDim.x >> 6); ++wv) {
            tr += red_re[wv];
            ti += red_im[wv];
            tn += red_n[wv];
            te |= red_e[wv];
            tp += red_p[wv];
        }
        S[j] = mk<double2>(tr, ti);
        N[j] = tn;
        cnt[j] = tp;
        if (!PUSH)
            E[j] = te;
    }
    // pending records, in (thread, candidate) order; for the first visit of a
    // point in the window, also whether the push chain can take it
    const bool first_visit = first_pos[s + j] == j;
    int irregular = 0, counted = 0;
    PendingRec* out = rec + (int64_t)j * cap + base;
    it = 0;
    for (int c = tid; c < ncand; c += blockDim.x, ++it) {
        if (!(pend & (1ull << it)))
            continue;
        const int vy = vmn1 + c / nvx, vx = vmn0 + c % nvx;
        const int ux = w.x - vx, uy = w.y - vy;
        const int32_t pu = pm_at(a, ux, uy), pv = pm_at(a, vx, vy);
        const bool fu = flags[pu] != 0, fv = flags[pv] != 0;
        const int64_t mu = map[pu], mv = map[pv];
        PendingRec r;
        r.ju = ((mu >> 32) == win && (int32_t)(mu & 0xffffffff) < j) ? (int32_t)(mu & 0xffffffff) : -1;
        r.jv = ((mv >> 32) == win && (int32_t)(mv & 0xffffffff) < j) ? (int32_t)(mv & 0xffffffff) : -1;
        r.f = (fu ? kPreU : 0) | (fv ? kPreV : 0);
        r.pad = 0;
        r.c = mk<double2>(0, 0);
        uint64_t addr;
        bool cj;
        if (resolve_element(d, ux, uy, vx, vy, &addr, &cj)) {
            r.f |= kErr;
        } else {
            double2 bh;
            if (beta_hat(load_beta(B, is_f32, addr, cj), &bh)) {
                r.f |= kValid;
                r.c = fu ? cmul(phases[pu], bh) : (fv ? cmul(phases[pv], bh) : bh);
            }
        }
        if (PUSH && first_visit) {
            if (r.f & kErr) {
                irregular = 1;
            } else if (push_counts(a, r, make_int2(ux, uy), make_int2(vx, vy))) {
                int young, old;
                push_ages(r, j, young, old);
                ++counted;
                irregular |= !push_fits(young, old);
            }
        }
        *out++ = r;
    }
    if (PUSH) {
        if (irregular)
            atomicOr(&push_irregular, 1);
        if (counted)
            atomicAdd(&push_counted, counted);
        __syncthreads();
        if (tid == 0) {
            // a first visit that would not flag breaks the push chain's static schedule too
            const bool irr = first_visit && (push_irregular || tn + push_counted == 0);
            E[j] = te | (irr ? kEIrregular : 0);
        }
    }
}

// Step scalars of one visit, 32 bytes: two 16-byte LDS loads.
struct StepInfo {
    int32_t n0;  // pre-window eligible count, -1: the visit does nothing
    int32_t fp;  // window slot of the first visit of this point
    int32_t cnt; // pending records
    int32_t err; // an eligible pre-window lookup was out of range
    double2 S;   // pre-window sum
};

typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
typedef int32_t i32x3 __attribute__((ext_vector_type(3)));

constexpr double kNotYet = 2.0;           // ppos marker: slot not flagged in this window
// Waves of a fixup workgroup: the chain wave and its helpers.  The push chain's
// helpers have more to do per visit than the general schedule's.
constexpr int kPushWaves = 12;
constexpr int kPushThreads = 64 * kPushWaves;
constexpr int kPushHelpers = kPushWaves - 1;
constexpr int kGenWaves = 8;
constexpr int kGenThreads = 64 * kGenWaves;
constexpr int kLead = 8;                  // visits between a helper's snapshot and its visit
constexpr int kRing = 16;                 // near-record slots, more than kLead
constexpr int kNearCap = 32;              // near records per visit; more: the chain takes every record
constexpr int32_t kNever = 0x7fffffff;    // ftime of a slot not flagged in this window
constexpr uint32_t kSpinLimit = 1u << 24; // polls (each after a short sleep) before a wait counts as stuck

// Sum over the first 16 lanes (row 0) with four DPP levels; the total ends in
// every lane of the row and is read from lane 0.  Used when only lanes 0-15
// hold terms.
__device__ inline void row_sum3(int& n, double& re, double& im)
{
#define SMIP_LEVEL(CTRL)                                                                                    \
    {                                                                                                       \
        const int tn = dpp_i<CTRL, 0xf>(n);                                                                 \
        const double tr = dpp_d<CTRL, 0xf>(re), ti = dpp_d<CTRL, 0xf>(im);                                  \
        n += tn;                                                                                            \
        re += tr;                                                                                           \
        im += ti;                                                                                           \
    }
    SMIP_LEVEL(0xB1)  // quad_perm [1,0,3,2]
    SMIP_LEVEL(0x4E)  // quad_perm [2,3,0,1]
    SMIP_LEVEL(0x141) // row_half_mirror
    SMIP_LEVEL(0x140) // row_mirror
#undef SMIP_LEVEL
    n = __builtin_amdgcn_readfirstlane(n);
    const long long br = __double_as_longlong(re), bi = __double_as_longlong(im);
    re = __longlong_as_double(((long long)__builtin_amdgcn_readfirstlane((int)(br >> 32)) << 32)
        | (unsigned int)__builtin_amdgcn_readfirstlane((int)(br & 0xffffffff)));
    im = __longlong_as_double(((long long)__builtin_amdgcn_readfirstlane((int)(bi >> 32)) << 32)
        | (unsigned int)__builtin_amdgcn_readfirstlane((int)(bi & 0xffffffff)));
}

// Complex sums over the first 16 lanes (row_sum2) or the whole wavefront
// (wave_sum2), as row_sum3 / wave_sum3 without the count.
__device__ inline double dpp_rfl(double v)
{
    const long long b = __double_as_longlong(v);
    return __longlong_as_double(((long long)__builtin_amdgcn_readfirstlane((int)(b >> 32)) << 32)
        | (unsigned int)__builtin_amdgcn_readfirstlane((int)(b & 0xffffffff)));
}
__device__ inline void row_sum2(double& re, double& im)
{
#define SMIP_LEVEL(CTRL)                                                                                    \
    {                                                                                                       \
        const double tr = dpp_row_d<CTRL>(re), ti = dpp_row_d<CTRL>(im);                                    \
        re += tr;                                                                                           \
        im += ti;                                                                                           \
    }
    SMIP_LEVEL(0xB1)  // quad_perm [1,0,3,2]
    SMIP_LEVEL(0x4E)  // quad_perm [2,3,0,1]
    SMIP_LEVEL(0x141) // row_half_mirror
    SMIP_LEVEL(0x140) // row_mirror
#undef SMIP_LEVEL
    re = dpp_rfl(re); // lane 0 holds the row's total
    im = dpp_rfl(im);
}
__device__ inline void wave_sum2(double& re, double& im)
{
    int n = 0;
    wave_sum3(n, re, im);
}

// Cross-wave signalling through LDS.  A wavefront's LDS instructions execute
// in issue order, so a flag written after its data and read before it orders
// them.  The flags are relaxed atomics (read afresh, never cached in a
// register) and signal fences keep the compiler from moving other accesses
// across them; unlike volatile accesses, neither adds a wait for the access to
// complete, which would cost the chain an LDS round trip each time.
__device__ inline int32_t lds_get(const int32_t* p)
{
    return __builtin_amdgcn_readfirstlane(
        __hip_atomic_load(const_cast<int32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
}
__device__ inline void lds_put(int32_t* p, int32_t v)
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// the ready word of a visit's helper header
__device__ inline int32_t* ready_flag(i32x4* head, int k) { return reinterpret_cast<int32_t*>(&head[k]) + 2; }

// Waits until *v >= want.  False when the workgroup aborted meanwhile, or when
// the poll count ran out: then the abort flag and status[0] (if still clear) are
// set, so every wave leaves its loop and the kernel ends.
__device__ inline bool wait_at_least(const int32_t* v, int32_t want, int32_t* abort_flag, int32_t* status, int lane)
{
    for (uint32_t it = 0;; ++it) {
        if (lds_get(v) >= want) {
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            return true;
        }
        if (lds_get(abort_flag))
            return false;
        if (it >= kSpinLimit) {
            if (lane == 0)
                atomicCAS(status, 0, SMIP_ERR_RUNTIME);
            lds_put(abort_flag, 1);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// One side of a pending record, as a helper sees it for visit k from the
// snapshot after visit p - 1 (p = k - kLead): 0 flagged before p (it stays
// flagged), 1 undecided (the point is visited in [q, k), q = max(p, 0), so it may
// be flagged by visit k), 2 not flagged at visit k.
__device__ inline int side_state(bool pre, int32_t j, int32_t p, int32_t q, const int32_t* ftime,
    const int32_t (&fpw)[kLead])
{
    if (pre)
        return 0;
    j = max(j, 0);
    if (ftime[j] < p)
        return 0;
    bool seen = j >= q; // first visit of the point inside [q, k)
#pragma unroll
    for (int i = 0; i < kLead; ++i)
        seen |= fpw[i] == j; // a later visit of it inside [q, k)
    return seen ? 1 : 2;
}

// LDS of the fixup workgroup: the part both schedules use, then the general
// schedule's state or the push chain's, which share memory.
struct FixCommon {
    StepInfo info[kWindow];
    // phase and consistency of each window slot once flagged in this window
    // (ppos x = kNotYet before: no unit or zero phase has x = 2); the window's
    // results go to global memory once, at the end
    double2 ppos[kWindow];
    double pcons[kWindow];
    int32_t progress, abort_flag, degenerate, pad;
};
struct FixGeneral {
    int32_t ftime[kWindow]; // visit that flagged each slot
    double2 far_s[kWindow];
    // per visit: far count + (errors << 16), near count (-1: the chain takes
    // every record of the visit), ready flag; written as one 16-byte store
    // after far_s, read as one 16-byte load before it
    i32x4 head[kWindow];
    PendingRec nearbuf[kRing][kNearCap];
};
// A helper's result for one visit of the push chain: A = pre-window sum plus the
// sum over the records whose points are both older than kPushLag visits; the
// eligible count n and thr = (DBL_EPSILON n)^2, below which |sum|^2 means a
// zero phase.  call = 0: the visit runs no calc_phase.
struct __align__(16) PushHead {
    double2 A;
    double thr;
    int32_t call, n;
};
// Q_k[a] lives in row (k - a) mod kPushLag, the step that adds it, at column
// k mod 64, the lane that holds visit k: the chain reads one contiguous row per
// step.  Rows are padded by one entry so that a helper's 64 entries of one
// visit, one per row, fall in different banks.
constexpr int kPushRow = kPushAges + 1;
__device__ inline int push_slot(int k, int a) { return ((k - a) % kPushLag) * kPushRow + (k & 63); }
constexpr int kReadyWords = kWindow + kPushAges + 8;
struct FixPush {
    PushHead head[kWindow + 2]; // + the step that pads an odd window
    // ready[k + 1]: head[k] and the entries Q_k are done (1 beyond the window)
    int32_t ready[kReadyWords];
    double2 Q[kPushLag * kPushRow];
    int32_t ncalls;
};
struct FixLds {
    FixCommon c;
    union {
        FixGeneral g;
        FixPush p;
    } u;
};
// the general schedule alone (grouped recoveries)
struct FixLdsGeneral {
    FixCommon c;
    union {
        FixGeneral g;
    } u;
};

__device__ inline double readlane_d(double v, int l)
{
    const long long b = __double_as_longlong(v);
    return __longlong_as_double(((long long)__builtin_amdgcn_readlane((int)(b >> 32), l) << 32)
        | (unsigned int)__builtin_amdgcn_readlane((int)(b & 0xffffffff), l));
}

// The push chain, for a window whose schedule is static: no visit of it has
// kEIrregular, so every first visit of a point that calc_phase would run flags
// that point, the visits that run are known in advance, and so is, for every
// pending record, the visit at which each of its points is flagged.
//
// Wave 0 (the chain) keeps in lane k mod 64 the running sum of visit k, which it
// finalises within the next 64 steps.  At step t it finalises visit t (lane
// t mod 64: that sum plus the helper's part A_t), hands the lane to visit t + 64,
// and adds phi_t Q_k[k - t] to the sum of every lane's visit k.  Q_k[a] is the
// sum of the unit phasors X of visit k's records whose younger point was flagged
// at visit k - a (a <= kPushAges) and whose other point is older than kPushLag
// visits or was flagged before the window: the record's term is then
// phi(younger) X, the reference's normalised product up to rounding, since both
// factors are unit phasors.  At most two records share an entry (the same two
// points as (u, v) and as (v, u)), so adding them into a zeroed entry with LDS
// atomics gives the same bits in either order.
// Helper waves fill the entries Q_k and A_k (every record of visit k whose two
// points are older than kPushLag visits) once the chain has finished visit
// k - kPushLag - 1, so that the phases they read are final; they have about
// kPushLag - kPushAges steps of the chain until it first reads an entry.
// The step is a few FP64 operations and no reduction, so the serial path is
// short.  A phase of zero (|mean| <= DBL_EPSILON, or NaN) makes the reference's
// later terms with that point NaN (0/0), which a linear term would not
// reproduce; the window is then redone by the general schedule.  Returns false
// in that case and when a wait gave up.
__device__ __forceinline__ bool reco_push_window(int nsteps, const PendingRec* __restrict__ rec, int cap,
    int32_t* __restrict__ status, FixLds& L, int& calls, int& flagged)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    StepInfo* info = L.c.info;
    double2* ppos = L.c.ppos;
    double* pcons = L.c.pcons; // |sum|^2 until the end of the window, then the consistency
    PushHead* head = L.u.p.head;
    double2* Q = L.u.p.Q;
    int32_t* ready = L.u.p.ready;
    int32_t* progress = &L.c.progress;
    int32_t* abort_flag = &L.c.abort_flag;
    for (int t = tid; t < kReadyWords; t += kPushThreads)
        ready[t] = t - 1 >= nsteps ? 1 : 0;
    if (tid < 2) { // the pad step of an odd window, and the one after it
        head[nsteps + tid].A = mk<double2>(0.0, 0.0);
        head[nsteps + tid].thr = 0.0;
        head[nsteps + tid].call = 0;
        head[nsteps + tid].n = 0;
    }
    if (tid == 0)
        L.u.p.ncalls = 0;
    __syncthreads();
    if (wave == 0) {
        __builtin_amdgcn_s_setprio(3);
        bool ok = true;
        // the entries and heads of visits 0 .. 64 are read at step 0
        for (int k = 0; k <= min(kPushAges, nsteps - 1) && ok; ++k)
            ok = wait_at_least(&ready[k + 1], 1, abort_flag, status, lane);
        if (ok) {
            double zr = 0.0, zi = 0.0;
            // What a step reads, loaded one step ahead: this lane's Q entry and the
            // visit's helper result.
            struct Ahead {
                double2 qv;
                double2 A;
                double thr;
                int32_t call;
            };
            bool bad = false; // a zero phase met (checked once, after the loop)
            int row1 = 0;     // (t + 1) mod kPushLag
            // One step: loads of step t + 1 (whose entries were checked ready),
            // then visit t (lane t mod 64) finalised and phi_t added to every lane.
            auto step = [&](int t, const Ahead& cur, Ahead& nxt, bool odd) {
                const int t1 = t + 1;
                row1 = row1 + 1 == kPushLag ? 0 : row1 + 1;
                nxt.qv = Q[row1 * kPushRow + lane];
                const PushHead* hn = &head[t1];
                nxt.A = hn->A;
                nxt.thr = hn->thr;
                nxt.call = hn->call;
                const bool call = __builtin_amdgcn_readfirstlane(cur.call) != 0;
                const int sl = t & 63;
                const double tr = readlane_d(zr, sl) + cur.A.x, ti = readlane_d(zi, sl) + cur.A.y;
                // a visit without a calc_phase has a zero sum: the 1e-300 keeps its
                // factor finite, so that it pushes exact zeros
                const double m2 = fma(ti, ti, fma(tr, tr, 1e-300));
                const double ia = rsqrt_nr1(m2);
                bad |= call && !(m2 > cur.thr);
                const double pr = tr * ia, pi = ti * ia;
                if (lane == 0) {
                    if (call) {
                        ppos[t] = mk<double2>(pr, pi);
                        pcons[t] = m2;
                    }
                    if (odd)
                        lds_put(progress, t1);
                }
                // lane t mod 64 now collects visit t + 64; its entry at this step is
                // Q_{t+64}[64].  A sum that is not finite only ever belongs to a
                // window that is redone.
                const double keep = lane == sl ? 0.0 : 1.0;
                zr = fma(pr, cur.qv.x, fma(-pi, cur.qv.y, zr * keep));
                zi = fma(pr, cur.qv.y, fma(pi, cur.qv.x, zi * keep));
            };
            Ahead A, B;
            A.qv = Q[lane];
            A.A = head[0].A;
            A.thr = head[0].thr;
            A.call = head[0].call;
            // two steps per iteration, so that the two prefetch sets swap roles
            // instead of being copied (an odd window ends with a pad step that
            // pushes only into visits beyond it); before steps t + 1 and t + 2 load
            // the entries of visits t + 65 and t + 66, one check that both are done
            int2 rdy = *reinterpret_cast<const int2*>(&ready[kPushAges + 2]);
            for (int t = 0; t < nsteps; t += 2) {
                if (__builtin_expect(!__builtin_amdgcn_readfirstlane(rdy.x & rdy.y), 0)) {
                    for (int k = t + kPushAges + 1; k <= t + kPushAges + 2; ++k)
                        if (!wait_at_least(&ready[k + 1], 1, abort_flag, status, lane))
                            ok = false;
                }
                rdy = *reinterpret_cast<const int2*>(&ready[t + kPushAges + 4]);
                step(t, A, B, false);
                step(t + 1, B, A, true);
            }
            if (lane == 0)
                lds_put(progress, nsteps);
            if (bad) {
                if (lane == 0)
                    L.c.degenerate = 1;
                ok = false;
            }
        }
        if (!ok)
            lds_put(abort_flag, 1);
    } else {
        // a helper: visits k = wave - 1 + i * kPushHelpers.  The first 128 records of
        // its next visit are loaded while it works on the current one.
        PendingRec pre[2];
        auto fetch = [&](int k) {
            const int cn = k < nsteps ? __builtin_amdgcn_readfirstlane(info[k].cnt) : 0;
            const PendingRec* rk = rec + (int64_t)k * cap;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                pre[c] = PendingRec {};
                if (64 * c + lane < cn)
                    pre[c] = rk[64 * c + lane];
            }
        };
        int ncalls = 0;
        fetch(wave - 1);
        for (int k = wave - 1; k < nsteps; k += kPushHelpers) {
            const int n0 = __builtin_amdgcn_readfirstlane(info[k].n0);
            const int fp = __builtin_amdgcn_readfirstlane(info[k].fp);
            PushHead* hk = &head[k];
            const int p = k - kPushLag; // the chain has finished visit p - 1: the entries' rows are free
            if (n0 < 0 || fp != k) { // no calc_phase at this visit: zero entries
                fetch(k + kPushHelpers);
                if (p > 0 && !wait_at_least(progress, p, abort_flag, status, lane))
                    break;
                if (lane < k) // entries of steps k - 64 .. k - 1 that exist
                    Q[push_slot(k, lane + 1)] = mk<double2>(0.0, 0.0);
                if (lane == 0) {
                    hk->A = mk<double2>(0.0, 0.0);
                    hk->thr = 0.0;
                    hk->call = 0;
                    hk->n = 0;
                    lds_put(&ready[k + 1], 1);
                }
                continue;
            }
            ++ncalls;
            const int cn = __builtin_amdgcn_readfirstlane(info[k].cnt);
            const PendingRec* rk = rec + (int64_t)k * cap;
            const PendingRec cur0 = pre[0], cur1 = pre[1];
            fetch(k + kPushHelpers);
            if (p > 0 && !wait_at_least(progress, p, abort_flag, status, lane))
                break;
            if (lane < k)
                Q[push_slot(k, lane + 1)] = mk<double2>(0.0, 0.0);
            wave_lds_order();
            double sr = 0.0, sim = 0.0;
            int ne = 0;
            for (int t0 = 0; t0 < cn; t0 += 64) {
                const int t = t0 + lane;
                PendingRec r;
                if (t0 == 0)
                    r = cur0;
                else if (t0 == 64)
                    r = cur1;
                else
                    r = t < cn ? rk[t] : PendingRec {};
                // an in-window point runs calc_phase at its first visit (info n0 >= 0) or never
                bool counts = false;
                int young = 0, old = 0;
                if (t < cn) {
                    counts = (r.f & kValid) && ((r.f & kPreU) || info[max(r.ju, 0)].n0 >= 0)
                        && ((r.f & kPreV) || info[max(r.jv, 0)].n0 >= 0);
                    push_ages(r, k, young, old);
                }
                const bool far = counts && young > kPushLag;
                const bool push = counts && young <= kPushAges;
                double2 q = r.c;
                if (far) {
                    if (!(r.f & kPreU))
                        q = cmul(ppos[r.ju], q);
                    if (!(r.f & kPreV))
                        q = cmul(q, ppos[r.jv]);
                } else if (push && old != kAgeNone) {
                    q = cmul(q, ppos[min(r.ju, r.jv)]); // the older in-window point
                }
                q = normalised(q);
                if (far) {
                    sr += q.x;
                    sim += q.y;
                }
                ne += far || push;
                if (push) { // X into Q_k[young]
                    double2* e = &Q[push_slot(k, young)];
                    __hip_atomic_fetch_add(&e->x, q.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_fetch_add(&e->y, q.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
            wave_sum3(ne, sr, sim);
            wave_lds_order(); // the entries' atomics and the head before the ready word
            if (lane == 0) {
                const double2 S0 = info[k].S;
                const int n = n0 + ne;
                hk->A = mk<double2>(S0.x + sr, S0.y + sim);
                hk->thr = (kEps * n) * (kEps * n);
                hk->call = 1;
                hk->n = n;
                lds_put(&ready[k + 1], 1);
            }
        }
        if (lane == 0)
            atomicAdd(&L.u.p.ncalls, ncalls);
    }
    __syncthreads();
    const bool ok = !lds_get(abort_flag);
    if (ok) { // consistency |sum| / n of each flagged slot
        for (int t = tid; t < nsteps; t += kPushThreads)
            if (ppos[t].x != kNotYet)
                pcons[t] = sqrt(pcons[t]) / (double)head[t].n;
        calls = flagged = L.u.p.ncalls;
    }
    return ok;
}

// The recursion over one window, by one workgroup: with PUSH (one recovery at a
// time), the push chain above where the window's schedule is static, else (and
// where the push chain met a zero phase) the general schedule below; without
// PUSH (grouped recoveries), the general schedule alone.
//
// General schedule.  Wave 0 (the chain) walks the visits in order.  For visit k
// it needs the sum over the visit's pending records whose points are flagged by
// now.  Helper waves take visits k = h, h + kHelpers, ... ahead of the chain:
// once the chain has finished visit p - 1 (p = k - kLead), a record both of whose
// in-window points were flagged before p is certain to count, and one with a
// point neither flagged before p nor visited in [p, k) is certain not to; the
// helper sums the first kind into far[k] and copies only the undecided rest
// ("near" records) to LDS.  The chain then adds the few near records whose
// points are flagged at visit k and finalises the point as calc_phase does.
// Which record lands where depends only on the visit list and on when each slot
// was flagged (ftime), so the summation order, and every bit of the result, is
// the same on every run.
template <bool PUSH>
__device__ __forceinline__ void reco_fixup_body(const RecoArgs& a, const int2* __restrict__ visits,
    const int32_t* __restrict__ first_pos, int64_t s, int64_t e, const double2* __restrict__ S,
    const int32_t* __restrict__ N, const int32_t* __restrict__ E, const int32_t* __restrict__ cnt,
    const PendingRec* __restrict__ rec, int cap, uint8_t* __restrict__ flags, double2* __restrict__ phases,
    double* __restrict__ cons, int32_t* __restrict__ status, int64_t win, int64_t* __restrict__ map)
{
    using Lds = typename std::conditional<PUSH, FixLds, FixLdsGeneral>::type;
    constexpr int kFixThreads = PUSH ? kPushThreads : kGenThreads;
    constexpr int kHelpers = kFixThreads / 64 - 1;
    __shared__ __align__(16) unsigned char fix_lds[sizeof(Lds)];
    Lds& L = *reinterpret_cast<Lds*>(fix_lds);
    StepInfo* info = L.c.info;
    double2* ppos = L.c.ppos;
    double* pcons = L.c.pcons;
    int32_t* ftime = L.u.g.ftime;
    double2* far_s = L.u.g.far_s;
    i32x4* head = L.u.g.head;
    PendingRec(*nearbuf)[kNearCap] = L.u.g.nearbuf;
    int32_t& progress = L.c.progress;
    int32_t& abort_flag = L.c.abort_flag;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nsteps = (int)(e - s);
    if (tid == 0) {
        progress = 0;
        abort_flag = status[0] != 0; // an earlier window raised
        L.c.degenerate = 0;
    }
    bool irregular = false;
    for (int t = tid; t < nsteps; t += kFixThreads) {
        StepInfo si;
        si.n0 = N[t];
        si.fp = first_pos[s + t];
        si.cnt = cnt[t];
        si.err = E[t];
        si.S = S[t];
        info[t] = si;
        irregular |= si.err != 0;
    }
    for (int t = tid; t < kWindow; t += kFixThreads)
        ppos[t] = mk<double2>(kNotYet, 0.0);
    irregular = __syncthreads_or(irregular) != 0;

    int calls = 0, flagged = 0;
    bool done = false;
    if constexpr (PUSH) {
        if (a.push && !irregular && !lds_get(&abort_flag)) {
            done = reco_push_window(nsteps, rec, cap, status, L, calls, flagged);
        if (!done && L.c.degenerate && !status[0]) { // redo the window by the general schedule
            for (int t = tid; t < kWindow; t += kFixThreads)
                ppos[t] = mk<double2>(kNotYet, 0.0);
            calls = flagged = 0;
            __syncthreads();
            if (tid == 0) {
                progress = 0;
                abort_flag = 0;
            }
        }
        }
    }
    if (!done) {
        for (int t = tid; t < kWindow; t += kFixThreads) {
            ftime[t] = kNever;
            head[t] = i32x4 { 0, 0, 0, 0 };
        }
        __syncthreads();
    }
    if (!done && !lds_get(&abort_flag)) {
        auto take = [&](const PendingRec& r, double2 pu, double2 pv, double& sr, double& sim, int& ne) {
            const bool uok = (r.f & kPreU) || (r.ju >= 0 && pu.x != kNotYet);
            const bool vok = (r.f & kPreV) || (r.jv >= 0 && pv.x != kNotYet);
            if (!(uok && vok))
                return;
            if (r.f & kErr) {
                ne += 1 << 16;
                return;
            }
            if (r.f & kValid) {
                double2 qv = r.c;
                if (!(r.f & kPreU))
                    qv = cmul(pu, qv);
                if (!(r.f & kPreV))
                    qv = cmul(qv, pv);
                qv = normalised(qv);
                sr += qv.x;
                sim += qv.y;
                ++ne;
            }
        };
        if (wave == 0) {
            // The chain, software-pipelined: while visit k is summed and
            // finalised, the loads of visit k + 1 (its step scalars, the helper's
            // results, its near records and the phases of their points as of
            // visit k - 1) are in flight; visit k's own result is patched in
            // from registers, so nothing on the step-to-step path waits for LDS.
            // A single wavefront pays the issue latency of every instruction on
            // this path, so the step is kept short: scalar loads for the step's
            // scalars, one 16-byte load for the helper's counts, selects rather
            // than branches.
            __builtin_amdgcn_s_setprio(3);
            struct Ahead {
                StepInfo si;              // step scalars (the same in every lane)
                i32x3 hd;                 // helper: far count, near count, ready
                double2 fs, pfp;          // helper sum; phase of the visit's own slot
                PendingRec r;             // this lane's near record (lane < near count)
                double2 pu, pv;           // phases of its points
            };
            // step scalars and helper results, then (in issue order behind the
            // ready flag) the helper's data
            auto load1 = [&](int k, Ahead& A) {
                A.si = info[k];
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                // one 12-byte load: the ready word with the counts (the fourth word
                // is not loaded, so no register of the load is reused before it lands)
                A.hd = *reinterpret_cast<const i32x3*>(&head[k]);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                A.fs = far_s[k];
                A.r = nearbuf[k % kRing][lane & (kNearCap - 1)];
            };
            // phases as written so far (the caller patches in later visits)
            auto load2 = [&](Ahead& A) {
                A.pfp = ppos[A.si.fp];
                A.pu = ppos[max(A.r.ju, 0)];
                A.pv = ppos[max(A.r.jv, 0)];
            };
            double2 ph_prev = mk<double2>(kNotYet, 0.0);
            int fp_prev = -1;
            bool stop = false;
            auto visit = [&](int k, Ahead& A, Ahead& B) {
                if (fp_prev >= 0) { // the previous visit flagged slot fp_prev
                    if (A.si.fp == fp_prev)
                        A.pfp = ph_prev;
                    if (A.r.ju == fp_prev)
                        A.pu = ph_prev;
                    if (A.r.jv == fp_prev)
                        A.pv = ph_prev;
                }
                const int n0 = __builtin_amdgcn_readfirstlane(A.si.n0), fp = __builtin_amdgcn_readfirstlane(A.si.fp);
                const bool run = n0 >= 0 && !(fp != k && __builtin_amdgcn_readfirstlane(A.pfp.x != kNotYet));
                if (__builtin_expect(run && !__builtin_amdgcn_readfirstlane(A.hd.z), 0)) {
                    // the helper was not done when the loads left: wait, then load
                    // again (every earlier visit's phase is in LDS by now)
                    if (!wait_at_least(ready_flag(head, k), 1, &abort_flag, status, lane)) {
                        stop = true;
                        return;
                    }
                    load1(k, A);
                    load2(A);
                }
                const bool more = k + 1 < nsteps;
                if (more)
                    load1(k + 1, B);
                fp_prev = -1;
                if (run) {
                    ++calls;
                    const int nn = __builtin_amdgcn_readfirstlane(A.hd.y);
                    double sr = 0.0, sim = 0.0;
                    int nval = 0, nerr = 0;
                    if (__builtin_expect(nn >= 0, 1)) {
                        // near records, one per lane, as take() does, with selects
                        const PendingRec& r = A.r;
                        const bool pre_u = r.f & kPreU, pre_v = r.f & kPreV;
                        const bool uok = pre_u || (r.ju >= 0 && A.pu.x != kNotYet);
                        const bool vok = pre_v || (r.jv >= 0 && A.pv.x != kNotYet);
                        const bool ok = lane < nn && uok && vok;
                        const bool is_err = ok && (r.f & kErr);
                        const bool is_val = ok && !(r.f & kErr) && (r.f & kValid);
                        const double2 qa = pre_u ? mk<double2>(1.0, 0.0) : A.pu;
                        const double2 qb = pre_v ? mk<double2>(1.0, 0.0) : A.pv;
                        const double2 q = cmul(cmul(qa, r.c), qb);
                        const double m2 = q.x * q.x + q.y * q.y;
                        double inv = 1.5 - 0.5 * m2; // normalised()'s one Newton step from 1
                        const bool slow = is_val && !(fabs(m2 - 1.0) < 1e-8);
                        if (__builtin_expect(__ballot(slow) != 0, 0)) {
                            if (slow)
                                inv = 1.0 / sqrt(m2);
                        }
                        sr = is_val ? q.x * inv : 0.0;
                        sim = is_val ? q.y * inv : 0.0;
                        nval = __popcll(__ballot(is_val));
                        nerr = __popcll(__ballot(is_err));
                        if (__builtin_expect(nn > 16, 0))
                            wave_sum2(sr, sim);
                        else
                            row_sum2(sr, sim);
                    } else {
                        // the helper left every record of this visit to the chain
                        int ne = 0;
                        const int cn = __builtin_amdgcn_readfirstlane(A.si.cnt);
                        for (int t = lane; t < cn; t += 64) {
                            const PendingRec r = rec[(int64_t)k * cap + t];
                            take(r, ppos[max(r.ju, 0)], ppos[max(r.jv, 0)], sr, sim, ne);
                        }
                        wave_sum3(ne, sr, sim);
                        nval = ne & 0xffff;
                        nerr = ne >> 16;
                    }
                    const int fn = __builtin_amdgcn_readfirstlane(A.hd.x);
                    if (__builtin_expect((fn >> 16) || nerr || (__builtin_amdgcn_readfirstlane(A.si.err) & kEOutOfRange), 0)) {
                        if (lane == 0)
                            status[0] = SMIP_ERR_OUT_OF_RANGE; // get_element throws std::out_of_range
                        lds_put(&abort_flag, 1);
                        stop = true;
                        return;
                    }
                    if (more)
                        load2(B); // phases as of visit k - 1
                    const int tn = n0 + (fn & 0xffff) + nval;
                    if (tn > 0) {
                        double ap;
                        const double2 ph = point_phase((A.si.S.x + A.fs.x) + sr, (A.si.S.y + A.fs.y) + sim, tn, &ap);
                        if (lane == 0) {
                            ppos[fp] = ph;
                            pcons[fp] = ap;
                            ftime[fp] = k;
                        }
                        ph_prev = ph;
                        fp_prev = fp;
                        ++flagged;
                    }
                } else if (more) {
                    load2(B);
                }
                if (lane == 0)
                    lds_put(&progress, k + 1);
            };
            // two visits per iteration, so the two prefetch sets swap roles
            // instead of being copied
            Ahead A, B;
            if (nsteps > 0) {
                load1(0, A);
                load2(A);
            }
            for (int k = 0; k < nsteps && !stop; k += 2) {
                visit(k, A, B);
                if (!stop && k + 1 < nsteps)
                    visit(k + 1, B, A);
            }
        } else {
            // a helper: visits k = wave - 1 + i * kHelpers
            for (int k = wave - 1; k < nsteps; k += kHelpers) {
                const int n0 = __builtin_amdgcn_readfirstlane(info[k].n0);
                if (n0 < 0) { // the chain does not look at it
                    if (lane == 0)
                        lds_put(ready_flag(head, k), 1);
                    continue;
                }
                const int cn = __builtin_amdgcn_readfirstlane(info[k].cnt);
                const PendingRec* rk = rec + (int64_t)k * cap;
                // the first 128 records travel while the helper waits
                PendingRec pre[2];
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    pre[c] = PendingRec {};
                    if (64 * c + lane < cn)
                        pre[c] = rk[64 * c + lane];
                }
                const int p = k - kLead, q = max(p, 0);
                if (p > 0 && !wait_at_least(&progress, p, &abort_flag, status, lane))
                    break;
                // the ring slot's previous visit is published (the chain may have
                // skipped that visit without waiting for it)
                if (k >= kRing && !wait_at_least(ready_flag(head, k - kRing), 1, &abort_flag, status, lane))
                    break;
                int32_t fpw[kLead];
#pragma unroll
                for (int i = 0; i < kLead; ++i)
                    fpw[i] = q + i < k ? __builtin_amdgcn_readfirstlane(info[q + i].fp) : -2;
                double sr = 0.0, sim = 0.0;
                int ne = 0, nnear = 0;
                PendingRec* nb = nearbuf[k % kRing];
                for (int t0 = 0; t0 < cn; t0 += 64) {
                    const int t = t0 + lane;
                    PendingRec r;
                    if (t0 < 128)
                        r = pre[t0 >> 6];
                    else
                        r = t < cn ? rk[t] : PendingRec {};
                    int cls = 2;
                    if (t < cn)
                        cls = max(side_state(r.f & kPreU, r.ju, p, q, ftime, fpw),
                            side_state(r.f & kPreV, r.jv, p, q, ftime, fpw));
                    if (cls == 0)
                        take(r, ppos[max(r.ju, 0)], ppos[max(r.jv, 0)], sr, sim, ne);
                    const uint64_t m = __ballot(cls == 1);
                    const int idx = nnear + __popcll(m & ((1ull << lane) - 1));
                    if (cls == 1 && idx < kNearCap)
                        nb[idx] = r;
                    nnear += __popcll(m);
                }
                wave_sum3(ne, sr, sim);
                if (nnear > kNearCap) { // the chain takes every record of this visit
                    sr = sim = 0.0;
                    ne = 0;
                    nnear = -1;
                }
                if (lane == 0) {
                    far_s[k] = mk<double2>(sr, sim);
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    head[k] = i32x4 { ne, nnear, 1, 0 }; // one 16-byte store, behind far_s
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                }
            }
        }
    }
    __syncthreads();
    const bool aborted = lds_get(&abort_flag) != 0;
    // the window's results, once
    for (int t = tid; t < nsteps; t += kFixThreads) {
        const double2 ph = ppos[t];
        if (ph.x != kNotYet) {
            const int2 w = visits[s + t];
            const int32_t pw = pm_at(a, w.x, w.y);
            flags[pw] = 1;
            cons[pw] = pcons[t];
            phases[pw] = ph;
        }
    }
    // the next window's point map, in place of a launch of its own before the next
    // bulk pass: nothing in this window reads the map any more
    if (!aborted) {
        const int64_t s1 = e, e1 = min(e + (int64_t)kWindow, a.nvisits);
        for (int64_t k = s1 + tid; k < e1; k += kFixThreads)
            if (first_pos[k] == (int32_t)(k - s1)) {
                const int2 w = visits[k];
                map[pm_at(a, w.x, w.y)] = ((win + 1) << 32) | (k - s1);
            }
    }
    if (tid == 0) {
        atomicAdd(&status[1], calls);
        atomicAdd(&status[2], flagged);
    }
}

__global__ void k_reco_map(const int2* __restrict__ visits, const int32_t* __restrict__ first_pos, RecoArgs a,
    int64_t s, int64_t e, int64_t win, int64_t* __restrict__ map)
{
    reco_map_body(blockIdx.x, visits, first_pos, a, s, e, win, map);
}

__global__ __launch_bounds__(kBulkThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_reco_bulk(RecoArgs a, const int2* __restrict__ visits,
    const int32_t* __restrict__ first_pos, int64_t s, int64_t win, const int64_t* __restrict__ map,
    const uint8_t* __restrict__ flags, const double2* __restrict__ phases, const void* __restrict__ B, int is_f32,
    double2* __restrict__ S, int32_t* __restrict__ N, int32_t* __restrict__ E, int32_t* __restrict__ cnt,
    PendingRec* __restrict__ rec, int cap, const int32_t* __restrict__ status)
{
    reco_bulk_body<true>(blockIdx.x, a, visits, first_pos, s, win, map, flags, phases, B, is_f32, S, N, E, cnt, rec,
        cap, status);
}

__global__ __launch_bounds__(kPushThreads) void k_reco_fixup(RecoArgs a, const int2* __restrict__ visits,
    const int32_t* __restrict__ first_pos, int64_t s, int64_t e, const double2* __restrict__ S,
    const int32_t* __restrict__ N, const int32_t* __restrict__ E, const int32_t* __restrict__ cnt,
    const PendingRec* __restrict__ rec, int cap, uint8_t* __restrict__ flags, double2* __restrict__ phases,
    double* __restrict__ cons, int32_t* __restrict__ status, int64_t win, int64_t* __restrict__ map)
{
    reco_fixup_body<true>(a, visits, first_pos, s, e, S, N, E, cnt, rec, cap, flags, phases, cons, status, win, map);
}

// ---------------------------------------------------------------------------
// many recoveries at once: the same bodies over a table of jobs, one launch per
// stage and window index for all of them (jobs without that window return)
// ---------------------------------------------------------------------------
struct RecoJob {
    RecoArgs a;
    const int2* visits;
    const int32_t* first_pos;
    int64_t* map;
    double2* S;
    int32_t *N, *E, *cnt;
    PendingRec* rec;
    int32_t* status;
    const void* B;
    uint8_t* flags;
    double2* phases;
    double* cons;
    int32_t cap, is_f32;
};

__global__ void k_reco_map_many(const RecoJob* __restrict__ jobs, int64_t win)
{
    const RecoJob& J = jobs[blockIdx.y];
    const RecoArgs a = J.a;
    const int64_t s0 = win * kWindow;
    if (s0 >= a.nvisits)
        return;
    reco_map_body(blockIdx.x, global_ptr(J.visits), global_ptr(J.first_pos), a, s0, min(s0 + kWindow, a.nvisits), win,
        global_ptr(J.map));
}

template <bool PUSH>
__global__ __launch_bounds__(kBulkThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_reco_bulk_many(
    const RecoJob* __restrict__ jobs, int64_t win)
{
    const RecoJob& J = jobs[blockIdx.y];
    const RecoArgs a = J.a;
    const int64_t s0 = win * kWindow;
    if (s0 + (int64_t)blockIdx.x >= a.nvisits)
        return;
    reco_bulk_body<PUSH>(blockIdx.x, a, global_ptr(J.visits), global_ptr(J.first_pos), s0, win, global_ptr(J.map),
        global_ptr(J.flags), global_ptr(J.phases), global_ptr(J.B), J.is_f32, global_ptr(J.S), global_ptr(J.N),
        global_ptr(J.E), global_ptr(J.cnt), global_ptr(J.rec), J.cap, global_ptr(J.status));
}

template <bool PUSH>
__global__ __launch_bounds__(PUSH ? kPushThreads : kGenThreads) void k_reco_fixup_many(
    const RecoJob* __restrict__ jobs, int64_t win)
{
    const RecoJob& J = jobs[blockIdx.x];
    const RecoArgs a = J.a;
    const int64_t s0 = win * kWindow;
    if (s0 >= a.nvisits)
        return;
    reco_fixup_body<PUSH>(a, global_ptr(J.visits), global_ptr(J.first_pos), s0, min(s0 + kWindow, a.nvisits),
        global_ptr(J.S), global_ptr(J.N), global_ptr(J.E), global_ptr(J.cnt), global_ptr(J.rec), J.cap,
        global_ptr(J.flags), global_ptr(J.phases), global_ptr(J.cons), global_ptr(J.status), win, global_ptr(J.map));
}

// ---------------------------------------------------------------------------
// one visit at a time (cross-check form)
// ---------------------------------------------------------------------------
constexpr int kSeqThreads = 1024;

__global__ __launch_bounds__(kSeqThreads) void k_reco_sequential(RecoArgs a, const int2* __restrict__ visits,
    const void* __restrict__ B, int is_f32, double2* __restrict__ phases, uint8_t* __restrict__ flags,
    double* __restrict__ cons, int32_t* __restrict__ status, int force)
{
    extern __shared__ uint32_t bits[];
    __shared__ double red_re[kSeqThreads / 64], red_im[kSeqThreads / 64];
    __shared__ int red_n[kSeqThreads / 64];
    __shared__ int err;
    const int tid = threadIdx.x;
    const int npts = a.X * a.Y;
    const int nwords = (npts + 31) >> 5;
    for (int wd = tid; wd < nwords; wd += blockDim.x) {
        uint32_t m = 0;
        for (int b = 0; b < 32; ++b) {
            const int p = (wd << 5) + b;
            if (p < npts && flags[p])
                m |= 1u << b;
        }
        bits[wd] = m;
    }
    if (tid == 0)
        err = 0;
    __syncthreads();
    const Desc& d = a.d;
    const int vmn0 = d.min_idx[2], vmn1 = d.min_idx[3];
    const int nvx = d.max_idx[2] - vmn0 + 1, nvy = d.max_idx[3] - vmn1 + 1;
    const int ncand = nvx * nvy;
    int calls = 0, flagged = 0;
    for (int64_t vi = 0; vi < a.nvisits; ++vi) {
        const int2 w = visits[vi];
        const int wpos = pm_at(a, w.x, w.y);
        if (!force && (bits[wpos >> 5] & (1u << (wpos & 31))))
            continue;
        if (calc_phase_returns_early(a, w))
            continue;
        ++calls;
        double sr = 0.0, si = 0.0;
        int cnt = 0;
        for (int c = tid; c < ncand; c += blockDim.x) {
            const int vy = vmn1 + c / nvx, vx = vmn0 + c % nvx;
            const int ux = w.x - vx, uy = w.y - vy;
            if (ux < a.pmn0 || ux > a.pmx0 || uy < a.pmn1 || uy > a.pmx1)
                continue;
            const int pu = pm_at(a, ux, uy), pv = pm_at(a, vx, vy);
            if (!(bits[pu >> 5] & (1u << (pu & 31))) || !(bits[pv >> 5] & (1u << (pv & 31))))
                continue;
            uint64_t addr;
            bool cj;
            if (resolve_element(d, ux, uy, vx, vy, &addr, &cj)) {
                err = SMIP_ERR_OUT_OF_RANGE;
                continue;
            }
            const double2 beta = load_beta(B, is_f32, addr, cj);
            const double ab = hypot(beta.x, beta.y);
            if (ab > kEps) {
                const double2 q = unit_phasor(phases[pu], phases[pv], mk<double2>(beta.x / ab, -(beta.y / ab)));
                sr += q.x;
                si += q.y;
                ++cnt;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            sr += __shfl_xor(sr, o);
            si += __shfl_xor(si, o);
            cnt += __shfl_xor(cnt, o);
        }
        if ((tid & 63) == 0) {
            red_re[tid >> 6] = sr;
            red_im[tid >> 6] = si;
            red_n[tid >> 6] = cnt;
        }
        __syncthreads();
        if (tid == 0) {
            double tr = 0.0, ti = 0.0;
            int tn = 0;
            for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
                tr += red_re[k];
                ti += red_im[k];
                tn += red_n[k];
            }
            if (tn > 0 && err == 0) {
                bits[wpos >> 5] |= 1u << (wpos & 31);
                finalize_point(tr, ti, tn, phases, flags, cons, wpos, nullptr);
                ++flagged;
            }
        }
        __syncthreads();
        if (err)
            break;
    }
    if (tid == 0) {
        status[0] = err;
        status[1] = calls;
        status[2] = flagged;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static size_t up256(size_t n) { return (n + 255) & ~(size_t)255; }

static RecoArgs reco_args(const Desc& d, int32_t xsize, int32_t ysize, int64_t nvisits)
{
    RecoArgs a;
    a.d = d;
    a.X = xsize;
    a.Y = ysize;
    a.pmn0 = -(xsize >> 1);
    a.pmn1 = -(ysize >> 1);
    a.pmx0 = -xsize / 2 + xsize - 1;
    a.pmx1 = -ysize / 2 + ysize - 1;
    a.nvisits = nvisits;
    // SMIP_RECO_PUSH=0 keeps every window on the general schedule (a cross-check)
    static const bool push = [] {
        const char* v = std::getenv("SMIP_RECO_PUSH");
        return !(v && v[0] == '0');
    }();
    a.push = push ? 1 : 0;
    a.pad = 0;
    return a;
}

static int reco_cap(const Desc& d)
{
    const int ncand = (d.max_idx[2] - d.min_idx[2] + 1) * (d.max_idx[3] - d.min_idx[3] + 1);
    if (ncand > 64 * kBulkThreads)
        fail(SMIP_ERR_INVALID_ARGUMENT, "bispectrum depth too large for phase recovery ((b+1)^2 <= 65536)");
    return std::min(ncand, 2 * kWindow);
}

// Byte offsets of the scratch arrays inside one block.
struct ScratchLayout {
    size_t map, S, N, E, cnt, rec, status, total;
};

static ScratchLayout scratch_layout(int32_t xsize, int32_t ysize, int cap)
{
    ScratchLayout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += up256(bytes);
        return at;
    };
    L.map = take(sizeof(int64_t) * (size_t)xsize * ysize);
    L.S = take(sizeof(double2) * kWindow);
    L.N = take(sizeof(int32_t) * kWindow);
    L.E = take(sizeof(int32_t) * kWindow);
    L.cnt = take(sizeof(int32_t) * kWindow);
    L.rec = take(sizeof(PendingRec) * (size_t)kWindow * cap);
    L.status = take(sizeof(int32_t) * 4);
    L.total = o;
    return L;
}

uint64_t reco_scratch_bytes(const Desc& d, int32_t xsize, int32_t ysize)
{
    return scratch_layout(xsize, ysize, reco_cap(d)).total;
}

// Window slot of the first visit of each visit's point within its window.
static std::vector<int32_t> first_positions(int32_t xsize, int32_t ysize, const int32_t* visits, int64_t nvisits)
{
    std::vector<int32_t> first(std::max<int64_t>(nvisits, 1));
    std::vector<int32_t> seen((size_t)xsize * ysize, -1);
    for (int64_t s = 0; s < nvisits; s += kWindow) {
        const int64_t e = std::min<int64_t>(s + kWindow, nvisits);
        for (int64_t k = s; k < e; ++k) {
            int32_t x = visits[2 * k], y = visits[2 * k + 1];
            const size_t p = (size_t)((y < 0 ? y + ysize : y) * xsize + (x < 0 ? x + xsize : x));
            if (seen[p] < s) // not yet visited in this window
                seen[p] = (int32_t)k;
            first[k] = seen[p] - (int32_t)s;
        }
    }
    return first;
}

static std::shared_ptr<const RecoGeom> geom_upload(int32_t xsize, int32_t ysize, const int32_t* visits,
    int64_t nvisits, hipStream_t s)
{
    const std::vector<int32_t> first = first_positions(xsize, ysize, visits, nvisits);
    const size_t n = (size_t)std::max<int64_t>(nvisits, 1);
    const size_t vb = up256(sizeof(int2) * n);
    auto g = std::make_shared<RecoGeom>();
    g->X = xsize;
    g->Y = ysize;
    g->nvisits = nvisits;
    g->bytes = vb + sizeof(int32_t) * n;
    g->block = dev_malloc(g->bytes);
    g->visits = static_cast<const int2*>(g->block);
    g->first_pos = reinterpret_cast<const int32_t*>(static_cast<const char*>(g->block) + vb);
    if (nvisits) { // one upload of both arrays, laid out as on the device
        std::vector<char> h(g->bytes, 0);
        std::memcpy(h.data(), visits, sizeof(int2) * nvisits);
        std::memcpy(h.data() + vb, first.data(), sizeof(int32_t) * nvisits);
        copy_sync(g->block, h.data(), g->bytes, hipMemcpyHostToDevice, s);
    }
    return g;
}

// Visit list of reconstruct_phases(radius) on an X x Y map, on the current device.
// Kept per (device, radius, X, Y) for the life of the process, up to 64
// geometries and 1 GiB; a geometry dropped from the cache lives on in the
// recoveries that hold it.
std::shared_ptr<const RecoGeom> reco_geom(double reco_radius, int32_t xsize, int32_t ysize, hipStream_t s)
{
    int dev = 0;
    SMIP_HIP(hipGetDevice(&dev));
    using Key = std::tuple<int, double, int32_t, int32_t>;
    static std::mutex mu;
    static auto* cache = new std::map<Key, std::shared_ptr<const RecoGeom>>; // never destroyed: no frees at exit
    static uint64_t cached_bytes = 0;
    const Key key { dev, reco_radius, xsize, ysize };
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache->find(key);
    if (it != cache->end())
        return it->second;
    const auto v = visit_list(reco_radius, xsize, ysize);
    auto g = geom_upload(xsize, ysize, v->data(), (int64_t)(v->size() / 2), s);
    if (cache->size() >= 64 || cached_bytes + g->bytes > (1ull << 30)) {
        cache->clear();
        cached_bytes = 0;
    }
    (*cache)[key] = g;
    cached_bytes += g->bytes;
    return g;
}

// Upper estimate of a geometry's device bytes: in-range visits number about
// 1.22 pi r^2, clipped to 1.23 X Y where the disc leaves the map.
uint64_t reco_geom_bytes_estimate(double reco_radius, int32_t xsize, int32_t ysize)
{
    const double r = std::max(reco_radius, 0.0) + 2.0;
    const double nv = std::min(1.25 * M_PI * r * r, 1.25 * (double)xsize * ysize) + 1024.0;
    return (uint64_t)(nv * (sizeof(int2) + sizeof(int32_t))) + 256;
}

RecoWork* reco_work_create(const Desc& d, std::shared_ptr<const RecoGeom> g)
{
    const int cap = reco_cap(d);
    auto* w = new RecoWork;
    w->args = reco_args(d, g->X, g->Y, g->nvisits);
    w->cap = cap;
    w->nwin = (g->nvisits + kWindow - 1) / kWindow;
    w->visits = g->visits;
    w->first_pos = g->first_pos;
    w->geom = std::move(g);
    const ScratchLayout L = scratch_layout(w->args.X, w->args.Y, cap);
    try {
        w->block = dev_malloc(L.total);
    } catch (...) {
        delete w;
        throw;
    }
    char* b = static_cast<char*>(w->block);
    w->map = reinterpret_cast<int64_t*>(b + L.map);
    w->S = reinterpret_cast<double2*>(b + L.S);
    w->N = reinterpret_cast<int32_t*>(b + L.N);
    w->E = reinterpret_cast<int32_t*>(b + L.E);
    w->cnt = reinterpret_cast<int32_t*>(b + L.cnt);
    w->rec = reinterpret_cast<PendingRec*>(b + L.rec);
    w->status = reinterpret_cast<int32_t*>(b + L.status);
    return w;
}

void reco_work_destroy(RecoWork* w)
{
    if (!w)
        return;
    dev_free(w->block);
    delete w;
}

// Seeds of reconstruct_phases (include/phasereco.h:39-54) on a zeroed map.
__global__ void k_reco_init(RecoArgs a, double2* phases, uint8_t* flags, double* cons, int64_t* map)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.X * a.Y)
        return;
    const int32_t col = i % a.X, row = i / a.X;
    const int32_t x = col > a.pmx0 ? col - a.X : col, y = row > a.pmx1 ? row - a.Y : row;
    const bool seed = (x == 0 && y == 0) || (abs(x) + abs(y) == 1 && (x == 0 || y == 0));
    // init_phase = (1, 0) at (0,0), (1,0), (0,1); std::conj(init_phase) = (1, -0) at (-1,0), (0,-1)
    phases[i] = seed ? mk<double2>(1.0, (x < 0 || y < 0) ? -0.0 : 0.0) : mk<double2>(0.0, 0.0);
    flags[i] = seed ? 1 : 0;
    cons[i] = seed ? 1.0 : 0.0;
    if (map)
        map[i] = -1;
}

__global__ void k_zero_status(int32_t* st)
{
    if (threadIdx.x < 4)
        st[threadIdx.x] = 0;
}

int run_reconstruct(RecoWork* w, const void* B, int is_f32, double2* phases, uint8_t* flags, double* consistency,
    hipStream_t s)
{
    const RecoArgs& a = w->args;
    const int npts = a.X * a.Y;
    hipLaunchKernelGGL(k_reco_init, dim3((npts + 255) / 256), dim3(256), 0, s, a, phases, flags, consistency,
        w->map);
    hipLaunchKernelGGL(k_zero_status, dim3(1), dim3(64), 0, s, w->status);
    SMIP_HIP(hipGetLastError());
    const char* seq = std::getenv("SMIP_RECO_SEQUENTIAL");
    if (seq && seq[0] == '1') {
        const size_t lds = sizeof(uint32_t) * ((npts + 31) / 32);
        if (lds > 96 * 1024)
            fail(SMIP_ERR_INVALID_ARGUMENT, "phase map too large for the sequential form");
        hipLaunchKernelGGL(k_reco_sequential, dim3(1), dim3(kSeqThreads), lds, s, a, w->visits, B, is_f32, phases,
            flags, consistency, w->status, 0);
        SMIP_HIP(hipGetLastError());
    } else {
        for (int64_t win = 0; win < w->nwin; ++win) {
            const int64_t s0 = win * kWindow, e0 = std::min<int64_t>(s0 + kWindow, a.nvisits);
            const int nk = (int)(e0 - s0);
            if (win == 0) // later windows' maps are written by the fixup before them
                hipLaunchKernelGGL(k_reco_map, dim3((nk + 255) / 256), dim3(256), 0, s, w->visits, w->first_pos, a,
                    s0, e0, win, w->map);
            hipLaunchKernelGGL(k_reco_bulk, dim3(nk), dim3(kBulkThreads), 0, s, a, w->visits, w->first_pos, s0, win,
                w->map, flags, phases, B, is_f32, w->S, w->N, w->E, w->cnt, w->rec, w->cap, w->status);
            hipLaunchKernelGGL(k_reco_fixup, dim3(1), dim3(kPushThreads), 0, s, a, w->visits, w->first_pos, s0, e0, w->S, w->N,
                w->E, w->cnt, w->rec, w->cap, flags, phases, consistency, w->status, win, w->map);
        }
        SMIP_HIP(hipGetLastError());
    }
    int32_t st[4] = { 0, 0, 0, 0 };
    SMIP_HIP(hipMemcpyAsync(st, w->status, sizeof(int32_t) * 3, hipMemcpyDeviceToHost, s));
    SMIP_HIP(hipStreamSynchronize(s));
    if (st[0])
        fail(st[0], reco_error_text(st[0]));
    return st[1];
}

const char* reco_error_text(int32_t code)
{
    if (code == SMIP_ERR_OUT_OF_RANGE)
        return "reconstruct_phases: bispectrum element address out of bounds";
    return "reconstruct_phases: the fixup's waves stopped waiting for each other (internal error)";
}

// Phase recovery of n bispectra at once: per window index one map, one bulk and
// one fixup launch cover every job that has that window, so the single-wave
// fixups of different jobs run side by side.  With push, each job's arithmetic
// and order are those of run_reconstruct; without, every window takes the
// general schedule.  calls[k] receives job k's calc_phase count and status[k] its
// status (SMIP_ERR_OUT_OF_RANGE where get_element would throw).
void run_reconstruct_many(RecoWork* const* w, int n, const void* const* B, const int* is_f32,
    double2* const* phases, uint8_t* const* flags, double* const* consistency, int32_t* calls, int32_t* status,
    hipStream_t s, bool push)
{
    if (n <= 0)
        return;
    // The table holds the jobs by window count, most first, so the jobs that still
    // have window `win` are a prefix of it: each window's launches cover those
    // jobs only, instead of every job with workgroups that return at once (each
    // empty fixup workgroup still waits for a CU with its full LDS free).
    std::vector<int> order(n);
    for (int k = 0; k < n; ++k)
        order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return w[x]->nwin > w[y]->nwin; });
    std::vector<RecoJob> tab(n);
    int64_t maxwin = 0;
    for (int i = 0; i < n; ++i) {
        const int k = order[i];
        const RecoArgs& a = w[k]->args;
        const int npts = a.X * a.Y;
        hipLaunchKernelGGL(k_reco_init, dim3((npts + 255) / 256), dim3(256), 0, s, a, phases[k], flags[k],
            consistency[k], w[k]->map);
        hipLaunchKernelGGL(k_zero_status, dim3(1), dim3(64), 0, s, w[k]->status);
        RecoJob& J = tab[i];
        J.a = a;
        J.visits = w[k]->visits;
        J.first_pos = w[k]->first_pos;
        J.map = w[k]->map;
        J.S = w[k]->S;
        J.N = w[k]->N;
        J.E = w[k]->E;
        J.cnt = w[k]->cnt;
        J.rec = w[k]->rec;
        J.status = w[k]->status;
        J.B = B[k];
        J.flags = flags[k];
        J.phases = phases[k];
        J.cons = consistency[k];
        J.cap = w[k]->cap;
        J.is_f32 = is_f32[k];
        maxwin = std::max(maxwin, w[k]->nwin);
    }
    SMIP_HIP(hipGetLastError());
    RecoJob* dtab = static_cast<RecoJob*>(dev_malloc(sizeof(RecoJob) * n));
    try {
        copy_sync(dtab, tab.data(), sizeof(RecoJob) * n, hipMemcpyHostToDevice, s);
        int na = n; // jobs with window `win`: a prefix of the table
        for (int64_t win = 0; win < maxwin; ++win) {
            while (na > 0 && w[order[na - 1]]->nwin <= win)
                --na;
            if (win == 0) // later windows' maps are written by the fixups before them
                hipLaunchKernelGGL(k_reco_map_many, dim3(1, n), dim3(kWindow), 0, s, dtab, win);
            if (push) {
                hipLaunchKernelGGL(k_reco_bulk_many<true>, dim3(kWindow, na), dim3(kBulkThreads), 0, s, dtab, win);
                hipLaunchKernelGGL(k_reco_fixup_many<true>, dim3(na), dim3(kPushThreads), 0, s, dtab, win);
            } else {
                hipLaunchKernelGGL(k_reco_bulk_many<false>, dim3(kWindow, na), dim3(kBulkThreads), 0, s, dtab, win);
                hipLaunchKernelGGL(k_reco_fixup_many<false>, dim3(na), dim3(kGenThreads), 0, s, dtab, win);
            }
        }
        SMIP_HIP(hipGetLastError());
        std::vector<int32_t> st(4 * (size_t)n);
        for (int k = 0; k < n; ++k)
            SMIP_HIP(hipMemcpyAsync(&st[4 * (size_t)k], w[k]->status, sizeof(int32_t) * 3, hipMemcpyDeviceToHost, s));
        SMIP_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < n; ++k) {
            status[k] = st[4 * (size_t)k]; // per job: an out-of-range lookup stops that job only
            calls[k] = st[4 * (size_t)k + 1];
        }
    } catch (...) {
        dev_free(dtab);
        throw;
    }
    IdleRelease idle; // s is synchronised
    dev_free(dtab);
}

// One calc_phase(w) on caller-provided state (include/phasereco.h:75-123): the
// point is evaluated whatever its flag, as calc_phase itself does not look at it.
void run_calc_phase(const Desc& d, int32_t xsize, int32_t ysize, const void* B, int is_f32, double2* phases,
    uint8_t* flags, double* consistency, int32_t wx, int32_t wy, hipStream_t s)
{
    const int npts = xsize * ysize;
    const size_t lds = sizeof(uint32_t) * ((npts + 31) / 32);
    if (lds > 96 * 1024)
        fail(SMIP_ERR_INVALID_ARGUMENT, "phase map too large for calc_phase");
    (void)reco_cap(d);
    // one block: the visit, then the status words
    const int32_t head[4] = { wx, wy, 0, 0 };
    void* blk = dev_malloc(512);
    try {
        int2* visit = static_cast<int2*>(blk);
        int32_t* status = reinterpret_cast<int32_t*>(static_cast<char*>(blk) + 256);
        SMIP_HIP(hipMemcpyAsync(visit, head, sizeof(int2), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_zero_status, dim3(1), dim3(64), 0, s, status);
        hipLaunchKernelGGL(k_reco_sequential, dim3(1), dim3(kSeqThreads), lds, s, reco_args(d, xsize, ysize, 1),
            visit, B, is_f32, phases, flags, consistency, status, 1);
        SMIP_HIP(hipGetLastError());
        int32_t st[4] = { 0, 0, 0, 0 };
        SMIP_HIP(hipMemcpyAsync(st, status, sizeof(int32_t) * 3, hipMemcpyDeviceToHost, s));
        SMIP_HIP(hipStreamSynchronize(s));
        if (st[0])
            fail(st[0], "calc_phase: bispectrum element address out of bounds");
    } catch (...) {
        dev_free(blk);
        throw;
    }
    dev_free(blk);
}

} // namespace smip_amd
