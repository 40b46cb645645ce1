// reco.hip -- closure-phase recovery on the device.
//
// Replaces reconstruct_phases / calc_phase (include/phasereco.h:33-123).  The
// spiral visit order depends on libm cos/sin and is generated on the host by a
// literal restatement of NextRecoIndex (src/phasereco.cpp:17-36); only visits
// inside the PhaseMap range are uploaded (the others do nothing).
//
// The recursion is sequential: each calc_phase reads the phases of points set by
// earlier calls.  One workgroup walks the visit list; for every unflagged point w
// its threads evaluate the candidate pairs (u, v = w - u), v in the stored v range,
// in parallel, reduce the unit phasors in FP64 with a fixed tree, and thread 0
// sets flag, consistency and phase exactly as calc_phase does.  Eligibility
// (both flags set, |beta| > DBL_EPSILON) is evaluated exactly, so the flag mask
// is bit-exact; only the summation order of the FP64 mean differs.
#include <cmath>
#include <vector>

#include "smip_complex.h"
#include "smip_internal.h"

namespace smip_amd {

constexpr double kEps = 2.220446049250313e-16; // constants::c_epsilon<double>

struct RecoArgs {
    Desc d;
    int32_t X, Y;
    int32_t pmn0, pmn1, pmx0, pmx1; // PhaseMap signed range (Array2::range)
    int64_t nvisits;
};

struct RecoWork {
    RecoArgs args;
    int2* visits = nullptr;
    int32_t* status = nullptr;  // device: [0] error code, [1] calls, [2] flagged
};

__device__ inline int32_t pm_at(const RecoArgs& a, int32_t x, int32_t y)
{
    const int32_t xx = x < 0 ? x + a.X : x;
    const int32_t yy = y < 0 ? y + a.Y : y;
    return yy * a.X + xx;
}

constexpr int kRecoThreads = 1024;

__global__ __launch_bounds__(kRecoThreads) void k_reco_simple(RecoArgs a, const int2* __restrict__ visits,
    const void* __restrict__ B, int is_f32, double2* __restrict__ phases, uint8_t* __restrict__ flags,
    double* __restrict__ cons, int32_t* __restrict__ status)
{
    extern __shared__ uint32_t bits[]; // PhaseMap flags, one bit per point
    __shared__ double red_re[kRecoThreads / 64], red_im[kRecoThreads / 64];
    __shared__ int red_n[kRecoThreads / 64];
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
    const int vmn0 = d.min_idx[2], vmx0 = d.max_idx[2], vmn1 = d.min_idx[3], vmx1 = d.max_idx[3];
    const int nvx = vmx0 - vmn0 + 1, nvy = vmx1 - vmn1 + 1;
    const int ncand = nvx * nvy;
    int calls = 0, flagged = 0;

    for (int64_t vi = 0; vi < a.nvisits; ++vi) {
        const int2 w = visits[vi];
        const int wpos = pm_at(a, w.x, w.y);
        if (bits[wpos >> 5] & (1u << (wpos & 31)))
            continue;
        // calc_phase's early return, include/phasereco.h:84-86
        if (abs(w.x) + abs(w.y) <= 1 || w.x < d.min_idx[0] || w.x > d.max_idx[0] || w.y < d.min_idx[1]
            || w.y > d.max_idx[1])
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
                err = SMIP_ERR_OUT_OF_RANGE; // get_element throws std::out_of_range
                continue;
            }
            double2 beta;
            if (is_f32) {
                const float2 f = static_cast<const float2*>(B)[addr];
                beta = mk<double2>(f.x, f.y);
            } else {
                beta = static_cast<const double2*>(B)[addr];
            }
            if (cj)
                beta.y = -beta.y;
            // ph = phases(u) * phases(v); if |beta| > eps: ph *= conj(beta/|beta|); push ph/|ph|
            const double2 ph = cmul(phases[pu], phases[pv]);
            const double ab = hypot(beta.x, beta.y);
            if (ab > kEps) {
                const double2 t = mk<double2>(beta.x / ab, -(beta.y / ab));
                const double2 q = cmul(ph, t);
                const double aq = hypot(q.x, q.y);
                sr += q.x / aq;
                si += q.y / aq;
                ++cnt;
            }
        }
        // fixed-order reduction: wave butterflies, then the waves in order
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
                flags[wpos] = 1;
                tr /= (double)tn;
                ti /= (double)tn;
                const double ap = hypot(tr, ti);
                cons[wpos] = ap;
                phases[wpos] = ap > kEps ? mk<double2>(tr / ap, ti / ap) : mk<double2>(0.0, 0.0);
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

RecoWork* reco_work_create(const Desc& d, int32_t xsize, int32_t ysize, const int32_t* visits_host, int64_t nvisits)
{
    auto* w = new RecoWork;
    RecoArgs& a = w->args;
    a.d = d;
    a.X = xsize;
    a.Y = ysize;
    a.pmn0 = -(xsize >> 1);
    a.pmn1 = -(ysize >> 1);
    a.pmx0 = -xsize / 2 + xsize - 1;
    a.pmx1 = -ysize / 2 + ysize - 1;
    a.nvisits = nvisits;
    try {
        SMIP_HIP(hipMalloc(&w->visits, sizeof(int2) * std::max<int64_t>(nvisits, 1)));
        SMIP_HIP(hipMalloc(&w->status, sizeof(int32_t) * 4));
        if (nvisits)
            SMIP_HIP(hipMemcpy(w->visits, visits_host, sizeof(int2) * nvisits, hipMemcpyHostToDevice));
    } catch (...) {
        reco_work_destroy(w);
        throw;
    }
    return w;
}

void reco_work_destroy(RecoWork* w)
{
    if (!w)
        return;
    (void)hipFree(w->visits);
    (void)hipFree(w->status);
    delete w;
}

// Seeds of reconstruct_phases (include/phasereco.h:42-54).
__global__ void k_reco_seed(RecoArgs a, double2* phases, uint8_t* flags, double* cons)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int npts = a.X * a.Y;
    if (i < npts) {
        phases[i] = mk<double2>(0, 0);
        flags[i] = 0;
        cons[i] = 0;
    }
}

__global__ void k_reco_seed2(RecoArgs a, double2* phases, uint8_t* flags, double* cons)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    const int sx[5] = { 0, 1, 0, -1, 0 }, sy[5] = { 0, 0, 1, 0, -1 };
    for (int s = 0; s < 5; ++s) {
        const int p = pm_at(a, sx[s], sy[s]);
        phases[p] = mk<double2>(1.0, 0.0);
        flags[p] = 1;
        cons[p] = 1.0;
    }
}

int run_reconstruct(RecoWork* w, const void* B, int is_f32, double2* phases, uint8_t* flags, double* consistency,
    hipStream_t s)
{
    const RecoArgs& a = w->args;
    const int npts = a.X * a.Y;
    hipLaunchKernelGGL(k_reco_seed, dim3((npts + 255) / 256), dim3(256), 0, s, a, phases, flags, consistency);
    hipLaunchKernelGGL(k_reco_seed2, dim3(1), dim3(1), 0, s, a, phases, flags, consistency);
    SMIP_HIP(hipGetLastError());
    const size_t lds = sizeof(uint32_t) * ((npts + 31) / 32);
    if (lds > 96 * 1024)
        fail(SMIP_ERR_INVALID_ARGUMENT, "phase map too large for the LDS flag bitmap");
    hipLaunchKernelGGL(k_reco_simple, dim3(1), dim3(kRecoThreads), lds, s, a, w->visits, B, is_f32, phases, flags,
        consistency, w->status);
    SMIP_HIP(hipGetLastError());
    int32_t st[4] = { 0, 0, 0, 0 };
    SMIP_HIP(hipMemcpyAsync(st, w->status, sizeof(int32_t) * 3, hipMemcpyDeviceToHost, s));
    SMIP_HIP(hipStreamSynchronize(s));
    if (st[0])
        fail(st[0], "reconstruct_phases: bispectrum element address out of bounds");
    return st[1];
}

} // namespace smip_amd
