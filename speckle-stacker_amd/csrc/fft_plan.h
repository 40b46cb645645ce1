// fft_plan.h -- the transform plan shared by fft.hip (power-of-two kernels) and
// fft_mixed.hip (any length).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "smip_internal.h"

namespace smip_amd {

// Factorisation of one transform length n into Stockham stages: radix 4 while
// 4 divides, then 2, 3, 5 and the remaining primes in ascending order.
constexpr int kMaxStages = 16;
struct MixedRadix {
    int32_t n = 0, nst = 0;
    int32_t radix[kMaxStages] = {};
};
MixedRadix factorize(int32_t n);

struct FftPlan {
    int32_t X, Y, lx, ly;
    bool pow2 = true;   // both lengths powers of two: the fft.hip kernels
    MixedRadix mx, my;  // otherwise the fft_mixed.hip kernels
    // exp(-2 pi i k / n), k < n, per length and precision
    float2* twx32 = nullptr;
    float2* twy32 = nullptr;
    double2* twx64 = nullptr;
    double2* twy64 = nullptr;
    bool cached = false; // kept by fft_plan_create for later jobs: fft_plan_destroy leaves it
};

// The launchers of fft.hip, for a plan with pow2 == false.
void mixed_forward_f32(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval,
    const HalfPlane& hp, float2* out, float2* work, hipStream_t s);
void mixed_forward_f64(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval,
    const HalfPlane& hp, double2* out, double2* work, hipStream_t s);
void mixed_rfft_f64(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval, double2* out,
    double2* work, hipStream_t s);
void mixed_shift_f64(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval,
    const double2* ref_spec, RowBest* best, double2* work, int32_t* shifts, hipStream_t s);
void mixed_xcorr_f64(const FftPlan* p, const double2* ref_spec, const double2* spec, int64_t nf, double* corr,
    double2* work, hipStream_t s);
void mixed_c2c_inverse_f64(const FftPlan* p, double2* data, hipStream_t s);
// entries of the per-workgroup maxima buffer mixed_shift_f64 needs per frame
int mixed_shift_blocks(const FftPlan* p);
// merges nblk per-workgroup maxima per frame into the frame's displacement
void launch_best_to_shift(const RowBest* best, int nblk, int64_t nf, int X, int Y, int32_t* shifts, hipStream_t s);

} // namespace smip_amd
