// post.hip -- shift-and-add and image synthesis on gfx950.
//
//  k_argmax_shift  calculate_displacement (include/crosscorrel.h:166-179): first
//                  maximum of the correlation in row-major order, then the signed
//                  wrap x >= (n+1)/2 -> x - n.  Wavefront reduction keeps the
//                  lowest index among equal maxima, so ties resolve as
//                  std::max_element does.
//  k_saa_add       sumarray += |frame|.shifted(-shift) (smip-cli/main.cpp:289-294,
//                  include/array2.h:603-622), frames added in order per pixel so
//                  the FP64 sum has the reference's association.
//  k_synth_prepare result = sqrt(Re P) * phase * Hann(2p) (smip-cli/main.cpp:329-338,
//                  include/window_function.h:69-84)
//  k_absmax_*      result /= max |result| (smip-cli/main.cpp:371-376)
#include <algorithm>
#include <cmath>

#include "smip_complex.h"
#include "smip_internal.h"

namespace smip_amd {

constexpr int kArgThreads = 1024;

__device__ inline void better(double& bv, int64_t& bi, double v, int64_t i)
{
    // (v, i) replaces (bv, bi) when it is larger, or equal with a lower index
    if (bv < v || (!(v < bv) && i < bi)) {
        bv = v;
        bi = i;
    }
}

__global__ __launch_bounds__(kArgThreads) void k_argmax_shift(const double* __restrict__ corr, int32_t X, int32_t Y,
    int32_t* __restrict__ shifts)
{
    __shared__ double sv[kArgThreads / 64];
    __shared__ int64_t si[kArgThreads / 64];
    const int64_t n = (int64_t)X * Y;
    const double* c = corr + (int64_t)blockIdx.x * n;
    double bv = c[0];
    int64_t bi = 0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const double v = c[i];
        if (bv < v) { // strict: the first maximum of this lane's stride wins
            bv = v;
            bi = i;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(bv, o);
        const int64_t oi = __shfl_xor(bi, o);
        better(bv, bi, ov, oi);
    }
    if ((threadIdx.x & 63) == 0) {
        sv[threadIdx.x >> 6] = bv;
        si[threadIdx.x >> 6] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bv = sv[0];
        bi = si[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k)
            better(bv, bi, sv[k], si[k]);
        int32_t sx = (int32_t)(bi % X), sy = (int32_t)(bi / X);
        if (!(sx < (X + 1) / 2))
            sx -= X;
        if (!(sy < (Y + 1) / 2))
            sy -= Y;
        shifts[2 * blockIdx.x] = sx;
        shifts[2 * blockIdx.x + 1] = sy;
    }
}

void launch_argmax_shift(const double* corr, int64_t nf, int32_t ncols, int32_t nrows, int32_t* shifts,
    hipStream_t s)
{
    if (nf <= 0)
        return;
    hipLaunchKernelGGL(k_argmax_shift, dim3((unsigned)nf), dim3(kArgThreads), 0, s, corr, ncols, nrows, shifts);
    SMIP_HIP(hipGetLastError());
}

// |frames[f]| shifted, in the sample type (0 outside the window): the absolute
// value and the widening to double are exact in either order.
template <typename In>
__device__ inline In shifted_abs_raw(const In* __restrict__ frames, int64_t f, int32_t X, int32_t Y, int32_t col,
    int32_t row, const int32_t* __restrict__ shifts)
{
    // distance = -shift; shifted(d)(col, row) = in(col - d0, row - d1) on the
    // clamped window, zero elsewhere
    const int32_t dx = -shifts[2 * f], dy = -shifts[2 * f + 1];
    const int32_t sx = min(max(dx, 0), X - 1), ex = min(max(dx + X, 1), X);
    const int32_t sy = min(max(dy, 0), Y - 1), ey = min(max(dy + Y, 1), Y);
    // the load is unconditional (from a clamped, in-frame position) and masked
    // afterwards, so that a thread's loads of many frames are all in flight
    // together instead of each waiting inside its own branch
    const bool in = col >= sx && col < ex && row >= sy && row < ey;
    const int32_t r = min(max(row - dy, 0), Y - 1), c = min(max(col - dx, 0), X - 1);
    const In x = frames[f * (int64_t)X * Y + (int64_t)r * X + c];
    const In a = x < (In)0 ? -x : x;
    return in ? a : (In)0;
}

template <typename In>
__device__ __forceinline__ void saa_add_body(const In* __restrict__ frames, int64_t nf, int32_t X, int32_t Y,
    const int32_t* __restrict__ shifts, double* __restrict__ sum, int64_t p)
{
    if (p >= (int64_t)X * Y)
        return;
    const int32_t col = (int32_t)(p % X), row = (int32_t)(p / X);
    double acc = sum[p];
    int64_t f = 0;
    // sixteen frames' loads in flight at once, added in frame order
    constexpr int U = 16;
    for (; f + U <= nf; f += U) {
        In v[U];
#pragma unroll
        for (int k = 0; k < U; ++k)
            v[k] = shifted_abs_raw(frames, f + k, X, Y, col, row, shifts);
#pragma unroll
        for (int k = 0; k < U; ++k)
            acc += (double)v[k];
    }
    for (; f < nf; ++f)
        acc += (double)shifted_abs_raw(frames, f, X, Y, col, row, shifts);
    sum[p] = acc;
}

template <typename In>
__global__ void k_saa_add(const In* __restrict__ frames, int64_t nf, int32_t X, int32_t Y,
    const int32_t* __restrict__ shifts, double* __restrict__ sum)
{
    saa_add_body(frames, nf, X, Y, shifts, sum, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Integer samples: |x| as a double is an integer, and the per-pixel sums stay far
// below 2^53, so every partial sum is exact and their order cannot change a bit.
// The frames are split into `parts` slices, one per grid row, each summed by its
// own threads and added to the pixel with an FP64 atomic: sixteen times the
// threads of the frame-ordered form on small frames, where that form leaves most
// CUs idle.
template <typename In>
__device__ __forceinline__ void saa_add_int_body(const In* __restrict__ frames, int64_t nf, int32_t X, int32_t Y,
    const int32_t* __restrict__ shifts, double* __restrict__ sum, int64_t p, unsigned slice, unsigned parts)
{
    if (p >= (int64_t)X * Y)
        return;
    const int32_t col = (int32_t)(p % X), row = (int32_t)(p / X);
    const int64_t per = (nf + parts - 1) / parts;
    const int64_t f0 = (int64_t)slice * per, f1 = min(nf, f0 + per);
    double acc = 0.0;
    int64_t f = f0;
    for (; f + 8 <= f1; f += 8) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            v[k] = (double)shifted_abs_raw(frames, f + k, X, Y, col, row, shifts);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            acc += v[k];
    }
    for (; f < f1; ++f)
        acc += (double)shifted_abs_raw(frames, f, X, Y, col, row, shifts);
    if (f1 > f0)
        atomicAdd(sum + p, acc);
}

template <typename In>
__global__ void k_saa_add_int(const In* __restrict__ frames, int64_t nf, int32_t X, int32_t Y,
    const int32_t* __restrict__ shifts, double* __restrict__ sum)
{
    saa_add_int_body(frames, nf, X, Y, shifts, sum, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y,
        gridDim.y);
}

unsigned saa_add_parts(int dtype, int64_t nf, int32_t ncols, int32_t nrows)
{
    if (dtype != SMIP_DTYPE_U8 && dtype != SMIP_DTYPE_U16)
        return 1;
    // about 2048 workgroups in all, 8 frames or more per slice
    const int64_t blocks = ((int64_t)ncols * nrows + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>({ 16, nf / 8, (2048 + blocks - 1) / blocks }));
}

// The shift-and-add of many jobs' chunks: per job ceil(N^2 / 256) blocks times
// its slices; float frames keep the frame-ordered sum (one slice).
__global__ void k_group_saa_add(const GroupSaa* __restrict__ tab, const uint32_t* __restrict__ off, int n)
{
    const uint32_t g = blockIdx.x;
    const int j = group_job(off, n, g);
    const GroupSaa& J = tab[j];
    const uint32_t local = g - off[j];
    const uint32_t pb = (uint32_t)(((int64_t)J.X * J.Y + 255) / 256);
    const int64_t p = (int64_t)(local % pb) * blockDim.x + threadIdx.x;
    const unsigned slice = local / pb;
    switch (J.dtype) {
    case SMIP_DTYPE_U8:
        saa_add_int_body((const uint8_t*)global_ptr(J.frames), J.nf, J.X, J.Y, global_ptr(J.shifts), global_ptr(J.sum),
            p, slice, (unsigned)J.parts);
        break;
    case SMIP_DTYPE_U16:
        saa_add_int_body((const uint16_t*)global_ptr(J.frames), J.nf, J.X, J.Y, global_ptr(J.shifts),
            global_ptr(J.sum), p, slice, (unsigned)J.parts);
        break;
    default:
        saa_add_body((const float*)global_ptr(J.frames), J.nf, J.X, J.Y, global_ptr(J.shifts), global_ptr(J.sum), p);
    }
}

void launch_group_saa_add(const GroupSaa* tab, const uint32_t* off, int n, uint32_t nblocks, hipStream_t s)
{
    if (n <= 0 || nblocks == 0)
        return;
    hipLaunchKernelGGL(k_group_saa_add, dim3(nblocks), dim3(256), 0, s, tab, off, n);
    SMIP_HIP(hipGetLastError());
}

void launch_saa_add(const void* frames, int dtype, int64_t nf, int32_t ncols, int32_t nrows, const int32_t* shifts,
    double* sum, hipStream_t s)
{
    if (nf <= 0)
        return;
    const int64_t n = (int64_t)ncols * nrows;
    dim3 grid((unsigned)((n + 255) / 256));
    const unsigned parts = saa_add_parts(dtype, nf, ncols, nrows);
    switch (dtype) {
    case SMIP_DTYPE_U8:
        hipLaunchKernelGGL(k_saa_add_int<uint8_t>, dim3(grid.x, parts), dim3(256), 0, s, (const uint8_t*)frames, nf,
            ncols, nrows, shifts, sum);
        break;
    case SMIP_DTYPE_U16:
        hipLaunchKernelGGL(k_saa_add_int<uint16_t>, dim3(grid.x, parts), dim3(256), 0, s, (const uint16_t*)frames, nf,
            ncols, nrows, shifts, sum);
        break;
    case SMIP_DTYPE_F32:
        hipLaunchKernelGGL(k_saa_add<float>, grid, dim3(256), 0, s, (const float*)frames, nf, ncols, nrows, shifts,
            sum);
        break;
    case SMIP_DTYPE_F64:
        hipLaunchKernelGGL(k_saa_add<double>, grid, dim3(256), 0, s, (const double*)frames, nf, ncols, nrows,
            shifts, sum);
        break;
    default:
        fail(SMIP_ERR_INVALID_ARGUMENT, "unsupported frame dtype");
    }
    SMIP_HIP(hipGetLastError());
}

// P(x, y) from the stored half plane Ph[c][r] (c = x + X/2 for x <= 0,
// r = y + Y/2); x > 0 uses P(x, y) = P(-x, -y) of a real frame.
__device__ inline double power_at(const double* Ph, int32_t X, int32_t Y, int32_t x, int32_t y)
{
    if (x > 0) {
        x = -x;
        y = -y;
        if (y > Y - 1 - (Y >> 1))
            y -= Y;
        if (y < -(Y >> 1))
            y += Y;
    }
    return Ph[(int64_t)(x + (X >> 1)) * Y + (y + (Y >> 1))];
}

__global__ void k_power_expand(const double* __restrict__ Ph, int32_t X, int32_t Y, double norm,
    double* __restrict__ out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)X * Y)
        return;
    const int32_t col = (int32_t)(p % X), row = (int32_t)(p / X);
    const int32_t x = col > X - 1 - (X >> 1) ? col - X : col;
    const int32_t y = row > Y - 1 - (Y >> 1) ? row - Y : row;
    out[p] = power_at(Ph, X, Y, x, y) / norm;
}

void launch_power_expand(const double* power_half, int32_t ncols, int32_t nrows, double norm, double* out,
    hipStream_t s)
{
    const int64_t n = (int64_t)ncols * nrows;
    hipLaunchKernelGGL(k_power_expand, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, power_half, ncols, nrows,
        norm, out);
    SMIP_HIP(hipGetLastError());
}

__global__ void k_synth_prepare(const double* __restrict__ Ph, const double2* __restrict__ phases, int32_t X,
    int32_t Y, double trigarg, double norm, double2* __restrict__ out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)X * Y)
        return;
    const int32_t col = (int32_t)(p % X), row = (int32_t)(p / X);
    const int32_t x = col > X - 1 - (X >> 1) ? col - X : col;
    const int32_t y = row > Y - 1 - (Y >> 1) ? row - Y : row;
    // GeneralHamming with alpha = 0.5 on signed indices
    const double t1 = 0.5 + 0.5 * cos(trigarg * (double)x);
    const double t2 = 0.5 + 0.5 * cos(trigarg * (double)y);
    const double w = t1 * t2;
    const double2 ph = phases[p];
    // phases *= (w, 0); result = (sqrt(Re P), 0) * phases
    const double pr = ph.x * w - ph.y * 0.0;
    const double pi = ph.x * 0.0 + ph.y * w;
    const double s = sqrt(power_at(Ph, X, Y, x, y) / norm);
    out[p] = mk<double2>(s * pr - 0.0 * pi, s * pi + 0.0 * pr);
}

void launch_synth_prepare(const double* power_half, const double2* phases, int32_t ncols, int32_t nrows,
    double reco_radius, double norm, double2* out, hipStream_t s)
{
    const double aperture = reco_radius * 2;
    if (!(aperture > 0.))
        fail(SMIP_ERR_DOMAIN, "WindowFunction aperture size <= 0");
    const double trigarg = 6.283185307179586476925286766559 / aperture;
    const int64_t n = (int64_t)ncols * nrows;
    hipLaunchKernelGGL(k_synth_prepare, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, power_half, phases,
        ncols, nrows, trigarg, norm, out);
    SMIP_HIP(hipGetLastError());
}

__global__ __launch_bounds__(1024) void k_absmax(const double2* __restrict__ d, int64_t n, double* __restrict__ out)
{
    __shared__ double sm[16];
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
        m = fmax(m, hypot(d[i].x, d[i].y));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0)
        sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k)
            m = fmax(m, sm[k]);
        out[0] = m;
    }
}

__global__ void k_div_by(double2* __restrict__ d, int64_t n, const double* __restrict__ by)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double m = by[0];
        d[i] = mk<double2>(d[i].x / m, d[i].y / m);
    }
}

void launch_abs_max_normalize(double2* data, int64_t n, double* scratch, hipStream_t s)
{
    hipLaunchKernelGGL(k_absmax, dim3(1), dim3(1024), 0, s, data, n, scratch);
    hipLaunchKernelGGL(k_div_by, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, data, n, scratch);
    SMIP_HIP(hipGetLastError());
}

} // namespace smip_amd
