// smip_complex.h -- complex arithmetic on HIP vector types (float2 / double2).
#pragma once

#include <hip/hip_runtime.h>

namespace smip_amd {

template <typename V>
struct scalar_of;
template <>
struct scalar_of<float2> {
    using type = float;
};
template <>
struct scalar_of<double2> {
    using type = double;
};

template <typename V>
__host__ __device__ inline V mk(typename scalar_of<V>::type x, typename scalar_of<V>::type y)
{
    V r;
    r.x = x;
    r.y = y;
    return r;
}
template <typename V>
__host__ __device__ inline V cadd(V a, V b) { return mk<V>(a.x + b.x, a.y + b.y); }
template <typename V>
__host__ __device__ inline V csub(V a, V b) { return mk<V>(a.x - b.x, a.y - b.y); }
// (a.x b.x - a.y b.y, a.x b.y + a.y b.x): the operand order of std::complex's
// operator*= (ac - bd, ad + bc).
template <typename V>
__host__ __device__ inline V cmul(V a, V b) { return mk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename V>
__host__ __device__ inline V cconj(V a) { return mk<V>(a.x, -a.y); }
// a * conj(b)
template <typename V>
__host__ __device__ inline V cmulc(V a, V b) { return mk<V>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
// multiply by -i (forward) or +i (inverse)
template <typename V>
__host__ __device__ inline V cmul_mi(V a) { return mk<V>(a.y, -a.x); }
template <typename V>
__host__ __device__ inline V cmul_pi(V a) { return mk<V>(-a.y, a.x); }

} // namespace smip_amd
