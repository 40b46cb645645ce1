// types.h -- concepts and complex types of the drop-in surface
// (reference: include/types.h:7-35).
#pragma once

#include <complex>
#include <concepts>
#include <type_traits>
#include <valarray>

template <typename T>
concept concept_integral = std::is_integral_v<T>;
template <typename T>
concept concept_floating = std::is_floating_point_v<T>;
template <typename T>
concept concept_arithmetic = std::is_arithmetic_v<T>;

// std::valarray<T> of an arithmetic T, or a type derived from one
template <typename T>
concept concept_valarray_of_arithmetic = requires { typename std::decay_t<T>::value_type; }
    && std::is_arithmetic_v<typename std::decay_t<T>::value_type>
    && std::is_base_of_v<std::valarray<typename std::decay_t<T>::value_type>, std::decay_t<T>>;

template <typename T>
concept concept_complex = requires { typename std::decay_t<T>::value_type; }
    && std::is_same_v<std::decay_t<T>, std::complex<typename std::decay_t<T>::value_type>>;

template <typename T>
concept concept_arithmetic_or_valarray_or_complex
    = concept_arithmetic<T> || concept_valarray_of_arithmetic<T> || concept_complex<T>;

namespace smip {
typedef std::complex<double> complex_t;
typedef std::complex<float> bispec_complex_t;
} // namespace smip
