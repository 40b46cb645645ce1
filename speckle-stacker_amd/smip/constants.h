// constants.h -- numeric constants (reference: include/constants.h:9-21).
#pragma once

#include <limits>
#include <numbers>

#include "types.h"

namespace smip::constants {
template <concept_floating T = double>
inline constexpr T pi = std::numbers::pi_v<T>;
template <concept_floating T = double>
inline constexpr T c_0 { static_cast<T>(299'792'458.0) };
template <concept_floating T = double>
inline constexpr T c_2pi { pi<T> * 2 };
template <concept_floating T = double>
inline constexpr T c_sqrt_2pi { static_cast<T>(2.5070831590532E+00) };
template <concept_floating T = double>
inline constexpr T c_epsilon { std::numeric_limits<T>::epsilon() };
} // namespace smip::constants
