// utility.h -- small helpers of the reference surface (include/utility.h:28-41).
#pragma once

#include <cmath>
#include <complex>

#include "constants.h"
#include "types.h"

namespace smip {

struct PhaseMapElement;

template <concept_arithmetic T>
T complex_phase(const std::complex<T>& a)
{
    return (std::arg(a) + constants::pi<T>) / constants::c_2pi<T>;
}
template <concept_arithmetic T>
T complex_abs(const std::complex<T>& z)
{
    return std::abs(z);
}
double get_phase_consistency(const PhaseMapElement& pme);

} // namespace smip
