// phasemap.h -- per-frequency recovery flag and consistency
// (reference interface: include/phasemap.h:10-20).
#pragma once

#include "array2.h"
#include "utility.h"

namespace smip {

struct PhaseMapElement {
    PhaseMapElement() = default;
    PhaseMapElement(bool a_flag)
        : flag(a_flag)
    {
    }
    PhaseMapElement(bool a_flag, double a_consistency)
        : flag(a_flag)
        , consistency(a_consistency)
    {
    }
    double operator()() const { return consistency; }
    operator double() const { return consistency; }
    bool flag { false };
    double consistency { 0. };
};

typedef Array2<PhaseMapElement> PhaseMap;

inline double get_phase_consistency(const PhaseMapElement& pme) { return pme.consistency; }

} // namespace smip
