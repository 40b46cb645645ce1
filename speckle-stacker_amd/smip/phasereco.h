// phasereco.h -- closure-phase recovery on the GPU (reference interface:
// include/phasereco.h:15-27).  reconstruct_phases runs the whole spiral on the
// device (libsmip_mi355x.so); calc_phase runs one recursion step on
// caller-held phases and phase map.
#pragma once

#include <complex>
#include <vector>

#include "array2.h"
#include "bispectrum.h"
#include "engine.h"
#include "phasemap.h"

namespace smip {

// NextRecoIndex, src/phasereco.cpp:17-36
inline void NextRecoIndex(double& r, double& phi, int& i, int& j)
{
    int32_t ii = i, jj = j;
    smip_next_reco_index(&r, &phi, &ii, &jj);
    i = ii;
    j = jj;
}

// reconstruct_phases<T, U>, include/phasereco.h:33-73
template <typename T, typename U>
Array2<T> reconstruct_phases(const Bispectrum<U>& bispec, std::size_t xsize, std::size_t ysize, double reco_radius,
    PhaseMap* phasemap = nullptr)
{
    const std::size_t n = xsize * ysize;
    std::vector<double> ph(2 * n);
    std::vector<uint8_t> fl(n);
    std::vector<double> co(n);
    detail::check(smip_reconstruct_phases(bispec.device_handle(), xsize, ysize, reco_radius, ph.data(), fl.data(),
        co.data()));
    Array2<T> phases(xsize, ysize);
    for (std::size_t k = 0; k < n; ++k)
        phases[0][k] = T(static_cast<typename T::value_type>(ph[2 * k]), static_cast<typename T::value_type>(ph[2 * k + 1]));
    if (phasemap != nullptr) {
        PhaseMap pm(xsize, ysize);
        for (std::size_t k = 0; k < n; ++k)
            pm[0][k] = PhaseMapElement(fl[k] != 0, co[k]);
        *phasemap = pm;
    }
    return phases;
}

// calc_phase<T, U>, include/phasereco.h:75-123
template <typename T, typename U>
void calc_phase(const Bispectrum<U>& bispec, Array2<T>& phases, PhaseMap& pm, DimVector<int, 2> w)
{
    const std::size_t xs = phases.ncols(), ys = phases.nrows(), n = xs * ys;
    if (pm.ncols() != xs || pm.nrows() != ys)
        throw std::invalid_argument("calc_phase: phase array and phase map differ in size");
    std::vector<double> ph(2 * n);
    std::vector<uint8_t> fl(n);
    std::vector<double> co(n);
    for (std::size_t k = 0; k < n; ++k) {
        ph[2 * k] = static_cast<double>(phases[0][k].real());
        ph[2 * k + 1] = static_cast<double>(phases[0][k].imag());
        fl[k] = pm[0][k].flag ? 1 : 0;
        co[k] = pm[0][k].consistency;
    }
    detail::check(smip_calc_phase(bispec.device_handle(), xs, ys, ph.data(), fl.data(), co.data(), w[0], w[1]));
    for (std::size_t k = 0; k < n; ++k) {
        phases[0][k] = T(static_cast<typename T::value_type>(ph[2 * k]), static_cast<typename T::value_type>(ph[2 * k + 1]));
        pm[0][k] = PhaseMapElement(fl[k] != 0, co[k]);
    }
}

} // namespace smip
