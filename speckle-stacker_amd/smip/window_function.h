// window_function.h -- apodisation windows (reference interface:
// include/window_function.h:11-84).  Values from libsmip_mi355x.so's host
// routine: w(i, j) = (a + (1-a) cos(2 pi i / A)) (a + (1-a) cos(2 pi j / A)) at
// every signed FFT-order index (i, j) of the array, with no cut-off at |i| = A/2:
// beyond it the cosine is evaluated as it is inside, as the reference does.
#pragma once

#include <stdexcept>
#include <vector>

#include "array2.h"
#include "engine.h"

namespace smip {

template <typename T>
class WindowFunction : public Array2<T> {
public:
    WindowFunction() = delete;
    WindowFunction(std::size_t a_xsize, std::size_t a_ysize, double window_aperture)
        : Array2<T>(a_xsize, a_ysize)
        , m_aperture(window_aperture)
    {
        if (!(window_aperture > 0.))
            throw std::domain_error("WindowFunction aperture size <= 0");
    }

protected:
    double m_aperture {};
};

template <typename T>
class GeneralHamming : public WindowFunction<T> {
public:
    GeneralHamming(std::size_t a_xsize, std::size_t a_ysize, double window_aperture, double a_alpha)
        : WindowFunction<T>(a_xsize, a_ysize, window_aperture)
        , m_alpha(a_alpha)
    {
        std::vector<double> w(a_xsize * a_ysize);
        detail::check(smip_general_hamming(a_xsize, a_ysize, window_aperture, a_alpha, w.data()));
        for (std::size_t k = 0; k < w.size(); ++k)
            this->begin()[k] = static_cast<T>(w[k]);
    }

protected:
    double m_alpha {};
};

template <typename T>
class Hann : public GeneralHamming<T> {
public:
    Hann(std::size_t a_xsize, std::size_t a_ysize, double window_aperture)
        : GeneralHamming<T>(a_xsize, a_ysize, window_aperture, 0.5)
    {
    }
};

template <typename T>
class Hamming : public GeneralHamming<T> {
public:
    Hamming(std::size_t a_xsize, std::size_t a_ysize, double window_aperture)
        : GeneralHamming<T>(a_xsize, a_ysize, window_aperture, 0.54)
    {
    }
};

template <typename T, std::size_t xsize, std::size_t ysize>
class GeneralWindowFunction : public Array2<T> {
public:
    GeneralWindowFunction()
        : Array2<T>(xsize, ysize)
    {
    }
};

} // namespace smip
