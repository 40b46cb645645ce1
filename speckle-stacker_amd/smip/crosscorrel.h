// crosscorrel.h -- CrossCorrelation<T> on the GPU (reference interface:
// include/crosscorrel.h:43-189).  The correlation is computed in FP64 by
// libsmip_mi355x.so: r2c FFTs of reference and frame, conj(R_ref) * R_frame,
// c2r; the displacement is the first maximum in storage order, wrapped to
// signed offsets.
#pragma once

#include <concepts>
#include <functional>
#include <stdexcept>

#include "array2.h"
#include "dimvector.h"
#include "engine.h"

namespace smip {

template <typename T>
    requires std::floating_point<T>
class CrossCorrelation {
public:
    CrossCorrelation() = delete;
    CrossCorrelation(const Array2<T>& ref)
        : m_X(ref.ncols())
        , m_Y(ref.nrows())
    {
        const Array2<double> r(ref);
        detail::check(smip_xcorr_create(detail::context(), r.begin(), m_X, m_Y, &m_dev));
    }
    CrossCorrelation(const CrossCorrelation&) = delete;
    CrossCorrelation& operator=(const CrossCorrelation&) = delete;
    ~CrossCorrelation()
    {
        if (m_dev)
            smip_xcorr_destroy(m_dev);
    }

    void correlate(const Array2<T>& frame)
    {
        const Array2<double> f(frame);
        detail::check(smip_xcorr_correlate(m_dev, f.begin(), f.ncols(), f.nrows()));
        m_have_corr = false;
    }
    auto get_correlation_array() -> const Array2<T>&
    {
        if (!m_have_corr) {
            Array2<double> c(m_X, m_Y);
            detail::check(smip_xcorr_get_correlation(m_dev, c.begin()));
            m_correlation = Array2<T>(c);
            m_have_corr = true;
        }
        return m_correlation;
    }
    auto get_displacement() -> DimVector<int, 2>
    {
        int32_t s[2];
        detail::check(smip_xcorr_get_displacement(m_dev, s));
        return { s[0], s[1] };
    }
    auto operator()(const Array2<T>& frame) -> DimVector<int, 2>
    {
        const Array2<double> f(frame);
        int32_t s[2];
        detail::check(smip_xcorr_displacement(m_dev, f.begin(), f.ncols(), f.nrows(), s));
        m_have_corr = false;
        return { s[0], s[1] };
    }
    static auto get_displacement(const Array2<T>& a, const Array2<T>& b) -> DimVector<int, 2>
    {
        CrossCorrelation<T> correl(a);
        return correl(b);
    }

private:
    smip_xcorr* m_dev = nullptr;
    std::size_t m_X, m_Y;
    Array2<T> m_correlation;
    bool m_have_corr = false;
};

template <typename T>
    requires std::floating_point<T>
CrossCorrelation(const Array2<T>& ref) -> CrossCorrelation<T>;

} // namespace smip
