// dimvector.h -- fixed-rank index vector on std::valarray
// (reference interface: include/dimvector.h:23-102).
#pragma once

#include <algorithm>
#include <initializer_list>
#include <numeric>
#include <ostream>
#include <stdexcept>
#include <valarray>

#include "types.h"

namespace smip {

template <concept_arithmetic T, std::size_t NrDims = 0>
class DimVector : public std::valarray<T> {
    using base = std::valarray<T>;

public:
    using typename base::value_type;
    using base::base; // valarray constructors, incl. those from valarray expressions
    using base::operator=;
    using base::operator[];
    using base::size;
    using base::sum;
    using base::min;
    using base::operator+=;
    using base::operator-=;
    using base::operator*=;
    using base::operator/=;

    DimVector()
        : base(NrDims)
    {
    }
    DimVector(const DimVector& x)
        : base(NrDims)
    {
        if (x.size() != NrDims && x.size() != 0)
            throw std::domain_error("invalid initializer size");
        if (x.size())
            std::copy(std::begin(x), std::end(x), std::begin(*this));
    }
    // a list of another length leaves the vector zero-filled
    DimVector(std::initializer_list<T> l)
        : base(NrDims)
    {
        if (l.size() == NrDims)
            std::copy(l.begin(), l.end(), std::begin(*this));
    }
    DimVector& operator=(const DimVector& x)
    {
        if (x.size() != NrDims)
            throw std::domain_error("invalid initializer size");
        std::copy(std::begin(x), std::end(x), std::begin(*this));
        return *this;
    }
    bool operator==(const DimVector& o) const
    {
        return size() == o.size() && std::equal(std::begin(*this), std::end(*this), std::begin(o));
    }
    bool operator!=(const DimVector& o) const { return !(*this == o); }

    [[nodiscard]] T product() const { return std::accumulate(std::begin(*this), std::end(*this), T { 1 }, std::multiplies<T>()); }
    void fill(T value) { base::operator=(value); }

    [[nodiscard]] DimVector<T, NrDims + 1> appended_back(T v) const
    {
        DimVector<T, NrDims + 1> r;
        std::copy(std::begin(*this), std::end(*this), std::begin(r));
        r[NrDims] = v;
        return r;
    }
    [[nodiscard]] DimVector<T, NrDims + 1> appended_front(T v) const
    {
        DimVector<T, NrDims + 1> r;
        r[0] = v;
        std::copy(std::begin(*this), std::end(*this), std::begin(r) + 1);
        return r;
    }
    [[nodiscard]] DimVector<T, NrDims - 1> removed_back() const
    {
        DimVector<T, NrDims - 1> r;
        std::copy(std::begin(*this), std::begin(*this) + (NrDims - 1), std::begin(r));
        return r;
    }
    [[nodiscard]] DimVector<T, NrDims - 1> removed_front() const
    {
        DimVector<T, NrDims - 1> r;
        std::copy(std::begin(*this) + 1, std::end(*this), std::begin(r));
        return r;
    }
    [[nodiscard]] DimVector<T, NrDims + 1> operator,(T in) const { return appended_back(in); }

    template <concept_arithmetic U>
    operator DimVector<U, NrDims>() const
    {
        DimVector<U, NrDims> r;
        std::transform(std::begin(*this), std::end(*this), std::begin(r), [](const T& x) { return static_cast<U>(x); });
        return r;
    }
};

// rank-independent helpers
template <concept_arithmetic T>
class DimVector<T, 0> : public std::valarray<T> {
public:
    template <std::size_t D1, std::size_t D2>
    static auto merge(const DimVector<T, D1>& a, const DimVector<T, D2>& b) -> DimVector<T, D1 + D2>
    {
        DimVector<T, D1 + D2> r;
        std::copy(std::begin(a), std::end(a), std::begin(r));
        std::copy(std::begin(b), std::end(b), std::begin(r) + D1);
        return r;
    }
};

template <concept_arithmetic T, std::size_t N>
std::ostream& operator<<(std::ostream& os, const DimVector<T, N>& v)
{
    os << "[ ";
    for (auto x : v)
        os << x << " ";
    return os << "]";
}

} // namespace smip
