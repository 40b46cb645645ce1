// range.h -- inclusive value ranges with row-major (dimension 0 fastest)
// iteration over index vectors (reference interface: include/range.h:16-192).
#pragma once

#include <cmath>
#include <iterator>
#include <limits>
#include <ostream>
#include <stdexcept>

#include "types.h"

namespace smip {

template <typename T>
    requires concept_arithmetic<T> || concept_valarray_of_arithmetic<T>
class Range {
public:
    T low {};
    T high {};

    Range(T lo, T hi)
        : low(lo)
        , high(hi)
    {
    }

    bool contains(const T& value) const
    {
        if constexpr (concept_arithmetic<T>) {
            return value >= low && value <= high;
        } else {
            for (std::size_t k = 0; k < value.size(); ++k)
                if (value[k] < low[k] || value[k] > high[k])
                    return false;
            return true;
        }
    }

    auto extent() const -> T
    {
        if constexpr (concept_arithmetic<T>) {
            if constexpr (std::is_floating_point_v<T>)
                return static_cast<T>(std::abs(high - low) + std::numeric_limits<T>::epsilon());
            else
                return static_cast<T>((high > low ? high - low : low - high) + T(1));
        } else {
            using V = typename std::decay_t<T>::value_type;
            T r = high;
            for (std::size_t k = 0; k < r.size(); ++k) {
                const V d = high[k] > low[k] ? high[k] - low[k] : low[k] - high[k];
                if constexpr (std::is_floating_point_v<V>)
                    r[k] = d + std::numeric_limits<V>::epsilon();
                else
                    r[k] = d + V(1);
            }
            return r;
        }
    }

    class iterator {
        T m_cur, m_low, m_high;

    public:
        using iterator_category = std::input_iterator_tag;
        using value_type = T;
        using difference_type = std::ptrdiff_t;
        using pointer = T*;
        using reference = T&;
        iterator() = default;
        iterator(T cur, T lo, T hi)
            : m_cur(cur)
            , m_low(lo)
            , m_high(hi)
        {
        }
        value_type operator*() const { return m_cur; }
        iterator& operator++()
        {
            if constexpr (concept_arithmetic<T>) {
                if (m_cur < m_high)
                    ++m_cur;
                else
                    m_cur = m_high + 1;
            } else {
                // dimension 0 runs fastest; past the last element the iterator is high + 1
                std::size_t k = 0;
                for (; k < m_cur.size(); ++k) {
                    if (m_cur[k] < m_high[k]) {
                        ++m_cur[k];
                        break;
                    }
                    m_cur[k] = m_low[k];
                }
                if (k == m_cur.size())
                    for (std::size_t d = 0; d < m_cur.size(); ++d)
                        m_cur[d] = m_high[d] + 1;
            }
            return *this;
        }
        iterator operator++(int)
        {
            iterator t = *this;
            ++(*this);
            return t;
        }
        bool operator==(const iterator& o) const
        {
            if constexpr (concept_arithmetic<T>) {
                return m_cur == o.m_cur;
            } else {
                for (std::size_t k = 0; k < m_cur.size(); ++k)
                    if (m_cur[k] != o.m_cur[k])
                        return false;
                return true;
            }
        }
        bool operator!=(const iterator& o) const { return !(*this == o); }
    };

    iterator begin() const { return iterator(low, low, high); }
    iterator end() const
    {
        T past = high;
        if constexpr (concept_arithmetic<T>)
            past = high + 1;
        else
            for (std::size_t k = 0; k < past.size(); ++k)
                past[k] = high[k] + 1;
        return iterator(past, low, high);
    }
};

template <typename T>
    requires concept_arithmetic<T> || concept_valarray_of_arithmetic<T>
std::ostream& operator<<(std::ostream& os, const Range<T>& r)
{
    return os << "(" << r.low << "," << r.high << ")";
}

} // namespace smip
