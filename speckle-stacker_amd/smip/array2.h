// array2.h -- 2-D row-major container, x (column) fastest, signed indices that
// wrap like FFT order (reference interface: include/array2.h:41-164).
#pragma once

#include <algorithm>
#include <cassert>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "array_base.h"
#include "dimvector.h"
#include "range.h"

namespace smip {

template <typename T>
class Array2 : public Array_base<T> {
    using Array_base<T>::m_data;
    using Array_base<T>::m_size;

public:
    using Array_base<T>::begin;
    using Array_base<T>::end;
    typedef DimVector<std::size_t, 2> extents;
    typedef DimVector<int, 2> s_indices;
    typedef DimVector<std::size_t, 2> u_indices;

    Array2() = default;
    Array2(const Array2& src) = default;
    template <concept_arithmetic U>
    Array2(const Array2<U>& src)
        : Array_base<T>(src.xsize() * src.ysize())
        , m_xsize(src.xsize())
        , m_ysize(src.ysize())
    {
        std::transform(src.begin(), src.end(), begin(), [](const U& x) { return static_cast<T>(x); });
    }
    Array2(Array2&& o) noexcept
        : Array_base<T>(std::move(o))
        , m_xsize(o.m_xsize)
        , m_ysize(o.m_ysize)
    {
        o.m_xsize = o.m_ysize = 0;
    }
    Array2(std::size_t xsize, std::size_t ysize)
        : Array_base<T>(xsize * ysize)
        , m_xsize(xsize)
        , m_ysize(ysize)
    {
    }
    Array2(std::size_t xsize, std::size_t ysize, const T& init)
        : Array_base<T>(xsize * ysize, init)
        , m_xsize(xsize)
        , m_ysize(ysize)
    {
    }
    Array2(const extents& e)
        : Array2(e[0], e[1])
    {
    }
    Array2(const extents& e, const T& init)
        : Array2(e[0], e[1], init)
    {
    }
    Array2(std::initializer_list<std::initializer_list<T>> l)
        : Array_base<T>(l.size() * (l.size() ? l.begin()->size() : 0))
        , m_xsize(l.size() ? l.begin()->size() : 0)
        , m_ysize(l.size())
    {
        std::size_t r = 0;
        for (const auto& row : l)
            std::copy(row.begin(), row.end(), (*this)[(int)r++]);
    }

    // a refused resize (reference-backed storage) surfaces as std::runtime_error,
    // as in include/array2.h:341-357
    Array2& operator=(const Array2& o)
    {
        if (this == &o)
            return *this;
        resize_for_assign(o.size());
        m_xsize = o.m_xsize;
        m_ysize = o.m_ysize;
        std::copy(o.begin(), o.end(), begin());
        return *this;
    }
    template <concept_arithmetic U>
    Array2& operator=(const Array2<U>& o)
    {
        resize_for_assign(o.size());
        m_xsize = o.xsize();
        m_ysize = o.ysize();
        std::transform(o.begin(), o.end(), begin(), [](const U& x) { return static_cast<T>(x); });
        return *this;
    }
    Array2& operator=(Array2&& o) noexcept
    {
        if (this != &o) {
            Array_base<T>::operator=(std::move(o));
            m_xsize = o.m_xsize;
            m_ysize = o.m_ysize;
            o.m_xsize = o.m_ysize = 0;
        }
        return *this;
    }
    Array2& operator=(const T& v)
    {
        std::fill(begin(), end(), v);
        return *this;
    }
    template <concept_arithmetic U>
    Array2& operator=(const U& v)
    {
        std::fill(begin(), end(), static_cast<T>(v));
        return *this;
    }

    Array2& operator+=(const Array2& x) { return zip(x, std::plus<T>()); }
    Array2& operator-=(const Array2& x) { return zip(x, std::minus<T>()); }
    Array2& operator*=(const Array2& x) { return zip(x, std::multiplies<T>()); }
    Array2& operator/=(const Array2& x) { return zip(x, std::divides<T>()); }
    template <concept_arithmetic U>
    Array2& operator+=(const Array2<U>& x) { return zipc(x, [](const T& a, const T& b) { return a + b; }); }
    template <concept_arithmetic U>
    Array2& operator-=(const Array2<U>& x) { return zipc(x, [](const T& a, const T& b) { return a - b; }); }
    template <concept_arithmetic U>
    Array2& operator*=(const Array2<U>& x) { return zipc(x, [](const T& a, const T& b) { return a * b; }); }
    template <concept_arithmetic U>
    Array2& operator/=(const Array2<U>& x) { return zipc(x, [](const T& a, const T& b) { return a / b; }); }
    Array2& operator+=(const T& v) { return each([&](T& a) { a += v; }); }
    Array2& operator-=(const T& v) { return each([&](T& a) { a -= v; }); }
    Array2& operator*=(const T& v) { return each([&](T& a) { a *= v; }); }
    Array2& operator/=(const T& v) { return each([&](T& a) { a /= v; }); }
    template <concept_arithmetic U>
    Array2& operator+=(const U& v) { return each([&](T& a) { a += static_cast<T>(v); }); }
    template <concept_arithmetic U>
    Array2& operator-=(const U& v) { return each([&](T& a) { a -= static_cast<T>(v); }); }
    template <concept_arithmetic U>
    Array2& operator*=(const U& v) { return each([&](T& a) { a *= static_cast<T>(v); }); }
    template <concept_arithmetic U>
    Array2& operator/=(const U& v) { return each([&](T& a) { a /= static_cast<T>(v); }); }
    Array2 operator+() const { return *this; }
    Array2 operator-() const
    {
        Array2 r(m_xsize, m_ysize);
        std::transform(begin(), end(), r.begin(), [](const T& a) { return -a; });
        return r;
    }

    // row pointer; negative rows wrap
    const T* operator[](int row) const { return m_data.get() + checked_row(row) * m_xsize; }
    T* operator[](int row) { return m_data.get() + checked_row(row) * m_xsize; }
    T& operator()(std::size_t col, std::size_t row) { return m_data[checked(col, row)]; }
    const T& operator()(std::size_t col, std::size_t row) const { return m_data[checked(col, row)]; }
    // signed indices: negative values wrap by the extent (FFT order)
    const T& at(s_indices idx) const { return m_data[wrapped(idx)]; }
    T& at(s_indices idx) { return m_data[wrapped(idx)]; }

    std::vector<T> get_col(std::size_t col) const
    {
        std::vector<T> v(m_ysize);
        for (std::size_t r = 0; r < m_ysize; ++r)
            v[r] = m_data[col + r * m_xsize];
        return v;
    }
    std::vector<T> get_row(std::size_t row) const { return std::vector<T>(begin() + row * m_xsize, begin() + (row + 1) * m_xsize); }

    // non-periodic integer shift with zero fill: out(c, r) = in(c - d0, r - d1)
    Array2 shifted(const DimVector<int, 2>& d)
    {
        Array2 out(m_xsize, m_ysize, T {});
        const int X = static_cast<int>(m_xsize), Y = static_cast<int>(m_ysize);
        const int sx = std::clamp(d[0], 0, X - 1), sy = std::clamp(d[1], 0, Y - 1);
        const int ex = std::clamp(d[0] + X, 1, X), ey = std::clamp(d[1] + Y, 1, Y);
        for (int r = sy; r < ey; ++r)
            for (int c = sx; c < ex; ++c)
                out(c, r) = (*this)(c - d[0], r - d[1]);
        return out;
    }
    void shift(const DimVector<int, 2>& d)
    {
        Array2 s = shifted(d);
        std::swap(s.data(), this->data());
    }

    std::size_t xsize() const { return m_xsize; }
    std::size_t ysize() const { return m_ysize; }
    std::size_t ncols() const { return m_xsize; }
    std::size_t nrows() const { return m_ysize; }
    s_indices min_sindices() const { return { -static_cast<int>(m_xsize >> 1), -static_cast<int>(m_ysize >> 1) }; }
    s_indices max_sindices() const
    {
        return { -static_cast<int>(m_xsize) / 2 + static_cast<int>(m_xsize) - 1,
            -static_cast<int>(m_ysize) / 2 + static_cast<int>(m_ysize) - 1 };
    }
    Range<DimVector<int, 2>> range() const { return Range<DimVector<int, 2>>(min_sindices(), max_sindices()); }

    void print() const
    {
        std::cout << "object: Array2\nextents: x=" << m_xsize << ", y=" << m_ysize << "\nsize of datatype: "
                  << sizeof(T) << " bytes\nnr. of elements: " << m_size << std::endl;
    }

    template <typename U>
    void import(const Array2<U>& src, std::function<T(const U&)> conversion)
    {
        m_xsize = src.xsize();
        m_ysize = src.ysize();
        if (m_size != src.size())
            Array_base<T>::resize(src.size());
        std::transform(src.begin(), src.end(), begin(), conversion);
    }
    template <concept_arithmetic U>
    void import(const Array2<U>& src)
    {
        import<U>(src, [](const U& x) { return static_cast<T>(x); });
    }
    template <typename U>
    static Array2 convert(const Array2<U>& src, std::function<T(const U&)> conversion)
    {
        Array2 a;
        a.import(src, conversion);
        return a;
    }
    template <concept_arithmetic U>
    static Array2 convert(const Array2<U>& src) { return Array2(src); }

private:
    void resize_for_assign(std::size_t n)
    {
        if (m_size == n)
            return;
        try {
            Array_base<T>::resize(n);
        } catch (const std::exception& e) {
            throw std::runtime_error(std::string("Array2 assignment failed: ") + e.what());
        }
    }
    template <typename F>
    Array2& zip(const Array2& x, F f)
    {
        std::transform(begin(), end(), x.begin(), begin(), f);
        return *this;
    }
    template <typename U, typename F>
    Array2& zipc(const Array2<U>& x, F f)
    {
        std::transform(begin(), end(), x.begin(), begin(), [&](const T& a, const U& b) { return f(a, static_cast<T>(b)); });
        return *this;
    }
    template <typename F>
    Array2& each(F f)
    {
        std::for_each(begin(), end(), f);
        return *this;
    }
    std::size_t checked_row(int row) const
    {
        const std::size_t r = row < 0 ? m_ysize + row : static_cast<std::size_t>(row);
        if (r >= m_ysize)
            throw std::out_of_range("Row index out of bounds");
        return r;
    }
    std::size_t checked(std::size_t col, std::size_t row) const
    {
        if (col >= m_xsize)
            throw std::out_of_range("X index out of bounds");
        if (row >= m_ysize)
            throw std::out_of_range("Y index out of bounds");
        return row * m_xsize + col;
    }
    std::size_t wrapped(const s_indices& idx) const
    {
        const long x = idx[0] < 0 ? idx[0] + static_cast<long>(m_xsize) : idx[0];
        const long y = idx[1] < 0 ? idx[1] + static_cast<long>(m_ysize) : idx[1];
        return static_cast<std::size_t>(x) + static_cast<std::size_t>(y) * m_xsize;
    }

    std::size_t m_xsize {};
    std::size_t m_ysize {};
};

template <typename T>
Array2<T> operator+(const Array2<T>& x, const Array2<T>& y) { Array2<T> z { x }; return z += y; }
template <typename T>
Array2<T> operator-(const Array2<T>& x, const Array2<T>& y) { Array2<T> z { x }; return z -= y; }
template <typename T>
Array2<T> operator*(const Array2<T>& x, const Array2<T>& y) { Array2<T> z { x }; return z *= y; }
template <typename T>
Array2<T> operator/(const Array2<T>& x, const Array2<T>& y) { Array2<T> z { x }; return z /= y; }
template <typename T>
Array2<T> operator+(const Array2<T>& x, const T& y) { Array2<T> z { x }; return z += y; }
template <typename T>
Array2<T> operator-(const Array2<T>& x, const T& y) { Array2<T> z { x }; return z -= y; }
template <typename T>
Array2<T> operator*(const Array2<T>& x, const T& y) { Array2<T> z { x }; return z *= y; }
template <typename T>
Array2<T> operator/(const Array2<T>& x, const T& y) { Array2<T> z { x }; return z /= y; }

} // namespace smip

template <typename T>
std::ostream& operator<<(std::ostream& o, const std::vector<T>& v)
{
    o << "(";
    for (const auto& x : v)
        o << x << " ";
    return o << ")";
}

template <typename T>
std::ostream& operator<<(std::ostream& o, const smip::Array2<T>& a)
{
    for (std::size_t r = 0; r < a.nrows(); ++r) {
        for (std::size_t c = 0; c < a.ncols(); ++c)
            o << a(c, r) << " ";
        o << "\n";
    }
    return o;
}
