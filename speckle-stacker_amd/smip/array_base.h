// array_base.h -- owning 1-D storage with shared ownership
// (reference interface: include/array_base.h:24-173).
#pragma once

#include <algorithm>
#include <cstddef>
#include <initializer_list>
#include <memory>
#include <stdexcept>

#include "types.h"

namespace smip {

template <typename T>
class Array_base {
public:
    typedef T value_type;
    typedef T* ptr_type;
    typedef T& reference;
    typedef const T& const_reference;
    using iterator = T*;
    using const_iterator = const T*;

    Array_base() = default;
    explicit Array_base(std::size_t n)
        : m_size(n)
    {
        if (n)
            m_data.reset(new T[n]()); // value-initialised, as make_unique<T[]> does
    }
    Array_base(std::size_t n, const T& def)
        : Array_base(n)
    {
        std::fill_n(m_data.get(), n, def);
    }
    Array_base(const Array_base& src)
        : Array_base(src.m_size)
    {
        if (m_size)
            std::copy(src.m_data.get(), src.m_data.get() + m_size, m_data.get());
    }
    template <concept_arithmetic U>
    Array_base(const Array_base<U>& src)
        : Array_base(src.size())
    {
        std::transform(src.begin(), src.end(), begin(), [](const U& x) { return static_cast<T>(x); });
    }
    Array_base(Array_base&& src) noexcept
        : m_size(src.m_size)
        , m_is_reference(src.m_is_reference)
        , m_data(std::move(src.m_data))
    {
        src.m_size = 0;
        src.m_is_reference = false;
    }
    Array_base(std::shared_ptr<T[]> data, std::size_t n)
        : m_size(n)
        , m_data(std::move(data))
    {
    }
    Array_base(std::initializer_list<T> l)
        : Array_base(l.size())
    {
        std::copy(l.begin(), l.end(), begin());
    }
    virtual ~Array_base() = default;

    Array_base& operator=(Array_base&& o) noexcept
    {
        if (this != &o) {
            m_size = o.m_size;
            m_is_reference = o.m_is_reference;
            m_data = std::move(o.m_data);
            o.m_size = 0;
            o.m_is_reference = false;
        }
        return *this;
    }
    template <concept_arithmetic U>
    Array_base& operator=(const Array_base<U>& o)
    {
        Array_base tmp(o);
        *this = std::move(tmp);
        return *this;
    }

    template <typename U = T>
        requires concept_arithmetic_or_valarray_or_complex<U>
    Array_base& operator+=(const U& v)
    {
        for (auto& x : *this)
            x += v;
        return *this;
    }
    template <typename U = T>
        requires concept_arithmetic_or_valarray_or_complex<U>
    Array_base& operator-=(const U& v)
    {
        for (auto& x : *this)
            x -= v;
        return *this;
    }
    template <typename U = T>
        requires concept_arithmetic_or_valarray_or_complex<U>
    Array_base& operator*=(const U& v)
    {
        for (auto& x : *this)
            x *= v;
        return *this;
    }
    template <typename U = T>
        requires concept_arithmetic_or_valarray_or_complex<U>
    Array_base& operator/=(const U& v)
    {
        if (v == U {})
            throw std::runtime_error("Array_base<T>::operator/=(T) : division by zero");
        for (auto& x : *this)
            x /= v;
        return *this;
    }

    iterator begin() { return m_data.get(); }
    iterator end() { return m_data.get() + m_size; }
    const_iterator begin() const { return m_data.get(); }
    const_iterator end() const { return m_data.get() + m_size; }

    std::shared_ptr<T[]>& data() { return m_data; }
    std::shared_ptr<const T[]> data() const { return std::const_pointer_cast<const T[]>(m_data); }

    reference operator[](std::size_t i) { return m_data[i]; }
    const_reference operator[](std::size_t i) const { return m_data[i]; }
    reference at(std::size_t i) { return m_data[i]; }
    const_reference at(std::size_t i) const { return m_data[i]; }

    std::size_t size() const { return m_size; }
    std::size_t typesize() const { return sizeof(T); }

    // installs an external, non-resizable buffer
    void set_at(std::shared_ptr<T[]> data, std::size_t n)
    {
        if (!data && n != 0)
            throw std::invalid_argument("set_at: non-zero size with null data pointer.");
        m_size = n;
        m_data = std::move(data);
        m_is_reference = true;
    }
    void resize(std::size_t n)
    {
        if (m_is_reference)
            throw std::logic_error("Cannot resize a reference-backed Array_base.");
        if (n == m_size)
            return;
        m_data.reset(n ? new T[n]() : nullptr);
        m_size = n;
    }

protected:
    std::size_t m_size { 0 };
    bool m_is_reference { false };
    std::shared_ptr<T[]> m_data {};
};

} // namespace smip
