// bispectrum.h -- Bispectrum<T> on the GPU (reference interface:
// include/bispectrum.h:28-125).
//
// The cells live on the device in the reference layout; accumulation, element
// lookup, arithmetic, file I/O and phase recovery run through libsmip_mi355x.so.
// The Array_base<T> part is a host mirror, materialised on first host access
// (begin(), data(), operator[]).  A const host access downloads the cells if the
// device copy is newer.  A non-const host access hands out raw pointers whose
// writes cannot be tracked, so it pins the mirror as the authoritative copy:
// while pinned, every device operation first uploads the mirror and, if it
// modifies cells, downloads them again afterwards.  host_release() uploads once
// and ends the pinned state; afterwards device operations run without host
// copies.  Code that only accumulates, looks up and recovers phases never
// touches the mirror and never moves the cells over PCIe.
#pragma once

#include <complex>
#include <iostream>
#include <string>
#include <type_traits>
#include <vector>

#include "array2.h"
#include "engine.h"
#include "types.h"

namespace smip {

template <concept_complex T>
class Bispectrum : public Array_base<T> {
    static_assert(std::is_same_v<T, std::complex<float>> || std::is_same_v<T, std::complex<double>>,
        "the device bispectrum holds complex<float> or complex<double>");
    static constexpr int32_t kDtype = std::is_same_v<T, std::complex<float>> ? SMIP_DTYPE_C64 : SMIP_DTYPE_C128;

public:
    using extents = DimVector<std::size_t, 4>;
    using s_indices = DimVector<int, 4>;
    using u_indices = DimVector<std::size_t, 4>;

    Bispectrum() { create({ 0, 0, 0, 0 }); }
    Bispectrum(const extents& dimsizes) { create(dimsizes); }
    Bispectrum(const Bispectrum& o)
    {
        create(o.m_dims);
        o.push();
        detail::check(smip_bispec_assign(m_dev, o.m_dev));
    }
    Bispectrum(Bispectrum&& o) noexcept
        : Array_base<T>(std::move(o))
        , m_dev(o.m_dev)
        , m_dims(o.m_dims)
        , m_info(o.m_info)
        , m_state(o.m_state)
    {
        o.m_dev = nullptr;
    }
    Bispectrum& operator=(Bispectrum&& o) noexcept
    {
        if (this != &o) {
            release();
            Array_base<T>::operator=(std::move(o));
            m_dev = o.m_dev;
            m_dims = o.m_dims;
            m_info = o.m_info;
            m_state = o.m_state;
            o.m_dev = nullptr;
        }
        return *this;
    }
    Bispectrum& operator=(const Bispectrum& o)
    {
        if (this != &o) {
            o.push();
            const std::size_t old_base = base_size();
            detail::check(smip_bispec_assign(m_dev, o.m_dev));
            m_dims = o.m_dims;
            refresh_info();
            resized(old_base);
        }
        return *this;
    }
    ~Bispectrum() override { release(); }

    void print() const
    {
        std::cout << "Object: Bispectrum\nsize: " << totalsize() << "\nsizes: " << sizes() << "\nbase_sizes: "
                  << base_sizes() << "\nmin indices: " << min_indices() << "\nmax indices: " << max_indices()
                  << "\nsize of datatype: " << sizeof(T) << " bytes\nreal memory size: "
                  << sizeof(T) * base_size() / 1024 / 1024 << " MB" << std::endl;
    }
    void write_to_file(const std::string& filename) const
    {
        push();
        detail::check(smip_bispec_write_file(m_dev, filename.c_str()));
    }
    void read_from_file(const std::string& filename)
    {
        if (this->m_is_reference)
            throw std::logic_error("Cannot resize a reference-backed Array_base.");
        const std::size_t old_base = base_size();
        detail::check(smip_bispec_read_file(m_dev, filename.c_str()));
        refresh_info();
        m_dims = { m_info.dims[0], m_info.dims[1], m_info.dims[2], m_info.dims[3] };
        resized(old_base);
    }

    [[nodiscard]] T get_element(s_indices indices) const
    {
        push();
        const int32_t idx[4] = { indices[0], indices[1], indices[2], indices[3] };
        double v[2];
        detail::check(smip_bispec_get_elements(m_dev, idx, 1, v));
        return T(static_cast<typename T::value_type>(v[0]), static_cast<typename T::value_type>(v[1]));
    }
    T operator()(s_indices indices) const { return get_element(indices); }
    void get_elements(const std::vector<s_indices>& idx_list, std::vector<T>& results) const
    {
        push();
        std::vector<int32_t> idx(4 * idx_list.size());
        for (std::size_t i = 0; i < idx_list.size(); ++i)
            for (int k = 0; k < 4; ++k)
                idx[4 * i + k] = idx_list[i][k];
        std::vector<double> v(2 * idx_list.size());
        detail::check(smip_bispec_get_elements(m_dev, idx.data(), idx_list.size(), v.data()));
        results.clear();
        results.reserve(idx_list.size());
        for (std::size_t i = 0; i < idx_list.size(); ++i)
            results.emplace_back(static_cast<typename T::value_type>(v[2 * i]),
                static_cast<typename T::value_type>(v[2 * i + 1]));
    }
    void put_element(s_indices indices, const T& value)
    {
        push();
        const int32_t idx[4] = { indices[0], indices[1], indices[2], indices[3] };
        const double v[2] = { static_cast<double>(value.real()), static_cast<double>(value.imag()) };
        detail::check(smip_bispec_put_element(m_dev, idx, v));
        device_changed();
    }

    Bispectrum& operator+=(const Bispectrum& x) { return array_op(x, SMIP_OP_ADD); }
    Bispectrum& operator-=(const Bispectrum& x) { return array_op(x, SMIP_OP_SUB); }
    Bispectrum& operator*=(const Bispectrum& x) { return array_op(x, SMIP_OP_MUL); }
    Bispectrum& operator/=(const Bispectrum& x) { return array_op(x, SMIP_OP_DIV); }
    template <typename U>
        requires concept_arithmetic_or_valarray_or_complex<U>
    Bispectrum& operator+=(const U& v) { return scalar_op(v, SMIP_OP_ADD); }
    template <typename U>
        requires concept_arithmetic_or_valarray_or_complex<U>
    Bispectrum& operator-=(const U& v) { return scalar_op(v, SMIP_OP_SUB); }
    template <typename U>
        requires concept_arithmetic_or_valarray_or_complex<U>
    Bispectrum& operator*=(const U& v) { return scalar_op(v, SMIP_OP_MUL); }
    template <typename U>
        requires concept_arithmetic_or_valarray_or_complex<U>
    Bispectrum& operator/=(const U& v) { return scalar_op(v, SMIP_OP_DIV); }

    [[nodiscard]] s_indices calc_indices(std::size_t addr) const
    {
        int32_t idx[4];
        detail::check(smip_bispec_calc_indices(m_dev, addr, idx));
        return { idx[0], idx[1], idx[2], idx[3] };
    }
    [[nodiscard]] std::size_t calc_offset(s_indices indices) const noexcept
    {
        const int32_t idx[4] = { indices[0], indices[1], indices[2], indices[3] };
        return smip_bispec_calc_offset(m_dev, idx);
    }

    // Bispectrum<T>::accumulate_from_fft (include/bispectrum.h:356-409)
    template <concept_complex U>
    void accumulate_from_fft(const Array2<U>& fft)
    {
        push();
        if constexpr (std::is_same_v<U, std::complex<double>>) {
            detail::check(smip_bispec_accumulate_from_fft(m_dev, reinterpret_cast<const double*>(fft.begin()),
                fft.ncols(), fft.nrows()));
        } else {
            Array2<std::complex<double>> f(fft.ncols(), fft.nrows());
            std::transform(fft.begin(), fft.end(), f.begin(), [](const U& z) { return std::complex<double>(z); });
            detail::check(smip_bispec_accumulate_from_fft(m_dev, reinterpret_cast<const double*>(f.begin()),
                f.ncols(), f.nrows()));
        }
        device_changed();
    }

    [[nodiscard]] std::size_t size() const noexcept { return m_info.base_size; }
    [[nodiscard]] extents sizes() const noexcept { return ext(m_info.sizes); }
    [[nodiscard]] extents base_sizes() const noexcept { return ext(m_info.base_sizes); }
    [[nodiscard]] std::size_t base_size() const noexcept { return m_info.base_size; }
    [[nodiscard]] std::size_t totalsize() const noexcept { return m_info.totalsize; }
    [[nodiscard]] s_indices min_indices() const noexcept { return idx4(m_info.min_indices); }
    [[nodiscard]] s_indices max_indices() const noexcept { return idx4(m_info.max_indices); }
    Range<DimVector<int, 4>> range() const { return Range<DimVector<int, 4>>(min_indices(), max_indices()); }
    Range<DimVector<int, 4>> true_range() const
    {
        s_indices hi = max_indices();
        hi[0] = 0;
        hi[2] = 0;
        return Range<DimVector<int, 4>>(min_indices(), hi);
    }

    // host mirror access (Array_base interface)
    T* begin() { return mirror(true); }
    T* end() { return mirror(true) + base_size(); }
    const T* begin() const { return mirror(false); }
    const T* end() const { return mirror(false) + base_size(); }
    std::shared_ptr<T[]>& data()
    {
        mirror(true);
        return this->m_data;
    }
    std::shared_ptr<const T[]> data() const
    {
        mirror(false);
        return std::const_pointer_cast<const T[]>(this->m_data);
    }
    T& operator[](std::size_t i) { return mirror(true)[i]; }
    const T& operator[](std::size_t i) const { return mirror(false)[i]; }

    // the device object, for callers that mix this surface with the C ABI
    smip_bispec* device_handle() const
    {
        push();
        return m_dev;
    }
    // call after writing cells through device_handle()
    void device_changed() const
    {
        if (m_state == kHostPinned) {
            if (base_size())
                detail::check(smip_bispec_download(m_dev, this->m_data.get()));
        } else if (m_state != kNoMirror) {
            m_state = kDeviceNewer;
        }
    }
    // end the pinned state of the host mirror (see above)
    void host_release()
    {
        if (m_state == kHostPinned) {
            if (base_size())
                detail::check(smip_bispec_upload(m_dev, this->m_data.get()));
            m_state = kSynced;
        }
    }

private:
    enum : int { kSynced = 0, kHostPinned = 1, kDeviceNewer = 2, kNoMirror = 3 };

    void create(const extents& dims)
    {
        const uint64_t d[4] = { dims[0], dims[1], dims[2], dims[3] };
        detail::check(smip_bispec_create(detail::context(), d, kDtype, &m_dev));
        m_dims = dims;
        refresh_info();
        m_state = kNoMirror;
    }
    void release()
    {
        if (m_dev)
            smip_bispec_destroy(m_dev);
        m_dev = nullptr;
    }
    void refresh_info()
    {
        detail::check(smip_bispec_info_get(m_dev, &m_info));
        this->m_size = m_info.base_size;
    }
    // after a device operation that may have changed the extents
    void resized(std::size_t old_base)
    {
        if (base_size() != old_base || !this->m_data) {
            this->m_data.reset();
            m_state = kNoMirror;
        } else {
            device_changed();
        }
    }
    T* mirror(bool writes) const
    {
        auto* self = const_cast<Bispectrum*>(this);
        if (!self->m_data && base_size())
            self->m_data.reset(new T[base_size()]());
        if (m_state == kDeviceNewer || m_state == kNoMirror) {
            if (base_size())
                detail::check(smip_bispec_download(m_dev, self->m_data.get()));
            m_state = kSynced;
        }
        if (writes)
            m_state = kHostPinned;
        return self->m_data.get();
    }
    void push() const
    {
        if (m_state == kHostPinned && base_size())
            detail::check(smip_bispec_upload(m_dev, this->m_data.get()));
    }
    Bispectrum& array_op(const Bispectrum& x, int op)
    {
        x.push();
        push();
        detail::check(smip_bispec_array_op(m_dev, x.m_dev, op));
        device_changed();
        return *this;
    }
    template <typename U>
    Bispectrum& scalar_op(const U& v, int op)
    {
        push();
        const std::complex<double> z(v);
        detail::check(smip_bispec_scalar_op(m_dev, op, z.real(), z.imag()));
        device_changed();
        return *this;
    }
    static extents ext(const uint64_t* v) { return { v[0], v[1], v[2], v[3] }; }
    static s_indices idx4(const int32_t* v) { return { v[0], v[1], v[2], v[3] }; }

    smip_bispec* m_dev = nullptr;
    extents m_dims {};
    smip_bispec_info m_info {};
    mutable int m_state = kNoMirror;
};

} // namespace smip
