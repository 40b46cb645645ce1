// dropin_test.cpp -- tests of the drop-in C++ surface (speckle-stacker_amd/smip),
// modelled on the reference's own test programs: tests/array2_test.cpp,
// tests/range_test.cpp, tests/bispectrum_test.cpp, tests/correl_test.cpp.
//
//   smip_dropin_test cpu | gpu | <case name> ...
//
// "cpu" cases touch no device; "gpu" cases create a device context.  Exit status
// is the number of failed checks (capped at 100), 0 when all pass.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <functional>
#include <iostream>
#include <map>
#include <numeric>
#include <random>
#include <ranges>
#include <string>
#include <vector>

#include "smip/smip.h"

using namespace smip;
using cd = std::complex<double>;
using cf = std::complex<float>;

namespace {

int g_fail = 0, g_pass = 0;

#define CHECK(expr)                                                                                        \
    do {                                                                                                   \
        if (expr) {                                                                                        \
            ++g_pass;                                                                                      \
        } else {                                                                                           \
            ++g_fail;                                                                                      \
            std::cerr << "  FAIL " << __FILE__ << ":" << __LINE__ << ": " #expr << std::endl;              \
        }                                                                                                  \
    } while (0)

#define CHECK_THROWS(stmt, E)                                                                              \
    do {                                                                                                   \
        bool thrown_ = false;                                                                              \
        try {                                                                                              \
            stmt;                                                                                          \
        } catch (const E&) {                                                                               \
            thrown_ = true;                                                                                \
        } catch (...) {                                                                                    \
        }                                                                                                  \
        if (thrown_) {                                                                                     \
            ++g_pass;                                                                                      \
        } else {                                                                                           \
            ++g_fail;                                                                                      \
            std::cerr << "  FAIL " << __FILE__ << ":" << __LINE__ << ": " #stmt " did not throw " #E       \
                      << std::endl;                                                                        \
        }                                                                                                  \
    } while (0)

struct Case {
    bool gpu;
    std::function<void()> fn;
};
std::map<std::string, Case>& registry()
{
    static std::map<std::string, Case> r;
    return r;
}
struct Reg {
    Reg(const char* name, bool gpu, std::function<void()> fn) { registry()[name] = { gpu, std::move(fn) }; }
};
#define CPU_CASE(name)                                                                                     \
    void name();                                                                                           \
    Reg reg_##name(#name, false, name);                                                                    \
    void name()
#define GPU_CASE(name)                                                                                     \
    void name();                                                                                           \
    Reg reg_##name(#name, true, name);                                                                     \
    void name()

// ---------------------------------------------------------------- Array2 (cpu)
template <typename T>
void array2_arith_typed()
{
    Array2<T> a = { { T(1), T(2) }, { T(3), T(4) } };
    Array2<T> b = { { T(5), T(6) }, { T(7), T(8) } };
    auto c = a + b;
    CHECK(c.ncols() == 2 && c.nrows() == 2);
    CHECK(c(0, 0) == T(6) && c(1, 0) == T(8) && c(0, 1) == T(10) && c(1, 1) == T(12));
    c = b - a;
    CHECK(c(0, 0) == T(4) && c(1, 1) == T(4));
    c = a * b;
    CHECK(c(0, 0) == T(5) && c(1, 0) == T(12) && c(0, 1) == T(21) && c(1, 1) == T(32));
    c = b / a;
    CHECK(c(0, 0) == T(5) && c(1, 1) == T(2));
    c = a * T(2);
    CHECK(c(1, 1) == T(8));
    c = a + T(5);
    CHECK(c(0, 0) == T(6) && c(1, 1) == T(9));
    c = b - T(5);
    CHECK(c(0, 0) == T(0) && c(1, 1) == T(3));
    Array2<T> d = { { T(2), T(4) }, { T(6), T(8) } };
    c = d / T(2);
    CHECK(c(0, 0) == T(1) && c(1, 0) == T(2) && c(0, 1) == T(3) && c(1, 1) == T(4));
}

CPU_CASE(array2_construction)
{
    Array2<double> arr(3, 2);
    CHECK(arr.xsize() == 3 && arr.ysize() == 2 && arr.size() == 6);
    CHECK(arr(0, 0) == 0.0);
    Array2<int> z(0, 0);
    CHECK(z.size() == 0);
    CHECK_THROWS(z(0, 0), std::out_of_range);
}

CPU_CASE(array2_bounds)
{
    Array2<int> arr(2, 2);
    arr(1, 1) = 5;
    CHECK(arr(1, 1) == 5);
    CHECK_THROWS(arr(2, 0), std::out_of_range);
    CHECK_THROWS(arr(0, 2), std::out_of_range);
    CHECK_THROWS(arr(100, 100), std::out_of_range);
}

CPU_CASE(array2_copy_move_resize)
{
    Array2<int> a(2, 2);
    a(0, 0) = 10;
    a(1, 1) = 20;
    Array2<int> b(a);
    CHECK(b(0, 0) == 10 && b(1, 1) == 20);
    Array2<int> c;
    c = a;
    CHECK(c(0, 0) == 10 && c(1, 1) == 20);
    a(0, 0) = 11;
    CHECK(b(0, 0) == 10); // deep copy
    Array2<int> m(std::move(b));
    CHECK(m(1, 1) == 20);
    Array2<int> r(2, 2);
    r = Array2<int>(3, 3, 99);
    CHECK(r.xsize() == 3 && r.ysize() == 3 && r(0, 0) == 99 && r(1, 1) == 99);
    r = Array2<int>(3, 3);
    CHECK(r(2, 2) == 0);
}

CPU_CASE(array2_reference_assignment)
{
    Array2<int> a(5, 5);
    Array2<int> b;
    b.set_at(a.data(), a.size());
    Array2<int> c(5, 4);
    CHECK_THROWS(b = c, std::runtime_error);
}

CPU_CASE(array2_fill_index)
{
    Array2<double> a(3, 3);
    std::fill(a.begin(), a.end(), 99.0);
    CHECK(a(0, 0) == 99.0 && a(2, 2) == 99.0);
    Array2<int> b(3, 3);
    b[0][0] = 1;
    b[1][1] = 2;
    b[2][2] = 3;
    CHECK(b(0, 0) == 1 && b(1, 1) == 2 && b(2, 2) == 3);
    // signed FFT-order access wraps
    b.at({ -1, -1 }) = 7;
    CHECK(b(2, 2) == 7);
}

CPU_CASE(array2_arithmetic)
{
    array2_arith_typed<int>();
    array2_arith_typed<float>();
    array2_arith_typed<double>();
    array2_arith_typed<cd>();
}

CPU_CASE(array2_convert_import)
{
    Array2<int> a = { { 1, 2 }, { 3, 4 } };
    auto f = Array2<float>::convert<int>(a);
    CHECK(f.ncols() == 2 && f(0, 0) == 1.0f && f(1, 0) == 2.0f && f(0, 1) == 3.0f && f(1, 1) == 4.0f);
    Array2<double> d;
    d.import(a);
    CHECK(d.ncols() == 2 && d.nrows() == 2 && d(1, 1) == 4.0);
}

CPU_CASE(array2_shift)
{
    Array2<int> arr { { 1, 2, 3 }, { 4, 5, 6 }, { 7, 8, 9 } };
    auto s = arr.shifted({ -1, 1 });
    CHECK(s.get_row(1) == std::vector<int>({ 2, 3, 0 }));
    CHECK(s.get_col(1) == std::vector<int>({ 0, 3, 6 }));
}

CPU_CASE(array2_ranges)
{
    CHECK(std::ranges::range<Array_base<int>>);
    CHECK(std::ranges::range<Array2<int>>);
    Array2<int> arr2 { { 1, 2, 3 }, { 4, 5, 6 }, { 7, 8, 9 } };
    int s = 0;
    for (int v : std::views::all(arr2) | std::views::transform([](int x) { return x * x; }))
        s += v;
    CHECK(s == 285);
}

CPU_CASE(range_contains)
{
    Range<int> ir { -4, 5 };
    CHECK(ir.extent() == 10);
    Range<DimVector<int, 2>> vr { { -3, 3 }, { 5, 10 } };
    CHECK(vr.contains(DimVector<int, 2> { 0, 6 }));
    CHECK(!vr.contains(DimVector<int, 2> { -4, 6 }));
    CHECK(!vr.contains(DimVector<int, 2> { 0, 11 }));
    CHECK(!vr.contains(DimVector<int, 2> { -4, 11 }));
    auto lo = DimVector<int>::merge(DimVector<int, 2> { -1, -2 }, DimVector<int, 2> { -2, -1 });
    CHECK((lo == DimVector<int, 4> { -1, -2, -2, -1 }));
}

CPU_CASE(next_reco_index_spiral)
{
    // NextRecoIndex visits every integer point of radius <= r0 before leaving it
    double r = 0., phi = 0.;
    int i = 0, j = 0;
    std::vector<std::pair<int, int>> seen;
    while (r <= 6.0) {
        NextRecoIndex(r, phi, i, j);
        seen.emplace_back(i, j);
    }
    int inside = 0;
    for (int x = -5; x <= 5; ++x)
        for (int y = -5; y <= 5; ++y)
            if (x * x + y * y <= 25)
                inside += std::find(seen.begin(), seen.end(), std::make_pair(x, y)) != seen.end();
    int want = 0;
    for (int x = -5; x <= 5; ++x)
        for (int y = -5; y <= 5; ++y)
            want += x * x + y * y <= 25;
    CHECK(inside == want);
}

CPU_CASE(window_functions)
{
    Hann<double> h(16, 16, 16.);
    CHECK(std::fabs(h.at({ 0, 0 }) - 1.0) < 1e-15);
    CHECK(h.at({ 4, 0 }) > 0.49 && h.at({ 4, 0 }) < 0.51);
    CHECK(std::fabs(h.at({ 3, -5 }) - h.at({ -3, 5 })) < 1e-15);
    Hamming<float> hm(8, 8, 8.);
    CHECK(std::fabs(hm.at({ 0, 0 }) - 1.0f) < 1e-6f);
    CHECK_THROWS((Hann<double>(8, 8, 0.)), std::domain_error);
    CHECK_THROWS((Hann<double>(8, 8, -1.)), std::domain_error);
}

// ----------------------------------------------------------- Bispectrum (gpu)
template <typename T>
void bispec_typed()
{
    // complex conjugation symmetry
    {
        Bispectrum<T> B({ 4, 4, 4, 4 });
        B.put_element({ -1, 2, -1, 1 }, T(1.0, 2.0));
        CHECK(B.get_element({ 1, -2, 1, -1 }) == std::conj(B.get_element({ -1, 2, -1, 1 })));
    }
    // multiple get
    {
        Bispectrum<T> b({ 2, 2, 2, 2 });
        std::vector<typename Bispectrum<T>::s_indices> idx = { { 0, 0, 0, 0 }, { -1, 1, -1, 1 } };
        b.put_element(idx[0], T(1.5, 1.0));
        b.put_element(idx[1], T(2.5, -1.0));
        std::vector<T> res;
        b.get_elements(idx, res);
        CHECK(res.size() == 2 && res[0] == T(1.5, 1.0) && res[1] == T(2.5, -1.0));
    }
    // fill through the host mirror
    {
        Bispectrum<T> b({ 3, 4, 5, 6 });
        const T val(42.0, -1.);
        std::fill(b.begin(), b.end(), val);
        CHECK(b.get_element({ 0, 0, 0, 0 }) == val);
        CHECK(b[b.size() - 1] == val);
        b += T(1.0, 0.0); // device op while the mirror is pinned
        CHECK(b[0] == T(43.0, -1.0));
        b.host_release();
        b *= 2.0;
        CHECK(b.get_element({ 0, 0, 0, 0 }) == T(86.0, -2.0));
        CHECK(b[1] == T(86.0, -2.0));
    }
    // file round trip
    {
        Bispectrum<T> b({ 2, 2, 2, 2 });
        auto dims = b.sizes();
        b.put_element({ -1, 1, -1, 1 }, T(99.9, -1.));
        const std::string fn = "dropin_bispectrum_io.dat";
        b.write_to_file(fn);
        Bispectrum<T> b2;
        b2.read_from_file(fn);
        CHECK(b2.sizes() == dims);
        CHECK(b2.get_element({ -1, 1, -1, 1 }) == T(99.9, -1.));
        std::remove(fn.c_str());
    }
}

GPU_CASE(bispectrum_construction)
{
    Bispectrum<cd> bs(Bispectrum<cd>::extents { 2, 3, 4, 5 });
    CHECK(bs.sizes()[0] == 3 && bs.sizes()[1] == 3 && bs.sizes()[2] == 5 && bs.sizes()[3] == 5);
    CHECK(bs.totalsize() == 3 * 3 * 5 * 5);
    CHECK(bs.size() > 0);
    CHECK(bs.get_element({ 0, 0, 0, 0 }) == cd {});
    Bispectrum<cd> zero;
    CHECK(zero.size() == 0);
    CHECK_THROWS((void)zero.get_element({ 0, 0, 0, 0 }), std::out_of_range);
    CHECK_THROWS((void)bs.get_element({ 0, 9, 0, 0 }), std::out_of_range);
}

GPU_CASE(bispectrum_element_access)
{
    Bispectrum<cd> b(Bispectrum<cd>::extents { 3, 4, 5, 6 });
    const Bispectrum<cd>::s_indices idx { -1, 1, -1, 1 };
    b.put_element(idx, cd(42.0, 1.0));
    CHECK(b.get_element(idx) == cd(42.0, 1.0));
    int bad = 0;
    for (auto it = b.begin(); it != b.end(); ++it) {
        const auto k = std::distance(b.begin(), it);
        auto ind = b.calc_indices(k);
        CHECK(b.calc_offset(ind) == static_cast<std::size_t>(k));
        *it = cd(static_cast<double>(k), 0.);
        bad += b.get_element(ind) != cd(static_cast<double>(k), 0.);
    }
    CHECK(bad == 0);
}

GPU_CASE(bispectrum_arithmetic)
{
    Bispectrum<cd> a(Bispectrum<cd>::extents { 2, 2, 2, 2 }), b(Bispectrum<cd>::extents { 2, 2, 2, 2 });
    CHECK(a.sizes() == b.sizes());
    a.put_element({ 0, 0, 0, 0 }, cd(2.0));
    b.put_element({ 0, 0, 0, 0 }, cd(3.0));
    a += b;
    CHECK(a.get_element({ 0, 0, 0, 0 }) == cd(5.0));
    a -= b;
    CHECK(a.get_element({ 0, 0, 0, 0 }) == cd(2.0));
    a *= b;
    CHECK(a.get_element({ 0, 0, 0, 0 }) == cd(6.0));
    a /= b;
    CHECK(a.get_element({ 0, 0, 0, 0 }) == cd(2.0));
    Bispectrum<cd> c(a);
    CHECK(c.get_element({ 0, 0, 0, 0 }) == cd(2.0));
    Bispectrum<cd> d;
    d = a;
    CHECK(d.sizes() == a.sizes() && d.get_element({ 0, 0, 0, 0 }) == cd(2.0));
}

GPU_CASE(bispectrum_typed)
{
    bispec_typed<cf>();
    bispec_typed<cd>();
}

// naive 2-D DFT of a real image (test input only)
Array2<cd> dft2(const Array2<double>& img)
{
    const int X = static_cast<int>(img.ncols()), Y = static_cast<int>(img.nrows());
    Array2<cd> out(X, Y);
    for (int ky = 0; ky < Y; ++ky)
        for (int kx = 0; kx < X; ++kx) {
            cd s {};
            for (int y = 0; y < Y; ++y)
                for (int x = 0; x < X; ++x) {
                    const double a = -2.0 * M_PI * (double(kx * x % X) / X + double(ky * y % Y) / Y);
                    s += img(x, y) * cd(std::cos(a), std::sin(a));
                }
            out(kx, ky) = s;
        }
    return out;
}

Array2<double> random_image(int X, int Y, unsigned seed)
{
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    Array2<double> img(X, Y);
    for (auto& v : img)
        v = u(gen);
    return img;
}

GPU_CASE(bispectrum_accumulate_from_fft)
{
    // one frame: every stored cell is F(u) F(v) conj(F(u+v)), bit for bit
    const int N = 16;
    auto F = dft2(random_image(N, N, 7));
    Bispectrum<cd> B(Bispectrum<cd>::extents { N, N, 6, 6 });
    B.accumulate_from_fft(F);
    int checked = 0, bad = 0;
    for (int ux = -3; ux <= 0; ++ux)
        for (int uy = -3; uy <= 3; ++uy)
            for (int vx = -3; vx <= 0; ++vx)
                for (int vy = -3; vy <= 3; ++vy) {
                    if (ux + vx < -N / 2 || uy + vy < -N / 2 || uy + vy >= N / 2)
                        continue;
                    cd t = F.at({ ux, uy });
                    t *= F.at({ vx, vy });
                    t *= std::conj(F.at({ ux + vx, uy + vy }));
                    ++checked;
                    bad += B.get_element({ ux, uy, vx, vy }) != t;
                }
    CHECK(checked > 0);
    CHECK(bad == 0);
    // accumulating twice doubles every cell exactly
    Bispectrum<cd> B2(B);
    B2.accumulate_from_fft(F);
    B *= 2.0;
    std::vector<cd> a(B.begin(), B.end()), b(B2.begin(), B2.end());
    CHECK(a == b);
}

GPU_CASE(phase_recovery_stepwise)
{
    // The reference's spiral (include/phasereco.h:33-73) driven step by step
    // through calc_phase equals reconstruct_phases in its one-visit-at-a-time
    // schedule bit for bit, and the windowed schedule (pre-window contributions
    // summed in bulk, so another FP64 summation order) with identical flags
    // and phases within 1e-10.
    const int N = 32;
    const double radius = 12.0;
    Bispectrum<cd> B(Bispectrum<cd>::extents { N, N, 24, 24 });
    for (unsigned f = 0; f < 3; ++f)
        B.accumulate_from_fft(dft2(random_image(N, N, 100 + f)));
    PhaseMap pm_dev, pm_seq;
    auto ph_dev = reconstruct_phases<cd>(B, N, N, radius, &pm_dev);
    setenv("SMIP_RECO_SEQUENTIAL", "1", 1);
    auto ph_seq = reconstruct_phases<cd>(B, N, N, radius, &pm_seq);
    unsetenv("SMIP_RECO_SEQUENTIAL");

    Array2<cd> ph(N, N);
    PhaseMap pm(N, N);
    const cd init(1., 0.);
    for (auto w : { DimVector<int, 2> { 0, 0 }, DimVector<int, 2> { 1, 0 }, DimVector<int, 2> { 0, 1 } }) {
        ph.at(w) = init;
        pm.at(w) = { true, 1.0 };
    }
    for (auto w : { DimVector<int, 2> { -1, 0 }, DimVector<int, 2> { 0, -1 } }) {
        ph.at(w) = std::conj(init);
        pm.at(w) = { true, 1.0 };
    }
    double r = 0., phi = 0.;
    DimVector<int, 2> w {};
    int steps = 0;
    while (r <= radius) {
        NextRecoIndex(r, phi, w[0], w[1]);
        if (pm.range().contains(w) && !pm.at(w).flag) {
            calc_phase(B, ph, pm, w);
            ++steps;
        }
    }
    CHECK(steps > 100);
    int bad_ph = 0, bad_fl = 0, bad_co = 0, flagged = 0, bad_fl_w = 0;
    double dph = 0., dco = 0.;
    for (int y = 0; y < N; ++y)
        for (int x = 0; x < N; ++x) {
            const cd a = ph(x, y), b = ph_seq(x, y);
            bad_ph += std::memcmp(&a, &b, sizeof(cd)) != 0;
            bad_fl += pm(x, y).flag != pm_seq(x, y).flag;
            bad_co += std::memcmp(&pm(x, y).consistency, &pm_seq(x, y).consistency, sizeof(double)) != 0;
            flagged += pm(x, y).flag;
            bad_fl_w += pm(x, y).flag != pm_dev(x, y).flag;
            dph = std::max(dph, std::abs(a - ph_dev(x, y)));
            dco = std::max(dco, std::fabs(pm(x, y).consistency - pm_dev(x, y).consistency));
        }
    std::cout << "  steps " << steps << ", flagged " << flagged << ", windowed max |dphase| " << dph
              << ", max |dconsistency| " << dco << std::endl;
    CHECK(bad_ph == 0);
    CHECK(bad_fl == 0);
    CHECK(bad_co == 0);
    CHECK(bad_fl_w == 0);
    CHECK(dph < 1e-10);
    CHECK(dco < 1e-10);
}

double gauss(double x, double y, double cx, double cy, double s)
{
    return std::exp(-0.5 * ((x - cx) * (x - cx) + (y - cy) * (y - cy)) / (s * s));
}

GPU_CASE(cross_correlation_shifts)
{
    // shifted Gaussians are found at their shift (tests/correl_test.cpp:8-58)
    const int X = 128, Y = 128;
    const double sigma = 5.;
    Array2<double> g(X, Y);
    for (int r = 0; r < Y; ++r)
        for (int c = 0; c < X; ++c)
            g(c, r) = gauss(c, r, 0.5 * X, 0.5 * Y, sigma);
    std::mt19937 gen(1234);
    std::uniform_int_distribution<int> d(g.min_sindices().max() + int(sigma), g.max_sindices().min() - int(sigma) - 1);
    CrossCorrelation<double> xc(g);
    Array2<double> sum(X, Y, 0.);
    int exact = 0;
    const int trials = 40;
    for (int t = 0; t < trials; ++t) {
        const DimVector<int, 2> s { d(gen), d(gen) };
        auto sg = g.shifted(s);
        const DimVector<int, 2> m = xc(sg);
        CHECK(std::abs(m[0] - s[0]) <= 1 && std::abs(m[1] - s[1]) <= 1);
        exact += m[0] == s[0] && m[1] == s[1];
        sum += sg.shifted(-m);
    }
    CHECK(exact == trials);
    sum /= double(trials);
    double ss = 0.;
    for (std::size_t k = 0; k < g.size(); ++k)
        ss += (sum.begin()[k] - g.begin()[k]) * (sum.begin()[k] - g.begin()[k]);
    CHECK(ss / g.size() < 1e-3);
    // static form, correlation array and size mismatch
    const auto s2 = CrossCorrelation<double>::get_displacement(g, g.shifted({ 3, -2 }));
    CHECK(s2[0] == 3 && s2[1] == -2);
    xc.correlate(g.shifted({ 3, -2 }));
    const auto& ca = xc.get_correlation_array();
    CHECK(ca.ncols() == X && ca.nrows() == Y);
    CHECK(*std::max_element(ca.begin(), ca.end()) == ca.at({ 3, -2 }));
    CHECK(xc.get_displacement()[0] == 3);
    CHECK_THROWS(xc.correlate(Array2<double>(64, 64)), std::invalid_argument);
}

} // namespace

int main(int argc, char** argv)
{
    std::vector<std::string> names;
    for (int a = 1; a < argc; ++a) {
        const std::string s = argv[a];
        if (s == "cpu" || s == "gpu" || s == "all") {
            for (auto& [n, c] : registry())
                if (s == "all" || c.gpu == (s == "gpu"))
                    names.push_back(n);
        } else if (s == "list") {
            for (auto& [n, c] : registry())
                std::cout << n << (c.gpu ? " gpu" : " cpu") << "\n";
            return 0;
        } else {
            names.push_back(s);
        }
    }
    if (names.empty()) {
        std::cerr << "usage: smip_dropin_test cpu|gpu|all|list|<case>...\n";
        return 2;
    }
    for (const auto& n : names) {
        auto it = registry().find(n);
        if (it == registry().end()) {
            std::cerr << "unknown case " << n << "\n";
            return 2;
        }
        const int f0 = g_fail;
        std::cout << "[case] " << n << std::endl;
        try {
            it->second.fn();
        } catch (const std::exception& e) {
            ++g_fail;
            std::cerr << "  FAIL uncaught exception: " << e.what() << std::endl;
        }
        std::cout << "[" << (g_fail == f0 ? "ok" : "FAILED") << "] " << n << std::endl;
    }
    std::cout << g_pass << " checks passed, " << g_fail << " failed" << std::endl;
    return std::min(g_fail, 100);
}
