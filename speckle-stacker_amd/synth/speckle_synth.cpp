// speckle_synth.cpp -- deterministic synthetic speckle frame stacks (host only).
//
// Each frame is an object convolved with a fresh speckle PSF
//   PSF = |FFT2(P exp(i phi))|^2 / sum,
// P a circular pupil of diameter D_p pixels centred in the N x N grid and phi a
// Kolmogorov phase screen (white complex Gaussian noise filtered by the
// k^(-11/3) phase power spectrum 0.023 r0^(-5/3) f^(-11/3), scaled to D/r0),
// scaled to the photons per frame, Poisson-sampled, plus Gaussian read noise,
// rounded and clipped for integer sample types.
//
// Random numbers come from counter-based streams keyed by (seed, frame,
// purpose), so frame f is identical whatever the thread count or the range
// requested.  Gaussian deviates use Box-Muller; Poisson deviates use inversion
// for small means and the PTRS transformed-rejection method of W. Hoermann
// ("The transformed rejection method for generating Poisson random variables",
// Insurance: Mathematics and Economics 12, 1993) otherwise.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

extern "C" {

typedef struct smip_synth_params {
    uint64_t seed;
    int32_t ncols, nrows;    // powers of two
    int32_t dtype;           // 1 u8, 2 u16, 3 f32 (SMIP_DTYPE_*)
    int32_t object;          // 0 binary, 1 triple, 2 extended
    double pupil_diameter;   // pixels
    double d_over_r0;        // seeing strength
    double photons;          // per frame
    double read_noise;       // DN, Gaussian sigma
    double clip_max;         // largest sample value of integer types
} smip_synth_params;

int smip_synth_frames(const smip_synth_params* p, uint64_t first, uint64_t count, void* out, int32_t threads);
int smip_synth_object(const smip_synth_params* p, double* out);
}

namespace {

using cd = std::complex<double>;
constexpr double kPi = 3.14159265358979323846;

// --- counter-based random numbers -------------------------------------------
inline uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

struct Stream {
    uint64_t key, ctr = 0;
    Stream(uint64_t seed, uint64_t a, uint64_t b)
        : key(mix64(seed ^ mix64(a * 0x632be59bd9b4e019ULL + mix64(b + 0x1234567ULL))))
    {
    }
    uint64_t next() { return mix64(key + 0x9e3779b97f4a7c15ULL * (++ctr)); }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); } // (0,1)
    double gauss()
    {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
    int64_t poisson(double lam)
    {
        if (!(lam > 0))
            return 0;
        if (lam < 30.0) { // inversion
            double u = uniform(), pk = std::exp(-lam), cdf = pk;
            int64_t k = 0;
            while (u > cdf && k < 1000) {
                ++k;
                pk *= lam / (double)k;
                cdf += pk;
            }
            return k;
        }
        const double slam = std::sqrt(lam), loglam = std::log(lam);
        const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
        const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
        for (;;) {
            const double U = uniform() - 0.5, V = uniform();
            const double us = 0.5 - std::fabs(U);
            const double k = std::floor((2.0 * a / us + b) * U + lam + 0.43);
            if (us >= 0.07 && V <= vr)
                return (int64_t)k;
            if (k < 0 || (us < 0.013 && V > us))
                continue;
            if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b)
                <= -lam + k * loglam - std::lgamma(k + 1.0))
                return (int64_t)k;
        }
    }
};

enum Purpose : uint64_t { kScreen = 1, kNoise = 2, kObject = 3 };

// --- radix-2 FFT (FP64), unnormalised, sign -1 forward --------------------
// Twiddles of every stage of a length-n transform: the stage of span len holds
// exp(sign 2 pi i m / len), m < len/2, at offset len/2 - 1.
std::vector<cd> twiddles(int n, int sign)
{
    std::vector<cd> t(std::max(n - 1, 1));
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1;
        for (int m = 0; m < half; ++m) {
            const double ang = 2.0 * kPi * m / len;
            t[(size_t)(half - 1 + m)] = cd(std::cos(ang), sign * std::sin(ang));
        }
    }
    return t;
}

// in-place transform of contiguous w[0, n): bit reversal, then the radix-2 stages
void fft_core(cd* w, int n, const cd* tw)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1)
            j ^= bit;
        j ^= bit;
        if (i < j)
            std::swap(w[i], w[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1;
        for (int m = 0; m < half; ++m) {
            const cd t = tw[half - 1 + m];
            for (int s = 0; s < n; s += len) {
                const cd q = w[s + m + half] * t;
                w[s + m + half] = w[s + m] - q;
                w[s + m] += q;
            }
        }
    }
}

// twiddle tables of one frame shape, both directions
struct FftTables {
    std::vector<cd> x[2], y[2]; // [0]: sign -1, [1]: sign +1
    FftTables(int X, int Y)
    {
        x[0] = twiddles(X, -1);
        x[1] = twiddles(X, +1);
        y[0] = twiddles(Y, -1);
        y[1] = twiddles(Y, +1);
    }
};

void fft2(std::vector<cd>& a, int X, int Y, int sign, const FftTables& T)
{
    const int d = sign < 0 ? 0 : 1;
    for (int r = 0; r < Y; ++r)
        fft_core(a.data() + (size_t)r * X, X, T.x[d].data());
    // columns in groups of kCols, gathered into contiguous buffers: each gathered
    // row segment is one or two cache lines instead of one line per element
    constexpr int kCols = 8;
    std::vector<cd> w((size_t)kCols * Y);
    for (int c0 = 0; c0 < X; c0 += kCols) {
        const int nc = std::min(kCols, X - c0);
        for (int yy = 0; yy < Y; ++yy)
            for (int c = 0; c < nc; ++c)
                w[(size_t)c * Y + yy] = a[(size_t)yy * X + c0 + c];
        for (int c = 0; c < nc; ++c)
            fft_core(w.data() + (size_t)c * Y, Y, T.y[d].data());
        for (int yy = 0; yy < Y; ++yy)
            for (int c = 0; c < nc; ++c)
                a[(size_t)yy * X + c0 + c] = w[(size_t)c * Y + yy];
    }
}

struct Source {
    double x, y, flux;
};

// Object definition: point sources (binary, triple) or an extended image.
struct Object {
    std::vector<Source> points;
    std::vector<double> image; // extended objects
};

Object make_object(const smip_synth_params& p)
{
    Object o;
    Stream s(p.seed, 0, kObject);
    const double cx = p.ncols / 2, cy = p.nrows / 2;
    if (p.object == 0) { // binary: flux 1 : 0.6, separation 9 px, seeded position angle
        const double pa = 2 * kPi * s.uniform();
        o.points = { { cx, cy, 1.0 }, { cx + std::round(9 * std::cos(pa)), cy + std::round(9 * std::sin(pa)), 0.6 } };
    } else if (p.object == 1) { // triple: fluxes 1 / 0.5 / 0.3, separations 6-20 px
        o.points = { { cx, cy, 1.0 } };
        const double fl[2] = { 0.5, 0.3 };
        for (double f : fl) {
            const double sep = 6 + 14 * s.uniform(), pa = 2 * kPi * s.uniform();
            o.points.push_back({ cx + std::round(sep * std::cos(pa)), cy + std::round(sep * std::sin(pa)), f });
        }
    } else { // limb-darkened disk r = min(150, 0.3 N) plus 200 Gaussian blobs
        const int X = p.ncols, Y = p.nrows;
        o.image.assign((size_t)X * Y, 0.0);
        const double R = std::min(150.0, 0.3 * std::min(X, Y));
        for (int yy = 0; yy < Y; ++yy)
            for (int xx = 0; xx < X; ++xx) {
                const double r = std::hypot(xx - cx, yy - cy) / R;
                if (r < 1)
                    o.image[(size_t)yy * X + xx] = 1.0 - 0.6 * (1.0 - std::sqrt(1.0 - r * r));
            }
        for (int k = 0; k < 200; ++k) {
            const double rr = R * std::sqrt(s.uniform()), th = 2 * kPi * s.uniform();
            const double bx = cx + rr * std::cos(th), by = cy + rr * std::sin(th);
            const double sig = 1 + 7 * s.uniform();
            const double flux = std::pow(10.0, 2.0 * s.uniform()) * 0.5;
            const int w = (int)std::ceil(4 * sig);
            for (int yy = (int)by - w; yy <= (int)by + w; ++yy)
                for (int xx = (int)bx - w; xx <= (int)bx + w; ++xx) {
                    const int px = ((xx % X) + X) % X, py = ((yy % Y) + Y) % Y;
                    const double d2 = (xx - bx) * (xx - bx) + (yy - by) * (yy - by);
                    o.image[(size_t)py * X + px] += flux * std::exp(-d2 / (2 * sig * sig)) / (2 * kPi * sig * sig);
                }
        }
    }
    return o;
}

// What every frame of one parameter set shares: the transform tables, the phase
// screen's filter amplitude per frequency and the pupil's pixels.
struct Setup {
    FftTables fft;
    std::vector<double> amp; // sqrt of the phase power spectrum; < 0 at f = 0
    double den = 1.0;        // sqrt(L Ly), the screen's normalisation
    std::vector<size_t> pupil;
    explicit Setup(const smip_synth_params& p)
        : fft(p.ncols, p.nrows)
    {
        const int X = p.ncols, Y = p.nrows;
        // phase screen by FFT filtering; pupil diameter D = 1 spans D_p pixels
        const double L = X / p.pupil_diameter, Ly = Y / p.pupil_diameter;
        const double r0 = 1.0 / p.d_over_r0;
        amp.resize((size_t)X * Y);
        for (int ky = 0; ky < Y; ++ky)
            for (int kx = 0; kx < X; ++kx) {
                const double fx = (kx <= X / 2 ? kx : kx - X) / L, fy = (ky <= Y / 2 ? ky : ky - Y) / Ly;
                const double f = std::hypot(fx, fy);
                const double psd = 0.023 * std::pow(r0, -5.0 / 3.0) * std::pow(f, -11.0 / 3.0);
                amp[(size_t)ky * X + kx] = f == 0 ? -1.0 : std::sqrt(psd);
            }
        den = std::sqrt(L * Ly);
        const double cx = X / 2.0, cy = Y / 2.0, rp = p.pupil_diameter / 2.0;
        for (int yy = 0; yy < Y; ++yy)
            for (int xx = 0; xx < X; ++xx)
                if (std::hypot(xx + 0.5 - cx, yy + 0.5 - cy) <= rp)
                    pupil.push_back((size_t)yy * X + xx);
    }
};

std::vector<double> speckle_psf(const smip_synth_params& p, const Setup& su, uint64_t frame)
{
    const int X = p.ncols, Y = p.nrows;
    const size_t n = (size_t)X * Y;
    Stream s(p.seed, frame + 1, kScreen);
    std::vector<cd> g(n);
    for (size_t i = 0; i < n; ++i) {
        const double re = s.gauss(), im = s.gauss();
        g[i] = su.amp[i] < 0 ? cd(0.0) : cd(re, im) * su.amp[i] / su.den;
    }
    fft2(g, X, Y, +1, su.fft);
    std::vector<cd> e(n, 0.0);
    for (size_t i : su.pupil)
        e[i] = std::polar(1.0, g[i].real());
    fft2(e, X, Y, -1, su.fft);
    std::vector<double> psf(n);
    double tot = 0;
    for (size_t i = 0; i < n; ++i) {
        psf[i] = std::norm(e[i]);
        tot += psf[i];
    }
    // centre the PSF core on pixel (0, 0) of a circular convolution
    for (auto& v : psf)
        v /= tot;
    return psf;
}

void make_frame(const smip_synth_params& p, const Setup& su, const Object& o, const std::vector<cd>* obj_ft,
    uint64_t f, void* out)
{
    const int X = p.ncols, Y = p.nrows;
    const size_t n = (size_t)X * Y;
    const std::vector<double> psf = speckle_psf(p, su, f);
    std::vector<double> img(n, 0.0);
    if (o.image.empty()) {
        double ftot = 0;
        for (const auto& s : o.points)
            ftot += s.flux;
        for (const auto& s : o.points) {
            const int sx = (int)s.x, sy = (int)s.y;
            for (int yy = 0; yy < Y; ++yy)
                for (int xx = 0; xx < X; ++xx) {
                    const int px = (xx + sx) % X, py = (yy + sy) % Y;
                    img[(size_t)py * X + px] += (s.flux / ftot) * psf[(size_t)yy * X + xx];
                }
        }
    } else {
        std::vector<cd> a(psf.begin(), psf.end());
        fft2(a, X, Y, -1, su.fft);
        for (size_t i = 0; i < n; ++i)
            a[i] *= (*obj_ft)[i];
        fft2(a, X, Y, +1, su.fft);
        double tot = 0;
        for (size_t i = 0; i < n; ++i) {
            img[i] = std::max(0.0, a[i].real());
            tot += img[i];
        }
        for (auto& v : img)
            v /= tot;
    }
    Stream s(p.seed, f + 1, kNoise);
    for (size_t i = 0; i < n; ++i) {
        double v = (double)s.poisson(p.photons * img[i]) + p.read_noise * s.gauss();
        switch (p.dtype) {
        case 1:
            static_cast<uint8_t*>(out)[i] = (uint8_t)std::clamp(std::round(v), 0.0, std::min(p.clip_max, 255.0));
            break;
        case 2:
            static_cast<uint16_t*>(out)[i] = (uint16_t)std::clamp(std::round(v), 0.0, std::min(p.clip_max, 65535.0));
            break;
        default:
            static_cast<float*>(out)[i] = (float)std::max(v, 0.0);
        }
    }
}

bool pow2(int n) { return n >= 8 && n <= 4096 && (n & (n - 1)) == 0; }

} // namespace

extern "C" int smip_synth_frames(const smip_synth_params* p, uint64_t first, uint64_t count, void* out,
    int32_t threads)
{
    if (!p || !out || !pow2(p->ncols) || !pow2(p->nrows) || p->dtype < 1 || p->dtype > 3 || p->object < 0
        || p->object > 2)
        return 1;
    const Object o = make_object(*p);
    const Setup su(*p);
    std::vector<cd> obj_ft;
    if (!o.image.empty()) {
        obj_ft.assign(o.image.begin(), o.image.end());
        fft2(obj_ft, p->ncols, p->nrows, -1, su.fft);
    }
    const size_t fb = (size_t)p->ncols * p->nrows * (p->dtype == 1 ? 1 : p->dtype == 2 ? 2 : 4);
    const int nt = std::max(1, std::min<int>(threads > 0 ? threads : 1, (int)std::max<uint64_t>(count, 1)));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            for (uint64_t k = (uint64_t)t; k < count; k += (uint64_t)nt)
                make_frame(*p, su, o, &obj_ft, first + k, static_cast<uint8_t*>(out) + k * fb);
        });
    for (auto& th : pool)
        th.join();
    return 0;
}

extern "C" int smip_synth_object(const smip_synth_params* p, double* out)
{
    if (!p || !out || !pow2(p->ncols) || !pow2(p->nrows))
        return 1;
    const Object o = make_object(*p);
    const size_t n = (size_t)p->ncols * p->nrows;
    if (!o.image.empty()) {
        std::copy(o.image.begin(), o.image.end(), out);
        return 0;
    }
    std::fill(out, out + n, 0.0);
    for (const auto& s : o.points)
        out[(size_t)((int)s.y % p->nrows) * p->ncols + (int)s.x % p->ncols] += s.flux;
    return 0;
}
