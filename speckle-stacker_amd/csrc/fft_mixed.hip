// fft_mixed.hip -- batched 2-D FFTs of any length in [2, 2048] on gfx950.
//
// FFTW plans transforms of any size (smip-cli/main.cpp:256-259,340-346,
// include/crosscorrel.h:112-141); the power-of-two kernels of fft.hip cover
// lengths 2^k >= 8.  A plan whose two lengths are not both such powers runs the
// kernels here instead, with the same inputs, outputs and conventions (forward
// exp(-2 pi i k x / n), backward exp(+...), both unnormalised).
//
// Each 1-D transform is a Stockham autosort sequence in LDS, one stage per factor
// of n (fft_plan.h: radix 4 while 4 divides, then 2, 3, 5 and the remaining
// primes ascending).  Radices 2, 3, 4 and 5 are butterflies in registers; any
// other prime R is a stage in which each thread forms one output as an R-term
// sum (accumulated in FP64), so any length works and a length with a large prime
// factor costs about n * R per transform.  Twiddles come from the plan's tables
// exp(-2 pi i k / n).  A workgroup holds as many rows (columns) as fill 2048
// elements of each of its two LDS buffers.
#include <algorithm>
#include <cmath>

#include "fft_plan.h"
#include "smip_complex.h"

namespace smip_amd {

MixedRadix factorize(int32_t n)
{
    if (n < 2)
        fail(SMIP_ERR_INVALID_ARGUMENT, "transform length must be at least 2");
    MixedRadix m;
    m.n = n;
    int32_t r = n;
    auto push = [&](int32_t f) {
        if (m.nst >= kMaxStages)
            fail(SMIP_ERR_LOGIC, "transform length has too many factors");
        m.radix[m.nst++] = f;
    };
    while (r % 4 == 0) {
        push(4);
        r /= 4;
    }
    if (r % 2 == 0) {
        push(2);
        r /= 2;
    }
    for (int32_t f = 3; f * f <= r; f += 2)
        while (r % f == 0) {
            push(f);
            r /= f;
        }
    if (r > 1)
        push(r);
    return m;
}

namespace {

constexpr int kMxThreads = 256;
constexpr int kMxElems = 2048; // elements of one LDS buffer per workgroup

template <typename V>
__device__ inline V twiddle(const V* __restrict__ tw, int e, bool inv)
{
    V w = tw[e];
    if (inv)
        w.y = -w.y;
    return w;
}

// i * a
template <typename V>
__device__ inline V times_i(V a) { return mk<V>(-a.y, a.x); }

// In-register DFT of R points: y_q = sum_r v_r exp(-+2 pi i q r / R)
template <typename V, int R>
__device__ inline void dft_small(V* v, bool inv);

template <>
__device__ inline void dft_small<float2, 2>(float2* v, bool)
{
    const float2 a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
}
template <>
__device__ inline void dft_small<double2, 2>(double2* v, bool)
{
    const double2 a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
}

template <typename V>
__device__ inline void dft3(V* v, bool inv)
{
    using S = typename scalar_of<V>::type;
    const S c = (S)-0.5, s = (S)0.86602540378443864676; // cos, sin of 2 pi / 3
    const V t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    const V m = mk<V>(v[0].x + c * t.x, v[0].y + c * t.y);
    const V isd = times_i(mk<V>(s * d.x, s * d.y)); // i s (v1 - v2)
    v[0] = cadd(v[0], t);
    // forward: y1 = m - i s d, y2 = m + i s d
    v[1] = inv ? cadd(m, isd) : csub(m, isd);
    v[2] = inv ? csub(m, isd) : cadd(m, isd);
}
template <>
__device__ inline void dft_small<float2, 3>(float2* v, bool inv) { dft3(v, inv); }
template <>
__device__ inline void dft_small<double2, 3>(double2* v, bool inv) { dft3(v, inv); }

template <typename V>
__device__ inline void dft4(V* v, bool inv)
{
    const V a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]), a2 = cadd(v[1], v[3]);
    const V a3 = inv ? cmul_pi(csub(v[1], v[3])) : cmul_mi(csub(v[1], v[3]));
    v[0] = cadd(a0, a2);
    v[1] = cadd(a1, a3);
    v[2] = csub(a0, a2);
    v[3] = csub(a1, a3);
}
template <>
__device__ inline void dft_small<float2, 4>(float2* v, bool inv) { dft4(v, inv); }
template <>
__device__ inline void dft_small<double2, 4>(double2* v, bool inv) { dft4(v, inv); }

template <typename V>
__device__ inline void dft5(V* v, bool inv)
{
    using S = typename scalar_of<V>::type;
    const S c1 = (S)0.30901699437494742410, c2 = (S)-0.80901699437494742410; // cos 2pi/5, cos 4pi/5
    const S s1 = (S)0.95105651629515357212, s2 = (S)0.58778525229247312917;  // sin 2pi/5, sin 4pi/5
    const V a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    const V a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    const V m1 = mk<V>(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
    const V m2 = mk<V>(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
    const V n1 = times_i(mk<V>(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
    const V n2 = times_i(mk<V>(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
    v[0] = cadd(v[0], cadd(a1, a2));
    // forward: y1 = m1 - i n1, y4 = m1 + i n1, y2 = m2 - i n2, y3 = m2 + i n2
    v[1] = inv ? cadd(m1, n1) : csub(m1, n1);
    v[4] = inv ? csub(m1, n1) : cadd(m1, n1);
    v[2] = inv ? cadd(m2, n2) : csub(m2, n2);
    v[3] = inv ? csub(m2, n2) : cadd(m2, n2);
}
template <>
__device__ inline void dft_small<float2, 5>(float2* v, bool inv) { dft5(v, inv); }
template <>
__device__ inline void dft_small<double2, 5>(double2* v, bool inv) { dft5(v, inv); }

// Stockham stage of radix R and span Ns over nb transforms of length n:
// butterfly j (k = j mod Ns) reads x[j + r n/R], multiplies by
// exp(-2 pi i r k / (Ns R)) = tw[r k n / (Ns R)], and writes its R-point DFT
// to y[(j - k) R + k + q Ns].
template <typename V, int R>
__device__ inline void stage_small(const V* __restrict__ src, V* __restrict__ dst, int n, int nb, int Ns, bool inv,
    const V* __restrict__ tw)
{
    const int m = n / R, tstep = n / (Ns * R);
    for (int t = threadIdx.x; t < nb * m; t += blockDim.x) {
        const int f = t / m, j = t - f * m, k = j % Ns;
        const V* s = src + f * n;
        V* d = dst + f * n;
        V v[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[r] = s[j + r * m];
        if (k) {
#pragma unroll
            for (int r = 1; r < R; ++r)
                v[r] = cmul(v[r], twiddle(tw, r * k * tstep, inv));
        }
        dft_small<V, R>(v, inv);
        const int o = (j - k) * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q)
            d[o + q * Ns] = v[q];
    }
}

// The same stage for any radix R, one output per thread:
// y[g Ns R + q Ns + k] = sum_r x[g Ns + k + r n/R] exp(-2 pi i r (k + q Ns) / (Ns R)),
// accumulated in FP64 (R may be in the thousands).
template <typename V>
__device__ inline void stage_any(const V* __restrict__ src, V* __restrict__ dst, int n, int nb, int Ns, int R,
    bool inv, const V* __restrict__ tw)
{
    const int m = n / R, tstep = n / (Ns * R), span = Ns * R;
    for (int t = threadIdx.x; t < nb * n; t += blockDim.x) {
        const int f = t / n, o = t - f * n;
        const int g = o / span, rem = o - g * span, q = rem / Ns, k = rem - q * Ns;
        const V* x = src + f * n + g * Ns + k;
        const int e = (k + q * Ns) * tstep; // < n
        double ar = 0.0, ai = 0.0;
        int idx = 0; // r e mod n
        for (int r = 0; r < R; ++r) {
            const V a = x[r * m];
            const V w = twiddle(tw, idx, inv);
            ar += (double)a.x * (double)w.x - (double)a.y * (double)w.y;
            ai += (double)a.x * (double)w.y + (double)a.y * (double)w.x;
            idx += e;
            if (idx >= n)
                idx -= n;
        }
        dst[t] = mk<V>((typename scalar_of<V>::type)ar, (typename scalar_of<V>::type)ai);
    }
}

// nb transforms of length P.n stored back to back in a; b is scratch of the same
// size.  Returns the buffer holding the result, in natural order.
template <typename V>
__device__ V* lds_fft_mr(V* a, V* b, const MixedRadix& P, int nb, const V* __restrict__ tw, bool inv)
{
    const int n = P.n;
    int Ns = 1;
    for (int s = 0; s < P.nst; ++s) {
        const int R = P.radix[s];
        switch (R) {
        case 2:
            stage_small<V, 2>(a, b, n, nb, Ns, inv, tw);
            break;
        case 3:
            stage_small<V, 3>(a, b, n, nb, Ns, inv, tw);
            break;
        case 4:
            stage_small<V, 4>(a, b, n, nb, Ns, inv, tw);
            break;
        case 5:
            stage_small<V, 5>(a, b, n, nb, Ns, inv, tw);
            break;
        default:
            stage_any<V>(a, b, n, nb, Ns, R, inv, tw);
        }
        __syncthreads();
        V* t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

template <typename S, typename In>
__device__ inline S sample(const In* p, int64_t i, bool absval)
{
    S v = (S)p[i];
    return absval ? (v < (S)0 ? -v : v) : v;
}

// ---------------------------------------------------------------------------
// Rows: real rows of frame f -> G[f][y][kx], kx < hx = X/2 + 1
// ---------------------------------------------------------------------------
template <typename V, typename In>
__global__ __launch_bounds__(kMxThreads) void k_mx_rows_fwd(const In* __restrict__ frames, int X, int Y, int rb,
    MixedRadix P, const V* __restrict__ tw, V* __restrict__ G, int hx, int absval)
{
    using S = typename scalar_of<V>::type;
    extern __shared__ char smem[];
    V* A = reinterpret_cast<V*>(smem);
    V* B = A + rb * X;
    const int64_t f = blockIdx.y;
    const int row0 = blockIdx.x * rb, nr = min(rb, Y - row0);
    const In* fr = frames + f * (int64_t)X * Y + (int64_t)row0 * X;
    for (int t = threadIdx.x; t < nr * X; t += blockDim.x)
        A[t] = mk<V>(sample<S>(fr, t, absval != 0), (S)0);
    __syncthreads();
    const V* R = lds_fft_mr<V>(A, B, P, nr, tw, false);
    for (int t = threadIdx.x; t < nr * hx; t += blockDim.x) {
        const int p = t / hx, k = t - p * hx;
        G[(f * Y + row0 + p) * hx + k] = R[p * X + k];
    }
}

// Columns kx0 .. kx0 + nc - 1 of G[f] into A[c * Y + y].
template <typename V>
__device__ inline void load_cols(const V* __restrict__ G, int64_t f, int Y, int hx, int kx0, int nc, V* A)
{
    for (int t = threadIdx.x; t < nc * Y; t += blockDim.x) {
        const int y = t / nc, c = t - y * nc;
        A[c * Y + y] = G[(f * Y + y) * hx + kx0 + c];
    }
}

// ---------------------------------------------------------------------------
// Columns: G[f][y][kx] -> transform along y.
//   MODE 0: half-plane layout H[f][X/2 - kx][pad + r] = conj(R[(Y/2 - r) mod Y][kx]),
//           the spectrum at ux = -kx <= 0, uy = r - Y/2 of the real frame
//   MODE 1: natural r2c layout R[f][ky][kx]
// ---------------------------------------------------------------------------
template <typename V, int MODE>
__global__ __launch_bounds__(kMxThreads) void k_mx_cols_fwd(const V* __restrict__ G, int X, int Y, int hx, int cb,
    MixedRadix P, const V* __restrict__ tw, V* __restrict__ out, HalfPlane hp)
{
    extern __shared__ char smem[];
    V* A = reinterpret_cast<V*>(smem);
    V* B = A + cb * Y;
    const int64_t f = blockIdx.y;
    const int kx0 = blockIdx.x * cb, nc = min(cb, hx - kx0);
    load_cols(G, f, Y, hx, kx0, nc, A);
    __syncthreads();
    const V* R = lds_fft_mr<V>(A, B, P, nc, tw, false);
    if (MODE == 0) {
        const int h = Y >> 1;
        for (int t = threadIdx.x; t < nc * Y; t += blockDim.x) {
            const int c = t / Y, r = t - c * Y;
            const V v = R[c * Y + (h - r + Y) % Y];
            out[f * hp.fs + (int64_t)((X >> 1) - (kx0 + c)) * hp.rs + hp.pad + r] = cconj(v);
        }
    } else {
        for (int t = threadIdx.x; t < nc * Y; t += blockDim.x) {
            const int ky = t / nc, c = t - ky * nc;
            out[(f * Y + ky) * hx + kx0 + c] = R[c * Y + ky];
        }
    }
}

// Column passes of the cross-correlation in one launch, in place on G: forward
// transform of each column, C = conj(Rref) * R (include/crosscorrel.h:126-130),
// inverse transform.
__global__ __launch_bounds__(kMxThreads) void k_mx_xcorr_cols(const double2* __restrict__ Rref, int X, int Y, int hx,
    int cb, MixedRadix P, const double2* __restrict__ tw, double2* __restrict__ G)
{
    extern __shared__ char smem[];
    double2* A = reinterpret_cast<double2*>(smem);
    double2* B = A + cb * Y;
    const int64_t f = blockIdx.y;
    const int kx0 = blockIdx.x * cb, nc = min(cb, hx - kx0);
    load_cols(G, f, Y, hx, kx0, nc, A);
    __syncthreads();
    double2* R = lds_fft_mr<double2>(A, B, P, nc, tw, false);
    double2* other = R == A ? B : A;
    for (int t = threadIdx.x; t < nc * Y; t += blockDim.x) {
        const int c = t / Y, ky = t - c * Y;
        // std::conj(a) * b with the (ac - bd, ad + bc) operand order
        R[t] = cmul(cconj(Rref[(int64_t)ky * hx + kx0 + c]), R[t]);
    }
    __syncthreads();
    const double2* Z = lds_fft_mr<double2>(R, other, P, nc, tw, true);
    for (int t = threadIdx.x; t < nc * Y; t += blockDim.x) {
        const int y = t / nc, c = t - y * nc;
        G[(f * Y + y) * hx + kx0 + c] = Z[c * Y + y];
    }
}

// conj(Rref) * Rf along each column, inverse transform, into G.
__global__ __launch_bounds__(kMxThreads) void k_mx_xcorr_cols_inv(const double2* __restrict__ Rref,
    const double2* __restrict__ Rf, int X, int Y, int hx, int cb, MixedRadix P, const double2* __restrict__ tw,
    double2* __restrict__ G)
{
    extern __shared__ char smem[];
    double2* A = reinterpret_cast<double2*>(smem);
    double2* B = A + cb * Y;
    const int64_t f = blockIdx.y;
    const int kx0 = blockIdx.x * cb, nc = min(cb, hx - kx0);
    for (int t = threadIdx.x; t < nc * Y; t += blockDim.x) {
        const int y = t / nc, c = t - y * nc;
        const double2 a = Rref[(int64_t)y * hx + kx0 + c];
        A[c * Y + y] = cmul(cconj(a), Rf[(f * Y + y) * hx + kx0 + c]);
    }
    __syncthreads();
    const double2* Z = lds_fft_mr<double2>(A, B, P, nc, tw, true);
    for (int t = threadIdx.x; t < nc * Y; t += blockDim.x) {
        const int y = t / nc, c = t - y * nc;
        G[(f * Y + y) * hx + kx0 + c] = Z[c * Y + y];
    }
}

// Element k of a Hermitian row stored as its first hx = X/2 + 1 entries; the
// imaginary parts of the kx = 0 and (X even) kx = X/2 terms are dropped, as a
// c2r transform does.
__device__ inline double2 hermitian_at(const double2* __restrict__ g, int k, int X)
{
    if (2 * k <= X) {
        double2 h = g[k];
        if (k == 0 || 2 * k == X)
            h.y = 0.0;
        return h;
    }
    return cconj(g[X - k]);
}

__global__ __launch_bounds__(kMxThreads) void k_mx_rows_c2r(const double2* __restrict__ G, int X, int Y, int rb,
    int hx, MixedRadix P, const double2* __restrict__ tw, double* __restrict__ out)
{
    extern __shared__ char smem[];
    double2* A = reinterpret_cast<double2*>(smem);
    double2* B = A + rb * X;
    const int64_t f = blockIdx.y;
    const int row0 = blockIdx.x * rb, nr = min(rb, Y - row0);
    for (int t = threadIdx.x; t < nr * X; t += blockDim.x) {
        const int p = t / X, k = t - p * X;
        A[t] = hermitian_at(G + (f * Y + row0 + p) * hx, k, X);
    }
    __syncthreads();
    const double2* Z = lds_fft_mr<double2>(A, B, P, nr, tw, true);
    for (int t = threadIdx.x; t < nr * X; t += blockDim.x)
        out[(f * Y + row0) * X + t] = Z[t].x;
}

__device__ inline void better_of(double& bv, int64_t& bi, double v, int64_t i)
{
    // (v, i) replaces (bv, bi) when it is larger, or equal with a lower index
    if (bv < v || (!(v < bv) && i < bi)) {
        bv = v;
        bi = i;
    }
}

// k_mx_rows_c2r with calculate_displacement's search (include/crosscorrel.h:166-179)
// fused in: each workgroup writes the first maximum of its rows, in row-major
// order, to best[f][blockIdx.x]; the correlation never reaches memory.
__global__ __launch_bounds__(kMxThreads) void k_mx_rows_c2r_argmax(const double2* __restrict__ G, int X, int Y,
    int rb, int hx, MixedRadix P, const double2* __restrict__ tw, RowBest* __restrict__ best)
{
    extern __shared__ char smem[];
    __shared__ double wv[kMxThreads / 64];
    __shared__ int64_t wi[kMxThreads / 64];
    double2* A = reinterpret_cast<double2*>(smem);
    double2* B = A + rb * X;
    const int64_t f = blockIdx.y;
    const int row0 = blockIdx.x * rb, nr = min(rb, Y - row0);
    for (int t = threadIdx.x; t < nr * X; t += blockDim.x) {
        const int p = t / X, k = t - p * X;
        A[t] = hermitian_at(G + (f * Y + row0 + p) * hx, k, X);
    }
    __syncthreads();
    const double2* Z = lds_fft_mr<double2>(A, B, P, nr, tw, true);
    // the scan starts from element 0 (max_element's first candidate) in the
    // workgroup that holds it, from nothing elsewhere
    double bv = -INFINITY;
    int64_t bi = INT64_MAX;
    if (blockIdx.x == 0) {
        bv = Z[0].x;
        bi = 0;
    }
    for (int t = threadIdx.x; t < nr * X; t += blockDim.x)
        better_of(bv, bi, Z[t].x, (int64_t)row0 * X + t);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(bv, o);
        const int64_t oi = __shfl_xor(bi, o);
        better_of(bv, bi, ov, oi);
    }
    if ((threadIdx.x & 63) == 0) {
        wv[threadIdx.x >> 6] = bv;
        wi[threadIdx.x >> 6] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k)
            better_of(bv, bi, wv[k], wi[k]);
        best[f * gridDim.x + blockIdx.x] = RowBest { bv, bi };
    }
}

// Complex 2-D transform passes in place on one nrows x ncols image (synthesis).
__global__ __launch_bounds__(kMxThreads) void k_mx_rows_c2c(double2* __restrict__ data, int X, int Y, int rb,
    MixedRadix P, const double2* __restrict__ tw, int inv)
{
    extern __shared__ char smem[];
    double2* A = reinterpret_cast<double2*>(smem);
    double2* B = A + rb * X;
    const int row0 = blockIdx.x * rb, nr = min(rb, Y - row0);
    double2* base = data + (int64_t)row0 * X;
    for (int t = threadIdx.x; t < nr * X; t += blockDim.x)
        A[t] = base[t];
    __syncthreads();
    const double2* Z = lds_fft_mr<double2>(A, B, P, nr, tw, inv != 0);
    for (int t = threadIdx.x; t < nr * X; t += blockDim.x)
        base[t] = Z[t];
}

__global__ __launch_bounds__(kMxThreads) void k_mx_cols_c2c(double2* __restrict__ data, int X, int Y, int cb,
    MixedRadix P, const double2* __restrict__ tw, int inv)
{
    extern __shared__ char smem[];
    double2* A = reinterpret_cast<double2*>(smem);
    double2* B = A + cb * Y;
    const int x0 = blockIdx.x * cb, nc = min(cb, X - x0);
    for (int t = threadIdx.x; t < nc * Y; t += blockDim.x) {
        const int y = t / nc, c = t - y * nc;
        A[c * Y + y] = data[(int64_t)y * X + x0 + c];
    }
    __syncthreads();
    const double2* Z = lds_fft_mr<double2>(A, B, P, nc, tw, inv != 0);
    for (int t = threadIdx.x; t < nc * Y; t += blockDim.x) {
        const int y = t / nc, c = t - y * nc;
        data[(int64_t)y * X + x0 + c] = Z[c * Y + y];
    }
}

// transforms per workgroup: as many as fill kMxElems elements, at most `cap`
int per_block(int n, int cap) { return std::max(1, std::min(kMxElems / n, cap)); }

template <typename V>
size_t lds_bytes(int per, int n) { return 2 * sizeof(V) * (size_t)per * (size_t)n; }

template <typename V>
const V* twx(const FftPlan* p);
template <>
const float2* twx<float2>(const FftPlan* p) { return p->twx32; }
template <>
const double2* twx<double2>(const FftPlan* p) { return p->twx64; }
template <typename V>
const V* twy(const FftPlan* p);
template <>
const float2* twy<float2>(const FftPlan* p) { return p->twy32; }
template <>
const double2* twy<double2>(const FftPlan* p) { return p->twy64; }

template <typename V>
void rows_fwd(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval, V* G, hipStream_t s)
{
    const int X = p->X, Y = p->Y, hx = X / 2 + 1;
    const int rb = per_block(X, Y);
    const dim3 grid((unsigned)((Y + rb - 1) / rb), (unsigned)nf);
    const size_t lds = lds_bytes<V>(rb, X);
    const int av = absval ? 1 : 0;
    switch (dtype) {
    case SMIP_DTYPE_U8:
        hipLaunchKernelGGL((k_mx_rows_fwd<V, uint8_t>), grid, dim3(kMxThreads), lds, s, (const uint8_t*)frames, X, Y,
            rb, p->mx, twx<V>(p), G, hx, av);
        break;
    case SMIP_DTYPE_U16:
        hipLaunchKernelGGL((k_mx_rows_fwd<V, uint16_t>), grid, dim3(kMxThreads), lds, s, (const uint16_t*)frames, X,
            Y, rb, p->mx, twx<V>(p), G, hx, av);
        break;
    case SMIP_DTYPE_F32:
        hipLaunchKernelGGL((k_mx_rows_fwd<V, float>), grid, dim3(kMxThreads), lds, s, (const float*)frames, X, Y, rb,
            p->mx, twx<V>(p), G, hx, av);
        break;
    case SMIP_DTYPE_F64:
        hipLaunchKernelGGL((k_mx_rows_fwd<V, double>), grid, dim3(kMxThreads), lds, s, (const double*)frames, X, Y,
            rb, p->mx, twx<V>(p), G, hx, av);
        break;
    default:
        fail(SMIP_ERR_INVALID_ARGUMENT, "unsupported frame dtype");
    }
    SMIP_HIP(hipGetLastError());
}

template <typename V, int MODE>
void cols_fwd(const FftPlan* p, const V* G, int64_t nf, V* out, const HalfPlane& hp, hipStream_t s)
{
    const int X = p->X, Y = p->Y, hx = X / 2 + 1;
    const int cb = per_block(Y, hx);
    const dim3 grid((unsigned)((hx + cb - 1) / cb), (unsigned)nf);
    hipLaunchKernelGGL((k_mx_cols_fwd<V, MODE>), grid, dim3(kMxThreads), lds_bytes<V>(cb, Y), s, G, X, Y, hx, cb,
        p->my, twy<V>(p), out, hp);
    SMIP_HIP(hipGetLastError());
}

} // namespace

int mixed_shift_blocks(const FftPlan* p)
{
    const int rb = per_block(p->X, p->Y);
    return (p->Y + rb - 1) / rb;
}

void mixed_forward_f32(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval,
    const HalfPlane& hp, float2* out, float2* work, hipStream_t s)
{
    rows_fwd<float2>(p, frames, dtype, nf, absval, work, s);
    cols_fwd<float2, 0>(p, work, nf, out, hp, s);
}

void mixed_forward_f64(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval,
    const HalfPlane& hp, double2* out, double2* work, hipStream_t s)
{
    rows_fwd<double2>(p, frames, dtype, nf, absval, work, s);
    cols_fwd<double2, 0>(p, work, nf, out, hp, s);
}

void mixed_rfft_f64(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval, double2* out,
    double2* work, hipStream_t s)
{
    rows_fwd<double2>(p, frames, dtype, nf, absval, work, s);
    cols_fwd<double2, 1>(p, work, nf, out, HalfPlane {}, s);
}

void mixed_shift_f64(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval,
    const double2* ref_spec, RowBest* best, double2* work, int32_t* shifts, hipStream_t s)
{
    const int X = p->X, Y = p->Y, hx = X / 2 + 1;
    rows_fwd<double2>(p, frames, dtype, nf, absval, work, s);
    const int cb = per_block(Y, hx);
    hipLaunchKernelGGL(k_mx_xcorr_cols, dim3((unsigned)((hx + cb - 1) / cb), (unsigned)nf), dim3(kMxThreads),
        lds_bytes<double2>(cb, Y), s, ref_spec, X, Y, hx, cb, p->my, p->twy64, work);
    SMIP_HIP(hipGetLastError());
    const int rb = per_block(X, Y);
    const int nblk = (Y + rb - 1) / rb;
    hipLaunchKernelGGL(k_mx_rows_c2r_argmax, dim3((unsigned)nblk, (unsigned)nf), dim3(kMxThreads),
        lds_bytes<double2>(rb, X), s, work, X, Y, rb, hx, p->mx, p->twx64, best);
    SMIP_HIP(hipGetLastError());
    launch_best_to_shift(best, nblk, nf, X, Y, shifts, s);
}

void mixed_xcorr_f64(const FftPlan* p, const double2* ref_spec, const double2* spec, int64_t nf, double* corr,
    double2* work, hipStream_t s)
{
    const int X = p->X, Y = p->Y, hx = X / 2 + 1;
    const int cb = per_block(Y, hx);
    hipLaunchKernelGGL(k_mx_xcorr_cols_inv, dim3((unsigned)((hx + cb - 1) / cb), (unsigned)nf), dim3(kMxThreads),
        lds_bytes<double2>(cb, Y), s, ref_spec, spec, X, Y, hx, cb, p->my, p->twy64, work);
    SMIP_HIP(hipGetLastError());
    const int rb = per_block(X, Y);
    hipLaunchKernelGGL(k_mx_rows_c2r, dim3((unsigned)((Y + rb - 1) / rb), (unsigned)nf), dim3(kMxThreads),
        lds_bytes<double2>(rb, X), s, work, X, Y, rb, hx, p->mx, p->twx64, corr);
    SMIP_HIP(hipGetLastError());
}

void mixed_c2c_inverse_f64(const FftPlan* p, double2* data, hipStream_t s)
{
    const int X = p->X, Y = p->Y;
    const int rb = per_block(X, Y);
    hipLaunchKernelGGL(k_mx_rows_c2c, dim3((unsigned)((Y + rb - 1) / rb)), dim3(kMxThreads),
        lds_bytes<double2>(rb, X), s, data, X, Y, rb, p->mx, p->twx64, 1);
    SMIP_HIP(hipGetLastError());
    const int cb = per_block(Y, X);
    hipLaunchKernelGGL(k_mx_cols_c2c, dim3((unsigned)((X + cb - 1) / cb)), dim3(kMxThreads),
        lds_bytes<double2>(cb, Y), s, data, X, Y, cb, p->my, p->twy64, 1);
    SMIP_HIP(hipGetLastError());
}

} // namespace smip_amd
