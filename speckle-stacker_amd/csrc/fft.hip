// fft.hip -- batched 2-D FFTs for frame stacks on gfx950.
//
// Replaces the FFTW calls of the reference path:
//   forward c2c of each frame            smip-cli/main.cpp:256-259,271,296
//   r2c / c2r of the cross-correlation   include/crosscorrel.h:112-141
//   backward c2c of the synthesis        smip-cli/main.cpp:340-346
// Conventions are FFTW's: forward exp(-2 pi i k x / n), backward exp(+...),
// both unnormalised.
//
// Each 1-D transform runs in LDS; one workgroup holds several transforms so
// every stage keeps all 256 lanes busy.  Column passes, and the FP64 row passes,
// use an in-place decimation-in-frequency form (one buffer; radix-8 stages at 64
// and 512 points, else radix-4 stages after a radix-2 one when log2(n) is odd)
// whose digit-reversed output order is absorbed where the result is read; the
// FP32 row passes use a Stockham autosort pair of buffers (radix-4 stages, a
// final radix-2 one).  Lengths 32..2048 run forms compiled for their length.
// Real frames are transformed two rows at a time (z = row0 + i row1) and split
// afterwards.  Twiddles come from tables computed on the host in double
// precision.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <cmath>
#include <vector>

#include "fft_plan.h"
#include "smip_complex.h"
#include "smip_internal.h"

namespace smip_amd {

static int ilog2(int n)
{
    int l = 0;
    while ((1 << l) < n)
        ++l;
    return l;
}

template <typename V>
static V* upload_twiddles(int n, hipStream_t stream)
{
    std::vector<V> h(n);
    const double tp = 6.283185307179586476925286766559;
    for (int k = 0; k < n; ++k) {
        // exp(-2 pi i k / n); the angle is reduced to the first half turn so the
        // table is symmetric to the last bit
        double c = std::cos(tp * (double)k / (double)n);
        double s = std::sin(tp * (double)k / (double)n);
        h[k].x = (typename scalar_of<V>::type)c;
        h[k].y = (typename scalar_of<V>::type)(-s);
    }
    V* d = static_cast<V*>(dev_malloc(sizeof(V) * n));
    try {
        copy_sync(d, h.data(), sizeof(V) * n, hipMemcpyHostToDevice, stream);
    } catch (...) {
        dev_free(d);
        throw;
    }
    return d;
}

static FftPlan* fft_plan_build(int32_t ncols, int32_t nrows, hipStream_t stream)
{
    auto pow2 = [](int n) { return n >= 8 && (n & (n - 1)) == 0; };
    auto* p = new FftPlan;
    p->X = ncols;
    p->Y = nrows;
    p->lx = ilog2(ncols);
    p->ly = ilog2(nrows);
    p->pow2 = pow2(ncols) && pow2(nrows);
    p->mx = factorize(ncols);
    p->my = factorize(nrows);
    try {
        p->twx32 = upload_twiddles<float2>(ncols, stream);
        p->twy32 = upload_twiddles<float2>(nrows, stream);
        p->twx64 = upload_twiddles<double2>(ncols, stream);
        p->twy64 = upload_twiddles<double2>(nrows, stream);
    } catch (...) {
        fft_plan_destroy(p);
        throw;
    }
    return p;
}

// Plans are read-only once built, so one per (device, ncols, nrows) is kept and
// shared by every job and correlator of that shape: a job then starts without
// the four synchronous twiddle uploads.  Up to kCachedPlans are kept (a few
// hundred KiB of tables at most); beyond that a plan is built for its owner and
// destroyed with it.  Like the other caches, kept plans are never destroyed.
FftPlan* fft_plan_create(int32_t ncols, int32_t nrows, hipStream_t stream)
{
    if (ncols < 2 || nrows < 2 || ncols > kMaxFftLength || nrows > kMaxFftLength)
        fail(SMIP_ERR_INVALID_ARGUMENT, "frame extents must be in [2, 2048]");
    constexpr size_t kCachedPlans = 64;
    int dev = 0;
    SMIP_HIP(hipGetDevice(&dev));
    static std::mutex mu;
    static auto* cache = new std::map<std::tuple<int, int32_t, int32_t>, FftPlan*>;
    const auto key = std::make_tuple(dev, ncols, nrows);
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache->find(key);
    if (it != cache->end())
        return it->second;
    FftPlan* p = fft_plan_build(ncols, nrows, stream);
    if (cache->size() < kCachedPlans) {
        p->cached = true;
        (*cache)[key] = p;
    }
    return p;
}

void fft_plan_destroy(FftPlan* p)
{
    if (!p || p->cached)
        return;
    dev_free(p->twx32);
    dev_free(p->twy32);
    dev_free(p->twx64);
    dev_free(p->twy64);
    delete p;
}

template <typename V>
static const V* tw_x(const FftPlan* p);
template <>
const float2* tw_x<float2>(const FftPlan* p) { return p->twx32; }
template <>
const double2* tw_x<double2>(const FftPlan* p) { return p->twx64; }
template <typename V>
static const V* tw_y(const FftPlan* p);
template <>
const float2* tw_y<float2>(const FftPlan* p) { return p->twy32; }
template <>
const double2* tw_y<double2>(const FftPlan* p) { return p->twy64; }

// ---------------------------------------------------------------------------
// LDS Stockham FFT of nb transforms of length n = 2^log2n stored back to back.
// Stage with radix R and span Ns: butterfly j reads x[j + r n/R], applies the
// twiddles w^(r k n/(Ns R)) with k = j mod Ns, and writes the R-point DFT to
// y[(j - k) R + k + r Ns].  Returns the buffer holding the result.
// ---------------------------------------------------------------------------
template <typename V>
__device__ V* lds_fft(V* a, V* b, int n, int log2n, int nb, const V* __restrict__ tw, bool inv)
{
    const int nthreads = blockDim.x;
    const int tid = threadIdx.x;
    int Ns = 1;
    const int q4 = n >> 2, lq4 = log2n - 2;
    for (int s = 0; s < (log2n >> 1); ++s) {
        const int ltw = log2n - 2 - 2 * s; // log2(n / (4 Ns))
        for (int t = tid; t < nb * q4; t += nthreads) {
            const int f = t >> lq4;
            const int j = t & (q4 - 1);
            const V* src = a + f * n;
            V* dst = b + f * n;
            const int k = j & (Ns - 1);
            V v0 = src[j], v1 = src[j + q4], v2 = src[j + 2 * q4], v3 = src[j + 3 * q4];
            if (Ns > 1) {
                const int e = k << ltw;
                V w1 = tw[e], w2 = tw[2 * e], w3 = tw[3 * e];
                if (inv) {
                    w1 = cconj(w1);
                    w2 = cconj(w2);
                    w3 = cconj(w3);
                }
                v1 = cmul(v1, w1);
                v2 = cmul(v2, w2);
                v3 = cmul(v3, w3);
            }
            V a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), a3 = csub(v1, v3);
            a3 = inv ? cmul_pi(a3) : cmul_mi(a3);
            const int d = ((j - k) << 2) + k;
            dst[d] = cadd(a0, a2);
            dst[d + Ns] = cadd(a1, a3);
            dst[d + 2 * Ns] = csub(a0, a2);
            dst[d + 3 * Ns] = csub(a1, a3);
        }
        __syncthreads();
        V* t = a;
        a = b;
        b = t;
        Ns <<= 2;
    }
    if (log2n & 1) {
        const int q2 = n >> 1, lq2 = log2n - 1;
        for (int t = tid; t < nb * q2; t += nthreads) {
            const int f = t >> lq2;
            const int j = t & (q2 - 1);
            const V* src = a + f * n;
            V* dst = b + f * n;
            const int k = j & (Ns - 1);
            V v0 = src[j], v1 = src[j + q2];
            if (Ns > 1) {
                V w = tw[k];
                if (inv)
                    w = cconj(w);
                v1 = cmul(v1, w);
            }
            const int d = ((j - k) << 1) + k;
            dst[d] = cadd(v0, v1);
            dst[d + Ns] = csub(v0, v1);
        }
        __syncthreads();
        V* t = a;
        a = b;
        b = t;
    }
    return a;
}

// Column tiles in LDS.  Element y of column c sits at c * col_ld<V>(Y) +
// col_at<V>(y): one padding element after every 2^col_shift elements (every 64
// bytes of float2, every 256 bytes of double2) and one more per column.  Without
// padding, a stride of Y elements is a multiple of the 64 banks' 256 bytes: the
// transposing stores and loads of a wavefront that walks the columns all hit one
// bank, and the late radix-4 stages, whose lanes step over runs of Q < 16
// elements, hit a few; with it, a model of the 64 banks puts every stage of the
// 256- and 512-point double2 transforms at the conflict-free minimum and the
// float2 ones within 12% of it.
template <typename V>
__host__ __device__ constexpr int col_shift() { return sizeof(V) == 16 ? 4 : 3; }
template <typename V>
__host__ __device__ constexpr int col_ld(int Y) { return Y + (Y >> col_shift<V>()) + 1; }
template <typename V>
__device__ inline int col_at(int i) { return i + (i >> col_shift<V>()); }

// ---------------------------------------------------------------------------
// In-place LDS FFT (decimation in frequency) of nb transforms of length
// n = 2^log2n, element i of transform f at a[f * ld + col_at<V>(i)] (a column
// tile as laid out above): a radix-2 stage first when log2n is odd,
// then radix-4 stages.  Every butterfly reads and writes the same four (two)
// locations, so one buffer and one barrier per stage suffice -- half the LDS of
// the Stockham form, which lets a column pass hold twice the columns.  The
// result is in digit-reversed order: X[k] is at dif_pos(k).
//   radix-4 stage of span L (quarter Q = L/4), butterfly j in [0, Q) of group g:
//     x_r = a[g L + j + r Q]; DFT-4 of (x_0..x_3); output r scaled by w_L^(r j)
//   radix-2 stage: x_0 + x_1, (x_0 - x_1) w_n^j
// ---------------------------------------------------------------------------
template <typename V>
__device__ void lds_fft_dif(V* a, int n, int log2n, int nb, int ld, const V* __restrict__ tw, bool inv)
{
    const int nthreads = blockDim.x;
    const int tid = threadIdx.x;
    int L = n, lL = log2n;
    if (log2n & 1) {
        const int h = n >> 1;
        for (int t = tid; t < nb * h; t += nthreads) {
            const int f = t >> (log2n - 1), j = t & (h - 1);
            V* x0p = a + f * ld + col_at<V>(j);
            V* x1p = a + f * ld + col_at<V>(j + h);
            const V x0 = *x0p, x1 = *x1p;
            V w = tw[j];
            if (inv)
                w = cconj(w);
            *x0p = cadd(x0, x1);
            *x1p = cmul(csub(x0, x1), w);
        }
        __syncthreads();
        L = h;
        lL = log2n - 1;
    }
    while (L >= 4) {
        const int Q = L >> 2, lQ = lL - 2;
        const int step = n >> lL; // w_L = tw[step]
        for (int t = tid; t < nb * (n >> 2); t += nthreads) {
            const int f = t >> (log2n - 2), b = t & ((n >> 2) - 1);
            const int g = b >> lQ, j = b & (Q - 1);
            V* col = a + f * ld;
            const int i0 = g * L + j;
            V* p0 = col + col_at<V>(i0);
            V* p1 = col + col_at<V>(i0 + Q);
            V* p2 = col + col_at<V>(i0 + 2 * Q);
            V* p3 = col + col_at<V>(i0 + 3 * Q);
            const V x0 = *p0, x1 = *p1, x2 = *p2, x3 = *p3;
            const V a0 = cadd(x0, x2), a1 = csub(x0, x2), a2 = cadd(x1, x3);
            const V a3 = inv ? cmul_pi(csub(x1, x3)) : cmul_mi(csub(x1, x3));
            V y1 = cadd(a1, a3), y2 = csub(a0, a2), y3 = csub(a1, a3);
            if (j > 0) {
                V w1 = tw[step * j], w2 = tw[2 * step * j], w3 = tw[3 * step * j];
                if (inv) {
                    w1 = cconj(w1);
                    w2 = cconj(w2);
                    w3 = cconj(w3);
                }
                y1 = cmul(y1, w1);
                y2 = cmul(y2, w2);
                y3 = cmul(y3, w3);
            }
            *p0 = cadd(a0, a2);
            *p1 = y1;
            *p2 = y2;
            *p3 = y3;
        }
        __syncthreads();
        L = Q;
        lL = lQ;
    }
}

// Column transforms of 64 and 512 points run as radix-8 stages (lds_fft_dif_ct);
// the other lengths as radix-4 stages after a radix-2 one when log2n is odd.
__host__ __device__ constexpr bool dif_radix8(int log2n) { return log2n == 6 || log2n == 9; }

// position of X[k] after the in-place DIF: for radix-8 lengths the base-8
// digits of k reverse; otherwise the radix-2 split (odd log2n) selects the half
// by k's lowest bit, then the base-4 digits of the rest reverse
__device__ inline int dif_pos(int k, int log2n)
{
    if (dif_radix8(log2n)) {
        const int lo = k & 7, mid = (k >> 3) & 7, hi = k >> 6; // hi = 0 for 64 points
        return log2n == 9 ? (lo << 6) | (mid << 3) | hi : (lo << 3) | mid;
    }
    // base-4 digit reversal of the m = log2n - odd bits: a bit reversal, then the
    // two bits of each digit swapped back
    const int odd = log2n & 1;
    const int m = log2n - odd;
    const int half = (k & odd) << (log2n - 1);
    const uint32_t q = (uint32_t)k >> odd;
    uint32_t r = m ? (__builtin_bitreverse32(q) >> (32 - m)) : 0u;
    r = ((r & 0x55555555u) << 1) | ((r >> 1) & 0x55555555u);
    return half | (int)r;
}

// Twiddle tables of at most 4 KiB are copied into LDS behind a kernel's tile, so
// that each FFT stage reads its twiddles beside its data instead of waiting on a
// memory round trip (5-6% on the 256- and 512-point FP32 passes and the 256-point
// FP64 ones).  Larger tables stay in memory: 8 KiB more per workgroup took the
// 512-point FP64 column pass from four workgroups per CU to three, and made it
// slower.  tw_lds_bytes is the LDS a launch adds for a transform of length n.
template <typename V>
__host__ __device__ constexpr bool tw_staged(int n) { return sizeof(V) * (size_t)n <= 4096; }
template <typename V>
__host__ __device__ constexpr size_t tw_lds_bytes(int n) { return tw_staged<V>(n) ? sizeof(V) * (size_t)n : 0; }
// Copies tw[0, n) to `at` when tw_staged(n) and returns the table to use; the
// caller's next barrier publishes it.
template <typename V>
__device__ inline const V* stage_twiddles(const V* __restrict__ tw, int n, char* at)
{
    if (!tw_staged<V>(n))
        return tw;
    V* T = reinterpret_cast<V*>(at);
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        T[i] = tw[i];
    return T;
}

template <typename S, typename In>
__device__ inline S sample_value(In x, bool absval)
{
    S v = (S)x;
    return absval ? (v < (S)0 ? -v : v) : v;
}

// LDS budget per workgroup (two ping-pong buffers): 64 KiB.
constexpr int kLdsBytes = 65536;
// Data bytes of a column-pass tile (the padding of col_ld comes on top).
// Measured against 64 KiB: the shift-and-add at 512x512 and 256x256 is 9-10%
// slower there (half the workgroups per CU), though its rows fill cache lines.
constexpr int kColLds = 32768;
constexpr int kThreads = 256;
// elements per thread of a column tile (cb * Y * sizeof(V) <= kColLds) and of a
// row tile (ppw * X <= 2048, see rows_per_wg)
template <typename V>
constexpr int kColPer = kColLds / (int)sizeof(V) / kThreads;
constexpr int kRowElems = 2048;
constexpr int kRowPer = kRowElems / kThreads;

// Columns per workgroup of a column pass (cols_per_wg on the host): as many as
// kColLds holds, at most 64, a power of two; 0 when one column does not fit.
template <typename V>
__host__ __device__ constexpr int col_count(int log2n)
{
    const long k0 = (long)kColLds / ((long)sizeof(V) << log2n);
    const int k = (int)(k0 > 64 ? 64 : k0);
    int p = 1;
    while (p * 2 <= k)
        p *= 2;
    return k < 1 ? 0 : p;
}

// lds_fft_dif of nb <= NBMAX transforms (columns of a column tile, rows of a
// row tile; EXACT: nb = NBMAX) at stride col_ld, with the length, NBMAX and the
// direction fixed at compile time: the same butterflies with the same arithmetic in the
// same order, so the results are identical, but with the index arithmetic
// folded, the loads of all of a thread's butterflies of a stage issued before
// any of them is computed, and the twiddle table's address space known (LDS
// when it is staged, global memory otherwise) instead of generic loads that
// wait on both counters one butterfly at a time.
template <typename V, int LOG2N, int NBMAX, bool EXACT, bool INV, int LL>
__device__ __forceinline__ void dif_radix4_ct(V* a, int nb, const V* T)
{
    if constexpr (LL >= 2) {
        constexpr int N = 1 << LOG2N, LD = col_ld<V>(N);
        constexpr int Q = 1 << (LL - 2), L = 1 << LL, STEP = N >> LL;
        constexpr int IT = (NBMAX * (N >> 2) + kThreads - 1) / kThreads;
        constexpr bool FULL = EXACT && (NBMAX * (N >> 2)) % kThreads == 0;
        const int TOT = (EXACT ? NBMAX : nb) * (N >> 2);
        const int tid = threadIdx.x;
        V x[IT][4], w[IT][3];
        int at0[IT], jj[IT];
#pragma unroll
        for (int q = 0; q < IT; ++q) {
            const int t = tid + q * kThreads;
            if (FULL || t < TOT) {
                const int f = t >> (LOG2N - 2), b = t & ((N >> 2) - 1);
                const int g = b >> (LL - 2), j = b & (Q - 1);
                jj[q] = j;
                at0[q] = f * LD;
                const int i0 = g * L + j;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    x[q][r] = a[at0[q] + col_at<V>(i0 + r * Q)];
                if constexpr (Q > 1) {
                    w[q][0] = T[STEP * j];
                    w[q][1] = T[2 * STEP * j];
                    w[q][2] = T[3 * STEP * j];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < IT; ++q) {
            const int t = tid + q * kThreads;
            if (FULL || t < TOT) {
                const int b = t & ((N >> 2) - 1);
                const int g = b >> (LL - 2), j = jj[q];
                const int i0 = g * L + j;
                const V x0 = x[q][0], x1 = x[q][1], x2 = x[q][2], x3 = x[q][3];
                const V a0 = cadd(x0, x2), a1 = csub(x0, x2), a2 = cadd(x1, x3);
                const V a3 = INV ? cmul_pi(csub(x1, x3)) : cmul_mi(csub(x1, x3));
                V y1 = cadd(a1, a3), y2 = csub(a0, a2), y3 = csub(a1, a3);
                if constexpr (Q > 1) {
                    if (j > 0) {
                        const V w1 = INV ? cconj(w[q][0]) : w[q][0];
                        const V w2 = INV ? cconj(w[q][1]) : w[q][1];
                        const V w3 = INV ? cconj(w[q][2]) : w[q][2];
                        y1 = cmul(y1, w1);
                        y2 = cmul(y2, w2);
                        y3 = cmul(y3, w3);
                    }
                }
                a[at0[q] + col_at<V>(i0)] = cadd(a0, a2);
                a[at0[q] + col_at<V>(i0 + Q)] = y1;
                a[at0[q] + col_at<V>(i0 + 2 * Q)] = y2;
                a[at0[q] + col_at<V>(i0 + 3 * Q)] = y3;
            }
        }
        __syncthreads();
        dif_radix4_ct<V, LOG2N, NBMAX, EXACT, INV, LL - 2>(a, nb, T);
    }
}

// 4-point DFT of (v0, v1, v2, v3) with w_4 = -i (+i when INV), in the operation
// order of the radix-4 butterflies above.
template <typename V, bool INV>
__device__ __forceinline__ void dft4(const V& v0, const V& v1, const V& v2, const V& v3, V& y0, V& y1, V& y2, V& y3)
{
    const V a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3);
    const V a3 = INV ? cmul_pi(csub(v1, v3)) : cmul_mi(csub(v1, v3));
    y0 = cadd(a0, a2);
    y1 = cadd(a1, a3);
    y2 = csub(a0, a2);
    y3 = csub(a1, a3);
}

// Radix-8 stage of span L (E = L/8) of the in-place DIF, butterfly j in [0, E)
// of group g: x_r = a[g L + j + r E]; X_s = sum_r x_r w_8^(r s), as two 4-point
// DFTs, of u_r = x_r + x_(r+4) (even s) and of (x_r - x_(r+4)) w_8^r (odd s);
// output s, scaled by w_L^(s j), goes to a[g L + j + s E].  Three such stages
// make a 512-point transform where radix-4 needs five, with fewer LDS round
// trips and barriers per point.
template <typename V, int LOG2N, int NBMAX, bool EXACT, bool INV, int LL>
__device__ __forceinline__ void dif_radix8_ct(V* a, int nb, const V* T)
{
    if constexpr (LL >= 3) {
        using S = typename scalar_of<V>::type;
        constexpr int N = 1 << LOG2N, LD = col_ld<V>(N);
        constexpr int E = 1 << (LL - 3), L = 1 << LL, STEP = N >> LL;
        constexpr int IT = (NBMAX * (N >> 3) + kThreads - 1) / kThreads;
        constexpr bool FULL = EXACT && (NBMAX * (N >> 3)) % kThreads == 0;
        const int TOT = (EXACT ? NBMAX : nb) * (N >> 3);
        const S h = (S)0.70710678118654752440084436210485; // 1/sqrt(2)
        const int tid = threadIdx.x;
        V x[IT][8], w[IT][7];
        int jj[IT];
#pragma unroll
        for (int q = 0; q < IT; ++q) {
            const int t = tid + q * kThreads;
            if (FULL || t < TOT) {
                const int f = t >> (LOG2N - 3), b = t & ((N >> 3) - 1);
                const int g = b >> (LL - 3), j = b & (E - 1);
                jj[q] = j;
#pragma unroll
                for (int r = 0; r < 8; ++r)
                    x[q][r] = a[f * LD + col_at<V>(g * L + j + r * E)];
                if constexpr (E > 1) {
#pragma unroll
                    for (int r = 1; r < 8; ++r)
                        w[q][r - 1] = T[STEP * r * j];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < IT; ++q) {
            const int t = tid + q * kThreads;
            if (FULL || t < TOT) {
                const int f = t >> (LOG2N - 3), b = t & ((N >> 3) - 1);
                const int g = b >> (LL - 3), j = jj[q];
                V u[4], d[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    u[r] = cadd(x[q][r], x[q][r + 4]);
                    d[r] = csub(x[q][r], x[q][r + 4]);
                }
                // d_r w_8^r: w_8 = (1 - i)/sqrt(2), w_8^2 = -i, w_8^3 = -(1 + i)/sqrt(2)
                // (conjugates when INV)
                if (INV) {
                    d[1] = mk<V>((d[1].x - d[1].y) * h, (d[1].x + d[1].y) * h);
                    d[2] = cmul_pi(d[2]);
                    d[3] = mk<V>(-(d[3].x + d[3].y) * h, (d[3].x - d[3].y) * h);
                } else {
                    d[1] = mk<V>((d[1].x + d[1].y) * h, (d[1].y - d[1].x) * h);
                    d[2] = cmul_mi(d[2]);
                    d[3] = mk<V>((d[3].y - d[3].x) * h, -(d[3].x + d[3].y) * h);
                }
                V X8[8];
                dft4<V, INV>(u[0], u[1], u[2], u[3], X8[0], X8[2], X8[4], X8[6]);
                dft4<V, INV>(d[0], d[1], d[2], d[3], X8[1], X8[3], X8[5], X8[7]);
                if constexpr (E > 1) {
                    if (j > 0) {
#pragma unroll
                        for (int r = 1; r < 8; ++r)
                            X8[r] = cmul(X8[r], INV ? cconj(w[q][r - 1]) : w[q][r - 1]);
                    }
                }
#pragma unroll
                for (int r = 0; r < 8; ++r)
                    a[f * LD + col_at<V>(g * L + j + r * E)] = X8[r];
            }
        }
        __syncthreads();
        dif_radix8_ct<V, LOG2N, NBMAX, EXACT, INV, LL - 3>(a, nb, T);
    }
}

template <typename V, int LOG2N, int NBMAX, bool EXACT, bool INV>
__device__ __forceinline__ void lds_fft_dif_ct(V* a, int nb, const V* T)
{
    if constexpr (dif_radix8(LOG2N)) {
        dif_radix8_ct<V, LOG2N, NBMAX, EXACT, INV, LOG2N>(a, nb, T);
    } else {
        if constexpr (LOG2N & 1) {
            constexpr int N = 1 << LOG2N, LD = col_ld<V>(N);
            constexpr int H = N >> 1, IT = (NBMAX * H + kThreads - 1) / kThreads;
            constexpr bool FULL = EXACT && (NBMAX * H) % kThreads == 0;
            const int TOT = (EXACT ? NBMAX : nb) * H;
            const int tid = threadIdx.x;
            V x0[IT], x1[IT], w[IT];
#pragma unroll
            for (int q = 0; q < IT; ++q) {
                const int t = tid + q * kThreads;
                if (FULL || t < TOT) {
                    const int f = t >> (LOG2N - 1), j = t & (H - 1);
                    x0[q] = a[f * LD + col_at<V>(j)];
                    x1[q] = a[f * LD + col_at<V>(j + H)];
                    w[q] = T[j];
                }
            }
#pragma unroll
            for (int q = 0; q < IT; ++q) {
                const int t = tid + q * kThreads;
                if (FULL || t < TOT) {
                    const int f = t >> (LOG2N - 1), j = t & (H - 1);
                    const V ww = INV ? cconj(w[q]) : w[q];
                    a[f * LD + col_at<V>(j)] = cadd(x0[q], x1[q]);
                    a[f * LD + col_at<V>(j + H)] = cmul(csub(x0[q], x1[q]), ww);
                }
            }
            __syncthreads();
        }
        dif_radix4_ct<V, LOG2N, NBMAX, EXACT, INV, LOG2N - (LOG2N & 1)>(a, nb, T);
    }
}

// lds_fft_dif of a column tile of col_count(log2n) columns, through the
// compile-time form for lengths 32..2048.  With STAGED, the twiddles sit at `at`
// in LDS when tw_staged(n) (stage_twiddles put them there), else they are read
// from tw.
template <typename V, bool INV, bool STAGED>
__device__ __forceinline__ void lds_fft_cols(V* a, int n, int log2n, int nb, int ld, const V* __restrict__ tw,
    const char* at)
{
#define SMIP_DIF_CASE(LN)                                                                                             \
    case LN:                                                                                                          \
        if constexpr (col_count<V>(LN) > 0) {                                                                        \
            lds_fft_dif_ct<V, LN, col_count<V>(LN), true, INV>(                                                            \
                a, nb, STAGED && tw_staged<V>(1 << LN) ? reinterpret_cast<const V*>(at) : tw);                       \
            return;                                                                                                   \
        }                                                                                                             \
        break;
    if (nb == col_count<V>(log2n) && ld == col_ld<V>(n)) {
        switch (log2n) {
            SMIP_DIF_CASE(5)
            SMIP_DIF_CASE(6)
            SMIP_DIF_CASE(7)
            SMIP_DIF_CASE(8)
            SMIP_DIF_CASE(9)
            SMIP_DIF_CASE(10)
            SMIP_DIF_CASE(11)
        default:
            break;
        }
    }
#undef SMIP_DIF_CASE
    lds_fft_dif<V>(a, n, log2n, nb, ld, STAGED && tw_staged<V>(n) ? reinterpret_cast<const V*>(at) : tw, INV);
}

// Transforms per workgroup of a row pass when the frame has rows to spare
// (rows_per_wg on the host): enough that a radix-4 stage has a butterfly for
// each of the 256 lanes.
// Row passes on the in-place DIF (one LDS buffer) for FP64, on the Stockham
// ping-pong pair for FP32.
template <typename V>
__host__ __device__ constexpr bool rows_dif() { return sizeof(V) == 16; }

__host__ __device__ constexpr int row_fill(int log2n)
{
    return (4 * kThreads) >> log2n > 1 ? (4 * kThreads) >> log2n : 1;
}

// The in-place DIF on a row tile: nb <= row_fill(log2n) transforms of length n,
// transform p at a[p * col_ld(n) + col_at(i)], through the compile-time form for
// lengths 32..2048 (and for 64 and 512 points always, whose radix-8 output order
// dif_pos assumes).  With STAGED, the twiddles sit at `at` in LDS when
// tw_staged(n), else they are read from tw.
template <typename V, bool INV, bool STAGED>
__device__ __forceinline__ void lds_fft_rows_dif(V* a, int n, int log2n, int nb, const V* __restrict__ tw,
    const char* at)
{
#define SMIP_ROWDIF_CASE(LN)                                                                                          \
    case LN:                                                                                                          \
        lds_fft_dif_ct<V, LN, row_fill(LN), false, INV>(                                                                     \
            a, nb, STAGED && tw_staged<V>(1 << LN) ? reinterpret_cast<const V*>(at) : tw);                           \
        return;
    switch (log2n) {
        SMIP_ROWDIF_CASE(5)
        SMIP_ROWDIF_CASE(6)
        SMIP_ROWDIF_CASE(7)
        SMIP_ROWDIF_CASE(8)
        SMIP_ROWDIF_CASE(9)
        SMIP_ROWDIF_CASE(10)
        SMIP_ROWDIF_CASE(11)
    default:
        break;
    }
#undef SMIP_ROWDIF_CASE
    lds_fft_dif<V>(a, n, log2n, nb, col_ld<V>(n), STAGED && tw_staged<V>(n) ? reinterpret_cast<const V*>(at) : tw,
        INV);
}


// lds_fft (Stockham, ping-pong) with the length, the transform count and the
// direction fixed at compile time, as lds_fft_dif_ct is for the column form:
// identical arithmetic, folded indices, loads of a stage issued together.
// Stage S is the radix-4 stage of span Ns = 4^S; after the radix-4 stages, a
// radix-2 stage when log2(n) is odd.  Returns the buffer holding the result.
template <typename V, int LOG2N, int NB, bool INV, int S>
__device__ __forceinline__ V* stockham_ct(V* a, V* b, const V* T)
{
    constexpr int N = 1 << LOG2N;
    const int tid = threadIdx.x;
    if constexpr (2 * S + 2 <= LOG2N) {
        constexpr int Ns = 1 << (2 * S), Q4 = N >> 2, LQ4 = LOG2N - 2, LTW = LOG2N - 2 - 2 * S;
        constexpr int TOT = NB * Q4, IT = (TOT + kThreads - 1) / kThreads;
        constexpr bool FULL = TOT % kThreads == 0;
        V v[IT][4], w[IT][3];
#pragma unroll
        for (int q = 0; q < IT; ++q) {
            const int t = tid + q * kThreads;
            if (FULL || t < TOT) {
                const int f = t >> LQ4, j = t & (Q4 - 1);
                const V* src = a + f * N;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    v[q][r] = src[j + r * Q4];
                if constexpr (Ns > 1) {
                    const int e = (j & (Ns - 1)) << LTW;
                    w[q][0] = T[e];
                    w[q][1] = T[2 * e];
                    w[q][2] = T[3 * e];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < IT; ++q) {
            const int t = tid + q * kThreads;
            if (FULL || t < TOT) {
                const int f = t >> LQ4, j = t & (Q4 - 1);
                const int k = j & (Ns - 1);
                V v0 = v[q][0], v1 = v[q][1], v2 = v[q][2], v3 = v[q][3];
                if constexpr (Ns > 1) {
                    const V w1 = INV ? cconj(w[q][0]) : w[q][0];
                    const V w2 = INV ? cconj(w[q][1]) : w[q][1];
                    const V w3 = INV ? cconj(w[q][2]) : w[q][2];
                    v1 = cmul(v1, w1);
                    v2 = cmul(v2, w2);
                    v3 = cmul(v3, w3);
                }
                V a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), a3 = csub(v1, v3);
                a3 = INV ? cmul_pi(a3) : cmul_mi(a3);
                V* dst = b + f * N;
                const int d = ((j - k) << 2) + k;
                dst[d] = cadd(a0, a2);
                dst[d + Ns] = cadd(a1, a3);
                dst[d + 2 * Ns] = csub(a0, a2);
                dst[d + 3 * Ns] = csub(a1, a3);
            }
        }
        __syncthreads();
        return stockham_ct<V, LOG2N, NB, INV, S + 1>(b, a, T);
    } else if constexpr (LOG2N & 1) {
        constexpr int Ns = 1 << (2 * S), Q2 = N >> 1, LQ2 = LOG2N - 1;
        constexpr int TOT = NB * Q2, IT = (TOT + kThreads - 1) / kThreads;
        constexpr bool FULL = TOT % kThreads == 0;
        V v0[IT], v1[IT], w[IT];
#pragma unroll
        for (int q = 0; q < IT; ++q) {
            const int t = tid + q * kThreads;
            if (FULL || t < TOT) {
                const int f = t >> LQ2, j = t & (Q2 - 1);
                v0[q] = a[f * N + j];
                v1[q] = a[f * N + j + Q2];
                if constexpr (Ns > 1)
                    w[q] = T[j & (Ns - 1)];
            }
        }
#pragma unroll
        for (int q = 0; q < IT; ++q) {
            const int t = tid + q * kThreads;
            if (FULL || t < TOT) {
                const int f = t >> LQ2, j = t & (Q2 - 1);
                const int k = j & (Ns - 1);
                V x1 = v1[q];
                if constexpr (Ns > 1)
                    x1 = cmul(x1, INV ? cconj(w[q]) : w[q]);
                V* dst = b + f * N;
                const int d = ((j - k) << 1) + k;
                dst[d] = cadd(v0[q], x1);
                dst[d + Ns] = csub(v0[q], x1);
            }
        }
        __syncthreads();
        return b;
    } else {
        return a;
    }
}

// lds_fft through the compile-time form for lengths 32..2048 when the
// workgroup holds row_fill(log2n) transforms.  With STAGED, the twiddles sit at
// `at` in LDS when tw_staged(n), else they are read from tw.
template <typename V, bool INV, bool STAGED>
__device__ __forceinline__ V* lds_fft_rows(V* a, V* b, int n, int log2n, int nb, const V* __restrict__ tw,
    const char* at)
{
#define SMIP_STOCKHAM_CASE(LN)                                                                                        \
    case LN:                                                                                                          \
        return stockham_ct<V, LN, row_fill(LN), INV, 0>(                                                              \
            a, b, STAGED && tw_staged<V>(1 << LN) ? reinterpret_cast<const V*>(at) : tw);
    if (nb == row_fill(log2n)) {
        switch (log2n) {
            SMIP_STOCKHAM_CASE(5)
            SMIP_STOCKHAM_CASE(6)
            SMIP_STOCKHAM_CASE(7)
            SMIP_STOCKHAM_CASE(8)
            SMIP_STOCKHAM_CASE(9)
            SMIP_STOCKHAM_CASE(10)
            SMIP_STOCKHAM_CASE(11)
        default:
            break;
        }
    }
#undef SMIP_STOCKHAM_CASE
    return lds_fft<V>(a, b, n, log2n, nb, STAGED && tw_staged<V>(n) ? reinterpret_cast<const V*>(at) : tw, INV);
}

// A workgroup's tile load, issued all at once.  A loop that stores each element
// to LDS as it arrives waits one full memory round trip per iteration; loading
// up to KPER elements per thread into registers first keeps them all in flight.
// load(t) and store(t, v) are called for t = threadIdx.x + q * kThreads < total.
template <int KPER, typename T, typename Load, typename Store>
__device__ inline void gather_tile(int total, Load&& load, Store&& store)
{
    T r[KPER];
#pragma unroll
    for (int q = 0; q < KPER; ++q) {
        const int t = (int)threadIdx.x + q * kThreads;
        if (t < total)
            r[q] = load(t);
    }
#pragma unroll
    for (int q = 0; q < KPER; ++q) {
        const int t = (int)threadIdx.x + q * kThreads;
        if (t < total)
            store(t, r[q]);
    }
}

// A tile load whose elements need arithmetic on the loaded values: fetch(t)
// only loads (raw values into registers), make(t, raw) converts them and stores
// to LDS.  In gather_tile, arithmetic inside the guarded load makes each
// element's loads wait there, one element at a time; here all of a round's loads
// are in flight together.  Rounds of KR elements per thread bound the registers;
// rounds past the tile are skipped (a uniform branch).
template <int KPER, int KR, typename R, typename Fetch, typename Make>
__device__ inline void gather_tile2(int total, Fetch&& fetch, Make&& make)
{
#pragma unroll
    for (int q0 = 0; q0 < KPER; q0 += KR) {
        if (q0 * kThreads >= total)
            break;
        R r[KR];
#pragma unroll
        for (int q = 0; q < KR; ++q) {
            const int t = (int)threadIdx.x + (q0 + q) * kThreads;
            if (t < total)
                r[q] = fetch(t);
        }
#pragma unroll
        for (int q = 0; q < KR; ++q) {
            const int t = (int)threadIdx.x + (q0 + q) * kThreads;
            if (t < total)
                make(t, r[q]);
        }
    }
}

template <typename T>
struct Pair {
    T a, b;
};

// ---------------------------------------------------------------------------
// Row pass: real rows y0 = 2p, y1 = 2p+1 of frame f -> G[f][y][kx], kx in [0, X/2]
// ---------------------------------------------------------------------------
// row-pair block bx of frame f; smem holds 2 ppw X elements of V
template <typename V, typename In>
__device__ __forceinline__ void rows_fwd_body(const In* __restrict__ frames, int X, int Y, int lx, int ppw,
    const V* __restrict__ tw, V* __restrict__ G, int hx, int absval, int bx, int64_t f, char* smem)
{
    using S = typename scalar_of<V>::type;
    // FP64 rows: one buffer, the in-place DIF (transform p at A[p * ld +
    // col_at(i)], X[k] at dif_pos(k)): half the LDS of a ping-pong pair, so more
    // workgroups fit a CU (the 512-point FP64 row passes of a 64-frame chunk took
    // 49-51 us instead of 73).  FP32 rows keep the Stockham pair, which measured
    // faster there.
    constexpr bool kDif = rows_dif<V>();
    const int ld = kDif ? col_ld<V>(X) : X;
    V* A = reinterpret_cast<V*>(smem);
    const int pair0 = bx * ppw;
    const In* fr = frames + f * (int64_t)X * Y;
    gather_tile2<kRowPer, kRowPer, Pair<In>>(
        ppw * X,
        [&](int t) {
            const int p = t >> lx, x = t & (X - 1);
            const int64_t y0 = 2 * (pair0 + p);
            return Pair<In> { fr[y0 * X + x], fr[(y0 + 1) * X + x] };
        },
        [&](int t, const Pair<In>& r) {
            const int at = kDif ? (t >> lx) * ld + col_at<V>(t & (X - 1)) : t;
            A[at] = mk<V>(sample_value<S>(r.a, absval), sample_value<S>(r.b, absval));
        });
    char* at = smem + sizeof(V) * ppw * ld * (kDif ? 1 : 2);
    (void)stage_twiddles<V>(tw, X, at);
    __syncthreads();
    const V* R = A;
    if constexpr (kDif)
        lds_fft_rows_dif<V, false, true>(A, X, lx, ppw, tw, at);
    else
        R = lds_fft_rows<V, false, true>(A, A + ppw * X, X, lx, ppw, tw, at);
    // position of element k of transform p
    auto pos = [&](int p, int k) { return kDif ? p * ld + col_at<V>(dif_pos(k, lx)) : p * X + k; };
    const S half = (S)0.5;
    // t / hx by a multiply with ceil(2^32 / hx): exact for t < 2^32 / hx
    const uint32_t hx_inv = 0xffffffffu / (uint32_t)hx + 1u;
    for (int t = threadIdx.x; t < ppw * hx; t += blockDim.x) {
        const int p = (int)__umulhi((uint32_t)t, hx_inv), k = t - p * hx;
        const V zk = R[pos(p, k)];
        const V zm = cconj(R[pos(p, (X - k) & (X - 1))]);
        const V s = cadd(zk, zm), dlt = cmul_mi(csub(zk, zm));
        const int64_t y0 = 2 * (pair0 + p);
        G[(f * Y + y0) * hx + k] = mk<V>(s.x * half, s.y * half);
        G[(f * Y + y0 + 1) * hx + k] = mk<V>(dlt.x * half, dlt.y * half);
    }
}

template <typename V, typename In>
__global__ __launch_bounds__(kThreads) void k_rows_fwd(const In* __restrict__ frames, int X, int Y, int lx, int ppw,
    const V* __restrict__ tw, V* __restrict__ G, int hx, int absval)
{
    extern __shared__ char smem[];
    rows_fwd_body<V, In>(frames, X, Y, lx, ppw, tw, G, hx, absval, blockIdx.x, blockIdx.y, smem);
}

// Column blocks of one row segment share 128-byte lines (a block holds 64 bytes
// of a row at the default tile), and workgroups go to the 8 XCDs, each with its
// own L2, round-robin in dispatch order.  The dispatch index is remapped so that
// each XCD takes one contiguous run of (frame, column block) indices: neighbours
// that share lines then share an L2.  A bijection on the first multiple of 8
// indices, the identity on the rest.
__device__ inline uint32_t xcd_remap(uint32_t d, uint32_t n)
{
    const uint32_t full = n & ~7u;
    return d < full ? (d & 7u) * (full >> 3) + (d >> 3) : d;
}
__device__ inline void xcd_tile(int& bx, int64_t& f)
{
    const uint32_t gx = gridDim.x, n = gx * gridDim.y;
    const uint32_t d = xcd_remap(blockIdx.x + blockIdx.y * gx, n);
    bx = (int)(d % gx);
    f = (int64_t)(d / gx);
}

// ---------------------------------------------------------------------------
// Column pass: G[f][y][kx] -> transform along y.
//   MODE 0: half-plane layout H[f][X/2 - kx][pad + r] = conj(R[(Y/2 - r) mod Y][kx])
//           (the spectrum at ux = -kx <= 0, uy = r - Y/2 of the real frame)
//   MODE 1: natural r2c layout R[f][ky][kx]
// ---------------------------------------------------------------------------
template <typename V, int MODE, typename Out>
__device__ __forceinline__ void cols_fwd_body(const V* __restrict__ G, int X, int Y, int ly, int hx, int cb,
    const V* __restrict__ tw, Out* __restrict__ out, HalfPlane hp, int bx, int64_t f, char* smem)
{
    const int lcb = __builtin_ctz(cb); // cb is a power of two
    const int ld = col_ld<V>(Y);
    V* A = reinterpret_cast<V*>(smem);
    const int kx0 = bx * cb;
    gather_tile<kColPer<V>, V>(
        cb * Y,
        [&](int t) {
            const int y = t >> lcb, c = t & (cb - 1);
            const int kx = kx0 + c;
            return kx < hx ? G[(f * Y + y) * hx + kx] : mk<V>(0, 0);
        },
        [&](int t, const V& v) {
            const int y = t >> lcb, c = t & (cb - 1);
            A[c * ld + col_at<V>(y)] = v;
        });
    char* at = smem + sizeof(V) * cb * ld;
    (void)stage_twiddles<V>(tw, Y, at);
    __syncthreads();
    lds_fft_cols<V, false, true>(A, Y, ly, cb, ld, tw, at);
    const V* R = A; // digit-reversed: X[k] at dif_pos(k)
    if (MODE == 0) {
        for (int t = threadIdx.x; t < cb * Y; t += blockDim.x) {
            const int c = t >> ly, r = t & (Y - 1);
            const int kx = kx0 + c;
            if (kx >= hx)
                continue;
            const V v = R[c * ld + col_at<V>(dif_pos((Y / 2 - r) & (Y - 1), ly))];
            Out o;
            o.x = v.x;
            o.y = -v.y;
            out[f * hp.fs + (int64_t)(X / 2 - kx) * hp.rs + hp.pad + r] = o;
        }
    } else {
        for (int t = threadIdx.x; t < cb * Y; t += blockDim.x) {
            const int ky = t >> lcb, c = t & (cb - 1);
            const int kx = kx0 + c;
            if (kx >= hx)
                continue;
            const V v = R[c * ld + col_at<V>(dif_pos(ky, ly))];
            Out o;
            o.x = v.x;
            o.y = v.y;
            out[(f * Y + ky) * hx + kx] = o;
        }
    }
}

template <typename V, int MODE, typename Out>
__global__ __launch_bounds__(kThreads) void k_cols_fwd(const V* __restrict__ G, int X, int Y, int ly, int hx, int cb,
    const V* __restrict__ tw, Out* __restrict__ out, HalfPlane hp)
{
    extern __shared__ char smem[];
    int bx;
    int64_t f;
    xcd_tile(bx, f);
    cols_fwd_body<V, MODE, Out>(G, X, Y, ly, hx, cb, tw, out, hp, bx, f, smem);
}

// P[c][r] += |H[f][c][pad + r]|^2 for f = 0..nf-1 in order (FP64), the power
// spectrum of smip-cli/main.cpp:274-279,300-305 on the stored half plane.
template <typename V>
__device__ __forceinline__ void power_body(const V* __restrict__ H, int64_t nf, const HalfPlane& hp,
    double* __restrict__ P, int64_t i)
{
    const int64_t n = (int64_t)hp.hx * hp.nrows;
    if (i >= n)
        return;
    const int c = (int)(i / hp.nrows), r = (int)(i - (int64_t)c * hp.nrows);
    const V* p = H + (int64_t)c * hp.rs + hp.pad + r;
    double acc = P[i];
    // eight frames' loads in flight at a time; the sum keeps frame order
    constexpr int kU = 8;
    int64_t f = 0;
    for (; f + kU <= nf; f += kU) {
        V v[kU];
#pragma unroll
        for (int q = 0; q < kU; ++q)
            v[q] = p[(f + q) * hp.fs];
#pragma unroll
        for (int q = 0; q < kU; ++q) {
            const double re = (double)v[q].x, im = (double)v[q].y;
            acc += re * re + im * im;
        }
    }
    for (; f < nf; ++f) {
        const V v = p[f * hp.fs];
        const double re = (double)v.x, im = (double)v.y;
        acc += re * re + im * im;
    }
    P[i] = acc;
}

template <typename V>
__global__ void k_power_acc(const V* __restrict__ H, int64_t nf, HalfPlane hp, double* __restrict__ P)
{
    power_body<V>(H, nf, hp, P, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// Column pass of the cross-correlation: C = conj(Rref) * R along ky, inverse
// transform, G[f][y][kx].  include/crosscorrel.h:126-141.
__global__ __launch_bounds__(kThreads) void k_xcorr_cols_inv(const double2* __restrict__ Rref,
    const double2* __restrict__ R, int X, int Y, int ly, int hx, int cb, const double2* __restrict__ tw,
    double2* __restrict__ G)
{
    const int lcb = __builtin_ctz(cb); // cb is a power of two
    const int ld = col_ld<double2>(Y);
    extern __shared__ char smem[];
    double2* A = reinterpret_cast<double2*>(smem);
    int bx;
    int64_t f;
    xcd_tile(bx, f);
    const int kx0 = bx * cb;
    gather_tile<kColPer<double2>, double2>(
        cb * Y,
        [&](int t) {
            const int y = t >> lcb, c = t & (cb - 1);
            const int kx = kx0 + c;
            double2 v = mk<double2>(0, 0);
            if (kx < hx) {
                const double2 a = Rref[(int64_t)y * hx + kx];
                const double2 b = R[(f * Y + y) * hx + kx];
                // std::conj(a) * b with the (ac - bd, ad + bc) operand order
                v = cmul(cconj(a), b);
            }
            return v;
        },
        [&](int t, const double2& v) {
            const int y = t >> lcb, c = t & (cb - 1);
            A[c * ld + col_at<double2>(y)] = v;
        });
    __syncthreads();
    lds_fft_cols<double2, true, false>(A, Y, ly, cb, ld, tw, nullptr);
    for (int t = threadIdx.x; t < cb * Y; t += blockDim.x) {
        const int y = t >> lcb, c = t & (cb - 1);
        const int kx = kx0 + c;
        if (kx < hx)
            G[(f * Y + y) * hx + kx] = A[c * ld + col_at<double2>(dif_pos(y, ly))];
    }
}

// The column passes of the cross-correlation in one launch: the forward
// transform of each column of G (as k_cols_fwd in MODE 1), C = conj(Rref) * R per
// element (as k_xcorr_cols_inv), and the inverse transform, written back over G.
// The spectrum stays in LDS: one read and one write of G instead of two each,
// with the same arithmetic in the same order as the two separate passes.
__device__ __forceinline__ void xcorr_cols_body(const double2* __restrict__ Rref, int X, int Y, int ly, int hx, int cb,
    const double2* __restrict__ tw, double2* __restrict__ G, int bx, int64_t f, char* smem)
{
    const int lcb = __builtin_ctz(cb); // cb is a power of two
    const int ld = col_ld<double2>(Y);
    constexpr int kPer = kColPer<double2>;
    double2* A = reinterpret_cast<double2*>(smem);
    const int kx0 = bx * cb;
    gather_tile<kPer, double2>(
        cb * Y,
        [&](int t) {
            const int y = t >> lcb, c = t & (cb - 1);
            const int kx = kx0 + c;
            return kx < hx ? G[(f * Y + y) * hx + kx] : mk<double2>(0, 0);
        },
        [&](int t, const double2& v) {
            const int y = t >> lcb, c = t & (cb - 1);
            A[c * ld + col_at<double2>(y)] = v;
        });
    char* at = smem + sizeof(double2) * cb * ld;
    (void)stage_twiddles<double2>(tw, Y, at);
    __syncthreads();
    lds_fft_cols<double2, false, true>(A, Y, ly, cb, ld, tw, at);
    // back to natural order through registers, multiplied on the way
    double2 v[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int t = threadIdx.x + q * kThreads;
        v[q] = mk<double2>(0, 0);
        if (t < cb * Y) {
            const int c = t >> ly, ky = t & (Y - 1);
            const int kx = kx0 + c;
            if (kx < hx) {
                const double2 a = Rref[(int64_t)ky * hx + kx];
                // std::conj(a) * b with the (ac - bd, ad + bc) operand order
                v[q] = cmul(cconj(a), A[c * ld + col_at<double2>(dif_pos(ky, ly))]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int t = threadIdx.x + q * kThreads;
        if (t < cb * Y)
            A[(t >> ly) * ld + col_at<double2>(t & (Y - 1))] = v[q];
    }
    __syncthreads();
    lds_fft_cols<double2, true, true>(A, Y, ly, cb, ld, tw, at);
    for (int t = threadIdx.x; t < cb * Y; t += blockDim.x) {
        const int y = t >> lcb, c = t & (cb - 1);
        const int kx = kx0 + c;
        if (kx < hx)
            G[(f * Y + y) * hx + kx] = A[c * ld + col_at<double2>(dif_pos(y, ly))];
    }
}

__global__ __launch_bounds__(kThreads) void k_xcorr_cols(const double2* __restrict__ Rref, int X, int Y, int ly,
    int hx, int cb, const double2* __restrict__ tw, double2* __restrict__ G)
{
    extern __shared__ char smem[];
    int bx;
    int64_t f;
    xcd_tile(bx, f);
    xcorr_cols_body(Rref, X, Y, ly, hx, cb, tw, G, bx, f, smem);
}

// Element k of H0 + i H1 for a c2r row pair: H0, H1 are the stored half
// spectra (kx in [0, X/2]) extended to the full length by Hermitian symmetry,
// with the imaginary parts of the kx = 0 and kx = X/2 terms dropped.  h0, h1 are
// the stored elements at kx = k for k <= X/2, at kx = X - k above.
__device__ inline double2 c2r_pair_input(double2 h0, double2 h1, int k, int X)
{
    const bool upper = k > X / 2;
    if (upper) {
        h0 = cconj(h0);
        h1 = cconj(h1);
    } else if (k == 0 || k == X / 2) {
        h0.y = 0.0;
        h1.y = 0.0;
    }
    return cadd(h0, cmul_pi(h1));
}

// Row pass of a c2r transform, two rows at a time: z = IDFT(H0 + i H1) gives
// row0 = Re z, row1 = Im z.  The imaginary parts of the kx = 0 and kx = X/2 terms
// are dropped, as a c2r transform of Hermitian data does.
__global__ __launch_bounds__(kThreads) void k_rows_c2r(const double2* __restrict__ G, int X, int Y, int lx, int ppw,
    int hx, const double2* __restrict__ tw, double* __restrict__ out)
{
    extern __shared__ char smem[];
    double2* A = reinterpret_cast<double2*>(smem);
    double2* B = A + ppw * X;
    const int64_t f = blockIdx.y;
    const int pair0 = blockIdx.x * ppw;
    gather_tile2<kRowPer, 4, Pair<double2>>(
        ppw * X,
        [&](int t) {
            const int p = t >> lx, k = t & (X - 1);
            const int64_t y0 = 2 * (pair0 + p);
            const double2* g0 = G + (f * Y + y0) * hx;
            const int i = k > X / 2 ? X - k : k;
            return Pair<double2> { g0[i], g0[hx + i] };
        },
        [&](int t, const Pair<double2>& r) { A[t] = c2r_pair_input(r.a, r.b, t & (X - 1), X); });
    __syncthreads();
    const double2* Z = lds_fft_rows<double2, true, false>(A, B, X, lx, ppw, tw, nullptr);
    for (int t = threadIdx.x; t < ppw * X; t += blockDim.x) {
        const int p = t >> lx, x = t & (X - 1);
        const int64_t y0 = 2 * (pair0 + p);
        out[(f * Y + y0) * X + x] = Z[t].x;
        out[(f * Y + y0 + 1) * X + x] = Z[t].y;
    }
}

// k_rows_c2r with calculate_displacement's search (include/crosscorrel.h:166-179)
// fused in: instead of writing its correlation rows, a workgroup writes the
// first maximum of its rows in row-major order to best[f][blockIdx.x], and
// k_best_to_shift merges those into the frame's first maximum.  The correlation
// never reaches memory.
__device__ inline void better_of(double& bv, int64_t& bi, double v, int64_t i)
{
    // (v, i) replaces (bv, bi) when it is larger, or equal with a lower index
    if (bv < v || (!(v < bv) && i < bi)) {
        bv = v;
        bi = i;
    }
}

// row-pair block bx of frame f; writes best[f nblk + bx]
__device__ __forceinline__ void rows_c2r_argmax_body(const double2* __restrict__ G, int X, int Y, int lx, int ppw,
    int hx, const double2* __restrict__ tw, RowBest* __restrict__ best, int bx, int64_t f, int nblk, char* smem)
{
    __shared__ double wv[kThreads / 64];
    __shared__ int64_t wi[kThreads / 64];
    // one buffer, as in rows_fwd_body: z(x) of pair p at A[p * ld + col_at(dif_pos(x))]
    const int ld = col_ld<double2>(X);
    double2* A = reinterpret_cast<double2*>(smem);
    const int pair0 = bx * ppw;
    gather_tile2<kRowPer, 4, Pair<double2>>(
        ppw * X,
        [&](int t) {
            const int p = t >> lx, k = t & (X - 1);
            const int64_t y0 = 2 * (pair0 + p);
            const double2* g0 = G + (f * Y + y0) * hx;
            const int i = k > X / 2 ? X - k : k;
            return Pair<double2> { g0[i], g0[hx + i] };
        },
        [&](int t, const Pair<double2>& r) {
            A[(t >> lx) * ld + col_at<double2>(t & (X - 1))] = c2r_pair_input(r.a, r.b, t & (X - 1), X);
        });
    char* at = smem + sizeof(double2) * ppw * ld;
    (void)stage_twiddles<double2>(tw, X, at);
    __syncthreads();
    lds_fft_rows_dif<double2, true, true>(A, X, lx, ppw, tw, at);
    // the scan starts from element 0 (max_element's first candidate) in the
    // workgroup that holds it, from nothing elsewhere
    double bv = -INFINITY;
    int64_t bi = INT64_MAX;
    if (bx == 0) {
        bv = A[0].x; // z(0) of pair 0: dif_pos(0) = 0
        bi = 0;
    }
    for (int t = threadIdx.x; t < ppw * X; t += blockDim.x) {
        const int p = t >> lx, x = t & (X - 1);
        const int64_t i0 = (2 * (int64_t)(pair0 + p)) * X + x; // Z.x is row y0, Z.y row y0 + 1
        const double2 z = A[p * ld + col_at<double2>(dif_pos(x, lx))];
        better_of(bv, bi, z.x, i0);
        better_of(bv, bi, z.y, i0 + X);
    }
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
        best[f * nblk + bx] = RowBest { bv, bi };
    }
}

__global__ __launch_bounds__(kThreads) void k_rows_c2r_argmax(const double2* __restrict__ G, int X, int Y, int lx,
    int ppw, int hx, const double2* __restrict__ tw, RowBest* __restrict__ best)
{
    extern __shared__ char smem[];
    rows_c2r_argmax_body(G, X, Y, lx, ppw, hx, tw, best, blockIdx.x, blockIdx.y, gridDim.x, smem);
}

// merges the workgroups' maxima of each frame in index order and applies the
// signed wrap x >= (n+1)/2 -> x - n
__device__ __forceinline__ void best_to_shift_body(const RowBest* __restrict__ best, int nblk, int64_t nf, int X, int Y,
    int32_t* __restrict__ shifts, int64_t f)
{
    // one wavefront per frame: each lane scans every 64th entry, then the lanes'
    // results merge; better_of keeps the larger value and, on ties, the lower
    // index, so the order of the merge does not change the result
    const int lane = threadIdx.x;
    if (f >= nf)
        return;
    const RowBest* b = best + f * nblk;
    double bv = -INFINITY;
    int64_t bi = INT64_MAX;
    for (int k = lane; k < nblk; k += 64)
        better_of(bv, bi, b[k].v, b[k].i);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(bv, o);
        const int64_t oi = __shfl_xor(bi, o);
        better_of(bv, bi, ov, oi);
    }
    if (lane != 0)
        return;
    int32_t sx = (int32_t)(bi % X), sy = (int32_t)(bi / X);
    if (!(sx < (X + 1) / 2))
        sx -= X;
    if (!(sy < (Y + 1) / 2))
        sy -= Y;
    shifts[2 * f] = sx;
    shifts[2 * f + 1] = sy;
}

__global__ __launch_bounds__(64) void k_best_to_shift(const RowBest* __restrict__ best, int nblk, int64_t nf, int X,
    int Y, int32_t* __restrict__ shifts)
{
    best_to_shift_body(best, nblk, nf, X, Y, shifts, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Grouped forms: the current batch of many jobs in one launch per stage (block
// -> job through the offset table, then the single-job body)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_group_rows_fwd(const GroupFwd* __restrict__ tab,
    const uint32_t* __restrict__ off, int n)
{
    extern __shared__ char smem[];
    const uint32_t g = blockIdx.x;
    const int j = group_job(off, n, g);
    const GroupFwd& J = tab[j];
    const uint32_t local = g - off[j];
    const uint32_t nbx = (uint32_t)((J.Y / 2) / J.ppw);
    const int bx = (int)(local % nbx);
    const int64_t f = local / nbx;
    switch (J.dtype) {
    case SMIP_DTYPE_U8:
        rows_fwd_body<float2, uint8_t>((const uint8_t*)global_ptr(J.frames), J.X, J.Y, J.lx, J.ppw, global_ptr(J.twx),
            global_ptr(J.G), J.hx, 0, bx, f, smem);
        break;
    case SMIP_DTYPE_U16:
        rows_fwd_body<float2, uint16_t>((const uint16_t*)global_ptr(J.frames), J.X, J.Y, J.lx, J.ppw, global_ptr(J.twx),
            global_ptr(J.G), J.hx, 0, bx, f, smem);
        break;
    default:
        rows_fwd_body<float2, float>((const float*)global_ptr(J.frames), J.X, J.Y, J.lx, J.ppw, global_ptr(J.twx),
            global_ptr(J.G), J.hx, 0, bx, f, smem);
    }
}

__global__ __launch_bounds__(kThreads) void k_group_cols_fwd(const GroupFwd* __restrict__ tab,
    const uint32_t* __restrict__ off, int n)
{
    extern __shared__ char smem[];
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const int j = group_job(off, n, g);
    const GroupFwd& J = tab[j];
    const uint32_t local = g - off[j];
    const uint32_t nbx = (uint32_t)((J.hx + J.cb - 1) / J.cb);
    cols_fwd_body<float2, 0, float2>(global_ptr(J.G), J.X, J.Y, J.ly, J.hx, J.cb, global_ptr(J.twy), global_ptr(J.H),
        J.hp, (int)(local % nbx), (int64_t)(local / nbx), smem);
}

__global__ void k_group_power(const GroupFwd* __restrict__ tab, const uint32_t* __restrict__ off, int n)
{
    const uint32_t g = blockIdx.x;
    const int j = group_job(off, n, g);
    const GroupFwd& J = tab[j];
    power_body<float2>(global_ptr(J.H), J.nf, J.hp, global_ptr(J.P), (int64_t)(g - off[j]) * blockDim.x + threadIdx.x);
}

void launch_group_fwd(const GroupFwd* tab, const uint32_t* off_rows, const uint32_t* off_cols,
    const uint32_t* off_pow, int n, uint32_t nb_rows, uint32_t nb_cols, uint32_t nb_pow, size_t lds_rows,
    size_t lds_cols, hipStream_t s)
{
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_group_rows_fwd, dim3(nb_rows), dim3(kThreads), lds_rows, s, tab, off_rows, n);
    SMIP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_group_cols_fwd, dim3(nb_cols), dim3(kThreads), lds_cols, s, tab, off_cols, n);
    SMIP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_group_power, dim3(nb_pow), dim3(256), 0, s, tab, off_pow, n);
    SMIP_HIP(hipGetLastError());
}

__global__ __launch_bounds__(kThreads) void k_group_saa_rows(const GroupSaa* __restrict__ tab,
    const uint32_t* __restrict__ off, int n)
{
    extern __shared__ char smem[];
    const uint32_t g = blockIdx.x;
    const int j = group_job(off, n, g);
    const GroupSaa& J = tab[j];
    const uint32_t local = g - off[j];
    const uint32_t nbx = (uint32_t)((J.Y / 2) / J.ppw);
    const int bx = (int)(local % nbx);
    const int64_t f = local / nbx;
    switch (J.dtype) {
    case SMIP_DTYPE_U8:
        rows_fwd_body<double2, uint8_t>((const uint8_t*)global_ptr(J.frames), J.X, J.Y, J.lx, J.ppw, global_ptr(J.twx),
            global_ptr(J.G), J.hx, 1, bx, f, smem);
        break;
    case SMIP_DTYPE_U16:
        rows_fwd_body<double2, uint16_t>((const uint16_t*)global_ptr(J.frames), J.X, J.Y, J.lx, J.ppw,
            global_ptr(J.twx), global_ptr(J.G), J.hx, 1, bx, f, smem);
        break;
    default:
        rows_fwd_body<double2, float>((const float*)global_ptr(J.frames), J.X, J.Y, J.lx, J.ppw, global_ptr(J.twx),
            global_ptr(J.G), J.hx, 1, bx, f, smem);
    }
}

__global__ __launch_bounds__(kThreads) void k_group_saa_cols(const GroupSaa* __restrict__ tab,
    const uint32_t* __restrict__ off, int n)
{
    extern __shared__ char smem[];
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const int j = group_job(off, n, g);
    const GroupSaa& J = tab[j];
    const uint32_t local = g - off[j];
    const uint32_t nbx = (uint32_t)((J.hx + J.cb - 1) / J.cb);
    xcorr_cols_body(global_ptr(J.ref), J.X, J.Y, J.ly, J.hx, J.cb, global_ptr(J.twy), global_ptr(J.G),
        (int)(local % nbx), (int64_t)(local / nbx), smem);
}

__global__ __launch_bounds__(kThreads) void k_group_saa_c2r(const GroupSaa* __restrict__ tab,
    const uint32_t* __restrict__ off, int n)
{
    extern __shared__ char smem[];
    const uint32_t g = blockIdx.x;
    const int j = group_job(off, n, g);
    const GroupSaa& J = tab[j];
    const uint32_t local = g - off[j];
    rows_c2r_argmax_body(global_ptr(J.G), J.X, J.Y, J.lx, J.ppw, J.hx, global_ptr(J.twx), global_ptr(J.best),
        (int)(local % (uint32_t)J.nblk), (int64_t)(local / (uint32_t)J.nblk), J.nblk, smem);
}

__global__ __launch_bounds__(64) void k_group_saa_merge(const GroupSaa* __restrict__ tab,
    const uint32_t* __restrict__ off, int n)
{
    const uint32_t g = blockIdx.x;
    const int j = group_job(off, n, g);
    const GroupSaa& J = tab[j];
    best_to_shift_body(global_ptr(J.best), J.nblk, J.nf, J.X, J.Y, global_ptr(J.shifts), (int64_t)(g - off[j]));
}

void launch_group_saa(const GroupSaa* tab, const uint32_t* off, int n, const uint32_t nb[5], size_t lds_rows,
    size_t lds_cols, hipStream_t s)
{
    if (n <= 0)
        return;
    const uint32_t* o = off;
    hipLaunchKernelGGL(k_group_saa_rows, dim3(nb[0]), dim3(kThreads), lds_rows, s, tab, o, n);
    SMIP_HIP(hipGetLastError());
    o += n + 1;
    hipLaunchKernelGGL(k_group_saa_cols, dim3(nb[1]), dim3(kThreads), lds_cols, s, tab, o, n);
    SMIP_HIP(hipGetLastError());
    o += n + 1;
    hipLaunchKernelGGL(k_group_saa_c2r, dim3(nb[2]), dim3(kThreads), lds_rows, s, tab, o, n);
    SMIP_HIP(hipGetLastError());
    o += n + 1;
    hipLaunchKernelGGL(k_group_saa_merge, dim3(nb[3]), dim3(64), 0, s, tab, o, n);
    SMIP_HIP(hipGetLastError());
    o += n + 1;
    launch_group_saa_add(tab, o, n, nb[4], s);
}

// Complex 2-D transform passes (in place), used by the synthesis.
__global__ __launch_bounds__(kThreads) void k_rows_c2c(double2* __restrict__ data, int X, int lx, int rpw,
    const double2* __restrict__ tw, int inv)
{
    extern __shared__ char smem[];
    double2* A = reinterpret_cast<double2*>(smem);
    double2* B = A + rpw * X;
    double2* base = data + (int64_t)blockIdx.x * rpw * X;
    gather_tile<kRowPer, double2>(
        rpw * X, [&](int t) { return base[t]; }, [&](int t, const double2& v) { A[t] = v; });
    __syncthreads();
    const double2* Z = inv ? lds_fft_rows<double2, true, false>(A, B, X, lx, rpw, tw, nullptr)
                           : lds_fft_rows<double2, false, false>(A, B, X, lx, rpw, tw, nullptr);
    for (int t = threadIdx.x; t < rpw * X; t += blockDim.x)
        base[t] = Z[t];
}

__global__ __launch_bounds__(kThreads) void k_cols_c2c(double2* __restrict__ data, int X, int Y, int ly, int cb,
    const double2* __restrict__ tw, int inv)
{
    const int lcb = __builtin_ctz(cb); // cb is a power of two
    const int ld = col_ld<double2>(Y);
    extern __shared__ char smem[];
    double2* A = reinterpret_cast<double2*>(smem);
    const int x0 = blockIdx.x * cb;
    gather_tile<kColPer<double2>, double2>(
        cb * Y,
        [&](int t) {
            const int y = t >> lcb, c = t & (cb - 1);
            return (x0 + c < X) ? data[(int64_t)y * X + x0 + c] : mk<double2>(0, 0);
        },
        [&](int t, const double2& v) {
            const int y = t >> lcb, c = t & (cb - 1);
            A[c * ld + col_at<double2>(y)] = v;
        });
    __syncthreads();
    if (inv)
        lds_fft_cols<double2, true, false>(A, Y, ly, cb, ld, tw, nullptr);
    else
        lds_fft_cols<double2, false, false>(A, Y, ly, cb, ld, tw, nullptr);
    for (int t = threadIdx.x; t < cb * Y; t += blockDim.x) {
        const int y = t >> lcb, c = t & (cb - 1);
        if (x0 + c < X)
            data[(int64_t)y * X + x0 + c] = A[c * ld + col_at<double2>(dif_pos(y, ly))];
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
template <typename V>
static int per_wg(int n, int cap, int budget = kLdsBytes)
{
    const int elems = budget / (2 * (int)sizeof(V));
    int k = elems / n;
    if (k < 1)
        fail(SMIP_ERR_INVALID_ARGUMENT, "transform length exceeds the LDS budget");
    if (k > cap)
        k = cap;
    // keep a power of two so every block gets the same count
    int p = 1;
    while (p * 2 <= k)
        p *= 2;
    return p;
}

// Columns per workgroup of a column pass (one LDS buffer: the in-place FFT):
// as many as kColLds holds, at most 64.
template <typename V>
static int cols_per_wg(int n)
{
    int k = kColLds / ((int)sizeof(V) * n);
    if (k < 1)
        fail(SMIP_ERR_INVALID_ARGUMENT, "transform length exceeds the LDS budget");
    k = std::min(k, 64);
    int p = 1;
    while (p * 2 <= k)
        p *= 2;
    // the column kernels take the compile-time transforms (and, for radix-8
    // lengths, their output order) only for this count
    if (p != col_count<V>(ilog2(n)))
        fail(SMIP_ERR_LOGIC, "column count differs from the compiled transforms");
    return p;
}

// Transforms per workgroup along contiguous rows: just enough that every radix-4
// stage keeps all 256 lanes busy (n/4 butterflies per transform), so small LDS
// footprints let several workgroups share a CU and hide each other's latency.
template <typename V>
static int rows_per_wg(int n, int cap)
{
    const int fill = std::max(1, (4 * kThreads) / n);
    const int r = std::min(per_wg<V>(n, cap), fill);
    if (r * n > kRowElems) // the row kernels' tile loads hold kRowPer elements per thread
        fail(SMIP_ERR_INVALID_ARGUMENT, "row tile exceeds the per-thread load budget");
    return r;
}

// LDS of a row pass (rows_fwd_body, rows_c2r_argmax_body) and its staged
// twiddles: ppw padded rows on the in-place DIF (FP64), a ping-pong pair of
// ppw rows on the Stockham form (FP32)
template <typename V>
static size_t rows_lds_bytes(int X, int ppw)
{
    const size_t data = rows_dif<V>() ? sizeof(V) * (size_t)ppw * col_ld<V>(X) : 2 * sizeof(V) * (size_t)ppw * X;
    return data + tw_lds_bytes<V>(X);
}

template <typename V>
static void launch_rows_fwd(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval, V* G,
    hipStream_t s)
{
    const int X = p->X, Y = p->Y, hx = X / 2 + 1;
    const int ppw = rows_per_wg<V>(X, Y / 2);
    dim3 grid((Y / 2) / ppw, (unsigned)nf);
    const size_t lds = rows_lds_bytes<V>(X, ppw);
    switch (dtype) {
    case SMIP_DTYPE_U8:
        hipLaunchKernelGGL((k_rows_fwd<V, uint8_t>), grid, dim3(kThreads), lds, s, (const uint8_t*)frames, X, Y, p->lx,
            ppw, tw_x<V>(p), G, hx, (int)absval);
        break;
    case SMIP_DTYPE_U16:
        hipLaunchKernelGGL((k_rows_fwd<V, uint16_t>), grid, dim3(kThreads), lds, s, (const uint16_t*)frames, X, Y,
            p->lx, ppw, tw_x<V>(p), G, hx, (int)absval);
        break;
    case SMIP_DTYPE_F32:
        hipLaunchKernelGGL((k_rows_fwd<V, float>), grid, dim3(kThreads), lds, s, (const float*)frames, X, Y, p->lx,
            ppw, tw_x<V>(p), G, hx, (int)absval);
        break;
    case SMIP_DTYPE_F64:
        hipLaunchKernelGGL((k_rows_fwd<V, double>), grid, dim3(kThreads), lds, s, (const double*)frames, X, Y, p->lx,
            ppw, tw_x<V>(p), G, hx, (int)absval);
        break;
    default:
        fail(SMIP_ERR_INVALID_ARGUMENT, "unsupported frame dtype");
    }
    SMIP_HIP(hipGetLastError());
}

template <typename V, int MODE, typename Out>
static void launch_cols_fwd(const FftPlan* p, const V* G, int64_t nf, Out* out, const HalfPlane& hp, hipStream_t s)
{
    const int X = p->X, Y = p->Y, hx = X / 2 + 1;
    const int cb = cols_per_wg<V>(Y);
    dim3 grid((hx + cb - 1) / cb, (unsigned)nf);
    const size_t lds = sizeof(V) * cb * col_ld<V>(Y) + tw_lds_bytes<V>(Y);
    hipLaunchKernelGGL((k_cols_fwd<V, MODE, Out>), grid, dim3(kThreads), lds, s, G, X, Y, p->ly, hx, cb, tw_y<V>(p),
        out, hp);
    SMIP_HIP(hipGetLastError());
}

template <typename V>
static void launch_power(const V* H, int64_t nf, const HalfPlane& hp, double* P, hipStream_t s)
{
    const int64_t n = (int64_t)hp.hx * hp.nrows;
    hipLaunchKernelGGL((k_power_acc<V>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, H, nf, hp, P);
    SMIP_HIP(hipGetLastError());
}

void launch_forward_f32(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval,
    const HalfPlane& hp, float2* out, float2* work, double* power_acc, hipStream_t s)
{
    if (!p->pow2) {
        mixed_forward_f32(p, frames, dtype, nf, absval, hp, out, work, s);
    } else {
        launch_rows_fwd<float2>(p, frames, dtype, nf, absval, work, s);
        launch_cols_fwd<float2, 0, float2>(p, work, nf, out, hp, s);
    }
    if (power_acc)
        launch_power<float2>(out, nf, hp, power_acc, s);
}

void launch_forward_f64(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval,
    const HalfPlane& hp, double2* out, double2* work, double* power_acc, hipStream_t s)
{
    if (!p->pow2) {
        mixed_forward_f64(p, frames, dtype, nf, absval, hp, out, work, s);
    } else {
        launch_rows_fwd<double2>(p, frames, dtype, nf, absval, work, s);
        launch_cols_fwd<double2, 0, double2>(p, work, nf, out, hp, s);
    }
    if (power_acc)
        launch_power<double2>(out, nf, hp, power_acc, s);
}

void launch_rfft_f64(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval, double2* out,
    double2* work, hipStream_t s)
{
    if (!p->pow2) {
        mixed_rfft_f64(p, frames, dtype, nf, absval, out, work, s);
        return;
    }
    HalfPlane hp {};
    launch_rows_fwd<double2>(p, frames, dtype, nf, absval, work, s);
    launch_cols_fwd<double2, 1, double2>(p, work, nf, out, hp, s);
}

void launch_shift_f64(const FftPlan* p, const void* frames, int dtype, int64_t nf, bool absval,
    const double2* ref_spec, RowBest* best, double2* work, int32_t* shifts, hipStream_t s)
{
    if (!p->pow2) {
        mixed_shift_f64(p, frames, dtype, nf, absval, ref_spec, best, work, shifts, s);
        return;
    }
    const int X = p->X, Y = p->Y, hx = X / 2 + 1;
    launch_rows_fwd<double2>(p, frames, dtype, nf, absval, work, s);
    const int cb = cols_per_wg<double2>(Y); // cb * Y * 16 <= kColLds, as k_xcorr_cols's registers assume
    hipLaunchKernelGGL(k_xcorr_cols, dim3((hx + cb - 1) / cb, (unsigned)nf), dim3(kThreads),
        sizeof(double2) * cb * col_ld<double2>(Y) + tw_lds_bytes<double2>(Y), s, ref_spec, X, Y, p->ly, hx, cb,
        p->twy64, work);
    SMIP_HIP(hipGetLastError());
    const int ppw = rows_per_wg<double2>(X, Y / 2);
    const int nblk = (Y / 2) / ppw;
    hipLaunchKernelGGL(k_rows_c2r_argmax, dim3(nblk, (unsigned)nf), dim3(kThreads),
        rows_lds_bytes<double2>(X, ppw), s, work, X, Y, p->lx, ppw, hx, p->twx64, best);
    SMIP_HIP(hipGetLastError());
    launch_best_to_shift(best, nblk, nf, X, Y, shifts, s);
}

void launch_best_to_shift(const RowBest* best, int nblk, int64_t nf, int X, int Y, int32_t* shifts, hipStream_t s)
{
    hipLaunchKernelGGL(k_best_to_shift, dim3((unsigned)nf), dim3(64), 0, s, best, nblk, nf, X, Y, shifts);
    SMIP_HIP(hipGetLastError());
}

int fft_shift_best_entries(const FftPlan* p)
{
    return p->pow2 ? p->Y / 2 : mixed_shift_blocks(p);
}

bool fft_plan_pow2(const FftPlan* p) { return p->pow2; }

FftTiling fft_tiling(const FftPlan* p)
{
    if (!p->pow2)
        fail(SMIP_ERR_LOGIC, "grouped transforms take power-of-two plans");
    FftTiling t;
    const int X = p->X, Y = p->Y;
    t.lx = p->lx;
    t.ly = p->ly;
    t.twx32 = p->twx32;
    t.twy32 = p->twy32;
    t.twx64 = p->twx64;
    t.twy64 = p->twy64;
    t.ppw32 = rows_per_wg<float2>(X, Y / 2);
    t.cb32 = cols_per_wg<float2>(Y);
    t.ppw64 = rows_per_wg<double2>(X, Y / 2);
    t.cb64 = cols_per_wg<double2>(Y);
    t.nblk64 = (Y / 2) / t.ppw64;
    t.lds_rows32 = rows_lds_bytes<float2>(X, t.ppw32);
    t.lds_cols32 = sizeof(float2) * t.cb32 * col_ld<float2>(Y) + tw_lds_bytes<float2>(Y);
    t.lds_rows64 = rows_lds_bytes<double2>(X, t.ppw64);
    t.lds_cols64 = sizeof(double2) * t.cb64 * col_ld<double2>(Y) + tw_lds_bytes<double2>(Y);
    return t;
}

void launch_xcorr_f64(const FftPlan* p, const double2* ref_spec, const double2* spec, int64_t nf, double* corr,
    double2* work, hipStream_t s)
{
    if (!p->pow2) {
        mixed_xcorr_f64(p, ref_spec, spec, nf, corr, work, s);
        return;
    }
    const int X = p->X, Y = p->Y, hx = X / 2 + 1;
    const int cb = cols_per_wg<double2>(Y);
    hipLaunchKernelGGL(k_xcorr_cols_inv, dim3((hx + cb - 1) / cb, (unsigned)nf), dim3(kThreads),
        sizeof(double2) * cb * col_ld<double2>(Y), s, ref_spec, spec, X, Y, p->ly, hx, cb, p->twy64, work);
    SMIP_HIP(hipGetLastError());
    const int ppw = rows_per_wg<double2>(X, Y / 2);
    hipLaunchKernelGGL(k_rows_c2r, dim3((Y / 2) / ppw, (unsigned)nf), dim3(kThreads), 2 * sizeof(double2) * ppw * X,
        s, work, X, Y, p->lx, ppw, hx, p->twx64, corr);
    SMIP_HIP(hipGetLastError());
}

void launch_c2c_inverse_f64(const FftPlan* p, double2* data, double2* work, hipStream_t s)
{
    (void)work;
    if (!p->pow2) {
        mixed_c2c_inverse_f64(p, data, s);
        return;
    }
    const int X = p->X, Y = p->Y;
    const int rpw = rows_per_wg<double2>(X, Y);
    hipLaunchKernelGGL(k_rows_c2c, dim3(Y / rpw), dim3(kThreads), 2 * sizeof(double2) * rpw * X, s, data, X, p->lx,
        rpw, p->twx64, 1);
    SMIP_HIP(hipGetLastError());
    const int cb = cols_per_wg<double2>(Y);
    hipLaunchKernelGGL(k_cols_c2c, dim3((X + cb - 1) / cb), dim3(kThreads), sizeof(double2) * cb * col_ld<double2>(Y), s, data, X,
        Y, p->ly, cb, p->twy64, 1);
    SMIP_HIP(hipGetLastError());
}

} // namespace smip_amd
