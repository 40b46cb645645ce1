// bispec.hip -- bispectrum accumulation and element access on gfx950.
//
// Replaces Bispectrum<T>::accumulate_from_fft (include/bispectrum.h:356-409),
// get_element / get_elements (:482-523) and the Array_base arithmetic used on the
// bispectrum (include/array_base.h:300-349; smip-cli/main.cpp:312).
//
// Two accumulation kernels share the reference layout of the result:
//  * k_accum_plain: one frame, one thread per stored cell, read-modify-write of
//    the FP64 (or FP32) cell in HBM -- the "plain" drop-in, HBM bound;
//  * k_accum_tile: a batch of frames, one wavefront per (ux, vx, 64 uy rows,
//    VYT vy values) tile.  Lanes own u = (ux, uy), registers hold VYT FP32
//    complex accumulators, F(v) is wave-uniform (scalar loads), F(u+v) comes from
//    a per-wave LDS strip of column ux+vx.  The FP32 sums are folded into the
//    FP64 bispectrum once per batch -- FP32 VALU bound.
// No MFMA: the triple product F(u)F(v)conj(F(u+v)) is not a dense contraction.
#include <algorithm>
#include <vector>

#include "smip_complex.h"
#include "smip_internal.h"

namespace smip_amd {

Desc make_desc(uint64_t ncols, uint64_t nrows, uint64_t b0, uint64_t b1)
{
    Desc d {};
    const uint64_t dims[4] = { ncols, nrows, b0, b1 };
    d.base_size = 1;
    d.totalsize = 1;
    for (int k = 0; k < 4; ++k) {
        d.dims[k] = dims[k];
        d.sizes[k] = (dims[k] / 2) * 2 + 1;
        d.base_sizes[k] = (k == 0 || k == 2) ? d.sizes[k] - d.sizes[k] / 2 : d.sizes[k];
        d.base_size *= d.base_sizes[k];
        d.totalsize *= d.sizes[k];
        d.min_idx[k] = -(int32_t)d.sizes[k] / 2;
        d.max_idx[k] = (int32_t)d.sizes[k] + d.min_idx[k] - 1;
    }
    return d;
}

LoopBounds loop_bounds(const Desc& d, int64_t ncols, int64_t nrows)
{
    // Array2 signed range, include/array2.h:675-688
    const int32_t fmn0 = -(int32_t)(ncols >> 1), fmn1 = -(int32_t)(nrows >> 1);
    const int32_t fmx0 = -(int32_t)ncols / 2 + (int32_t)ncols - 1;
    const int32_t fmx1 = -(int32_t)nrows / 2 + (int32_t)nrows - 1;
    LoopBounds L;
    L.min1 = std::max(fmn0, d.min_idx[0]);
    L.min2 = std::max(fmn1, d.min_idx[1]);
    L.min3 = std::max(fmn0, d.min_idx[2]);
    L.min4 = std::max(fmn1, d.min_idx[3]);
    L.max1 = std::min(fmx0, d.max_idx[0]);
    L.max2 = std::min(fmx1, d.max_idx[1]);
    L.max4 = std::min(fmx1, d.max_idx[3]);
    return L;
}

HalfPlane make_half_plane(int32_t ncols, int32_t nrows, int32_t pad, int32_t tail)
{
    HalfPlane h {};
    h.ncols = ncols;
    h.nrows = nrows;
    h.hx = (ncols >> 1) + 1;
    h.pad = pad;
    int rs = pad + nrows + tail;
    rs = (rs + 15) & ~15; // 128-byte aligned columns for FP64
    h.rs = rs;
    h.fs = (int64_t)h.hx * rs;
    return h;
}

// ---------------------------------------------------------------------------
// plain per-frame accumulation: one thread per stored cell (a0, a1, a2, a3)
// ---------------------------------------------------------------------------
struct PlainArgs {
    Desc d;
    LoopBounds L;
    HalfPlane hp;
    int32_t cx, cy; // ncols >> 1, nrows >> 1: column/row of index 0 in H
};

template <typename Acc>
__device__ inline const double2* hp_at(const double2* H, const PlainArgs& a, int32_t x, int32_t y)
{
    return H + (int64_t)(x + a.cx) * a.hp.rs + a.hp.pad + (y + a.cy);
}

template <typename T>
__global__ void k_accum_plain(PlainArgs a, const double2* __restrict__ H, T* __restrict__ B)
{
#pragma clang fp contract(off)
    const uint32_t inner = (uint32_t)(a.d.base_sizes[2] * a.d.base_sizes[3]);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= inner)
        return;
    const uint32_t bs1 = (uint32_t)a.d.base_sizes[1];
    const int32_t a0 = (int32_t)(blockIdx.y / bs1), a1 = (int32_t)(blockIdx.y - (uint32_t)a0 * bs1);
    const int32_t bs3 = (int32_t)a.d.base_sizes[3];
    const int32_t a2 = (int32_t)(t / (uint32_t)bs3), a3 = (int32_t)(t - (uint32_t)a2 * (uint32_t)bs3);
    const int32_t s1 = (int32_t)a.d.sizes[1], s3 = (int32_t)a.d.sizes[3];
    const int32_t i = -a0, j = a1 > s1 / 2 ? a1 - s1 : a1;
    const int32_t k = -a2, l = a3 > s3 / 2 ? a3 - s3 : a3;
    const LoopBounds& L = a.L;
    if (i < L.min1 || j < L.min2 || j > L.max2 || k < L.min3 || l < L.min4 || l > L.max4)
        return;
    if (i + k < L.min1 || j + l < L.min2 || i + k > L.max1 || j + l > L.max2)
        return;
    const double2 fu = *hp_at<T>(H, a, i, j);
    const double2 fv = *hp_at<T>(H, a, k, l);
    const double2 fw = *hp_at<T>(H, a, i + k, j + l);
    const uint64_t addr = calc_offset(a.d, i, j, k, l);
    // The products are written out here (not through cmul) so that the
    // contract(off) pragma covers them: (ac - bd, ad + bc) rounded per operation.
    if constexpr (sizeof(T) == sizeof(double2)) {
        // T t{F(u)}; t *= F(v); t *= conj(F(u+v)) in double
        const double tr = fu.x * fv.x - fu.y * fv.y;
        const double ti = fu.x * fv.y + fu.y * fv.x;
        const double cr = fw.x, ci = -fw.y;
        const double ur = tr * cr - ti * ci;
        const double ui = tr * ci + ti * cr;
        double2 b = B[addr];
        b.x += ur;
        b.y += ui;
        B[addr] = b;
    } else {
        // complex<float>: every factor narrowed to float before the multiply
        const float ar = (float)fu.x, ai = (float)fu.y;
        const float br = (float)fv.x, bi = (float)fv.y;
        const float cr = (float)fw.x, ci = (float)(-fw.y);
        const float tr = ar * br - ai * bi;
        const float ti = ar * bi + ai * br;
        const float ur = tr * cr - ti * ci;
        const float ui = tr * ci + ti * cr;
        float2 b = B[addr];
        b.x += ur;
        b.y += ui;
        B[addr] = b;
    }
}

template <typename T>
static void launch_plain(const Desc& d, const LoopBounds& lb, const HalfPlane& hp, const double2* H, T* B,
    hipStream_t s)
{
    PlainArgs a { d, lb, hp, hp.ncols >> 1, hp.nrows >> 1 };
    const uint64_t inner = d.base_sizes[2] * d.base_sizes[3];
    const uint64_t outer = d.base_sizes[0] * d.base_sizes[1];
    if (inner == 0 || outer == 0)
        return;
    if (outer > 0x7fffffffULL || inner > 0x7fffffffULL)
        fail(SMIP_ERR_INVALID_ARGUMENT, "bispectrum too large for the plain kernel grid");
    dim3 grid((unsigned)((inner + 255) / 256), (unsigned)outer);
    hipLaunchKernelGGL((k_accum_plain<T>), grid, dim3(256), 0, s, a, H, B);
    SMIP_HIP(hipGetLastError());
}

void launch_accum_plain_f64(const Desc& d, const LoopBounds& lb, const HalfPlane& hp, const double2* H,
    double2* B, hipStream_t s)
{
    launch_plain<double2>(d, lb, hp, H, B, s);
}

void launch_accum_plain_f32(const Desc& d, const LoopBounds& lb, const HalfPlane& hp, const double2* H, float2* B,
    hipStream_t s)
{
    launch_plain<float2>(d, lb, hp, H, B, s);
}

// ---------------------------------------------------------------------------
// batched tile accumulation
// ---------------------------------------------------------------------------
struct TilePlan {
    int vy_tile;
    int64_t nitems; // including padding items (cu < 0)
    int4* items;    // device: (c_u, vx, r0, vy0)
    TileArgs args;
    uint64_t cells; // cells computed per frame (including zero-padded edges)
};

int tile_plan_pad_needed(const Desc& d, int vy_tile)
{
    (void)vy_tile;
    return std::max<int>(d.max_idx[3], 0) + 1;
}

int tile_plan_tail_needed(const Desc& d, int nrows, int vy_tile)
{
    const int nblk = (nrows + 63) / 64;
    const int h1 = std::max<int>(d.max_idx[3], 0);
    // strip reads reach pad + (nblk*64 - 1) + (h1 + vy_tile - 1)
    return (nblk * 64 - nrows) + h1 + vy_tile + 1;
}

template <int VYT>
__device__ __forceinline__ void accum_tile_body(const int4 it, const TileArgs& a, const float2* __restrict__ H,
    int64_t nf, double2* __restrict__ B)
{
    __shared__ float2 strip[64 + VYT];
    const int cu = it.x, vx = it.y, r0 = it.z, vy0 = it.w;
    if (cu < 0)
        return;
    const int lane = threadIdx.x;
    const int r = r0 + lane;
    const int64_t off_u = (int64_t)cu * a.rs + a.pad + r;             // F(u)
    const int64_t off_s = (int64_t)(cu + vx) * a.rs + a.pad + r0 + vy0; // strip of F(u+v)
    const int64_t off_v = (int64_t)(vx + a.half_x) * a.rs + a.pad + a.half_y + vy0; // F(v)

    float ar[VYT], ai[VYT];
#pragma unroll
    for (int t = 0; t < VYT; ++t) {
        ar[t] = 0.f;
        ai[t] = 0.f;
    }
    float2 nu = H[off_u];
    float2 ns0 = H[off_s + lane];
    float2 ns1 = lane < VYT - 1 ? H[off_s + 64 + lane] : mk<float2>(0.f, 0.f);
    for (int64_t f = 0; f < nf; ++f) {
        const float2* Hf = H + f * a.fs;
        strip[lane] = ns0;
        if (lane < VYT - 1)
            strip[64 + lane] = ns1;
        const float2 u = nu;
        if (f + 1 < nf) {
            const float2* Hn = Hf + a.fs;
            nu = Hn[off_u];
            ns0 = Hn[off_s + lane];
            if (lane < VYT - 1)
                ns1 = Hn[off_s + 64 + lane];
        }
        __syncthreads();
        const float2* Fv = Hf + off_v;
#pragma unroll
        for (int t = 0; t < VYT; ++t) {
            const float2 v = Fv[t];
            const float2 w = strip[lane + t];
            // p = F(u) F(v); acc += p conj(F(u+v)): 2 mul + 6 fma per cell
            const float pr = fmaf(u.x, v.x, -(u.y * v.y));
            const float pi = fmaf(u.x, v.y, u.y * v.x);
            ar[t] = fmaf(pr, w.x, fmaf(pi, w.y, ar[t]));
            ai[t] = fmaf(pi, w.x, fmaf(-pr, w.y, ai[t]));
        }
        __syncthreads();
    }
    if (r >= a.Y)
        return;
    // fold into the FP64 bispectrum at calc_offset(ux, uy, vx, vy)
    const int uy = r - a.half_y;
    const int a0 = a.half_x - cu; // -ux
    const int a1 = uy < 0 ? uy + a.s1 : uy;
    const int64_t base = (((int64_t)a0 * a.bs1 + a1) * a.bs2 + (-vx)) * a.bs3;
#pragma unroll
    for (int t = 0; t < VYT; ++t) {
        const int vy = vy0 + t;
        if (vy < a.min4 || vy > a.max4)
            continue;
        const int a3 = vy < 0 ? vy + a.s3 : vy;
        double2* p = B + base + a3;
        double2 b = *p;
        b.x += (double)ar[t];
        b.y += (double)ai[t];
        *p = b;
    }
}

template <int VYT>
__global__ __launch_bounds__(64) void k_accum_tile(const int4* __restrict__ items, TileArgs a,
    const float2* __restrict__ H, int64_t nf, double2* __restrict__ B)
{
    accum_tile_body<VYT>(items[blockIdx.x], a, H, nf, B);
}

// The tile kernel over the items of many jobs: a job's items occupy a range of
// blocks starting at a multiple of 8, so each item runs on the XCD its position
// in the job's own list gives it.
template <int VYT>
__global__ __launch_bounds__(64) void k_group_tile(const GroupTile* __restrict__ tab, const uint32_t* __restrict__ off,
    int n)
{
    const uint32_t g = blockIdx.x;
    const int j = group_job(off, n, g);
    const GroupTile& J = tab[j];
    accum_tile_body<VYT>(load_const(J.items, g - off[j]), J.a, global_ptr(J.H), J.nf, global_ptr(J.B));
}

void launch_group_tile(int vy_tile, const GroupTile* tab, const uint32_t* off, int n, uint32_t nblocks, hipStream_t s)
{
    if (nblocks == 0 || n <= 0)
        return;
    switch (vy_tile) {
    case 9:
        hipLaunchKernelGGL(k_group_tile<9>, dim3(nblocks), dim3(64), 0, s, tab, off, n);
        break;
    case 13:
        hipLaunchKernelGGL(k_group_tile<13>, dim3(nblocks), dim3(64), 0, s, tab, off, n);
        break;
    case 17:
        hipLaunchKernelGGL(k_group_tile<17>, dim3(nblocks), dim3(64), 0, s, tab, off, n);
        break;
    case 25:
        hipLaunchKernelGGL(k_group_tile<25>, dim3(nblocks), dim3(64), 0, s, tab, off, n);
        break;
    case 33:
        hipLaunchKernelGGL(k_group_tile<33>, dim3(nblocks), dim3(64), 0, s, tab, off, n);
        break;
    default:
        fail(SMIP_ERR_LOGIC, "vy tile not instantiated");
    }
    SMIP_HIP(hipGetLastError());
}

static const int kVyTiles[] = { 9, 13, 17, 25, 33 };

// Default vy tile: least modelled time, chunks x tile width / (the width's
// throughput per computed cell x how fully the items fill the GPU).  Fitted to the
// kernel's measured rates at 128x128, 256x256 and 512x512 for b = 16..62:
// * per-cell throughput relative to 17 and 25: 0.9 for 9, 0.8 for 13, 0.7 for 33
//   (whose register footprint halves occupancy);
// * fill: items / 6144, at most 1 (smaller frames need narrow tiles to have
//   enough work items).
// Ties go to the wider tile.
int tile_default_vy(const LoopBounds& L, int X, int Y)
{
    const int nvy = L.max4 - L.min4 + 1;
    int64_t pairs = 0; // (ux, vx) columns of the reference loop
    for (int cu = 0; cu <= (X >> 1); ++cu)
        pairs += std::min(cu, -L.min3) + 1;
    const int64_t nblk = (Y + 63) / 64;
    int best = kVyTiles[0];
    double best_cost = 1e30;
    for (int c : kVyTiles) {
        const int64_t chunks = (nvy + c - 1) / c;
        const double eff = c == 9 ? 0.9 : c == 13 ? 0.8 : c == 33 ? 0.7 : 1.0;
        const double fill = std::min(1.0, (double)(pairs * nblk * chunks) / 6144.0);
        const double cost = (double)(chunks * c) / (eff * fill);
        if (cost < best_cost * (1.0 - 1e-9) || (cost <= best_cost * (1.0 + 1e-9) && c > best)) {
            best = c;
            best_cost = cost;
        }
    }
    return best;
}

int tile_plan_vy_tile(const TilePlan* tp) { return tp->vy_tile; }
uint64_t tile_plan_cells(const TilePlan* tp) { return tp->cells; }
const int4* tile_plan_items(const TilePlan* tp) { return tp->items; }
int64_t tile_plan_nitems(const TilePlan* tp) { return tp->nitems; }
const TileArgs& tile_plan_args(const TilePlan* tp) { return tp->args; }

TilePlan* tile_plan_create(const Desc& d, const LoopBounds& L, const HalfPlane& hp, int vy_tile, hipStream_t stream)
{
    const int nvy = L.max4 - L.min4 + 1;
    if (vy_tile <= 0)
        vy_tile = tile_default_vy(L, hp.ncols, hp.nrows);
    bool ok = false;
    for (int c : kVyTiles)
        ok |= (c == vy_tile);
    if (!ok)
        fail(SMIP_ERR_INVALID_ARGUMENT, "unsupported vy tile");
    const int X = hp.ncols, Y = hp.nrows;
    if (hp.pad < -L.min4 || hp.rs < hp.pad + ((Y + 63) / 64) * 64 + L.max4 + vy_tile)
        fail(SMIP_ERR_LOGIC, "half-plane padding too small for the tile kernel");
    // the tile kernel assumes the loop bounds of a full-size spectrum
    if (L.min1 != -(X >> 1) || L.min2 != -(Y >> 1) || L.max2 != Y - 1 - (Y >> 1))
        fail(SMIP_ERR_INVALID_ARGUMENT, "bispectrum extents do not match the frame extents");

    std::vector<int4> items;
    const int nblk = (Y + 63) / 64;
    const int nch = (nvy + vy_tile - 1) / vy_tile;
    uint64_t cells = 0;
    for (int cu = 0; cu <= (X >> 1); ++cu) // ux = cu - X/2 in [min1, 0]
        for (int vx = L.min3; vx <= 0; ++vx) {
            if (cu + vx < 0) // ux + vx >= min1
                continue;
            for (int b = 0; b < nblk; ++b)
                for (int c = 0; c < nch; ++c) {
                    items.push_back(make_int4(cu, vx, b * 64, L.min4 + c * vy_tile));
                    cells += (uint64_t)std::min(64, Y - b * 64) * std::min(vy_tile, nvy - c * vy_tile);
                }
        }
    // XCD-aware order: workgroup i runs on XCD i % 8, so give XCD k the k-th
    // contiguous slice of the (ux-major) item list; neighbouring ux columns share
    // spectrum columns in that XCD's L2.
    const int64_t n = (int64_t)items.size();
    const int64_t per = (n + 7) / 8;
    std::vector<int4> perm(per * 8, make_int4(-1, 0, 0, 0));
    for (int64_t i = 0; i < per * 8; ++i) {
        const int64_t src = (i % 8) * per + i / 8;
        if (src < n)
            perm[i] = items[src];
    }
    auto* tp = new TilePlan;
    tp->vy_tile = vy_tile;
    tp->nitems = (int64_t)perm.size();
    tp->cells = cells;
    tp->items = nullptr;
    try {
        tp->items = static_cast<int4*>(dev_malloc(sizeof(int4) * perm.size()));
        copy_sync(tp->items, perm.data(), sizeof(int4) * perm.size(), hipMemcpyHostToDevice, stream);
    } catch (...) {
        dev_free(tp->items);
        delete tp;
        throw;
    }
    TileArgs& a = tp->args;
    a.X = X;
    a.Y = Y;
    a.pad = hp.pad;
    a.rs = hp.rs;
    a.fs = hp.fs;
    a.half_x = X >> 1;
    a.half_y = Y >> 1;
    a.min4 = L.min4;
    a.max4 = L.max4;
    a.s1 = (int32_t)d.sizes[1];
    a.s3 = (int32_t)d.sizes[3];
    a.bs1 = (int32_t)d.base_sizes[1];
    a.bs2 = (int32_t)d.base_sizes[2];
    a.bs3 = (int32_t)d.base_sizes[3];
    return tp;
}

void tile_plan_destroy(TilePlan* tp)
{
    if (!tp)
        return;
    dev_free(tp->items);
    delete tp;
}

void launch_accum_tile(const TilePlan* tp, const HalfPlane& hp, const float2* H, int64_t nf, double2* B64,
    hipStream_t s)
{
    (void)hp;
    if (nf <= 0 || tp->nitems == 0)
        return;
    dim3 grid((unsigned)tp->nitems);
    switch (tp->vy_tile) {
    case 9:
        hipLaunchKernelGGL(k_accum_tile<9>, grid, dim3(64), 0, s, tp->items, tp->args, H, nf, B64);
        break;
    case 13:
        hipLaunchKernelGGL(k_accum_tile<13>, grid, dim3(64), 0, s, tp->items, tp->args, H, nf, B64);
        break;
    case 17:
        hipLaunchKernelGGL(k_accum_tile<17>, grid, dim3(64), 0, s, tp->items, tp->args, H, nf, B64);
        break;
    case 25:
        hipLaunchKernelGGL(k_accum_tile<25>, grid, dim3(64), 0, s, tp->items, tp->args, H, nf, B64);
        break;
    case 33:
        hipLaunchKernelGGL(k_accum_tile<33>, grid, dim3(64), 0, s, tp->items, tp->args, H, nf, B64);
        break;
    default:
        fail(SMIP_ERR_LOGIC, "vy tile not instantiated");
    }
    SMIP_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// element-wise arithmetic on the stored cells (Array_base operators)
// ---------------------------------------------------------------------------
template <typename V>
__global__ void k_scalar_op(V* __restrict__ B, uint64_t n, V z, int op)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    V b = B[i];
    switch (op) {
    case SMIP_OP_ADD:
        b = cadd(b, z);
        break;
    case SMIP_OP_SUB:
        b = csub(b, z);
        break;
    case SMIP_OP_MUL:
        b = cmul(b, z);
        break;
    default: // division by a complex scalar with zero imaginary part is exact
        if (z.y == 0) {
            b.x = b.x / z.x;
            b.y = b.y / z.x;
        } else {
            const auto den = z.x * z.x + z.y * z.y;
            b = mk<V>((b.x * z.x + b.y * z.y) / den, (b.y * z.x - b.x * z.y) / den);
        }
    }
    B[i] = b;
}

void launch_scale_c64(double2* B, uint64_t n, double re, double im, int op, hipStream_t s)
{
    if (!n)
        return;
    hipLaunchKernelGGL(k_scalar_op<double2>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, B, n,
        mk<double2>(re, im), op);
    SMIP_HIP(hipGetLastError());
}

void launch_scale_c32(float2* B, uint64_t n, float re, float im, int op, hipStream_t s)
{
    if (!n)
        return;
    hipLaunchKernelGGL(k_scalar_op<float2>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, B, n,
        mk<float2>(re, im), op);
    SMIP_HIP(hipGetLastError());
}

__global__ void k_c64_to_c32(const double2* __restrict__ in, float2* __restrict__ out, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = mk<float2>((float)in[i].x, (float)in[i].y);
}

void launch_convert_c64_to_c32(const double2* in, float2* out, uint64_t n, hipStream_t s)
{
    if (!n)
        return;
    hipLaunchKernelGGL(k_c64_to_c32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
    SMIP_HIP(hipGetLastError());
}

__global__ void k_c32_to_c64(const float2* __restrict__ in, double2* __restrict__ out, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = mk<double2>((double)in[i].x, (double)in[i].y);
}

void launch_convert_c32_to_c64(const float2* in, double2* out, uint64_t n, hipStream_t s)
{
    if (!n)
        return;
    hipLaunchKernelGGL(k_c32_to_c64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
    SMIP_HIP(hipGetLastError());
}

// get_element for a list of index tuples (device-side lookup with the reference's
// canonicalisation and address arithmetic, aliasing included).
__global__ void k_get_elements(Desc d, const void* __restrict__ B, int is_f32, const int32_t* __restrict__ idx,
    int64_t n, double2* __restrict__ out, int32_t* __restrict__ status)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint64_t addr = 0;
    bool conj = false;
    const int rc = resolve_element(d, idx[4 * i], idx[4 * i + 1], idx[4 * i + 2], idx[4 * i + 3], &addr, &conj);
    status[i] = rc;
    if (rc) {
        out[i] = mk<double2>(0, 0);
        return;
    }
    double2 v;
    if (is_f32) {
        const float2 f = static_cast<const float2*>(B)[addr];
        v = mk<double2>(f.x, f.y);
    } else {
        v = static_cast<const double2*>(B)[addr];
    }
    if (conj)
        v.y = -v.y;
    out[i] = v;
}

void launch_get_elements(const Desc& d, const void* B, int is_f32, const int32_t* idx, int64_t n, double2* out,
    int32_t* status, hipStream_t s)
{
    if (n <= 0)
        return;
    hipLaunchKernelGGL(k_get_elements, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, B, is_f32, idx, n, out,
        status);
    SMIP_HIP(hipGetLastError());
}

} // namespace smip_amd
