// Vector MPD-RDME kernels for gfx950 (wave64): each thread owns a group of 4
// consecutive x-sites, so every particle plane moves as 16-byte loads/stores
// and the site types as one 32-bit word.  Requires X % 4 == 0 (then a group
// never straddles a row, and every plane base stays 16-byte aligned).
//
//  sweep_x4_kernel     x sweep; the x-neighbours of a group's end sites come
//                      from the adjacent lanes through cross-lane shuffles
//                      (a 16-byte-aligned global load when the neighbour
//                      group is outside the wave)
//  sweep_march_kernel  y or z sweep; each thread walks a segment of rows along
//                      the axis keeping the previous, current and next rows
//                      (and one prefetched row) in registers, so each row is
//                      read once plus two apron rows per segment
//  reaction4_kernel    reaction phase, in place; groups with no particle and no
//                      zero-order propensity return without touching memory
//                      beyond their first load
//
// The per-site arithmetic is pull_site / react_site (lmx_site.h), shared with
// the one-site-per-thread kernels.
#include "lmx_device.h"
#include "lmx_kernels.h"
#include "lmx_site.h"

namespace lmx {

template <int W>
struct Group4 {
    uint4 w[W];   // plane k: sites 0..3 of the group in .x .y .z .w
    uint32_t t;   // site types of the 4 sites, byte k = site k
};

__device__ inline uint32_t lane_of(const uint4 &v, int k)
{
    return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

__device__ inline void set_lane(uint4 &v, int k, uint32_t x)
{
    if (k == 0) v.x = x;
    else if (k == 1) v.y = x;
    else if (k == 2) v.z = x;
    else v.w = x;
}

template <int W>
__device__ inline void load_group(const Params &p, const uint32_t *__restrict__ src, uint32_t g, Group4<W> &G)
{
#pragma unroll
    for (int k = 0; k < W; k++) G.w[k] = reinterpret_cast<const uint4 *>(src + (size_t)k * p.N)[g];
    G.t = reinterpret_cast<const uint32_t *>(p.sites)[g];
}

template <int W>
__device__ inline void zero_group(Group4<W> &G)
{
#pragma unroll
    for (int k = 0; k < W; k++) G.w[k] = make_uint4(0u, 0u, 0u, 0u);
    G.t = 0u;
}

template <int W>
__device__ inline uint32_t group_any(const Group4<W> &G)
{
    uint32_t a = 0;
#pragma unroll
    for (int k = 0; k < W; k++) a |= G.w[k].x | G.w[k].y | G.w[k].z | G.w[k].w;
    return a;
}

template <int W>
__device__ inline void site_words(const Group4<W> &G, int s, uint32_t *out)
{
#pragma unroll
    for (int k = 0; k < W; k++) out[k] = lane_of(G.w[k], s);
}

// ---------------------------------------------------------------------------
// x sweep
// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(256) void sweep_x4_kernel(Params p, const uint32_t *__restrict__ src,
                                                       uint32_t *__restrict__ dst)
{
    const uint32_t ngroups = p.N >> 2;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = g < ngroups;
    const uint32_t gc = active ? g : ngroups - 1;
    const int lane = (int)(threadIdx.x & 63u);
    const int wave_first = (int)g - lane;

    Group4<W> G;
    load_group<W>(p, src, gc, G);
    const uint32_t i0 = gc * 4u;
    const uint32_t x0 = i0 % p.X;
    const bool has_m = (x0 > 0) || (p.boundary == 0);
    const bool has_p = (x0 + 4 < p.X) || (p.boundary == 0);
    const uint32_t il = (x0 > 0) ? i0 - 1 : i0 + p.X - 1;
    const uint32_t ir = (x0 + 4 < p.X) ? i0 + 4 : i0 + 4 - p.X;
    const int ll = (int)(il >> 2) - wave_first;
    const int lr = (int)(ir >> 2) - wave_first;
    const bool shl = ll >= 0 && ll < 64;
    const bool shr = lr >= 0 && lr < 64;

    // Neighbour sites: the left neighbour is site 3 of its group, the right
    // neighbour site 0 of its group.
    uint32_t lw[W], rw[W], lt, rt;
#pragma unroll
    for (int k = 0; k < W; k++) {
        lw[k] = __shfl(G.w[k].w, shl ? ll : lane, 64);
        rw[k] = __shfl(G.w[k].x, shr ? lr : lane, 64);
    }
    lt = __shfl(G.t >> 24, shl ? ll : lane, 64);
    rt = __shfl(G.t & 0xFFu, shr ? lr : lane, 64);
    if (!active) return;
    if (!shl) {
#pragma unroll
        for (int k = 0; k < W; k++) lw[k] = src[(size_t)k * p.N + il];
        lt = p.sites[il];
    }
    if (!shr) {
#pragma unroll
        for (int k = 0; k < W; k++) rw[k] = src[(size_t)k * p.N + ir];
        rt = p.sites[ir];
    }
    if (!has_m) {
#pragma unroll
        for (int k = 0; k < W; k++) lw[k] = 0u;
    }
    if (!has_p) {
#pragma unroll
        for (int k = 0; k < W; k++) rw[k] = 0u;
    }
    uint32_t any = group_any<W>(G);
#pragma unroll
    for (int k = 0; k < W; k++) any |= lw[k] | rw[k];
    uint4 *out4[W];
#pragma unroll
    for (int k = 0; k < W; k++) out4[k] = reinterpret_cast<uint4 *>(dst + (size_t)k * p.N) + g;
    if (any == 0) {
#pragma unroll
        for (int k = 0; k < W; k++) *out4[k] = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    const unsigned long long h = timestep_hash(p.seed32, p.ctl->timestep, 0);
    uint4 res[W];
#pragma unroll
    for (int s = 0; s < 4; s++) {
        uint32_t wo[W], wm[W], wp[W], o[W];
        site_words<W>(G, s, wo);
        if (s == 0) {
#pragma unroll
            for (int k = 0; k < W; k++) wm[k] = lw[k];
        } else {
            site_words<W>(G, s - 1, wm);
        }
        if (s == 3) {
#pragma unroll
            for (int k = 0; k < W; k++) wp[k] = rw[k];
        } else {
            site_words<W>(G, s + 1, wp);
        }
        const uint32_t st = (G.t >> (8 * s)) & 0xFFu;
        const uint32_t stm = s == 0 ? lt : ((G.t >> (8 * (s - 1))) & 0xFFu);
        const uint32_t stp = s == 3 ? rt : ((G.t >> (8 * (s + 1))) & 0xFFu);
        const bool hm = s == 0 ? has_m : true;
        const bool hp = s == 3 ? has_p : true;
        const uint32_t i = i0 + s;
        const uint32_t im = s == 0 ? il : i - 1;
        const uint32_t ip = s == 3 ? ir : i + 1;
        pull_site<W>(p, i, im, ip, wo, wm, wp, st, stm, stp, hm, hp, h, 0, o);
#pragma unroll
        for (int k = 0; k < W; k++) set_lane(res[k], s, o[k]);
    }
#pragma unroll
    for (int k = 0; k < W; k++) *out4[k] = res[k];
}

// ---------------------------------------------------------------------------
// y / z sweep: march along the axis.
//   stride: X (y sweep) or X*Y (z sweep); len: Y or Z
//   ostride/nouter: the remaining axis (z planes for y, y rows for z)
// Thread t: group gx = t % (X/4), outer o = (t / (X/4)) % nouter,
// segment s = t / ((X/4) * nouter) covering rows [s*seg, min((s+1)*seg, len)).
// ---------------------------------------------------------------------------
template <int W>
__device__ inline void load_row(const Params &p, const uint32_t *__restrict__ src, uint32_t base, uint32_t stride,
                                uint32_t len, int r, Group4<W> &G)
{
    // r in [-1, len]; wraps periodically, empty beyond a reflecting edge.
    int rr = r;
    if (rr < 0) {
        if (p.boundary) { zero_group<W>(G); return; }
        rr += (int)len;
    } else if (rr >= (int)len) {
        if (p.boundary) { zero_group<W>(G); return; }
        rr -= (int)len;
    }
    load_group<W>(p, src, (base + (uint32_t)rr * stride) >> 2, G);
}

template <int W>
__global__ __launch_bounds__(256) void sweep_march_kernel(Params p, const uint32_t *__restrict__ src,
                                                          uint32_t *__restrict__ dst, uint32_t stride,
                                                          uint32_t len, uint32_t ostride, uint32_t nouter,
                                                          uint32_t seg, uint32_t substep)
{
    const uint32_t gx = p.X >> 2;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t g = t % gx;
    const uint32_t rest = t / gx;
    const uint32_t o = rest % nouter;
    const uint32_t sgi = rest / nouter;
    const uint32_t r0 = sgi * seg;
    if (r0 >= len) return;
    const uint32_t r1 = min(r0 + seg, len);
    const uint32_t base = o * ostride + g * 4u;
    const unsigned long long h = timestep_hash(p.seed32, p.ctl->timestep, substep);

    Group4<W> A, B, C, D;  // rows r-1, r, r+1, r+2
    zero_group<W>(D);
    load_row<W>(p, src, base, stride, len, (int)r0 - 1, A);
    load_row<W>(p, src, base, stride, len, (int)r0, B);
    load_row<W>(p, src, base, stride, len, (int)r0 + 1, C);
    for (uint32_t r = r0; r < r1; r++) {
        if (r + 1 < r1) load_row<W>(p, src, base, stride, len, (int)r + 2, D);
        const bool has_m = (r > 0) || (p.boundary == 0);
        const bool has_p = (r + 1 < len) || (p.boundary == 0);
        const uint32_t rowi = base + r * stride;
        const uint32_t rowm = (r > 0) ? rowi - stride : rowi + (len - 1) * stride;
        const uint32_t rowp = (r + 1 < len) ? rowi + stride : rowi - (len - 1) * stride;
        uint4 *out4[W];
#pragma unroll
        for (int k = 0; k < W; k++) out4[k] = reinterpret_cast<uint4 *>(dst + (size_t)k * p.N + rowi);
        const uint32_t any = group_any<W>(A) | group_any<W>(B) | group_any<W>(C);
        if (any == 0) {
#pragma unroll
            for (int k = 0; k < W; k++) *out4[k] = make_uint4(0u, 0u, 0u, 0u);
        } else {
            uint4 res[W];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                uint32_t wo[W], wm[W], wp[W], ov[W];
                site_words<W>(B, s, wo);
                site_words<W>(A, s, wm);
                site_words<W>(C, s, wp);
                const uint32_t st = (B.t >> (8 * s)) & 0xFFu;
                const uint32_t stm = (A.t >> (8 * s)) & 0xFFu;
                const uint32_t stp = (C.t >> (8 * s)) & 0xFFu;
                pull_site<W>(p, rowi + s, rowm + s, rowp + s, wo, wm, wp, st, stm, stp, has_m, has_p, h, substep, ov);
#pragma unroll
                for (int k = 0; k < W; k++) set_lane(res[k], s, ov[k]);
            }
#pragma unroll
            for (int k = 0; k < W; k++) *out4[k] = res[k];
        }
        A = B;
        B = C;
        C = D;
    }
}

// ---------------------------------------------------------------------------
// Reaction phase, 4 sites per thread, in place.
// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(256) void reaction4_kernel(Params p, uint32_t *__restrict__ lat)
{
    const uint32_t ngroups = p.N >> 2;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    uint4 w4[W];
#pragma unroll
    for (int k = 0; k < W; k++) w4[k] = reinterpret_cast<const uint4 *>(lat + (size_t)k * p.N)[g];
    uint32_t any = 0;
#pragma unroll
    for (int k = 0; k < W; k++) any |= w4[k].x | w4[k].y | w4[k].z | w4[k].w;
    if (any == 0 && !p.any_qp0) return;
    const uint32_t t4 = reinterpret_cast<const uint32_t *>(p.sites)[g];
    const unsigned long long h3 = timestep_hash(p.seed32, p.ctl->timestep, 3);
    bool changed = false;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        uint32_t w[W];
#pragma unroll
        for (int k = 0; k < W; k++) w[k] = lane_of(w4[k], s);
        uint32_t sany = 0;
#pragma unroll
        for (int k = 0; k < W; k++) sany |= w[k];
        if (sany == 0 && !p.any_qp0) continue;
        if (react_site<W>(p, g * 4u + s, (t4 >> (8 * s)) & 0xFFu, w, h3)) {
            changed = true;
#pragma unroll
            for (int k = 0; k < W; k++) set_lane(w4[k], s, w[k]);
        }
    }
    if (changed) {
#pragma unroll
        for (int k = 0; k < W; k++) reinterpret_cast<uint4 *>(lat + (size_t)k * p.N)[g] = w4[k];
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

bool fast_path_ok(const Params &p) { return (p.X & 3u) == 0 && p.N >= 4; }

hipError_t launch_sweep_fast(const Params &p, const uint32_t *src, uint32_t *dst, int axis, uint32_t seg,
                             hipStream_t st)
{
    const dim3 block(256);
    if (axis == 0) {
        const dim3 grid(blocks_for(p.N >> 2, 256));
        switch (p.W) {
        case 1: hipLaunchKernelGGL(sweep_x4_kernel<1>, grid, block, 0, st, p, src, dst); break;
        case 2: hipLaunchKernelGGL(sweep_x4_kernel<2>, grid, block, 0, st, p, src, dst); break;
        case 4: hipLaunchKernelGGL(sweep_x4_kernel<4>, grid, block, 0, st, p, src, dst); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    const uint32_t stride = axis == 1 ? p.X : p.XY;
    const uint32_t len = axis == 1 ? p.Y : p.Z;
    const uint32_t ostride = axis == 1 ? p.XY : p.X;
    const uint32_t nouter = axis == 1 ? p.Z : p.Y;
    if (seg == 0) seg = 1;
    const uint32_t nseg = (len + seg - 1) / seg;
    const size_t threads = (size_t)(p.X >> 2) * nouter * nseg;
    const dim3 grid(blocks_for(threads, 256));
    const uint32_t sub = (uint32_t)axis;
    switch (p.W) {
    case 1: hipLaunchKernelGGL(sweep_march_kernel<1>, grid, block, 0, st, p, src, dst, stride, len, ostride, nouter, seg, sub); break;
    case 2: hipLaunchKernelGGL(sweep_march_kernel<2>, grid, block, 0, st, p, src, dst, stride, len, ostride, nouter, seg, sub); break;
    case 4: hipLaunchKernelGGL(sweep_march_kernel<4>, grid, block, 0, st, p, src, dst, stride, len, ostride, nouter, seg, sub); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_reaction_fast(const Params &p, uint32_t *lat, hipStream_t st)
{
    const dim3 grid(blocks_for(p.N >> 2, 256)), block(256);
    switch (p.W) {
    case 1: hipLaunchKernelGGL(reaction4_kernel<1>, grid, block, 0, st, p, lat); break;
    case 2: hipLaunchKernelGGL(reaction4_kernel<2>, grid, block, 0, st, p, lat); break;
    case 4: hipLaunchKernelGGL(reaction4_kernel<4>, grid, block, 0, st, p, lat); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace lmx
