// MPD-RDME step kernels for gfx950 (CDNA4, wave64).
//
//  sweep_kernel     one diffusion sweep along x, y or z, one site per thread,
//                   any shape (mpd_{x,y,z}_kernel, MpdRdmeSolver.cu:1342-1490)
//  reaction_kernel  per-site reaction, one site per thread, any shape
//                   (precomp_reaction_kernel, MpdRdmeSolver.cu:1592-1675)
//  resolve_kernel   canonical-order overflow placement on the device (the
//                   classic host handler, MpdRdmeSolver.cu:1103-1189, or the
//                   relaxed correct_overflows, :1677-1761)
//  count_kernel     species counts (ByteLattice::getParticleCounts; the
//                   uint32 lattice's in count_words_kernel)
//
// Every per-particle decision depends only on (site index, slot, timestep,
// substep) through the counter-based hash, so the results do not depend on
// how sites are distributed over threads.
#include <algorithm>

#include "lmx_batch.h"
#include "lmx_device.h"
#include "lmx_kernels.h"
#include "lmx_site.h"

namespace lmx {

// ---------------------------------------------------------------------------
// Generic sweep: one site per thread, any lattice shape (the vector kernels in
// mpd_fast.hip need X % 4 == 0).
// ---------------------------------------------------------------------------
template <class L>
__global__ __launch_bounds__(256) void sweep_kernel(Params p, const uint32_t *__restrict__ src,
                                                    uint32_t *__restrict__ dst, uint32_t stride,
                                                    uint32_t len, uint32_t substep)
{
    constexpr int W = L::W;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.N) return;
    const unsigned long long h = timestep_hash(p.seed32, p.ctl->timestep, substep);
    const uint32_t c = (i / stride) % len;
    const bool has_m = (c > 0) || (p.boundary == 0);
    const bool has_p = (c + 1 < len) || (p.boundary == 0);
    const uint32_t im = (c > 0) ? i - stride : i + (len - 1) * stride;
    const uint32_t ip = (c + 1 < len) ? i + stride : i - (len - 1) * stride;

    // The site types and word 0 of the three sites first; word k >= 1 of a
    // site only if word k-1 leaves room for it (L::need_next: its last slot
    // occupied, or the input may be unpacked).
    const uint32_t sti = p.sites[i];
    const uint32_t stm = has_m ? p.sites[im] : 0u;
    const uint32_t stp = has_p ? p.sites[ip] : 0u;
    uint32_t wi[W], wm[W], wp[W], out[W];
    uint32_t any = 0;
#pragma unroll
    for (int k = 0; k < W; k++) {
        wi[k] = (k == 0 || L::need_next(p, wi[k > 0 ? k - 1 : 0])) ? src[(size_t)k * p.N + i] : 0u;
        wm[k] = has_m && (k == 0 || L::need_next(p, wm[k > 0 ? k - 1 : 0])) ? src[(size_t)k * p.N + im] : 0u;
        wp[k] = has_p && (k == 0 || L::need_next(p, wp[k > 0 ? k - 1 : 0])) ? src[(size_t)k * p.N + ip] : 0u;
        any |= wi[k] | wm[k] | wp[k];
    }
    if (any == 0) {
#pragma unroll
        for (int k = 0; k < W; k++) dst[(size_t)k * p.N + i] = 0u;
        return;
    }
    pull_site_l<L>(p, i, im, ip, wi, wm, wp, sti, stm, stp, has_m, has_p, h, substep, out);
#pragma unroll
    for (int k = 0; k < W; k++) dst[(size_t)k * p.N + i] = out[k];
}

// Generic reaction phase: one site per thread, in place.
template <class L>
__global__ __launch_bounds__(256) void reaction_kernel(Params p, uint32_t *__restrict__ lat)
{
    constexpr int W = L::W;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.N) return;
    const uint32_t st = p.sites[i];
    uint32_t w[W];
#pragma unroll
    for (int k = 0; k < W; k++) w[k] = (k == 0 || L::need_next(p, w[k > 0 ? k - 1 : 0])) ? lat[(size_t)k * p.N + i] : 0u;
    const unsigned long long h = timestep_hash(p.seed32, p.ctl->timestep, 3);
    if (react_site_l<L>(p, i, st, w, h)) {
#pragma unroll
        for (int k = 0; k < W; k++) lat[(size_t)k * p.N + i] = w[k];
    }
}

// ---------------------------------------------------------------------------
// Overflow resolution: one workgroup.  Entries are sorted into canonical
// order (phase, site, emission) and placed one after another; for each, the
// workgroup scans the reference's candidate list (Lattice::getNearbySites,
// Lattice.cpp:106-128: radius 0..4, shell max(R-1,0)^2 < d^2 <= R^2, x outer,
// z inner) 256 candidates at a time and takes the first site with room and
// the origin's site type.  The step counter advances at the end.
// ---------------------------------------------------------------------------
__device__ inline uint32_t load_agent(const uint32_t *ptr)
{
    return __hip_atomic_load(ptr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void store_agent(uint32_t *ptr, uint32_t v)
{
    __hip_atomic_store(ptr, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Particles up to the first empty slot (Lattice::getOccupancy).
__device__ inline uint32_t site_occupancy(const Params &p, const uint32_t *lat, uint32_t s)
{
    uint32_t occ = 0;
    for (uint32_t k = 0; k < p.W; k++) {
        const uint32_t word = load_agent(lat + (size_t)k * p.N + s);
        if (p.word_particles) {
            if (word == 0) return occ;
            occ++;
            continue;
        }
        for (uint32_t b = 0; b < 4; b++) {
            if (((word >> (8 * b)) & 0xFFu) == 0) return occ;
            occ++;
        }
    }
    return occ;
}

// A particle into slot occ (the site's first empty slot) of site s.
__device__ inline void place_particle(const Params &p, uint32_t *lat, uint32_t s, uint32_t occ, uint32_t particle)
{
    if (p.word_particles) {
        store_agent(lat + (size_t)occ * p.N + s, particle);
    } else {
        uint32_t *wp = lat + (size_t)(occ >> 2) * p.N + s;
        const uint32_t old = load_agent(wp);
        store_agent(wp, old | (particle << (8 * (occ & 3))));
        if ((occ >> 2) > 0) clear_chunk_flag(p, lat, occ >> 2, s >> 2);
    }
}

__device__ inline void set_error(Control *ctl, uint32_t code, uint32_t step)
{
    if (atomicCAS(&ctl->error, 0u, code) == 0u) ctl->error_step = step;
}

constexpr int kResolveThreads = 256;

__device__ void resolve_body(const Params &p, uint32_t *lat)
{
    __shared__ unsigned long long keys[kMaxOverflows];
    __shared__ int best;
    __shared__ uint32_t s_t;
    Control *ctl = p.ctl;
    const int tid = threadIdx.x;
    if (tid == 0) s_t = ctl->timestep;
    const uint32_t n = ctl->ovf_count;
    __syncthreads();
    // The step's reaction phase has consumed the reaction list.
    for (int k = tid; k < kListShards; k += (int)blockDim.x) ctl->list_count[k] = 0u;
    if (tid == 0) ctl->list_overflow = 0u;
    const uint32_t t = s_t;
    const uint32_t carried = p.relaxed ? ctl->carry_count : 0u;
    if (n == 0 && carried == 0) {
        if (tid == 0) ctl->timestep = t + 1;
        return;
    }
    if (n > kMaxOverflows || carried + n > kMaxOverflows) {
        if (tid == 0) {
            set_error(ctl, kErrOverflowCap, t);
            ctl->ovf_count = 0;
            ctl->timestep = t + 1;
        }
        return;
    }
    uint32_t M = 2;
    while (M < n) M <<= 1;
    for (uint32_t k = tid; k < M; k += kResolveThreads) keys[k] = (k < n) ? ctl->ovf_keys[k] : ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= M; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t a = tid; a < M; a += kResolveThreads) {
                const uint32_t b = a ^ j;
                if (b > a) {
                    const unsigned long long ka = keys[a], kb = keys[b];
                    const bool up = (a & k) == 0;
                    if ((ka > kb) == up) {
                        keys[a] = kb;
                        keys[b] = ka;
                    }
                }
            }
            __syncthreads();
        }
    }

    if (p.relaxed) {
        // Relaxed handler (correct_overflows, MpdRdmeSolver.cu:1677-1761): the
        // carried entries, then this step's in canonical order, each into its
        // own site's first empty slot; entries of one site go in that order
        // (the kernel's rounds) and different sites do not interact, so one
        // thread walks the list.  A full site keeps the entry for the next step.
        if (tid == 0) {
            uint32_t kept = 0;
            for (uint32_t e = 0; e < carried + n; e++) {
                uint32_t site, particle;
                if (e < carried) {
                    site = ctl->carry_site[e];
                    particle = ctl->carry_part[e];
                } else {
                    const unsigned long long key = keys[e - carried];
                    site = (uint32_t)((key >> 24) & 0xFFFFFFFFull);
                    particle = ctl->ovf_part[key & 0x7FFull];
                }
                const uint32_t occ = site_occupancy(p, lat, site);
                if (occ < p.P) {
                    place_particle(p, lat, site, occ, particle);
                } else {
                    ctl->carry_site[kept] = site;  // kept <= e: entry e has been read
                    ctl->carry_part[kept] = particle;
                    kept++;
                }
            }
            ctl->carry_count = kept;
            ctl->ovf_total += n;
            if (n > 0) ctl->ovf_steps += 1;
            ctl->ovf_count = 0;
            ctl->timestep = t + 1;
        }
        return;
    }

    const uint32_t X = p.X, Y = p.Y, Z = p.Z;
    bool failed = false;
    for (uint32_t e = 0; e < n && !failed; e++) {
        const unsigned long long key = keys[e];
        const uint32_t site = (uint32_t)((key >> 24) & 0xFFFFFFFFull);
        const uint32_t particle = ctl->ovf_part[key & 0x7FFull];
        const uint32_t xc = site % X, yc = (site / X) % Y, zc = site / p.XY;
        const uint32_t st0 = p.sites[site];
        int found = 0x7fffffff;
        for (uint32_t R = 0; R <= kMaxReplacementDist && found == 0x7fffffff; R++) {
            const uint32_t minD = R > 0 ? R - 1 : 0;
            const uint32_t x0 = xc >= R ? xc - R : 0, x1 = (xc <= X - 1 - R) ? xc + R : X - 1;
            const uint32_t y0 = yc >= R ? yc - R : 0, y1 = (yc <= Y - 1 - R) ? yc + R : Y - 1;
            const uint32_t z0 = zc >= R ? zc - R : 0, z1 = (zc <= Z - 1 - R) ? zc + R : Z - 1;
            const uint32_t ny = y1 - y0 + 1, nz = z1 - z0 + 1;
            const uint32_t total = (x1 - x0 + 1) * ny * nz;
            for (uint32_t base = 0; base < total; base += kResolveThreads) {
                if (tid == 0) best = 0x7fffffff;
                __syncthreads();
                const uint32_t c = base + tid;
                if (c < total) {
                    const uint32_t x = x0 + c / (ny * nz);
                    const uint32_t rem = c % (ny * nz);
                    const uint32_t y = y0 + rem / nz;
                    const uint32_t z = z0 + rem % nz;
                    const long long dx = (long long)x - xc, dy = (long long)y - yc, dz = (long long)z - zc;
                    const long long r2 = dx * dx + dy * dy + dz * dz;
                    const bool shell = r2 <= (long long)(R * R) && (r2 > (long long)(minD * minD) || R == 0);
                    if (shell) {
                        if (x >= X || y >= Y || z >= Z) {
                            atomicMin(&best, (int)(2 * c + 1));
                        } else {
                            const uint32_t s2 = x + X * y + p.XY * z;
                            if (p.sites[s2] == st0 && site_occupancy(p, lat, s2) < p.P)
                                atomicMin(&best, (int)(2 * c));
                        }
                    }
                }
                __syncthreads();
                const int b = best;
                __syncthreads();
                if (b != 0x7fffffff) {
                    found = b;
                    break;
                }
            }
            if (found != 0x7fffffff) {
                if (found & 1) {
                    if (tid == 0) set_error(ctl, kErrOutOfRange, t);
                    failed = true;
                } else if (tid == 0) {
                    const uint32_t c = (uint32_t)found >> 1;
                    const uint32_t x = x0 + c / (ny * nz);
                    const uint32_t rem = c % (ny * nz);
                    const uint32_t y = y0 + rem / nz;
                    const uint32_t z = z0 + rem % nz;
                    const uint32_t s2 = x + X * y + p.XY * z;
                    place_particle(p, lat, s2, site_occupancy(p, lat, s2), particle);
                }
            }
        }
        if (found == 0x7fffffff) {
            if (tid == 0) set_error(ctl, kErrUnplaceable, t);
            failed = true;
        }
        __syncthreads();
    }
    if (tid == 0) {
        ctl->ovf_total += n;
        ctl->ovf_steps += 1;
        ctl->ovf_count = 0;
        ctl->timestep = t + 1;
    }
}

__global__ __launch_bounds__(kResolveThreads) void resolve_kernel(Params p, uint32_t *lat)
{
    resolve_body(p, lat);
}

// Batched resolver: block r resolves member r's step on its post-step buffer
// (lat[1]) and swaps the member's two buffers for the next step.
__global__ __launch_bounds__(kResolveThreads) void batch_resolve_kernel(Member *__restrict__ m)
{
    Member &M = m[blockIdx.x];
    uint32_t *const before = M.lat[0], *const after = M.lat[1];
    resolve_body(M.p, after);
    __syncthreads();  // every wave has read both pointers (resolve_body returns block-uniformly)
    if (threadIdx.x == 0) {
        M.lat[0] = after;
        M.lat[1] = before;
        const uint32_t t = M.p.ctl->timestep;  // the next step's, written by this thread in resolve_body
        for (uint32_t s = 0; s < 3; s++) M.h[s] = timestep_hash(M.p.seed32, t, s);
    }
}

// ---------------------------------------------------------------------------
// Species counts (ByteLattice::getParticleCounts, ByteLattice.cpp:495-520):
// each word counts up to its first empty byte.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void count_kernel(const uint32_t *__restrict__ lat, size_t nwords,
                                                    uint32_t nS, uint32_t *__restrict__ counts)
{
    __shared__ uint32_t hist[256];
    for (uint32_t k = threadIdx.x; k < 256; k += blockDim.x) hist[k] = 0;
    __syncthreads();
    for (size_t wi = (size_t)blockIdx.x * blockDim.x + threadIdx.x; wi < nwords;
         wi += (size_t)gridDim.x * blockDim.x) {
        const uint32_t word = lat[wi];
        if (word == 0) continue;
        for (int b = 0; b < 4; b++) {
            const uint32_t s = (word >> (8 * b)) & 0xFFu;
            if (s == 0) break;
            atomicAdd(&hist[s], 1u);
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < 256; k += blockDim.x)
        if (k >= 1 && k <= nS && hist[k]) atomicAdd(&counts[k - 1], hist[k]);
}

// Species counts of the uint32 lattice (IntLattice::getParticleCounts,
// IntLattice.cpp:419-435): every nonzero word is one particle.
__global__ __launch_bounds__(256) void count_words_kernel(const uint32_t *__restrict__ lat, size_t nwords,
                                                          uint32_t nS, uint32_t *__restrict__ counts)
{
    for (size_t wi = (size_t)blockIdx.x * blockDim.x + threadIdx.x; wi < nwords;
         wi += (size_t)gridDim.x * blockDim.x) {
        const uint32_t s = lat[wi];
        if (s >= 1 && s <= nS) atomicAdd(&counts[s - 1], 1u);
    }
}

// Range check of a lattice changed on the device by a caller: particle
// types at most nS, site types below nST (the kernels index their tables with
// them).  A violation is a sticky error, reported by the next synchronize.
__global__ __launch_bounds__(256) void validate_kernel(Params p, const uint32_t *__restrict__ lat, size_t nwords)
{
    bool bad = false;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t wi = (size_t)blockIdx.x * blockDim.x + threadIdx.x; wi < nwords; wi += stride) {
        const uint32_t v = lat[wi];
        if (p.word_particles) {
            bad |= v > p.nS;
            continue;
        }
#pragma unroll
        for (int b = 0; b < 4; b++) bad |= ((v >> (8 * b)) & 0xFFu) > p.nS;
    }
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.N; i += stride) bad |= p.sites[i] >= p.nST;
    if (bad) set_error(p.ctl, kErrInvalidLattice, p.ctl->timestep);
}

hipError_t launch_validate(const Params &p, const uint32_t *lat, hipStream_t st)
{
    hipLaunchKernelGGL(validate_kernel, dim3(1024), dim3(256), 0, st, p, lat, (size_t)p.N * p.W);
    return hipGetLastError();
}

// STREAM copy: one 16-byte load and store per thread, one pass (no
// grid-stride loop: on gfx950 a flat grid of n/256 blocks measures ~6.2 TB/s
// where a grid-stride loop over 8-64 blocks per CU stays near 4.5-4.8 TB/s);
// the measured copy peak that the sweeps' HBM rates are quoted against beside
// the 8 TB/s figure.
__global__ __launch_bounds__(256) void copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

hipError_t launch_copy(const void *src, void *dst, size_t bytes, int num_cu, hipStream_t st)
{
    (void)num_cu;
    const size_t n = bytes / sizeof(uint4);
    const size_t nb = (n + 255) / 256;
    if (nb == 0 || nb > 0x7fffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)nb), dim3(256), 0, st, reinterpret_cast<const uint4 *>(src),
                       reinterpret_cast<uint4 *>(dst), n);
    return hipGetLastError();
}

// Fills n words with a value (used instead of a memset node, so that work
// enqueued on a stream someone captures holds kernels only).
__global__ __launch_bounds__(256) void fill_kernel(uint32_t *__restrict__ ptr, size_t n, uint32_t value)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) ptr[i] = value;
}

hipError_t launch_fill(uint32_t *ptr, size_t n, uint32_t value, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const size_t nb = std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)nb), dim3(256), 0, st, ptr, n, value);
    return hipGetLastError();
}

__global__ void exp_kernel(const double *x, double *y, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = exp_proj(x[i]);
}

__global__ void prob_kernel(const float *a, float *y, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = reaction_probability(a[i]);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

template <class L>
static void sweep_launch(const Params &p, const uint32_t *src, uint32_t *dst, uint32_t stride, uint32_t len,
                         int axis, hipStream_t st)
{
    hipLaunchKernelGGL(sweep_kernel<L>, dim3(blocks_for(p.N, 256)), dim3(256), 0, st, p, src, dst, stride, len,
                       (uint32_t)axis);
}

hipError_t launch_sweep(const Params &p, const uint32_t *src, uint32_t *dst, int axis, hipStream_t st)
{
    const uint32_t stride = axis == 0 ? 1u : (axis == 1 ? p.X : p.XY);
    const uint32_t len = axis == 0 ? p.X : (axis == 1 ? p.Y : p.Z);
    const int key = p.word_particles ? 100 + (int)p.P : (int)p.W;
    switch (key) {
    case 1: sweep_launch<ByteSites<1>>(p, src, dst, stride, len, axis, st); break;
    case 2: sweep_launch<ByteSites<2>>(p, src, dst, stride, len, axis, st); break;
    case 4: sweep_launch<ByteSites<4>>(p, src, dst, stride, len, axis, st); break;
    case 104: sweep_launch<WordSites<4>>(p, src, dst, stride, len, axis, st); break;
    case 108: sweep_launch<WordSites<8>>(p, src, dst, stride, len, axis, st); break;
    case 116: sweep_launch<WordSites<16>>(p, src, dst, stride, len, axis, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_reaction(const Params &p, uint32_t *lat, hipStream_t st)
{
    const dim3 grid(blocks_for(p.N, 256)), block(256);
    const int key = p.word_particles ? 100 + (int)p.P : (int)p.W;
    switch (key) {
    case 1: hipLaunchKernelGGL(reaction_kernel<ByteSites<1>>, grid, block, 0, st, p, lat); break;
    case 2: hipLaunchKernelGGL(reaction_kernel<ByteSites<2>>, grid, block, 0, st, p, lat); break;
    case 4: hipLaunchKernelGGL(reaction_kernel<ByteSites<4>>, grid, block, 0, st, p, lat); break;
    case 104: hipLaunchKernelGGL(reaction_kernel<WordSites<4>>, grid, block, 0, st, p, lat); break;
    case 108: hipLaunchKernelGGL(reaction_kernel<WordSites<8>>, grid, block, 0, st, p, lat); break;
    case 116: hipLaunchKernelGGL(reaction_kernel<WordSites<16>>, grid, block, 0, st, p, lat); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_resolve(const Params &p, uint32_t *lat, hipStream_t st)
{
    hipLaunchKernelGGL(resolve_kernel, dim3(1), dim3(kResolveThreads), 0, st, p, lat);
    return hipGetLastError();
}

hipError_t launch_batch_resolve(Member *m, uint32_t n, hipStream_t st)
{
    hipLaunchKernelGGL(batch_resolve_kernel, dim3(n), dim3(kResolveThreads), 0, st, m);
    return hipGetLastError();
}

hipError_t launch_counts(const uint32_t *lat, size_t nwords, uint32_t nS, bool words, uint32_t *counts,
                         hipStream_t st)
{
    unsigned g = blocks_for(nwords, 256);
    if (g > 4096) g = 4096;
    if (g == 0) g = 1;
    if (words)
        hipLaunchKernelGGL(count_words_kernel, dim3(g), dim3(256), 0, st, lat, nwords, nS, counts);
    else
        hipLaunchKernelGGL(count_kernel, dim3(g), dim3(256), 0, st, lat, nwords, nS, counts);
    return hipGetLastError();
}

hipError_t launch_exp(const double *x, double *y, size_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(exp_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, x, y, n);
    return hipGetLastError();
}

hipError_t launch_prob(const float *a, float *y, size_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(prob_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, a, y, n);
    return hipGetLastError();
}

}  // namespace lmx
