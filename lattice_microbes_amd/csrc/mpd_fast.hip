// Vector MPD-RDME kernels for gfx950 (wave64): each thread owns a group of 4
// consecutive x-sites, so every particle plane moves as 16-byte loads/stores
// and the site types as one 32-bit word.  Requires X % 4 == 0 (then a group
// never straddles a row, and every plane base stays 16-byte aligned).
//
// Particles are sparse (a few percent of sites), so the per-particle work is
// organised as per-lane loops over the set bits of an occupancy mask: a wave
// iterates as often as its busiest lane, not once per slot position any lane
// occupies.  Every particle's move decision is evaluated once per sweep:
//
//  sweep_x4_kernel     x sweep; a lane decides its own particles, then the
//                      decisions and words of the two neighbouring sites come
//                      from the adjacent lanes (cross-lane shuffles; decided
//                      locally when the neighbour group is outside the wave)
//  sweep_march_kernel  y or z sweep; each thread walks a segment of rows along
//                      the axis with rows r-1..r+2 in registers, decides row
//                      r+2 while writing row r, and carries the decisions
//  sweep_tile_kernel   y or z sweep over LDS row tiles (the default); the z
//                      sweep also lists its output sites that hold particles
//                      (reaction list)
//  reaction_list_kernel  reaction phase of the listed sites (default without
//                      zero-order reactions)
//  reaction4_kernel    reaction phase in place, looping over occupied sites;
//                      single-particle sites take their probability from the
//                      device-built table (react_site_table)
//
// Decision rule (makeDiffusionChoices, byte_diffusion_1d_dev.cu:86-121): a
// particle with uniform u moves minus iff u < qm and plus iff
// 0.5 <= u < qp + 0.5; assembly (performPropagation, :199-270): stayers, then
// the minus neighbour's plus movers, then the plus neighbour's minus movers,
// each in slot order, first P kept, the rest recorded as overflow.
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "lmx_batch.h"
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
    v.x = k == 0 ? x : v.x;
    v.y = k == 1 ? x : v.y;
    v.z = k == 2 ? x : v.z;
    v.w = k == 3 ? x : v.w;
}

// A group is read in two parts: plane 0 and the site types first, then each
// plane k >= 1 only if some site's word k-1 has its last slot occupied (or
// the input may be unpacked).  Callers issue every first part before any
// rest, so that the rests' wait on plane 0 does not hold back other loads.
template <int W>
__device__ inline void load_group_first(const Params &p, const uint32_t *__restrict__ src, uint32_t g,
                                        Group4<W> &G)
{
    G.w[0] = reinterpret_cast<const uint4 *>(src)[g];
#pragma unroll
    for (int k = 1; k < W; k++) G.w[k] = make_uint4(0u, 0u, 0u, 0u);
    G.t = reinterpret_cast<const uint32_t *>(p.sites)[g];
}

template <int W>
__device__ inline void load_group_rest(const Params &p, const uint32_t *__restrict__ src, uint32_t g,
                                       Group4<W> &G)
{
#pragma unroll
    for (int k = 1; k < W; k++) {
        const uint4 v = G.w[k - 1];
        if (need_next_byte_word(p, v.x | v.y | v.z | v.w))
            G.w[k] = reinterpret_cast<const uint4 *>(src + (size_t)k * p.N)[g];
    }
}

template <int W>
__device__ inline void load_group(const Params &p, const uint32_t *__restrict__ src, uint32_t g, Group4<W> &G)
{
    load_group_first<W>(p, src, g, G);
    load_group_rest<W>(p, src, g, G);
}

// Words of one site, plane k >= 1 read as in load_group.
// Stores a group's W planes from a sweep wave.  With chunk flags, a zero word
// of plane k >= 1 is not written if its chunk was known zero; whole_chunk
// (the wave's 64 lanes hold the 64 groups of one aligned chunk, inactive lanes
// only beyond the lattice) lets the wave also set the flag after writing only
// zeros.  Every lane of the wave that holds a group must call it.
// The chunk flags of group g's planes k >= 1 in dst (0 without flags).  A
// sweep reads them at the start of the group's work, so that their latency
// overlaps it rather than delaying the stores.  Within a sweep only the
// writers of a chunk change its flag: a wave that writes a whole chunk is its
// only writer, and otherwise a flag only goes from 1 (chunk known zero) to 0
// when another group of the chunk is written nonzero; so a 1 read early
// still says that this group is zero in dst.
template <int W>
__device__ inline void load_known(const Params &p, uint32_t *__restrict__ dst, uint32_t g, uint32_t *known)
{
#pragma unroll
    for (int k = 1; k < W; k++)
        known[k] = p.plane_flags ? __hip_atomic_load(chunk_flag(p, dst, k, g), __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)
                                 : 0u;
}

template <int W>
__device__ inline void store_group_known(const Params &p, uint32_t *__restrict__ dst, uint32_t g, const uint4 *v,
                                         bool whole_chunk, const uint32_t *known_k)
{
    reinterpret_cast<uint4 *>(dst)[g] = v[0];
#pragma unroll
    for (int k = 1; k < W; k++) {
        const bool nz = (v[k].x | v[k].y | v[k].z | v[k].w) != 0u;
        if (!p.plane_flags) {
            reinterpret_cast<uint4 *>(dst + (size_t)k * p.N)[g] = v[k];
            continue;
        }
        uint32_t *f = chunk_flag(p, dst, k, g);
        const uint32_t known = known_k[k];
        if (nz || known == 0u) reinterpret_cast<uint4 *>(dst + (size_t)k * p.N)[g] = v[k];
        if (whole_chunk) {
            const unsigned long long act = __ballot(1);
            const bool any_nz = __ballot(nz) != 0ull;
            const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
            if (lane == (uint32_t)__builtin_ctzll(act)) {
                if (any_nz && known != 0u) __hip_atomic_store(f, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (!any_nz && known == 0u) __hip_atomic_store(f, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        } else if (nz && known != 0u) {
            __hip_atomic_store(f, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <int W>
__device__ inline void store_group_sweep(const Params &p, uint32_t *__restrict__ dst, uint32_t g, const uint4 *v,
                                         bool whole_chunk)
{
    uint32_t known[W];
    load_known<W>(p, dst, g, known);
    store_group_known<W>(p, dst, g, v, whole_chunk, known);
}

template <int W>
__device__ inline void load_site(const Params &p, const uint32_t *__restrict__ src, uint32_t i, uint32_t *w)
{
    w[0] = src[i];
#pragma unroll
    for (int k = 1; k < W; k++) w[k] = need_next_byte_word(p, w[k - 1]) ? src[(size_t)k * p.N + i] : 0u;
}

template <int W>
__device__ inline void zero_group(Group4<W> &G)
{
#pragma unroll
    for (int k = 0; k < W; k++) G.w[k] = make_uint4(0u, 0u, 0u, 0u);
    G.t = 0u;
}

// 0x80 in each nonzero byte of v, 0 elsewhere.
__device__ inline uint32_t nonzero_hi(uint32_t v)
{
    return (((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v) & 0x80808080u;
}

// Bit b set iff byte b of v is nonzero (4 bits).  (A byte dot product can
// gather the four flags in one instruction, but measured slower here.)
__device__ inline uint32_t nonzero_bytes(uint32_t v)
{
    const uint32_t hi = nonzero_hi(v);
    return ((hi >> 7) & 1u) | ((hi >> 14) & 2u) | ((hi >> 21) & 4u) | ((hi >> 28) & 8u);
}

// Occupancy mask of a site's W words: bit j = slot j nonzero.
template <int W>
__device__ inline uint32_t site_occ(const uint32_t *w)
{
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < W; k++) m |= nonzero_bytes(w[k]) << (4 * k);
    return m;
}

// Occupancy of a group: bit s*P + j = slot j of site s nonzero.
template <int W>
__device__ inline unsigned long long group_occ(const Group4<W> &G)
{
    unsigned long long m = 0;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        uint32_t w[W];
#pragma unroll
        for (int k = 0; k < W; k++) w[k] = lane_of(G.w[k], s);
        m |= (unsigned long long)site_occ<W>(w) << (s * 4 * W);
    }
    return m;
}

// group_occ for a step input that may be packed: a packed site's particles
// fill slots 0..n-1, so its mask is the n lowest bits, with n counted from
// the highest nonzero byte of each word.
template <int W>
__device__ inline unsigned long long group_occ_in(const Params &p, const Group4<W> &G)
{
    if (!p.packed) return group_occ<W>(G);
    unsigned long long m = 0;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        uint32_t n = 0;
#pragma unroll
        for (int k = 0; k < W; k++) n += (39u - (uint32_t)__clz((int)lane_of(G.w[k], s))) >> 3;
        m |= (unsigned long long)((1u << n) - 1u) << (s * 4 * W);
    }
    return m;
}

template <int W>
__device__ inline void site_words(const Group4<W> &G, int s, uint32_t *out)
{
#pragma unroll
    for (int k = 0; k < W; k++) out[k] = lane_of(G.w[k], s);
}

// Slot j of a site (j may be a runtime value).
template <int W>
__device__ inline uint32_t slot_dyn(const uint32_t *w, uint32_t j)
{
    uint32_t word = w[0];
#pragma unroll
    for (int k = 1; k < W; k++) word = ((j >> 2) == (uint32_t)k) ? w[k] : word;
    return (word >> (8 * (j & 3u))) & 0xFFu;
}

// Slot j of site s of a group (s, j may be runtime values).
template <int W>
__device__ inline uint32_t group_slot(const Group4<W> &G, uint32_t s, uint32_t j)
{
    uint32_t word = 0;
#pragma unroll
    for (int k = 0; k < W; k++) {
        const uint32_t v = lane_of(G.w[k], (int)s);
        word = ((j >> 2) == (uint32_t)k) ? v : word;
    }
    return (word >> (8 * (j & 3u))) & 0xFFu;
}

// Lane s (0..3) of v.  SEL: two selects on the bits of s (a chain of
// equality tests compiles to branches, which cost the sweeps 4-6 us each).
// The fused step and the full reaction pass keep the chain: there the
// selects' live values cause register spills (fused step 36 -> 176 bytes
// per lane; the reaction kernels, which hold the full pass as their
// overflow fallback, 0 -> 144).
template <bool SEL = true>
__device__ inline uint32_t lane_dyn(const uint4 &v, uint32_t s)
{
    if constexpr (SEL) {
        const uint32_t lo = (s & 1u) ? v.y : v.x;
        const uint32_t hi = (s & 1u) ? v.w : v.z;
        return (s & 2u) ? hi : lo;
    } else {
        return s == 0 ? v.x : (s == 1 ? v.y : (s == 2 ? v.z : v.w));
    }
}

template <int W, bool SEL = true>
__device__ inline uint32_t group_slot_dyn(const Group4<W> &G, uint32_t s, uint32_t j)
{
    uint32_t word = 0;
#pragma unroll
    for (int k = 0; k < W; k++) {
        const uint32_t v = lane_dyn<SEL>(G.w[k], s);
        word = ((j >> 2) == (uint32_t)k) ? v : word;
    }
    return (word >> (8 * (j & 3u))) & 0xFFu;
}

// Decide every particle selected by pending.  Site s of the group has index
// base + s, type byte s of G.t, minus-neighbour type byte s of tm and
// plus-neighbour type byte s of tp; hm/hp (bit s) say whether that neighbour
// exists.  Bit b = s*P + j of minus/plus is set for a particle moving minus /
// plus.
// Tt is the transition table (p.T, or its copy in LDS).
template <int W, bool SEL = true>
__device__ inline void decide_t(const Params &p, const float *Tt, const Group4<W> &G, uint32_t tm, uint32_t tp,
                                uint32_t hm, uint32_t hp, uint32_t base, unsigned long long h,
                                unsigned long long pending, unsigned long long &minus, unsigned long long &plus)
{
    constexpr uint32_t P = 4 * W;
    constexpr uint32_t LOGP = W == 1 ? 2 : (W == 2 ? 3 : 4);
    unsigned long long mi = 0, pl = 0;
    while (pending) {
        const uint32_t b = (uint32_t)__builtin_ctzll(pending);
        pending &= pending - 1;
        const uint32_t s = b >> LOGP, j = b & (P - 1);
        const uint32_t part = group_slot_dyn<W, SEL>(G, s, j);
        const uint32_t st = (G.t >> (8 * s)) & 0xFFu;
        const uint32_t stm = (tm >> (8 * s)) & 0xFFu;
        const uint32_t stp = (tp >> (8 * s)) & 0xFFu;
        const float u = uniform01(base + s, p.pcb, j, h);
        const float qm = lookup_T_if(p, Tt, (hm >> s) & 1u, part, st, stm);
        const float qp = lookup_T_if(p, Tt, (hp >> s) & 1u, part, st, stp);
        if (u < qm) mi |= 1ull << b;
        else if (u >= 0.5f && u < (qp + 0.5f)) pl |= 1ull << b;
    }
    minus = mi;
    plus = pl;
}

// SEL as in lane_dyn: the P = 4 kernels and the register march and columns
// keep the chain (the selects cost them register spills).
template <int W, bool SEL = (W != 1)>
__device__ inline void decide(const Params &p, const Group4<W> &G, uint32_t tm, uint32_t tp, uint32_t hm,
                              uint32_t hp, uint32_t base, unsigned long long h, unsigned long long pending,
                              unsigned long long &minus, unsigned long long &plus)
{
    decide_t<W, SEL>(p, p.T, G, tm, tp, hm, hp, base, h, pending, minus, plus);
}

// New content of site i: own stayers, minus-source movers, plus-source movers.
// Particles beyond P are recorded as overflow when rec is set (a redundant
// copy of a site owned by another block drops them silently).
template <int W>
__device__ inline void assemble_rec(const Params &p, uint32_t i, uint32_t phase, bool rec, const uint32_t *own,
                                    uint32_t stay, uint32_t own_occ, const uint32_t *ms, uint32_t msel,
                                    const uint32_t *ps, uint32_t psel, uint32_t *out)
{
    // Nothing arrives or leaves and the site is packed: unchanged.
    if ((msel | psel) == 0 && stay == own_occ && (own_occ & (own_occ + 1)) == 0) {
#pragma unroll
        for (int k = 0; k < W; k++) out[k] = own[k];
        return;
    }
    SlotBuf<W> b;
    b.clear();
    while (stay) {
        const uint32_t j = (uint32_t)__builtin_ctz(stay);
        stay &= stay - 1;
        const uint32_t part = slot_dyn<W>(own, j);
        if (!b.push(part) && rec) record_overflow(p.ctl, phase, i, b.n - 1 - 4 * W, part);
    }
    while (msel) {
        const uint32_t j = (uint32_t)__builtin_ctz(msel);
        msel &= msel - 1;
        const uint32_t part = slot_dyn<W>(ms, j);
        if (!b.push(part) && rec) record_overflow(p.ctl, phase, i, b.n - 1 - 4 * W, part);
    }
    while (psel) {
        const uint32_t j = (uint32_t)__builtin_ctz(psel);
        psel &= psel - 1;
        const uint32_t part = slot_dyn<W>(ps, j);
        if (!b.push(part) && rec) record_overflow(p.ctl, phase, i, b.n - 1 - 4 * W, part);
    }
#pragma unroll
    for (int k = 0; k < W; k++) out[k] = b.w[k];
}

template <int W>
__device__ inline void assemble(const Params &p, uint32_t i, uint32_t phase, const uint32_t *own, uint32_t stay,
                                uint32_t own_occ, const uint32_t *ms, uint32_t msel, const uint32_t *ps,
                                uint32_t psel, uint32_t *out)
{
    assemble_rec<W>(p, i, phase, true, own, stay, own_occ, ms, msel, ps, psel, out);
}

template <int W>
__device__ inline uint32_t site_bits(unsigned long long m, int s)
{
    constexpr uint32_t P = 4 * W;
    return (uint32_t)(m >> (s * P)) & ((1u << P) - 1u);
}

// ---------------------------------------------------------------------------
// x sweep
// ---------------------------------------------------------------------------
// Body of the x sweep for 256-thread block number `block` (the kernel below,
// or one replicate's share of a batched launch); h is the step's x-sweep hash
// (timestep_hash substep 0), which a batched block reads from its member's
// descriptor for all the units it loops over.
template <int W>
__device__ inline void x4_body(const Params &p, const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                               unsigned long long h, uint32_t block)
{
    constexpr uint32_t P = 4 * W;
    const uint32_t ngroups = p.N >> 2;
    const uint32_t g = block * 256u + threadIdx.x;
    const bool active = g < ngroups;
    const uint32_t gc = active ? g : ngroups - 1;
    const int lane = (int)(threadIdx.x & 63u);
    const int wave_first = (int)g - lane;

    Group4<W> G;
    load_group_first<W>(p, src, gc, G);
    uint32_t known[W];  // the output group's chunk flags, read early: their latency overlaps the decisions
    if (active) load_known<W>(p, dst, g, known);
    const uint32_t i0 = gc * 4u;
    const uint32_t x0 = (p.X & (p.X - 1u)) == 0u ? (i0 & (p.X - 1u)) : i0 % p.X;
    const bool has_m = (x0 > 0) || (p.boundary == 0);
    const bool has_p = (x0 + 4 < p.X) || (p.boundary == 0);
    const uint32_t il = (x0 > 0) ? i0 - 1 : i0 + p.X - 1;
    const uint32_t ir = (x0 + 4 < p.X) ? i0 + 4 : i0 + 4 - p.X;
    const int ll = (int)(il >> 2) - wave_first;
    const int lr = (int)(ir >> 2) - wave_first;
    const bool shl = ll >= 0 && ll < 64;
    const bool shr = lr >= 0 && lr < 64;
    const int sl = shl ? ll : lane, sr = shr ? lr : lane;

    // Words and types of the neighbour sites (left: site 3 of its group,
    // right: site 0 of its group): from the neighbouring lanes, or read at
    // the wave's edges (word 0 and type with the group's first part).
    uint32_t lw0 = 0u, rw0 = 0u, ltd = 0u, rtd = 0u;
    if (active && !shl) {
        lw0 = src[il];
        ltd = p.sites[il];
    }
    if (active && !shr) {
        rw0 = src[ir];
        rtd = p.sites[ir];
    }
    load_group_rest<W>(p, src, gc, G);
    uint32_t lw[W], rw[W];
#pragma unroll
    for (int k = 0; k < W; k++) {
        lw[k] = __shfl(G.w[k].w, sl, 64);
        rw[k] = __shfl(G.w[k].x, sr, 64);
    }
    uint32_t lt = __shfl(G.t >> 24, sl, 64);
    uint32_t rt = __shfl(G.t & 0xFFu, sr, 64);
    if (active && !shl) {
        lw[0] = lw0;
#pragma unroll
        for (int k = 1; k < W; k++) lw[k] = need_next_byte_word(p, lw[k - 1]) ? src[(size_t)k * p.N + il] : 0u;
        lt = ltd;
    }
    if (active && !shr) {
        rw[0] = rw0;
#pragma unroll
        for (int k = 1; k < W; k++) rw[k] = need_next_byte_word(p, rw[k - 1]) ? src[(size_t)k * p.N + ir] : 0u;
        rt = rtd;
    }
    if (!has_m) {
#pragma unroll
        for (int k = 0; k < W; k++) lw[k] = 0u;
    }
    if (!has_p) {
#pragma unroll
        for (int k = 0; k < W; k++) rw[k] = 0u;
    }

    const unsigned long long occ = active ? group_occ_in<W>(p, G) : 0ull;
    unsigned long long minus = 0, plus = 0;
    if (occ) {
        const uint32_t tm = (G.t << 8) | lt;
        const uint32_t tp = (G.t >> 8) | (rt << 24);
        decide<W>(p, G, tm, tp, has_m ? 0xFu : 0xEu, has_p ? 0xFu : 0x7u, i0, h, occ, minus, plus);
    }
    // Plus movers of the left neighbour and minus movers of the right one.
    uint32_t lplus = __shfl(site_bits<W>(plus, 3), sl, 64);
    uint32_t rminus = __shfl(site_bits<W>(minus, 0), sr, 64);
    if (!active) return;
    const uint32_t st0 = G.t & 0xFFu, st3 = G.t >> 24;
    if (!shl) {
        lplus = 0;
        uint32_t m = site_occ<W>(lw);
        while (m) {
            const uint32_t j = (uint32_t)__builtin_ctz(m);
            m &= m - 1;
            const uint32_t part = slot_dyn<W>(lw, j);
            const float u = uniform01(il, p.pcb, j, h);
            const float qp = lookup_T(p, part, lt, st0);
            if (u >= 0.5f && u < (qp + 0.5f)) lplus |= 1u << j;
        }
    }
    if (!shr) {
        rminus = 0;
        uint32_t m = site_occ<W>(rw);
        while (m) {
            const uint32_t j = (uint32_t)__builtin_ctz(m);
            m &= m - 1;
            const uint32_t part = slot_dyn<W>(rw, j);
            const float u = uniform01(ir, p.pcb, j, h);
            const float qm = lookup_T(p, part, rt, st3);
            if (u < qm) rminus |= 1u << j;
        }
    }
    if (!has_m) lplus = 0;
    if (!has_p) rminus = 0;

    uint4 res[W];
#pragma unroll
    for (int k = 0; k < W; k++) res[k] = make_uint4(0u, 0u, 0u, 0u);
    const uint32_t any = (uint32_t)(occ != 0) | lplus | rminus;
    if (any != 0)
#pragma unroll
    for (int s = 0; s < 4; s++) {
        uint32_t wo[W], wm[W], wp[W], o[W];
        site_words<W>(G, s, wo);
        const uint32_t own_occ = site_bits<W>(occ, s);
        const uint32_t stay = own_occ & ~(site_bits<W>(minus, s) | site_bits<W>(plus, s));
        uint32_t msel, psel;
        if (s == 0) {
#pragma unroll
            for (int k = 0; k < W; k++) wm[k] = lw[k];
            msel = lplus;
        } else {
            site_words<W>(G, s - 1, wm);
            msel = site_bits<W>(plus, s - 1);
        }
        if (s == 3) {
#pragma unroll
            for (int k = 0; k < W; k++) wp[k] = rw[k];
            psel = rminus;
        } else {
            site_words<W>(G, s + 1, wp);
            psel = site_bits<W>(minus, s + 1);
        }
        assemble<W>(p, i0 + s, 0, wo, stay, own_occ, wm, msel, wp, psel, o);
#pragma unroll
        for (int k = 0; k < W; k++) set_lane(res[k], s, o[k]);
    }
    // A wave holds 64 consecutive groups starting at a multiple of 64.
    store_group_known<W>(p, dst, g, res, true, known);
    (void)P;
}

template <int W>
__global__ __launch_bounds__(256) void sweep_x4_kernel(Params p, const uint32_t *__restrict__ src,
                                                       uint32_t *__restrict__ dst)
{
    x4_body<W>(p, src, dst, timestep_hash(p.seed32, p.ctl->timestep, 0), blockIdx.x);
}

// ---------------------------------------------------------------------------
// y / z sweep: march along the axis.
//   stride: X (y sweep) or X*Y (z sweep); len: Y or Z
//   ostride/nouter: the remaining axis (z planes for y, y rows for z)
// Thread t: group gx = t % (X/4), outer o = (t / (X/4)) % nouter,
// segment s = t / ((X/4) * nouter) covering rows [s*seg, min((s+1)*seg, len)).
// ---------------------------------------------------------------------------
template <int W>
__device__ inline bool load_row(const Params &p, const uint32_t *__restrict__ src, uint32_t base, uint32_t stride,
                                uint32_t len, int r, Group4<W> &G)
{
    // r in [-1, len+1]; wraps periodically, empty beyond a reflecting edge.
    int rr = r;
    if (rr < 0) {
        if (p.boundary) { zero_group<W>(G); return false; }
        rr += (int)len;
    } else if (rr >= (int)len) {
        if (p.boundary) { zero_group<W>(G); return false; }
        rr -= (int)len;
        if (rr >= (int)len) rr -= (int)len;  // r = len + 1 with len = 1
    }
    load_group<W>(p, src, (base + (uint32_t)rr * stride) >> 2, G);
    return true;
}

// First part of load_row: returns the group index for load_group_rest, or
// ~0u when the row lies beyond a reflecting edge (G then stays zero).
template <int W>
__device__ inline uint32_t load_row_first(const Params &p, const uint32_t *__restrict__ src, uint32_t base,
                                          uint32_t stride, uint32_t len, int r, Group4<W> &G)
{
    int rr = r;
    if (rr < 0) {
        if (p.boundary) { zero_group<W>(G); return ~0u; }
        rr += (int)len;
    } else if (rr >= (int)len) {
        if (p.boundary) { zero_group<W>(G); return ~0u; }
        rr -= (int)len;
        if (rr >= (int)len) rr -= (int)len;  // r = len + 1 with len = 1
    }
    const uint32_t g = (base + (uint32_t)rr * stride) >> 2;
    load_group_first<W>(p, src, g, G);
    return g;
}

template <int W>
__device__ inline unsigned long long row_index(const Params &p, uint32_t base, uint32_t stride, uint32_t len, int r)
{
    int rr = r;
    if (rr < 0) rr += (int)len;
    if (rr >= (int)len) rr -= (int)len;
    if (rr >= (int)len) rr -= (int)len;  // r = len + 1 with len = 1
    return base + (uint32_t)rr * stride;
}

template <int W, int SEG>
__global__ __launch_bounds__(256) void sweep_march_kernel(Params p, const uint32_t *__restrict__ src,
                                                          uint32_t *__restrict__ dst, uint32_t stride,
                                                          uint32_t len, uint32_t ostride, uint32_t nouter,
                                                          uint32_t substep, uint32_t seg_first)
{
    constexpr uint32_t seg = SEG;
    const uint32_t gx = p.X >> 2;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t g = t % gx;
    const uint32_t rest = t / gx;
    const uint32_t nseg = (len + seg - 1) / seg;
    if (rest >= nouter * nseg) return;
    // Neighbouring waves take the work that is nearest in memory: the next
    // segment when segments are closer than outer rows (y sweep), else the
    // next outer row (z sweep).  Spreads concurrent accesses over channels.
    const uint32_t o = seg_first ? rest / nseg : rest % nouter;
    const uint32_t sgi = seg_first ? rest % nseg : rest / nouter;
    const uint32_t r0 = sgi * seg;
    if (r0 >= len) return;
    const uint32_t r1 = min(r0 + seg, len);
    const uint32_t base = o * ostride + g * 4u;
    const unsigned long long h = timestep_hash(p.seed32, p.ctl->timestep, substep);
    const bool periodic = p.boundary == 0;
    // Neighbour flags of row r, one bit per site (all four sites share them).
    auto hm_of = [&](int r) -> uint32_t { return (r > 0 || periodic) ? 0xFu : 0u; };
    auto hp_of = [&](int r) -> uint32_t { return (r + 1 < (int)len || periodic) ? 0xFu : 0u; };

    // Rows r-1 (A), r (B), r+1 (C), r+2 (D); E is the row being fetched.
    Group4<W> A, B, C, D, E;
    load_row<W>(p, src, base, stride, len, (int)r0 - 1, A);
    load_row<W>(p, src, base, stride, len, (int)r0, B);
    load_row<W>(p, src, base, stride, len, (int)r0 + 1, C);
    load_row<W>(p, src, base, stride, len, (int)r0 + 2, D);
    zero_group<W>(E);
    // Decisions: plus movers of A, both directions of B and C.  Rows beyond a
    // reflecting edge are empty, so they decide nothing.
    unsigned long long occA = group_occ<W>(A), occB = group_occ<W>(B), occC = group_occ<W>(C);
    unsigned long long mA, pA, mB, pB, mC, pC;
    {
        // A's minus neighbour (row r0-2) matters only for its minus movers,
        // which never reach row r0: pass "no minus neighbour".
        const int ra = (int)r0 - 1;
        decide<W, false>(p, A, 0u, B.t, 0u, hp_of(ra), (uint32_t)row_index<W>(p, base, stride, len, ra), h, occA, mA, pA);
        decide<W, false>(p, B, A.t, C.t, hm_of((int)r0), hp_of((int)r0), (uint32_t)row_index<W>(p, base, stride, len, (int)r0),
                  h, occB, mB, pB);
        decide<W, false>(p, C, B.t, D.t, hm_of((int)r0 + 1), hp_of((int)r0 + 1),
                  (uint32_t)row_index<W>(p, base, stride, len, (int)r0 + 1), h, occC, mC, pC);
    }
    (void)mA;
    // Fully unrolled: the row shift below is register renaming, and loads can
    // be issued ahead of the rows that need them.
#pragma unroll
    for (uint32_t rr = 0; rr < seg; rr++) {
        const uint32_t r = r0 + rr;
        if (r >= r1) break;
        const bool more = r + 1 < r1;
        if (more) load_row<W>(p, src, base, stride, len, (int)r + 3, E);
        const uint32_t rowi = base + r * stride;
        uint4 *out4[W];
#pragma unroll
        for (int k = 0; k < W; k++) out4[k] = reinterpret_cast<uint4 *>(dst + (size_t)k * p.N + rowi);
        if ((occA | occB | occC) == 0) {
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
                const uint32_t own_occ = site_bits<W>(occB, s);
                const uint32_t stay = own_occ & ~(site_bits<W>(mB, s) | site_bits<W>(pB, s));
                assemble<W>(p, rowi + s, substep, wo, stay, own_occ, wm, site_bits<W>(pA, s), wp,
                            site_bits<W>(mC, s), ov);
#pragma unroll
                for (int k = 0; k < W; k++) set_lane(res[k], s, ov[k]);
            }
#pragma unroll
            for (int k = 0; k < W; k++) *out4[k] = res[k];
        }
        if (!more) break;
        // Decide row r+2 (D) with its neighbours C and E, then shift.
        const unsigned long long occD = group_occ<W>(D);
        unsigned long long mD = 0, pD = 0;
        if (occD) {
            const int rd = (int)r + 2;
            decide<W, false>(p, D, C.t, E.t, hm_of(rd), hp_of(rd), (uint32_t)row_index<W>(p, base, stride, len, rd), h,
                      occD, mD, pD);
        }
        A = B;
        B = C;
        C = D;
        D = E;
        occA = occB;
        occB = occC;
        occC = occD;
        pA = pB;
        mB = mC;
        pB = pC;
        mC = mD;
        pC = pD;
    }
}

// ---------------------------------------------------------------------------
// y / z sweep on register columns.  A thread owns one group (4 x-sites) and a
// segment of SEG rows along the axis; it loads the SEG rows and the apron row
// on either side at once, so every neighbour along the axis is in its own
// registers: no LDS, no barriers, and each wave issues all of its loads
// before any decision waits on them.  The apron rows are the neighbouring
// segments' edge rows; segments are numbered so that the threads of a block
// (and consecutive blocks of an XCD) take adjacent segments, and those rows
// hit in cache.  Requires len % SEG == 0.
// ---------------------------------------------------------------------------
template <int W, int SEG>
__global__ __launch_bounds__(256) void sweep_col_kernel(Params p, const uint32_t *__restrict__ src,
                                                        uint32_t *__restrict__ dst, uint32_t stride, uint32_t len,
                                                        uint32_t ostride, uint32_t nouter, uint32_t substep)
{
    using mask_t = typename std::conditional<W == 4, unsigned long long, uint32_t>::type;
    const uint32_t gxn = p.X >> 2;
    const uint32_t nseg = len / SEG;
    const uint32_t nthreads = gxn * nseg * nouter;
    const uint32_t nblocks = (nthreads + 255u) / 256u;
    const uint32_t per_xcd = (nblocks + 7) / 8;
    const uint32_t logical = (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3);
    const uint32_t t = logical * 256u + threadIdx.x;
    if (logical >= nblocks || t >= nthreads) return;
    const uint32_t g = t % gxn;
    const uint32_t rest = t / gxn;
    const uint32_t sgi = rest % nseg, o = rest / nseg;
    const uint32_t r0 = sgi * SEG;
    const uint32_t base = o * ostride + g * 4u;
    const unsigned long long h = timestep_hash(p.seed32, p.ctl->timestep, substep);
    const bool periodic = p.boundary == 0;
    auto hm_of = [&](int r) -> uint32_t { return (r > 0 || periodic) ? 0xFu : 0u; };
    auto hp_of = [&](int r) -> uint32_t { return (r + 1 < (int)len || periodic) ? 0xFu : 0u; };

    // R[k] is row r0 - 1 + k: R[0] and R[SEG+1] are the apron rows.
    Group4<W> R[SEG + 2];
    load_row<W>(p, src, base, stride, len, (int)r0 - 1, R[0]);
#pragma unroll
    for (int k = 1; k <= SEG; k++) load_group<W>(p, src, (base + (r0 + k - 1) * stride) >> 2, R[k]);
    load_row<W>(p, src, base, stride, len, (int)(r0 + SEG), R[SEG + 1]);

    mask_t occ[SEG + 2], mi[SEG + 2], pl[SEG + 2];
#pragma unroll
    for (int k = 0; k < SEG + 2; k++) {
        const int r = (int)r0 - 1 + k;
        const unsigned long long oc = group_occ<W>(R[k]);
        unsigned long long m = 0, q = 0;
        if (oc) {
            // The apron rows only need the direction that enters the segment.
            const uint32_t hm = k == 0 ? 0u : hm_of(r);
            const uint32_t hp = k == SEG + 1 ? 0u : hp_of(r);
            const uint32_t tm = k == 0 ? 0u : R[k - 1].t;
            const uint32_t tp = k == SEG + 1 ? 0u : R[k + 1].t;
            const uint32_t ri = (k == 0 || k == SEG + 1) ? (uint32_t)row_index<W>(p, base, stride, len, r)
                                                         : base + (uint32_t)r * stride;
            decide<W, false>(p, R[k], tm, tp, hm, hp, ri, h, oc, m, q);
        }
        occ[k] = (mask_t)oc;
        mi[k] = (mask_t)m;
        pl[k] = (mask_t)q;
    }
#pragma unroll
    for (int k = 1; k <= SEG; k++) {
        const uint32_t rowi = base + (r0 + k - 1) * stride;
        uint4 res[W];
        if ((occ[k] | pl[k - 1] | mi[k + 1]) == 0) {
#pragma unroll
            for (int q = 0; q < W; q++) res[q] = make_uint4(0u, 0u, 0u, 0u);
        } else {
#pragma unroll
            for (int s = 0; s < 4; s++) {
                uint32_t wo[W], wm[W], wp[W], ov[W];
                site_words<W>(R[k], s, wo);
                site_words<W>(R[k - 1], s, wm);
                site_words<W>(R[k + 1], s, wp);
                const uint32_t own_occ = site_bits<W>(occ[k], s);
                const uint32_t stay = own_occ & ~(site_bits<W>(mi[k], s) | site_bits<W>(pl[k], s));
                assemble<W>(p, rowi + s, substep, wo, stay, own_occ, wm, site_bits<W>(pl[k - 1], s), wp,
                            site_bits<W>(mi[k + 1], s), ov);
#pragma unroll
                for (int q = 0; q < W; q++) set_lane(res[q], s, ov[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < W; q++) reinterpret_cast<uint4 *>(dst + (size_t)q * p.N + rowi)[0] = res[q];
    }
}

// ---------------------------------------------------------------------------
// y / z sweep on LDS tiles.  A block covers Gx groups along x by Ty rows along
// the axis (one group per thread), stages its rows plus one apron row on each
// side in LDS, decides its own rows' particles (the apron rows only in the
// direction that enters the tile), exchanges the decision masks through LDS
// and assembles its rows.  All row loads are issued at once.  Blocks are
// numbered so that the tiles a given XCD (blockIdx % 8) runs are adjacent
// along the axis, keeping shared apron rows in that XCD's L2.
// ---------------------------------------------------------------------------
// This lane's first slot in a list (LDS or global), for per-lane counts
// 0..7: a ballot prefix inside the wave and one atomic reservation per wave.
// Every lane of the wave must call it.
__device__ inline uint32_t wave_reserve(uint32_t *counter, uint32_t c)
{
    const unsigned long long b0 = __ballot(c & 1u), b1 = __ballot(c & 2u), b2 = __ballot(c & 4u);
    const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    const unsigned long long lt = (1ull << lane) - 1ull;
    const uint32_t pre = (uint32_t)(__popcll(b0 & lt) + 2 * __popcll(b1 & lt) + 4 * __popcll(b2 & lt));
    const uint32_t tot = (uint32_t)(__popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2));
    uint32_t base = 0;
    if (lane == 0 && tot) base = atomicAdd(counter, tot);
    return __shfl(base, 0, 64) + pre;
}

// Reserves c < 8 entries of the reaction list for the calling lane, in the
// shard of the calling wave (wave w of the launch's block `block` of this
// solver: shard w % kListShards); *shard receives the shard.  Every lane of
// the wave must call it.
__device__ inline uint32_t list_reserve(const Params &p, uint32_t c, uint32_t block, uint32_t *shard)
{
    *shard = (block * (blockDim.x >> 6) + (threadIdx.x >> 6)) % (uint32_t)kListShards;
    return wave_reserve(&p.ctl->list_count[*shard], c);
}

// Shards a launch of `blocks` blocks of `threads` threads fills.
__host__ __device__ inline uint32_t list_shards_used(uint32_t blocks, uint32_t threads)
{
    const unsigned long long waves = (unsigned long long)blocks * (threads / 64u);
    return waves < (unsigned long long)kListShards ? (uint32_t)waves : (uint32_t)kListShards;
}

// Writes one reaction-list entry; false (nothing written) when the shard is full.
template <int W>
__device__ inline bool list_put(const ReactionList &rl, uint32_t shard, uint32_t slot, uint32_t i,
                                const uint32_t *w, uint32_t st)
{
    if (slot >= rl.cap) return false;
    const size_t e = ((size_t)shard * rl.cap + slot) * list_units(W);
    rl.rec[e] = make_uint4(i, w[0], W > 1 ? w[W > 1 ? 1 : 0] : 0u, st);
    if constexpr (W == 4) rl.rec[e + 1] = make_uint4(w[2], w[3], 0u, 0u);
    return true;
}

template <int W>
constexpr int tile_threads() { return W == 4 ? 512 : 1024; }
static_assert(tile_threads<2>() == 1024 && tile_threads<4>() == 512, "tile_threads_for must agree");

// With REACT (z sweep) the output sites that hold particles are appended to
// the reaction list, and reaction_list_kernel evaluates their reactions after
// the sweep instead of a full reaction pass.
// LDS of a tile block of T threads: the rows' words and decisions, apron
// rows included (No * Gx <= 64, so the apron rows take at most 128 entries).
template <int W, int T>
struct TileLds {
    static constexpr int kEntries = T + 128;
    uint4 w[W][kEntries];
    unsigned long long m[kEntries], p[kEntries];
};

// Body of a y or z tile sweep (substep 1 or 2) for block number `block`; h
// is timestep_hash(seed32, timestep, substep).  A batched block reads h from
// its member's descriptor before its unit loop: computed inside the loop, its
// constants stayed in registers across the loop and cost the tile kernels a
// spill, and computed before it, the block waited for the timestep's load.
template <int W, bool REACT, int T>
__device__ inline void tile_body(const Params &p, const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                                 const TileGeom &geo, uint32_t substep, unsigned long long h, uint32_t block,
                                 const ReactionList &rl, TileLds<W, T> &lds)
{
    const uint32_t stride = geo.stride, len = geo.len, ostride = geo.ostride, nouter = geo.nouter, Gx = geo.Gx,
                   Ty = geo.Ty, nxchunk = geo.nxchunk, ntiles = geo.ntiles, tile_first = geo.tile_first,
                   No = geo.No, nob = geo.nob;
    auto &s_w = lds.w;
    auto &s_m = lds.m;
    auto &s_p = lds.p;

    const uint32_t nblocks = nxchunk * nob * ntiles;
    const uint32_t per_xcd = (nblocks + 7) / 8;
    const uint32_t logical = (block & 7u) * per_xcd + (block >> 3);
    if (logical >= nblocks) return;  // whole block: no barrier is skipped by part of it
    // Adjacent logical blocks take the work nearest in memory: the next row
    // tile when tiles are closer than outer rows (y sweep), else the next
    // outer rows (z sweep).
    uint32_t tile, xc, ob;
    if (tile_first) {
        const uint32_t rest = fast_div(logical, geo.div_ntiles);
        tile = logical - rest * ntiles;
        ob = fast_div(rest, geo.div_nxchunk);
        xc = rest - ob * nxchunk;
    } else {
        const uint32_t rest = fast_div(logical, geo.div_nob);
        ob = logical - rest * nob;
        tile = fast_div(rest, geo.div_nxchunk);
        xc = rest - tile * nxchunk;
    }

    // Thread: tile sub of the block (outer row ob * No + sub), row ly, group lx.
    const uint32_t tid = threadIdx.x;
    const uint32_t sub = fast_div(tid, geo.div_tile);
    const uint32_t q = tid - sub * (Gx * Ty);
    const uint32_t ly = fast_div(q, geo.div_Gx), lx = q - ly * Gx;
    const uint32_t o = ob * No + sub;
    const uint32_t gx = xc * Gx + lx;
    const bool in_block = sub < No;  // LDS rows exist for this thread's tile
    const bool xvalid = gx < (p.X >> 2) && in_block && o < nouter;
    const uint32_t r0 = tile * Ty;
    const uint32_t nrows = min(Ty, len - r0);
    const uint32_t r = r0 + ly;
    const bool valid = xvalid && ly < nrows;
    const uint32_t sb = sub * (Ty + 2u) * Gx;  // the tile's first LDS slot (its top apron row)
    const uint32_t base = o * ostride + gx * 4u;
    const bool periodic = p.boundary == 0;

    Group4<W> G;
    zero_group<W>(G);
    const uint32_t gG = (base + r * stride) >> 2;
    if (valid) load_group_first<W>(p, src, gG, G);
    uint32_t known[W];  // the output group's chunk flags, read early: their latency overlaps the tile's work
    if (valid) load_known<W>(p, dst, gG, known);
    const bool top = valid && ly == 0;
    const bool bottom = valid && ly == nrows - 1;
    // Site types of the neighbouring rows along the axis, read directly (the
    // block loads those rows itself, so these hit in cache) and issued with
    // the group's own loads: the decisions then need no barrier, and their
    // table loads overlap other waves' work.  For the edge rows of the tile
    // these are the apron rows' types.
    const bool hm_ok = valid && (r > 0 || periodic), hp_ok = valid && (r + 1 < len || periodic);
    const uint32_t *sites32 = reinterpret_cast<const uint32_t *>(p.sites);
    uint32_t tmy = 0, tpy = 0;
    if (hm_ok) tmy = sites32[(base + (r > 0 ? r - 1 : len - 1) * stride) >> 2];
    if (hp_ok) tpy = sites32[(base + (r + 1 < len ? r + 1 : 0u) * stride) >> 2];
    Group4<W> Ap, Bp;  // apron rows above and below the tile
    zero_group<W>(Ap);
    zero_group<W>(Bp);
    uint32_t gA = ~0u, gB = ~0u;
    if (top) gA = load_row_first<W>(p, src, base, stride, len, (int)r0 - 1, Ap);
    if (bottom) gB = load_row_first<W>(p, src, base, stride, len, (int)(r0 + nrows), Bp);
    if (valid) load_group_rest<W>(p, src, gG, G);
    if (gA != ~0u) load_group_rest<W>(p, src, gA, Ap);
    if (gB != ~0u) load_group_rest<W>(p, src, gB, Bp);

    const unsigned long long occ = valid ? group_occ_in<W>(p, G) : 0ull;
    unsigned long long mi = 0, pl = 0;
    if (occ) decide<W>(p, G, tmy, tpy, hm_ok ? 0xFu : 0u, hp_ok ? 0xFu : 0u, base + r * stride, h, occ, mi, pl);
    // Slot sb + (row+1)*Gx + lx; rows at or beyond nrows write nothing,
    // because slot nrows+1 holds the bottom apron row.
    const uint32_t e = sb + (ly + 1) * Gx + lx;
    if (in_block && ly < nrows) {
#pragma unroll
        for (int k = 0; k < W; k++) s_w[k][e] = G.w[k];
        s_m[e] = mi;
        s_p[e] = pl;
    }
    if (top) {
        // Row r0-1: only its plus movers enter the tile.
        unsigned long long am = 0, ap = 0;
        const unsigned long long ao = group_occ_in<W>(p, Ap);
        if (ao) {
            const int ra = (int)r0 - 1;
            decide<W>(p, Ap, 0u, G.t, 0u, 0xFu, (uint32_t)row_index<W>(p, base, stride, len, ra), h, ao, am, ap);
        }
#pragma unroll
        for (int k = 0; k < W; k++) s_w[k][sb + lx] = Ap.w[k];
        s_p[sb + lx] = ap;
    }
    if (bottom) {
        // Row r0+nrows: only its minus movers enter the tile.
        unsigned long long bm = 0, bpl = 0;
        const unsigned long long bo = group_occ_in<W>(p, Bp);
        if (bo) {
            const int rb = (int)(r0 + nrows);
            decide<W>(p, Bp, G.t, 0u, 0xFu, 0u, (uint32_t)row_index<W>(p, base, stride, len, rb), h, bo, bm, bpl);
        }
        const uint32_t eb = sb + (nrows + 1) * Gx + lx;
#pragma unroll
        for (int k = 0; k < W; k++) s_w[k][eb] = Bp.w[k];
        s_m[eb] = bm;
    }
    __syncthreads();

    uint4 res[W];
#pragma unroll
    for (int k = 0; k < W; k++) res[k] = make_uint4(0u, 0u, 0u, 0u);
    const uint32_t i0 = base + r * stride;
    if (valid) {
        const unsigned long long pa = s_p[e - Gx], mc = s_m[e + Gx];
        if ((occ | pa | mc) != 0) {
            Group4<W> A, C;
#pragma unroll
            for (int k = 0; k < W; k++) {
                A.w[k] = s_w[k][e - Gx];
                C.w[k] = s_w[k][e + Gx];
            }
#pragma unroll
            for (int s = 0; s < 4; s++) {
                uint32_t wo[W], wm[W], wp[W], ov[W];
                site_words<W>(G, s, wo);
                site_words<W>(A, s, wm);
                site_words<W>(C, s, wp);
                const uint32_t own_occ = site_bits<W>(occ, s);
                const uint32_t stay = own_occ & ~(site_bits<W>(mi, s) | site_bits<W>(pl, s));
                assemble<W>(p, i0 + s, substep, wo, stay, own_occ, wm, site_bits<W>(pa, s), wp, site_bits<W>(mc, s),
                            ov);
#pragma unroll
                for (int k = 0; k < W; k++) set_lane(res[k], s, ov[k]);
            }
        }
        // With X % 256 == 0 a wave is one aligned chunk of a row.
        store_group_known<W>(p, dst, i0 >> 2, res, Gx == kChunkGroups && (p.X & 255u) == 0u, known);
    }
    if (REACT && rl.rec) {
        // The output sites that hold particles go to the reaction list with
        // their words and types (without zero-order reactions an empty site
        // cannot react).  Output sites are packed (particles fill word 0
        // first), so a site holds particles exactly when its word 0 is
        // nonzero.  The wave's entries go to its shard ordered by site
        // position s, then lane: one ballot per s gives every lane its slot
        // (the ballot's popcount below the lane), and each position's
        // entries are consecutive records.
        const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
        const uint32_t shard = (block * (blockDim.x >> 6) + wave) % (uint32_t)kListShards;
        unsigned long long bs[4];
        uint32_t off[4], total = 0;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            bs[s] = __ballot(valid && lane_of(res[0], s) != 0u);
            off[s] = total;
            total += (uint32_t)__popcll(bs[s]);
        }
        if (total) {
            const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&p.ctl->list_count[shard], total);
            base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
            bool full = false;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                if (bs[s] == 0ull) continue;
                if ((bs[s] >> lane) & 1ull) {
                    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bs[s] >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)bs[s], 0u));
                    uint32_t w[W];
#pragma unroll
                    for (int k = 0; k < W; k++) w[k] = lane_of(res[k], s);
                    full |= !list_put<W>(rl, shard, base + off[s] + below, i0 + s, w, (G.t >> (8 * s)) & 0xFFu);
                }
            }
            if (full) p.ctl->list_overflow = 1u;
        }
    }
}

template <int W, bool REACT>
__global__ __launch_bounds__(kNarrowTileThreads) void sweep_tile_kernel(Params p, const uint32_t *__restrict__ src,
                                                                        uint32_t *__restrict__ dst, TileGeom geo,
                                                                        uint32_t substep, ReactionList rl)
{
    __shared__ TileLds<W, kNarrowTileThreads> lds;
    const unsigned long long h = timestep_hash(p.seed32, p.ctl->timestep, substep);
    tile_body<W, REACT, kNarrowTileThreads>(p, src, dst, geo, substep, h, blockIdx.x, rl, lds);
}

// ---------------------------------------------------------------------------
// y / z sweep over compacted work lists.
//
// Particles occupy a few percent of the sites, so the per-particle work of a
// sweep (move decisions and slot assembly) runs over block-wide lists rather
// than over every lane: a lane takes one list entry, so the busy lanes are
// packed into as few waves as the work needs.
//   1. Every thread loads two groups of the tile (Gx groups x Ty rows, plus
//      one group of the apron rows above and below), publishes their site
//      types and occupancy in LDS and appends its occupied sites to the
//      occupied list.
//   2. Each occupied site decides its particles into a per-site decision
//      word (apron rows only in the direction that enters the tile); owners
//      list their work sites (occupied, or with an occupied neighbour along
//      the axis) and store zeros for the others.
//   3. Each work site assembles its new content and stores it.
// The particle words themselves stay out of LDS: list entries re-read them
// (cache hits), which keeps a block at ~32 KB of LDS so that four blocks,
// each with two groups per thread in flight, share a compute unit.
//
// With REACT (z sweep) a work site that holds particles after the sweep is
// appended to the reaction list (see tile_body).
// ---------------------------------------------------------------------------
constexpr int kSparseThreads = 512;
constexpr int kSparseGpt = 2;  // owned groups per thread
template <int W>
using dec_t = typename std::conditional<W == 4, uint32_t, uint16_t>::type;

// Bit s set iff site s of the group holds a particle.
template <int W>
__device__ inline uint32_t group_nz(const Group4<W> &G)
{
    uint32_t m = 0;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        uint32_t a = 0;
#pragma unroll
        for (int k = 0; k < W; k++) a |= lane_of(G.w[k], s);
        m |= (a != 0u ? 1u : 0u) << s;
    }
    return m;
}

template <int W>
__device__ inline void global_site_words(const Params &p, const uint32_t *__restrict__ lat, uint32_t i, uint32_t *w)
{
    load_site<W>(p, lat, i, w);
}

template <int W, bool REACT>
__global__ __launch_bounds__(kSparseThreads) void sweep_sparse_kernel(
    Params p, const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, uint32_t stride, uint32_t len,
    uint32_t ostride, uint32_t nouter, uint32_t Gx, uint32_t Ty, uint32_t nxchunk, uint32_t ntiles,
    uint32_t substep, uint32_t tile_first, ReactionList rl)
{
    constexpr int NT = kSparseThreads;
    constexpr int GPT = kSparseGpt;
    constexpr int kOwn = NT * GPT;
    constexpr int kEnt = kOwn + 128;
    constexpr uint32_t P = 4 * W;
    constexpr uint32_t PM = (1u << P) - 1u;
    using D = dec_t<W>;
    __shared__ uint32_t s_t[kEnt];
    __shared__ uint8_t s_nz[kEnt];
    __shared__ D s_dec[4 * kEnt];
    __shared__ uint16_t s_occ[4 * kEnt];
    __shared__ uint16_t s_work[4 * kOwn];
    __shared__ uint32_t s_cnt[2];

    const uint32_t nblocks = nxchunk * nouter * ntiles;
    const uint32_t per_xcd = (nblocks + 7) / 8;
    const uint32_t logical = (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3);
    if (logical >= nblocks) return;  // whole block: no barrier is skipped by part of it
    // Adjacent logical blocks take the work nearest in memory: the next row
    // tile when tiles are closer than outer rows (y sweep), else the next
    // outer row (z sweep).
    uint32_t tile, xc, o;
    if (tile_first) {
        tile = logical % ntiles;
        const uint32_t rest = logical / ntiles;
        xc = rest % nxchunk;
        o = rest / nxchunk;
    } else {
        o = logical % nouter;
        const uint32_t rest = logical / nouter;
        xc = rest % nxchunk;
        tile = rest / nxchunk;
    }
    const uint32_t tid = threadIdx.x;
    const uint32_t gx_lim = min(Gx, (p.X >> 2) - xc * Gx);  // groups of this x chunk
    const uint32_t r0 = tile * Ty;
    const uint32_t nrows = min(Ty, len - r0);
    const uint32_t obase = o * ostride + xc * Gx * 4u;
    const bool periodic = p.boundary == 0;
    const uint32_t t = p.ctl->timestep;

    // ---- 1. groups, types and occupancy; occupied list -------------------------------
    Group4<W> G[GPT], A;
#pragma unroll
    for (int j = 0; j < GPT; j++) {
        const uint32_t q = tid + (uint32_t)j * NT;
        const uint32_t lx = q % Gx, ly = q / Gx;
        zero_group<W>(G[j]);
        if (lx < gx_lim && ly < nrows) load_group<W>(p, src, (obase + lx * 4u + (r0 + ly) * stride) >> 2, G[j]);
    }
    // Apron rows: threads [0, Gx) the row above the tile, [Gx, 2Gx) the row below.
    const bool top = tid < Gx, bot = tid >= Gx && tid < 2 * Gx;
    const uint32_t alx = top ? tid : tid - Gx;
    const bool apron = (top || bot) && alx < gx_lim;
    zero_group<W>(A);
    if (apron) load_row<W>(p, src, obase + alx * 4u, stride, len, top ? (int)r0 - 1 : (int)(r0 + nrows), A);
    if (tid < 2) s_cnt[tid] = 0u;
    __syncthreads();  // counters cleared

#pragma unroll
    for (int j = 0; j <= GPT; j++) {
        const bool is_apron = j == GPT;
        const Group4<W> &g = is_apron ? A : G[j];
        uint32_t e = 0;
        bool have;
        if (is_apron) {
            have = apron;
            e = top ? alx : (nrows + 1) * Gx + alx;
        } else {
            const uint32_t q = tid + (uint32_t)j * NT;
            const uint32_t lx = q % Gx, ly = q / Gx;
            have = lx < gx_lim && ly < nrows;
            e = (ly + 1) * Gx + lx;
        }
        uint32_t nz = 0;
        if (have) {
            s_t[e] = g.t;
            nz = group_nz<W>(g);
            s_nz[e] = (uint8_t)nz;
        }
        uint32_t slot = wave_reserve(&s_cnt[0], (uint32_t)__builtin_popcount(nz));
        while (nz) {
            const uint32_t sb = (uint32_t)__builtin_ctz(nz);
            nz &= nz - 1;
            s_occ[slot++] = (uint16_t)(e * 4u + sb);
        }
    }
    __syncthreads();

    // ---- 2a. decisions of the occupied sites ----------------------------------------
    const unsigned long long h = timestep_hash(p.seed32, t, substep);
    const uint32_t nocc = s_cnt[0];
    for (uint32_t idx = tid; idx < nocc; idx += NT) {
        const uint32_t c = s_occ[idx];
        const uint32_t e = c >> 2, sb = c & 3u;
        const uint32_t lx = e % Gx, k = e / Gx;
        const int r = (int)(r0 + k) - 1;
        const uint32_t rr = (uint32_t)((r + (int)len) % (int)len);
        const uint32_t i = obase + lx * 4u + sb + rr * stride;
        uint32_t w[W];
        global_site_words<W>(p, src, i, w);
        // The apron rows only need the direction that enters the tile.
        const bool hm = k > 0 && (periodic || r > 0);
        const bool hp = k <= nrows && (periodic || r + 1 < (int)len);
        const uint32_t st = (s_t[e] >> (8 * sb)) & 0xFFu;
        const uint32_t stm = hm ? (s_t[e - Gx] >> (8 * sb)) & 0xFFu : 0u;
        const uint32_t stp = hp ? (s_t[e + Gx] >> (8 * sb)) & 0xFFu : 0u;
        uint32_t m = site_occ<W>(w), mi = 0, pl = 0;
        while (m) {
            const uint32_t jb = (uint32_t)__builtin_ctz(m);
            m &= m - 1;
            const uint32_t part = slot_dyn<W>(w, jb);
            const float u = uniform01(i, p.pcb, jb, h);
            const float qm = lookup_T_if(p, p.T, hm, part, st, stm);
            const float qp = lookup_T_if(p, p.T, hp, part, st, stp);
            if (u < qm) mi |= 1u << jb;
            else if (u >= 0.5f && u < (qp + 0.5f)) pl |= 1u << jb;
        }
        s_dec[c] = (D)(mi | (pl << P));
    }

    // ---- 2b. work list; zeros for the sites nothing can reach -----------------------
#pragma unroll
    for (int j = 0; j < GPT; j++) {
        const uint32_t q = tid + (uint32_t)j * NT;
        const uint32_t lx = q % Gx, ly = q / Gx;
        const uint32_t e = (ly + 1) * Gx + lx;
        uint32_t wk = 0;
        if (lx < gx_lim && ly < nrows) {
            wk = (uint32_t)(s_nz[e] | s_nz[e - Gx] | s_nz[e + Gx]);
            const uint32_t i0 = obase + lx * 4u + (r0 + ly) * stride;
            if (wk == 0) {
#pragma unroll
                for (int k = 0; k < W; k++)
                    reinterpret_cast<uint4 *>(dst + (size_t)k * p.N)[i0 >> 2] = make_uint4(0u, 0u, 0u, 0u);
            } else {
#pragma unroll
                for (int sb = 0; sb < 4; sb++)
                    if (!((wk >> sb) & 1u)) {
#pragma unroll
                        for (int k = 0; k < W; k++) dst[(size_t)k * p.N + i0 + sb] = 0u;
                    }
            }
        }
        uint32_t slot = wave_reserve(&s_cnt[1], (uint32_t)__builtin_popcount(wk));
        while (wk) {
            const uint32_t sb = (uint32_t)__builtin_ctz(wk);
            wk &= wk - 1;
            s_work[slot++] = (uint16_t)(e * 4u + sb);
        }
    }
    __syncthreads();

    // ---- 3. assembly of the work sites (and the reaction test) ---------------------
    const uint32_t nwork = s_cnt[1];
    // Uniform trip count, so that every lane reaches the wave-wide list reservation.
    for (uint32_t base0 = 0; base0 < nwork; base0 += NT) {
        const uint32_t idx = base0 + tid;
        if (idx >= nwork) {
            if (REACT) {
                uint32_t shard;
                list_reserve(p, 0u, blockIdx.x, &shard);
            }
            continue;
        }
        const uint32_t c = s_work[idx];
        const uint32_t e = c >> 2, sb = c & 3u;
        const uint32_t lx = e % Gx, k = e / Gx;
        const uint32_t r = r0 + k - 1;
        const uint32_t i = obase + lx * 4u + sb + r * stride;
        const uint32_t nzo = s_nz[e], nzm = s_nz[e - Gx], nzp = s_nz[e + Gx];
        uint32_t wo[W], wm[W], wp[W], out[W];
#pragma unroll
        for (int q = 0; q < W; q++) wo[q] = wm[q] = wp[q] = 0u;
        if ((nzo >> sb) & 1u) global_site_words<W>(p, src, i, wo);
        if ((nzm >> sb) & 1u) {
            const uint32_t rm = r > 0 ? r - 1 : len - 1;  // the apron row wraps; never read when absent
            global_site_words<W>(p, src, obase + lx * 4u + sb + rm * stride, wm);
        }
        if ((nzp >> sb) & 1u) {
            const uint32_t rp = r + 1 < len ? r + 1 : 0u;
            global_site_words<W>(p, src, obase + lx * 4u + sb + rp * stride, wp);
        }
        const uint32_t own_occ = site_occ<W>(wo);
        const uint32_t dec_o = own_occ ? (uint32_t)s_dec[c] : 0u;
        const uint32_t dec_m = site_occ<W>(wm) ? (uint32_t)s_dec[c - 4u * Gx] : 0u;
        const uint32_t dec_p = site_occ<W>(wp) ? (uint32_t)s_dec[c + 4u * Gx] : 0u;
        const uint32_t stay = own_occ & ~((dec_o & PM) | (dec_o >> P));
        assemble<W>(p, i, substep, wo, stay, own_occ, wm, dec_m >> P, wp, dec_p & PM, out);
#pragma unroll
        for (int q = 0; q < W; q++) dst[(size_t)q * p.N + i] = out[q];
        if (REACT) {
            uint32_t any = 0, shard;
#pragma unroll
            for (int q = 0; q < W; q++) any |= out[q];
            const uint32_t slot = list_reserve(p, any != 0u ? 1u : 0u, blockIdx.x, &shard);
            if (any != 0u && !list_put<W>(rl, shard, slot, i, out, (s_t[e] >> (8 * sb)) & 0xFFu))
                p.ctl->list_overflow = 1u;
        }
    }
}

// ---------------------------------------------------------------------------
// Reaction phase in place.  Each thread owns kRG groups (4*kRG sites, groups
// 256 apart so every load stays coalesced) and issues all their loads before
// any per-site work; one loop then visits the occupied sites.  The
// single-particle table is staged in LDS when it fits.
// ---------------------------------------------------------------------------
constexpr int kRG = 4;
constexpr uint32_t kTabLds = 2048;  // float2 entries (16 KB)

template <int W>
__device__ inline void reaction4_body(const Params &p, uint32_t *__restrict__ lat, uint32_t block)
{
    __shared__ float2 s_tab[kTabLds];
    const uint32_t ntab = p.nST * p.nS;
    const bool in_lds = ntab <= kTabLds;
    if (in_lds) {
        for (uint32_t k = threadIdx.x; k < ntab; k += blockDim.x) s_tab[k] = reinterpret_cast<const float2 *>(p.prob1)[k];
    }
    __syncthreads();
    const float2 *tab = in_lds ? s_tab : reinterpret_cast<const float2 *>(p.prob1);

    const uint32_t ngroups = p.N >> 2;
    const uint32_t g0 = block * (256u * kRG) + threadIdx.x;
    uint4 w4[kRG][W];
    uint32_t t4[kRG];
    uint32_t todo = 0;
#pragma unroll
    for (int q = 0; q < kRG; q++) {
        const uint32_t g = g0 + 256u * q;
        if (g < ngroups) {
            w4[q][0] = reinterpret_cast<const uint4 *>(lat)[g];
#pragma unroll
            for (int k = 1; k < W; k++) w4[q][k] = make_uint4(0u, 0u, 0u, 0u);
            t4[q] = reinterpret_cast<const uint32_t *>(p.sites)[g];
        } else {
#pragma unroll
            for (int k = 0; k < W; k++) w4[q][k] = make_uint4(0u, 0u, 0u, 0u);
            t4[q] = 0u;
        }
    }
    // Planes k >= 1 after every group's plane 0 (read as in load_group_rest).
#pragma unroll
    for (int q = 0; q < kRG; q++) {
        const uint32_t g = g0 + 256u * q;
        if (g < ngroups) {
#pragma unroll
            for (int k = 1; k < W; k++) {
                const uint4 v = w4[q][k - 1];
                if (need_next_byte_word(p, v.x | v.y | v.z | v.w))
                    w4[q][k] = reinterpret_cast<const uint4 *>(lat + (size_t)k * p.N)[g];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kRG; q++) {
        const bool inside = g0 + 256u * q < ngroups;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            uint32_t a = 0;
#pragma unroll
            for (int k = 0; k < W; k++) a |= lane_of(w4[q][k], s);
            const bool need = inside && (p.any_qp0 || a != 0);
            todo |= (need ? 1u : 0u) << (4 * q + s);
        }
    }
    if (todo == 0) return;
    const unsigned long long h3 = timestep_hash(p.seed32, p.ctl->timestep, 3);
    uint32_t changed = 0;
    while (todo) {
        const uint32_t b = (uint32_t)__builtin_ctz(todo);
        todo &= todo - 1;
        const uint32_t q = b >> 2, s = b & 3u;
        uint32_t w[W], tq = 0;
#pragma unroll
        for (int k = 0; k < W; k++) w[k] = 0;
#pragma unroll
        for (int qq = 0; qq < kRG; qq++) {
            if (qq == (int)q) {
#pragma unroll
                for (int k = 0; k < W; k++) w[k] = lane_dyn<false>(w4[qq][k], s);
                tq = t4[qq];
            }
        }
        const uint32_t g = g0 + 256u * q;
        if (react_site_table<W>(p, g * 4u + s, (tq >> (8 * s)) & 0xFFu, w, h3, tab)) {
            changed |= 1u << q;
#pragma unroll
            for (int qq = 0; qq < kRG; qq++)
                if (qq == (int)q) {
#pragma unroll
                    for (int k = 0; k < W; k++) set_lane(w4[qq][k], (int)s, w[k]);
                }
        }
    }
#pragma unroll
    for (int q = 0; q < kRG; q++) {
        if (changed & (1u << q)) {
            const uint32_t g = g0 + 256u * q;
#pragma unroll
            for (int k = 0; k < W; k++) {
                reinterpret_cast<uint4 *>(lat + (size_t)k * p.N)[g] = w4[q][k];
                if (k > 0 && (w4[q][k].x | w4[q][k].y | w4[q][k].z | w4[q][k].w) != 0u)
                    clear_chunk_flag(p, lat, k, g);
            }
        }
    }
}

template <int W>
__global__ __launch_bounds__(256) void reaction4_kernel(Params p, uint32_t *__restrict__ lat)
{
    reaction4_body<W>(p, lat, blockIdx.x);
}

// Reaction phase of the sites a z sweep listed: one evaluation per entry,
// from the words and type in the list (the same per-site evaluation as
// reaction4_kernel; a site's reaction reads only that site, and an empty site
// cannot react without zero-order reactions, so the result equals the full
// pass).
// Block `block` of `nblocks` serves the shards below nshards (the shards the
// z sweep filled): one shard at a time with all its threads (shards of many
// entries), or with by_wave one shard per wave (short shards, as the small
// lattices of a batch fill them).
// (A shard that overflowed makes the reaction phase the full pass instead:
// the callers check Control::list_overflow.)
// One list entry between its two phases: the record's site, type and words,
// the particle count and first species, and (count <= 1) the tabulated total
// propensity and probability.
template <int W>
struct ListEntry {
    uint32_t i, st, w[W], n, first;
    float2 ap;
};

// Phase 1 of react_site_table for a list record (units a and, for W = 4,
// b): decode it, count its particles and load the tabulated propensity of a
// site with at most one.
template <int W>
__device__ inline void list_entry_load(const Params &p, const uint4 &a, const uint4 &b, const float2 *tab,
                                       ListEntry<W> &E)
{
    E.i = a.x;
    E.st = a.w;
    E.w[0] = a.y;
    if constexpr (W > 1) E.w[W > 1 ? 1 : 0] = a.z;
    if constexpr (W == 4) {
        E.w[2] = b.x;
        E.w[3] = b.y;
    }
    E.n = 0;
    E.first = 0;
#pragma unroll
    for (int j = 4 * W - 1; j >= 0; j--) {
        const uint32_t s = ByteSites<W>::slot(E.w, j);
        if (s) {
            E.n++;
            E.first = s;
        }
    }
    if (E.n == 1)
        E.ap = tab[E.st * p.nS + (E.first - 1)];
    else if (E.n == 0)
        E.ap = make_float2(p.qp0[E.st], p.prob0[E.st]);
}

// Phase 2: the reaction (react_site_table's result and in-place update of
// E.w), and the site written back if it changed.
template <int W>
__device__ inline void list_entry_react(const Params &p, uint32_t *__restrict__ lat, ListEntry<W> &E,
                                        unsigned long long h3)
{
    bool changed;
    if (E.n > 1)
        changed = react_site_l<ByteSites<W>>(p, E.i, E.st, E.w, h3);
    else
        changed = E.ap.x != 0.0f && react_apply<ByteSites<W>>(p, E.i, E.st, E.w, E.ap.x, E.ap.y, h3);
    if (changed) {
#pragma unroll
        for (int q = 0; q < W; q++) {
            lat[(size_t)q * p.N + E.i] = E.w[q];
            if (q > 0 && E.w[q] != 0u) clear_chunk_flag(p, lat, q, E.i >> 2);
        }
    }
}

template <int W>
__device__ inline void reaction_list_body(const Params &p, uint32_t *__restrict__ lat, const ReactionList &rl,
                                          uint32_t block, uint32_t nblocks, uint32_t nshards, bool by_wave)
{
    const float2 *tab = reinterpret_cast<const float2 *>(p.prob1);
    const unsigned long long h3 = timestep_hash(p.seed32, p.ctl->timestep, 3);
    const uint32_t units = by_wave ? (blockDim.x >> 6) : 1u, unit = by_wave ? (threadIdx.x >> 6) : 0u;
    const uint32_t first = by_wave ? (threadIdx.x & 63u) : threadIdx.x, width = by_wave ? 64u : blockDim.x;
    for (uint32_t shard = block * units + unit; shard < nshards; shard += nblocks * units) {
        const uint4 *rec = rl.rec + (size_t)shard * rl.cap * list_units(W);
        // The thread's first record is read before the shard's count
        // arrives (every slot below the capacity is allocated; a slot at or
        // past the count is not used), and each next one while the current
        // one is evaluated.  (Loading two or four entries before evaluating
        // any measured slower: 32.6-33.3 against 30.8 us at the whole-cell
        // size.)
        uint4 a = make_uint4(0u, 0u, 0u, 0u), b = a;
        if (first < rl.cap) {
            a = rec[(size_t)first * list_units(W)];
            if constexpr (W == 4) b = rec[(size_t)first * 2 + 1];
        }
        const uint32_t n = min(p.ctl->list_count[shard], rl.cap);
        for (uint32_t k = first; k < n; k += width) {
            uint4 na = make_uint4(0u, 0u, 0u, 0u), nb = na;
            if (k + width < n) {
                na = rec[(size_t)(k + width) * list_units(W)];
                if constexpr (W == 4) nb = rec[(size_t)(k + width) * 2 + 1];
            }
            ListEntry<W> E;
            list_entry_load<W>(p, a, b, tab, E);
            list_entry_react<W>(p, lat, E, h3);
            a = na;
            b = nb;
        }
    }
}

// At most 128 VGPRs: four waves per SIMD to hide the table and lattice
// latencies (measured 33 against 36 us with the compiler's 135).
template <int W>
__global__ __launch_bounds__(256, 4) void reaction_list_kernel(Params p, uint32_t *__restrict__ lat, ReactionList rl,
                                                            uint32_t nshards)
{
    if (p.ctl->list_overflow) {
        const uint32_t nblk = ((p.N >> 2) + 256u * kRG - 1u) / (256u * kRG);
        for (uint32_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) reaction4_body<W>(p, lat, blk);
        return;
    }
    reaction_list_body<W>(p, lat, rl, blockIdx.x, gridDim.x, nshards, false);
}

// Probability tables for sites with at most one particle, built with the
// same float sum and device exp as the reaction kernels.
__global__ void prob_table_kernel(Params p, float *prob1, float *prob0)
{
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n1 = p.nST * p.nS;
    if (idx < n1) {
        const uint32_t st = idx / p.nS;
        float a = p.qp0[st];
        a += p.qp1[idx];
        prob1[2 * idx] = a;
        prob1[2 * idx + 1] = a == 0.0f ? 0.0f : reaction_probability(a);
    }
    if (idx < p.nST) {
        const float a = p.qp0[idx];
        prob0[idx] = a == 0.0f ? 0.0f : reaction_probability(a);
    }
}

// ---------------------------------------------------------------------------
// Fused step: x, y and z sweeps and the reaction phase of one timestep in a
// single pass, reading each site's particles and type once and writing once.
//
// A block owns a tile of full x-rows: "work rows" k = 0..nrows+1 cover
// y0-1 .. y0+nrows (rows 0 and nrows+1 are halo rows owned by the
// neighbouring tiles).  A wave holds rpw rows of gx = X/4 lanes each, so
// x-neighbours are lanes of the same row (shuffles, periodic wrap inside
// the row).  The block marches along z over planes z0-1 .. z1 of its
// segment; for each input plane zi it
//   1. x-sweeps every work row of plane zi,
//   2. exchanges the x-swept rows through LDS (double buffered by plane
//      parity) and y-sweeps its owned rows -> YO(zi),
//   3. decides the z moves of YO(zi) (types of zi-1, zi, zi+1),
//   4. z-assembles plane zo = zi-1 from YO(zi-2), YO(zi-1), YO(zi), runs the
//      reaction phase on it and writes it.
// Halo rows and halo planes are computed redundantly and never recorded as
// overflow, so every overflow entry (phase, site, emission) is recorded once,
// by the owner, exactly as the separate kernels record it.
//
// Input planes are prefetched one plane step ahead, as the last memory
// operation of a step: vector memory operations complete in issue order, so
// any earlier wait (a table or spill load) would otherwise also wait for the
// prefetch.  The transition and single-particle reaction tables are copied to
// LDS (dynamic shared memory: T floats, then prob1 float2 pairs).
// ---------------------------------------------------------------------------
template <int W>
constexpr int fused_threads() { return W == 4 ? 512 : 1024; }

template <int W>
using ymask_t = typename std::conditional<W == 4, unsigned long long, uint32_t>::type;

template <int W>
constexpr size_t fused_static_lds()
{
    return (size_t)fused_threads<W>() * (2 * W * sizeof(uint4) + 2 * sizeof(uint32_t) + 2 * sizeof(ymask_t<W>));
}

template <int W>
__global__ __launch_bounds__(fused_threads<W>()) void fused_step_kernel(Params p, const uint32_t *__restrict__ src,
                                                                         uint32_t *__restrict__ dst, uint32_t gx,
                                                                         uint32_t rpw, uint32_t Ty, uint32_t nty,
                                                                         uint32_t Zs)
{
    constexpr int NT = fused_threads<W>();
    using mask_t = ymask_t<W>;
    __shared__ uint4 s_w[2][W][NT];
    __shared__ uint32_t s_t[2][NT];
    __shared__ mask_t s_m[NT], s_p[NT];
    extern __shared__ float s_dyn[];

    const uint32_t nT = p.nST * p.nST * p.nS, ntab = p.nST * p.nS;
    float *s_T = s_dyn;
    float2 *s_tab = reinterpret_cast<float2 *>(s_dyn + ((nT + 1u) & ~1u));
    for (uint32_t k = threadIdx.x; k < nT; k += NT) s_T[k] = p.T[k];
    for (uint32_t k = threadIdx.x; k < ntab; k += NT) s_tab[k] = reinterpret_cast<const float2 *>(p.prob1)[k];

    const uint32_t X = p.X, Y = p.Y, Z = p.Z, XY = p.XY;
    const bool periodic = p.boundary == 0;
    const uint32_t ty = blockIdx.x % nty, zsg = blockIdx.x / nty;
    const uint32_t y0 = ty * Ty, z0 = zsg * Zs;
    const uint32_t nrows = min(Ty, Y - y0);
    const uint32_t z1 = min(z0 + Zs, Z);
    const int lane = (int)(threadIdx.x & 63u);
    const uint32_t wv = threadIdx.x >> 6;
    const uint32_t r = (uint32_t)lane / gx, g = (uint32_t)lane % gx;
    const bool lane_ok = r < rpw;
    const uint32_t k = wv * rpw + r;                  // work row
    const bool work = lane_ok && k <= nrows + 1;
    const bool owned = work && k >= 1 && k <= nrows;
    const int yl = (int)y0 - 1 + (int)k;              // unwrapped y of the work row
    const bool row_exists = work && (periodic || (yl >= 0 && yl < (int)Y));
    const uint32_t y = (uint32_t)((yl + (int)Y) % (int)Y);
    const uint32_t e = k * gx + g;                    // LDS slot of this thread's group
    // x-neighbour lanes inside the row
    const int rb = (int)(r * gx);
    const int ll = lane_ok ? rb + (int)((g + gx - 1) % gx) : lane;
    const int lr = lane_ok ? rb + (int)((g + 1) % gx) : lane;
    const bool hmx = periodic || g > 0, hpx = periodic || g + 1 < gx;
    // y neighbours of an owned row exist unless a reflecting edge is crossed
    const bool hmy = periodic || yl > 0, hpy = periodic || yl + 1 < (int)Y;
    const uint32_t t = p.ctl->timestep;
    const unsigned long long h0 = timestep_hash(p.seed32, t, 0), h1 = timestep_hash(p.seed32, t, 1),
                             h2 = timestep_hash(p.seed32, t, 2), h3 = timestep_hash(p.seed32, t, 3);

    auto plane_exists = [&](int z) { return periodic || (z >= 0 && z < (int)Z); };
    // z in [-1, Z]: wrapped without a division
    auto row_base = [&](int z) {
        uint32_t zz = (uint32_t)(z + (int)Z);
        zz = zz >= Z ? zz - Z : zz;
        zz = zz >= Z ? zz - Z : zz;
        return zz * XY + y * X + 4u * g;
    };
    auto fetch = [&](int z, Group4<W> &G) {
        zero_group<W>(G);
        if (row_exists && plane_exists(z)) load_group<W>(p, src, row_base(z) >> 2, G);
    };

    Group4<W> G, Gn;  // input planes zi and zi+1
    fetch((int)z0 - 1, G);
    fetch((int)z0, Gn);

    // z window: Ya = YO(zi-2), Yb = YO(zi-1) (+ their z decisions)
    Group4<W> Ya, Yb;
    zero_group<W>(Ya);
    zero_group<W>(Yb);
    mask_t occA = 0, pA = 0, occB = 0, mB = 0, pB = 0;
    __syncthreads();  // tables ready

    for (int zi = (int)z0 - 1; zi <= (int)z1; zi++) {
        const int par = (zi - (int)z0 + 1) & 1;
        const bool pz = plane_exists(zi);
        const bool own_plane = zi >= (int)z0 && zi < (int)z1;
        const uint32_t base = row_base(zi);

        // ---- 1. x sweep of this work row -------------------------------------
        uint32_t lw[W], rw[W];
#pragma unroll
        for (int q = 0; q < W; q++) {
            lw[q] = __shfl(G.w[q].w, ll, 64);
            rw[q] = __shfl(G.w[q].x, lr, 64);
        }
        const uint32_t lt = __shfl(G.t >> 24, ll, 64), rt = __shfl(G.t & 0xFFu, lr, 64);
        if (!hmx) {
#pragma unroll
            for (int q = 0; q < W; q++) lw[q] = 0u;
        }
        if (!hpx) {
#pragma unroll
            for (int q = 0; q < W; q++) rw[q] = 0u;
        }
        const unsigned long long occ = (row_exists && pz) ? group_occ<W>(G) : 0ull;
        unsigned long long xm = 0, xp = 0;
        if (occ) {
            const uint32_t tm = (G.t << 8) | lt, tp = (G.t >> 8) | (rt << 24);
            decide_t<W, false>(p, s_T, G, tm, tp, hmx ? 0xFu : 0xEu, hpx ? 0xFu : 0x7u, base, h0, occ, xm, xp);
        }
        uint32_t lplus = __shfl(site_bits<W>(xp, 3), ll, 64);
        uint32_t rminus = __shfl(site_bits<W>(xm, 0), lr, 64);
        if (!hmx) lplus = 0;
        if (!hpx) rminus = 0;
        Group4<W> XO;
        XO.t = G.t;
        if ((occ | lplus | rminus) == 0) {
#pragma unroll
            for (int q = 0; q < W; q++) XO.w[q] = make_uint4(0u, 0u, 0u, 0u);
        } else {
            // Overflow of a halo row or halo plane belongs to another tile.
            const bool rec = owned && own_plane;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                uint32_t wo[W], wm[W], wp[W], o[W];
                site_words<W>(G, s, wo);
                const uint32_t own_occ = site_bits<W>(occ, s);
                const uint32_t stay = own_occ & ~(site_bits<W>(xm, s) | site_bits<W>(xp, s));
                uint32_t msel, psel;
                if (s == 0) {
#pragma unroll
                    for (int q = 0; q < W; q++) wm[q] = lw[q];
                    msel = lplus;
                } else {
                    site_words<W>(G, s - 1, wm);
                    msel = site_bits<W>(xp, s - 1);
                }
                if (s == 3) {
#pragma unroll
                    for (int q = 0; q < W; q++) wp[q] = rw[q];
                    psel = rminus;
                } else {
                    site_words<W>(G, s + 1, wp);
                    psel = site_bits<W>(xm, s + 1);
                }
                assemble_rec<W>(p, base + s, 0, rec, wo, stay, own_occ, wm, msel, wp, psel, o);
#pragma unroll
                for (int q = 0; q < W; q++) set_lane(XO.w[q], s, o[q]);
            }
        }

        // ---- 2. y sweep through LDS ------------------------------------------
        if (work) {
#pragma unroll
            for (int q = 0; q < W; q++) s_w[par][q][e] = XO.w[q];
            s_t[par][e] = XO.t;
        }
        __syncthreads();
        const unsigned long long yocc = (row_exists && pz) ? group_occ<W>(XO) : 0ull;
        unsigned long long ym = 0, yp = 0;
        if (yocc) {
            // Row k's y neighbours are work rows k-1 and k+1; the halo rows
            // only need the direction that enters the tile.
            const bool hm = k > 0 && hmy;
            const bool hp = k <= nrows && hpy;
            const uint32_t tmy = hm ? s_t[par][e - gx] : 0u;
            const uint32_t tpy = hp ? s_t[par][e + gx] : 0u;
            decide_t<W, false>(p, s_T, XO, tmy, tpy, hm ? 0xFu : 0u, hp ? 0xFu : 0u, base, h1, yocc, ym, yp);
        }
        if (work) {
            s_m[e] = (mask_t)ym;
            s_p[e] = (mask_t)yp;
        }
        __syncthreads();

        if (owned) {  // halo rows and idle lanes have no z state
        Group4<W> Yc;
        Yc.t = XO.t;
        {
            const unsigned long long pa = hmy ? (unsigned long long)s_p[e - gx] : 0ull;
            const unsigned long long mc = hpy ? (unsigned long long)s_m[e + gx] : 0ull;
            if ((yocc | pa | mc) == 0) {
#pragma unroll
                for (int q = 0; q < W; q++) Yc.w[q] = make_uint4(0u, 0u, 0u, 0u);
            } else {
                Group4<W> A, C;
#pragma unroll
                for (int q = 0; q < W; q++) {
                    A.w[q] = hmy ? s_w[par][q][e - gx] : make_uint4(0u, 0u, 0u, 0u);
                    C.w[q] = hpy ? s_w[par][q][e + gx] : make_uint4(0u, 0u, 0u, 0u);
                }
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    uint32_t wo[W], wm[W], wp[W], o[W];
                    site_words<W>(XO, s, wo);
                    site_words<W>(A, s, wm);
                    site_words<W>(C, s, wp);
                    const uint32_t own_occ = site_bits<W>(yocc, s);
                    const uint32_t stay = own_occ & ~(site_bits<W>(ym, s) | site_bits<W>(yp, s));
                    assemble_rec<W>(p, base + s, 1, own_plane, wo, stay, own_occ, wm, site_bits<W>(pa, s), wp,
                                    site_bits<W>(mc, s), o);
#pragma unroll
                    for (int q = 0; q < W; q++) set_lane(Yc.w[q], s, o[q]);
                }
            }
        }

        // ---- 3. z decisions of YO(zi) -----------------------------------------
        const mask_t occC = pz ? (mask_t)group_occ<W>(Yc) : (mask_t)0;
        mask_t mC = 0, pC = 0;
        if (occC) {
            // YO(z0-1) only sends plus movers in, YO(z1) only minus movers.
            const bool hm = zi > (int)z0 - 1 && (periodic || zi > 0);
            const bool hp = zi < (int)z1 && (periodic || zi + 1 < (int)Z);
            unsigned long long m64, p64;
            decide_t<W, false>(p, s_T, Yc, hm ? Yb.t : 0u, hp ? Gn.t : 0u, hm ? 0xFu : 0u, hp ? 0xFu : 0u, base, h2, occC,
                        m64, p64);
            mC = (mask_t)m64;
            pC = (mask_t)p64;
        }

        // ---- 4. z assembly and reaction of plane zi-1 ------------------------
        const int zo = zi - 1;
        if (zo >= (int)z0 && zo < (int)z1) {
            const uint32_t obase = row_base(zo);
            uint4 res[W];
            if ((occA | occB | occC) == 0) {
#pragma unroll
                for (int q = 0; q < W; q++) res[q] = make_uint4(0u, 0u, 0u, 0u);
            } else {
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    uint32_t wo[W], wm[W], wp[W], o[W];
                    site_words<W>(Yb, s, wo);
                    site_words<W>(Ya, s, wm);
                    site_words<W>(Yc, s, wp);
                    const uint32_t own_occ = site_bits<W>(occB, s);
                    const uint32_t stay = own_occ & ~(site_bits<W>(mB, s) | site_bits<W>(pB, s));
                    assemble<W>(p, obase + s, 2, wo, stay, own_occ, wm, site_bits<W>(pA, s), wp,
                                site_bits<W>(mC, s), o);
#pragma unroll
                    for (int q = 0; q < W; q++) set_lane(res[q], s, o[q]);
                }
            }
            if (p.R > 0) {
                uint32_t todo = 0;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    uint32_t a = 0;
#pragma unroll
                    for (int q = 0; q < W; q++) a |= lane_of(res[q], s);
                    todo |= ((p.any_qp0 || a != 0) ? 1u : 0u) << s;
                }
                while (todo) {
                    const uint32_t s = (uint32_t)__builtin_ctz(todo);
                    todo &= todo - 1;
                    uint32_t w[W];
#pragma unroll
                    for (int q = 0; q < W; q++) w[q] = lane_dyn<false>(res[q], s);
                    if (react_site_table<W>(p, obase + s, (Yb.t >> (8 * s)) & 0xFFu, w, h3, s_tab)) {
#pragma unroll
                        for (int q = 0; q < W; q++) set_lane(res[q], (int)s, w[q]);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < W; q++) reinterpret_cast<uint4 *>(dst + (size_t)q * p.N)[obase >> 2] = res[q];
        }
        Ya = Yb;
        occA = occB;
        pA = pB;
        Yb = Yc;
        occB = occC;
        mB = mC;
        pB = pC;
        }
        // Rotate the input planes; Gn was loaded a whole plane step ago.
        G = Gn;
        if (zi + 2 <= (int)z1) fetch(zi + 2, Gn);
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// Dynamic LDS of the fused step: T, then the single-particle table.
static size_t fused_dyn_lds(const Params &p)
{
    const size_t nT = (size_t)p.nST * p.nST * p.nS;
    return ((nT + 1) & ~(size_t)1) * sizeof(float) + (size_t)p.nST * p.nS * sizeof(float2);
}

static size_t fused_total_lds(const Params &p)
{
    const size_t st = p.W == 4 ? fused_static_lds<4>() : (p.W == 2 ? fused_static_lds<2>() : fused_static_lds<1>());
    return st + fused_dyn_lds(p);
}

// The fused step needs whole rows in a wave (X % 4 == 0, X <= 256) and its
// tables in LDS next to the row buffers, within gfx950's 160 KB per CU.
constexpr size_t kFusedLdsMax = 160u * 1024u;
bool fused_ok(const Params &p)
{
    return (p.X & 3u) == 0 && p.X >= 4 && p.X <= 256 && fused_total_lds(p) <= kFusedLdsMax;
}

// Owned rows per block: (waves x rows per wave) minus the two halo rows.
uint32_t fused_tile_rows(uint32_t X, uint32_t W)
{
    const uint32_t waves = (uint32_t)(W == 4 ? fused_threads<4>() : fused_threads<2>()) / 64u;
    return waves * (64u / (X >> 2)) - 2u;
}

hipError_t launch_fused(const Params &p, const uint32_t *src, uint32_t *dst, uint32_t zseg, hipStream_t st)
{
    if (!fused_ok(p)) return hipErrorInvalidValue;
    const uint32_t gx = p.X >> 2;
    const uint32_t rpw = 64u / gx;
    const uint32_t Ty = fused_tile_rows(p.X, p.W);
    const uint32_t nty = (p.Y + Ty - 1) / Ty;
    if (zseg == 0) zseg = 1;
    const uint32_t nzs = (p.Z + zseg - 1) / zseg;
    const dim3 grid(nty * nzs);
    const size_t dyn = fused_dyn_lds(p);
    hipError_t e = hipSuccess;
    switch (p.W) {
    case 1: e = hipFuncSetAttribute(reinterpret_cast<const void *>(fused_step_kernel<1>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn); break;
    case 2: e = hipFuncSetAttribute(reinterpret_cast<const void *>(fused_step_kernel<2>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn); break;
    case 4: e = hipFuncSetAttribute(reinterpret_cast<const void *>(fused_step_kernel<4>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    switch (p.W) {
    case 1: hipLaunchKernelGGL(fused_step_kernel<1>, grid, dim3(fused_threads<1>()), dyn, st, p, src, dst, gx, rpw, Ty, nty, zseg); break;
    case 2: hipLaunchKernelGGL(fused_step_kernel<2>, grid, dim3(fused_threads<2>()), dyn, st, p, src, dst, gx, rpw, Ty, nty, zseg); break;
    case 4: hipLaunchKernelGGL(fused_step_kernel<4>, grid, dim3(fused_threads<4>()), dyn, st, p, src, dst, gx, rpw, Ty, nty, zseg); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// The vector kernels hold four particles per word (the byte lattice).
bool fast_path_ok(const Params &p) { return !p.word_particles && (p.X & 3u) == 0 && p.N >= 4; }

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
    seg = seg == 4 ? 4u : 8u;  // compiled segment lengths
    const uint32_t nseg = (len + seg - 1) / seg;
    const size_t threads = (size_t)(p.X >> 2) * nouter * nseg;
    const dim3 grid(blocks_for(threads, 256));
    const uint32_t sub = (uint32_t)axis;
    // y sweep: the next segment is seg*X sites away, the next outer (z) plane
    // X*Y away; z sweep: the next outer (y) row is X away.
    const uint32_t seg_first = 0u;  // measured: outer rows first is faster for both axes
#define LMX_MARCH(WW, SS) hipLaunchKernelGGL((sweep_march_kernel<WW, SS>), grid, block, 0, st, p, src, dst, stride, len, ostride, nouter, sub, seg_first)
    if (seg == 4) {
        switch (p.W) {
        case 1: LMX_MARCH(1, 4); break;
        case 2: LMX_MARCH(2, 4); break;
        case 4: LMX_MARCH(4, 4); break;
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (p.W) {
        case 1: LMX_MARCH(1, 8); break;
        case 2: LMX_MARCH(2, 8); break;
        case 4: LMX_MARCH(4, 8); break;
        default: return hipErrorInvalidValue;
        }
    }
#undef LMX_MARCH
    return hipGetLastError();
}

template <int W>
static hipError_t launch_tile(const Params &p, const uint32_t *src, uint32_t *dst, int axis, bool react,
                              const ReactionList &rl, hipStream_t st)
{
    const TileGeom geo = tile_geom(p, axis, kNarrowTileThreads);
    const dim3 grid(tile_blocks(geo)), block(geo.Gx * geo.Ty * geo.No);
    if (react)
        hipLaunchKernelGGL((sweep_tile_kernel<W, true>), grid, block, 0, st, p, src, dst, geo, (uint32_t)axis, rl);
    else
        hipLaunchKernelGGL((sweep_tile_kernel<W, false>), grid, block, 0, st, p, src, dst, geo, (uint32_t)axis, rl);
    return hipGetLastError();
}

bool col_ok(const Params &p, int axis, uint32_t seg)
{
    const uint32_t len = axis == 1 ? p.Y : p.Z;
    return (seg == 4 || seg == 8) && len % seg == 0;
}

hipError_t launch_sweep_col(const Params &p, const uint32_t *src, uint32_t *dst, int axis, uint32_t seg,
                            hipStream_t st)
{
    if (axis < 1 || axis > 2 || !col_ok(p, axis, seg)) return hipErrorInvalidValue;
    const uint32_t stride = axis == 1 ? p.X : p.XY;
    const uint32_t len = axis == 1 ? p.Y : p.Z;
    const uint32_t ostride = axis == 1 ? p.XY : p.X;
    const uint32_t nouter = axis == 1 ? p.Z : p.Y;
    const size_t threads = (size_t)(p.X >> 2) * (len / seg) * nouter;
    const dim3 grid((unsigned)((blocks_for(threads, 256) + 7) / 8 * 8)), block(256);
    const uint32_t sub = (uint32_t)axis;
#define LMX_COL(WW, SS) hipLaunchKernelGGL((sweep_col_kernel<WW, SS>), grid, block, 0, st, p, src, dst, stride, len, ostride, nouter, sub)
    if (seg == 4) {
        switch (p.W) {
        case 1: LMX_COL(1, 4); break;
        case 2: LMX_COL(2, 4); break;
        case 4: LMX_COL(4, 4); break;
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (p.W) {
        case 1: LMX_COL(1, 8); break;
        case 2: LMX_COL(2, 8); break;
        case 4: LMX_COL(4, 8); break;
        default: return hipErrorInvalidValue;
        }
    }
#undef LMX_COL
    return hipGetLastError();
}

hipError_t launch_sweep_tile(const Params &p, const uint32_t *src, uint32_t *dst, int axis, bool react,
                             const ReactionList &rl, hipStream_t st)
{
    if (axis < 1 || axis > 2 || (react && (axis != 2 || p.any_qp0 || !rl.rec || !rl.cap))) return hipErrorInvalidValue;
    switch (p.W) {
    case 1: return launch_tile<1>(p, src, dst, axis, react, rl, st);
    case 2: return launch_tile<2>(p, src, dst, axis, react, rl, st);
    case 4: return launch_tile<4>(p, src, dst, axis, react, rl, st);
    }
    return hipErrorInvalidValue;
}

template <int W>
static hipError_t launch_sparse(const Params &p, const uint32_t *src, uint32_t *dst, int axis, bool react,
                                const ReactionList &rl, hipStream_t st)
{
    const uint32_t stride = axis == 1 ? p.X : p.XY;
    const uint32_t len = axis == 1 ? p.Y : p.Z;
    const uint32_t ostride = axis == 1 ? p.XY : p.X;
    const uint32_t nouter = axis == 1 ? p.Z : p.Y;
    const uint32_t gxn = p.X >> 2;
    const uint32_t Gx = gxn < 64u ? gxn : 64u;
    const uint32_t Ty = (uint32_t)(kSparseThreads * kSparseGpt) / Gx;
    const uint32_t nxchunk = (gxn + Gx - 1) / Gx;
    const uint32_t ntiles = (len + Ty - 1) / Ty;
    const uint32_t nblocks = nxchunk * nouter * ntiles;
    const uint32_t per_xcd = (nblocks + 7) / 8;
    const uint32_t tile_first = (size_t)Ty * stride < (size_t)ostride ? 1u : 0u;
    if (react)
        hipLaunchKernelGGL((sweep_sparse_kernel<W, true>), dim3(per_xcd * 8), dim3(kSparseThreads), 0, st, p, src,
                           dst, stride, len, ostride, nouter, Gx, Ty, nxchunk, ntiles, (uint32_t)axis, tile_first,
                           rl);
    else
        hipLaunchKernelGGL((sweep_sparse_kernel<W, false>), dim3(per_xcd * 8), dim3(kSparseThreads), 0, st, p, src,
                           dst, stride, len, ostride, nouter, Gx, Ty, nxchunk, ntiles, (uint32_t)axis, tile_first,
                           rl);
    return hipGetLastError();
}

hipError_t launch_sweep_sparse(const Params &p, const uint32_t *src, uint32_t *dst, int axis, bool react,
                               const ReactionList &rl, hipStream_t st)
{
    if (axis < 1 || axis > 2 || (react && (axis != 2 || p.any_qp0 || !rl.rec || !rl.cap))) return hipErrorInvalidValue;
    switch (p.W) {
    case 1: return launch_sparse<1>(p, src, dst, axis, react, rl, st);
    case 2: return launch_sparse<2>(p, src, dst, axis, react, rl, st);
    case 4: return launch_sparse<4>(p, src, dst, axis, react, rl, st);
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// The reaction list of lmx_run_steps_static, for callers that cannot allocate
// device memory: a file-scope device array (each GPU holds its own copy until
// the process ends) of kListShards shards x kStaticListCap records of up to
// two 16-byte units, laid out as a solver's own list.  Its device address comes from hipGetSymbolAddress, once
// per device with that device current; host code never passes the symbol
// itself to a kernel.  Every call on one GPU shares the array; the list
// counters live in each solver's control block and are cleared by the
// resolver after every step, so the array needs no clearing, but calls that
// overlap on one GPU race on it.
// ---------------------------------------------------------------------------
constexpr size_t kStaticListEntries = (size_t)kListShards * kStaticListCap;
__device__ __align__(256) static uint4 scratch[kStaticListEntries * 2];  // records of up to 2 units (W <= 4)

hipError_t static_reaction_list(int device, uint32_t W, ReactionList *out)
{
    static std::mutex mu;
    static uint4 *ptr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices || W < 1 || W > 4) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        ptr[device] = static_cast<uint4 *>(p);
    }
    out->rec = ptr[device];
    out->cap = kStaticListCap;
    return hipSuccess;
}

hipError_t launch_reaction_list(const Params &p, uint32_t *lat, const ReactionList &rl, int num_cu, hipStream_t st)
{
    const unsigned nb = (unsigned)std::min(kListShards, (num_cu > 0 ? num_cu : 64) * 4);
    const dim3 grid(nb), block(256);
    switch (p.W) {
    case 1: hipLaunchKernelGGL(reaction_list_kernel<1>, grid, block, 0, st, p, lat, rl, (uint32_t)kListShards); break;
    case 2: hipLaunchKernelGGL(reaction_list_kernel<2>, grid, block, 0, st, p, lat, rl, (uint32_t)kListShards); break;
    case 4: hipLaunchKernelGGL(reaction_list_kernel<4>, grid, block, 0, st, p, lat, rl, (uint32_t)kListShards); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_reaction_fast(const Params &p, uint32_t *lat, hipStream_t st)
{
    const dim3 grid(blocks_for(p.N >> 2, 256 * kRG)), block(256);
    switch (p.W) {
    case 1: hipLaunchKernelGGL(reaction4_kernel<1>, grid, block, 0, st, p, lat); break;
    case 2: hipLaunchKernelGGL(reaction4_kernel<2>, grid, block, 0, st, p, lat); break;
    case 4: hipLaunchKernelGGL(reaction4_kernel<4>, grid, block, 0, st, p, lat); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_prob_tables(const Params &p, float *prob1, float *prob0, hipStream_t st)
{
    const size_t n = (size_t)p.nST * p.nS > p.nST ? (size_t)p.nST * p.nS : p.nST;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(prob_table_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, prob1, prob0);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched launches: block b of a launch belongs to member map[b] (its blocks
// are start[r] .. start[r+1]-1) and runs that member's share with the
// single-solver body.  x reads lat[0], y writes it back, z and the reaction
// phase leave the step in lat[1].
// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(256) void batch_x4_kernel(const Member *__restrict__ m,
                                                       const uint32_t *__restrict__ start,
                                                       const uint32_t *__restrict__ map, uint32_t n)
{
    const uint32_t r = member_of(map, n, blockIdx.x);
    const Member &M = m[r];
    const uint32_t nb = start[r + 1] - start[r], units = ((M.p.N >> 2) + 255u) / 256u;  // x4_blocks
    const unsigned long long h = M.h[0];
    for (uint32_t u = blockIdx.x - start[r]; u < units; u += nb) x4_body<W>(M.p, M.lat[0], M.lat[1], h, u);
}

// y (AXIS 1) or z (AXIS 2) sweeps, each block looping over units of its
// member (batch_blocks); a z sweep of a member whose reaction phase runs over
// its list fills that list (the others pass no list).
template <int W, int T, int AXIS>
__global__ __launch_bounds__(T, 8) void batch_tile_kernel(const Member *__restrict__ m,
                                                          const uint32_t *__restrict__ start,
                                                          const uint32_t *__restrict__ map, uint32_t n)
{
    const uint32_t r = member_of(map, n, blockIdx.x);
    const Member &M = m[r];
    const uint32_t *src = M.lat[AXIS == 1 ? 1 : 0];
    uint32_t *dst = M.lat[AXIS == 1 ? 0 : 1];
    const uint32_t nb = start[r + 1] - start[r], units = tile_blocks(M.geo[AXIS - 1]);
    const ReactionList rl = AXIS == 2 && M.react == kReactList ? M.rl : ReactionList{};
    const unsigned long long h = M.h[AXIS];
    __shared__ TileLds<W, T> lds;
    for (uint32_t u = blockIdx.x - start[r]; u < units; u += nb) {
        tile_body<W, AXIS == 2, T>(M.p, src, dst, M.geo[AXIS - 1], AXIS, h, u, rl, lds);
        if (u + nb < units) __syncthreads();  // the next unit rewrites the LDS rows
    }
}

// Reaction phase: the full pass, or the member's list (its share of blocks
// walks the shards its z sweep filled).
template <int W>
__global__ __launch_bounds__(256, 4) void batch_reaction_kernel(const Member *__restrict__ m,
                                                             const uint32_t *__restrict__ start,
                                                             const uint32_t *__restrict__ map, uint32_t n)
{
    const uint32_t r = member_of(map, n, blockIdx.x);
    const Member &M = m[r];
    const uint32_t block = blockIdx.x - start[r], nb = start[r + 1] - start[r];
    uint32_t *lat = M.lat[1];
    if (M.react == kReactList && !M.p.ctl->list_overflow) {
        reaction_list_body<W>(M.p, lat, M.rl, block, nb, M.nshards, !M.list_by_block);
        return;
    }
    // the full pass (this member's blocks stride over its reaction blocks)
    const uint32_t nblk = ((M.p.N >> 2) + 256u * kRG - 1u) / (256u * kRG);
    for (uint32_t blk = block; blk < nblk; blk += nb) reaction4_body<W>(M.p, lat, blk);
}

uint32_t x4_blocks(const Params &p) { return blocks_for(p.N >> 2, 256); }
uint32_t reaction_blocks(const Params &p) { return p.R > 0 ? blocks_for(p.N >> 2, 256 * kRG) : 0u; }

// List shards a batched z sweep of this geometry fills (blocks of `threads`),
// and the list blocks that serve them: a block per shard for a member of at
// least kBatchListBlockSites sites (its shards hold many entries), else four
// waves of about eight shards each (small lattices fill many shards with a
// few entries each, and fewer blocks pay fewer member lookups).
uint32_t batch_list_shards(const TileGeom &zgeo, uint32_t threads)
{
    return list_shards_used(tile_blocks(zgeo), threads);
}
bool batch_list_by_block(uint32_t sites) { return sites >= kBatchListBlockSites; }
uint32_t batch_list_blocks(uint32_t nshards, uint32_t sites)
{
    return batch_list_by_block(sites) ? nshards : (nshards + 31u) / 32u;
}

hipError_t launch_batch_sweeps(const Member *m, const uint32_t *const start[4], const uint32_t *const map[4],
                               uint32_t n, const uint32_t *total, uint32_t W, uint32_t tile_threads_used,
                               hipStream_t st)
{
    if (tile_threads_used == 0 || tile_threads_used % 64u || tile_threads_used > tile_threads_for(W))
        return hipErrorInvalidValue;
    const dim3 tb(tile_threads_used);
    const bool narrow = tile_threads_used <= kNarrowTileThreads;  // LDS sized for the narrow tile
    switch (W) {
#define LMX_BATCH_CASE(WW)                                                                                        \
    case WW:                                                                                                      \
        hipLaunchKernelGGL(batch_x4_kernel<WW>, dim3(total[0]), dim3(256), 0, st, m, start[0], map[0], n);        \
        if (narrow) {                                                                                             \
            hipLaunchKernelGGL((batch_tile_kernel<WW, kNarrowTileThreads, 1>), dim3(total[1]), tb, 0, st, m,      \
                               start[1], map[1], n);                                                              \
            hipLaunchKernelGGL((batch_tile_kernel<WW, kNarrowTileThreads, 2>), dim3(total[2]), tb, 0, st, m,      \
                               start[2], map[2], n);                                                              \
        } else {                                                                                                  \
            hipLaunchKernelGGL((batch_tile_kernel<WW, tile_threads<WW>(), 1>), dim3(total[1]), tb, 0, st, m,      \
                               start[1], map[1], n);                                                              \
            hipLaunchKernelGGL((batch_tile_kernel<WW, tile_threads<WW>(), 2>), dim3(total[2]), tb, 0, st, m,      \
                               start[2], map[2], n);                                                              \
        }                                                                                                         \
        if (total[3])                                                                                             \
            hipLaunchKernelGGL(batch_reaction_kernel<WW>, dim3(total[3]), dim3(256), 0, st, m, start[3], map[3], n); \
        break;
        LMX_BATCH_CASE(1)
        LMX_BATCH_CASE(2)
        LMX_BATCH_CASE(4)
#undef LMX_BATCH_CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace lmx
