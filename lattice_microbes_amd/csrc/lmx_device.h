// Device-side building blocks of the MPD-RDME step: the counter-based hash
// RNG, the project double exp, slot packing and the per-launch parameter
// block.  Reference paths are relative to the Lattice Microbes 2.5.0 tree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lmx {

constexpr uint32_t kMaxOverflows = 2048;        // TUNE_MPD_MAX_PARTICLE_OVERFLOWS (options.cmake:88)
constexpr uint32_t kMaxReplacementDist = 4;     // TUNE_MPD_MAX_OVERFLOW_REPLACEMENT_DIST (:87)

// Reaction order codes of the MPD kernels (byte_reaction_dev.cu:56-59).
constexpr uint32_t kZeroOrder = 1, kFirstOrder = 2, kSecondOrder = 3, kSecondOrderSelf = 4;

// Sticky device error codes (mirrored by lmx_status).
constexpr uint32_t kErrOverflowCap = 3, kErrUnplaceable = 4, kErrOutOfRange = 5,
                   kErrReactionUnderflow = 6, kErrInvalidLattice = 9;

// Per-solver control block in device memory.  The step counter lives here so
// that the step sequence does not depend on host-side arguments.
constexpr int kListShards = 1024;

struct Control {
    uint32_t timestep;          // advanced by the resolver at the end of each step
    uint32_t ovf_count;         // entries appended this step (may exceed the cap)
    uint32_t error;             // first error code, 0 if none
    uint32_t error_step;        // timestep of the first error
    unsigned long long ovf_total;
    unsigned long long ovf_steps;
    // Relaxed handler: entries whose own site was full at the end of a step,
    // in list order, retried after the next step before its own entries.
    uint32_t carry_count;
    uint32_t carry_pad;
    unsigned long long ovf_keys[kMaxOverflows];
    uint32_t ovf_part[kMaxOverflows];   // the particle of list entry k
    uint32_t carry_site[kMaxOverflows];
    uint32_t carry_part[kMaxOverflows];
    // The reaction list a fused z sweep fills: entries per shard (may exceed
    // the capacity) and whether any shard overflowed.  The resolver clears
    // both at the end of each step.
    uint32_t list_overflow;
    uint32_t list_count[kListShards];
};

// Sites queued for the reaction phase by a z sweep that lists its output: one
// record per site of list_units(W) 16-byte units, {site index, word 0, word 1,
// type} and for W = 4 a second unit {word 2, word 3, 0, 0}, so that a list
// entry is one vector store and one vector load.  Shard s holds records
// [s*cap, (s+1)*cap); a wave appends to shard (wave index) % kListShards.
struct ReactionList {
    uint4 *rec;
    uint32_t cap;
};

__host__ __device__ constexpr uint32_t list_units(uint32_t W) { return W <= 2 ? 1u : 2u; }

// Overflow entry key: canonical order is phase, then site index, then
// emission order within the site; the low 11 bits hold the entry's position
// in the step's list, where Control::ovf_part keeps its particle (a byte, or
// a 32-bit word of the uint32 lattice).
__host__ __device__ inline unsigned long long overflow_key(uint32_t phase, uint32_t site, uint32_t emission,
                                                           uint32_t entry)
{
    return ((unsigned long long)phase << 62) | ((unsigned long long)site << 24) |
           ((unsigned long long)(emission & 0x1FFFu) << 11) | (unsigned long long)(entry & 0x7FFu);
}

// Everything a kernel needs about one replicate.
struct Params {
    uint32_t X, Y, Z, XY, N;
    uint32_t P, W, pcb;         // slots, 32-bit words per site, RNG key shift
    uint32_t word_particles;    // 1: one particle per 32-bit word (W == P), else 4 per word
    uint32_t packed;            // 1: every site of the step's input holds its particles in slots
                                // 0..n-1, so a word after one whose last slot is empty is zero
    uint32_t plane_flags;       // 1: the kernels keep the chunk flags of planes k >= 1 (below)
    uint32_t relaxed;           // overflow handler: 0 classic (nearby sites), 1 relaxed (own site)
    uint32_t nS, nST, R;
    uint32_t boundary;          // 0 periodic, 1 reflecting
    uint32_t seed32;
    uint32_t any_qp0;           // some site type has a zero-order propensity
    const float *T;             // [nST][nST][nS]
    const float *qp0, *qp1, *qp2;
    const float *rate;          // [R]
    const float *prob1;         // [nST][nS] float2 (total propensity, probability) of a one-particle site
    const float *prob0;         // [nST] probability of an empty site (zero-order only)
    const uint32_t *order, *D1, *D2;
    const int8_t *S;            // [R][nS]
    // Reactions by first reactant: rx_list[rx_off[s]..rx_off[s+1]) lists, in
    // ascending order, the reactions whose first reactant is species s; list
    // nS holds the zero-order reactions.
    const uint32_t *rx_off, *rx_list;
    // Nonzero stoichiometry of reaction r, ascending species:
    // (sto_sp[k], sto_val[k]) for k in [sto_off[r], sto_off[r+1]).
    const uint32_t *sto_off, *sto_sp;
    const int32_t *sto_val;
    const uint8_t *RL;          // [R][nST]
    const uint8_t *sites;       // [N]
    Control *ctl;
};

// ---------------------------------------------------------------------------
// RNG: the 64-bit multiply-xorshift hash of xor_random_dev.cu:46-53 keyed by
// (site, slot) under the per-substep seed of MpdRdmeSolver.cu:987-994.
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t hash32(unsigned long long v, unsigned long long seed)
{
    v ^= seed;
    v = v * 3935559000370003845ULL + 2691343689449507681ULL;
    v ^= v >> 21;
    v ^= v << 37;
    v ^= v >> 4;
    // High word of v * C (mod 2^64) in 32-bit parts, so that the result is a
    // 32-bit value to the compiler (its float conversion is then one
    // instruction, not the 64-bit conversion sequence).
    constexpr unsigned long long C = 2685821657736338717ULL;
    const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    const uint32_t clo = (uint32_t)C, chi = (uint32_t)(C >> 32);
    return (uint32_t)(((unsigned long long)lo * clo) >> 32) + lo * chi + hi * clo;
}

__host__ __device__ inline unsigned long long timestep_hash(uint32_t seed32, uint32_t t, uint32_t substep)
{
    unsigned long long h = (((((unsigned long long)seed32) << 30) + t) << 2) + substep;
    h = h * 3202034522624059733ULL + 4354685564936845319ULL;
    h ^= h >> 20;
    h ^= h << 41;
    h ^= h >> 5;
    h *= 7664345821815920749ULL;
    return h;
}

// (float)hash * 2^-32: conversion rounds to nearest, the scaling is exact.
__device__ inline float uniform01(uint32_t v1, uint32_t shift, uint32_t v2, unsigned long long seed)
{
    unsigned long long v = ((unsigned long long)v1 << shift) | (unsigned long long)v2;
    return (float)hash32(v, seed) * 0x1p-32f;
}

// ---------------------------------------------------------------------------
// Project double exp, identical bit for bit wherever it is compiled with
// -ffp-contract=off: Cody-Waite reduction x = k ln2 + r (ln2 head with 21
// trailing zero bits, so k*LN2_HI is exact for |k| < 2^11), degree-13 Taylor
// polynomial of exp(r)-1-r by Horner with fused multiply-adds, 1 + r_hi
// summed exactly (Fast2Sum), scaled by 2^k.
// ---------------------------------------------------------------------------
__device__ inline double exp_proj(double x)
{
    if (x != x) return x + x;
    if (x > 709.782712893384) return __builtin_huge_val();
    if (x < -745.1332191019412) return 0.0;
    const double kd = __builtin_rint(x * 0x1.71547652b82fep+0);
    const double rh = __builtin_fma(-kd, 0x1.62e42fee00000p-1, x);
    const double rl = -kd * 0x1.a39ef35793c76p-33;
    const double r = rh + rl;
    double q = 0x1.6124613a86d09p-33;
    q = __builtin_fma(q, r, 0x1.1eed8eff8d898p-29);
    q = __builtin_fma(q, r, 0x1.ae64567f544e4p-26);
    q = __builtin_fma(q, r, 0x1.27e4fb7789f5cp-22);
    q = __builtin_fma(q, r, 0x1.71de3a556c734p-19);
    q = __builtin_fma(q, r, 0x1.a01a01a01a01ap-16);
    q = __builtin_fma(q, r, 0x1.a01a01a01a01ap-13);
    q = __builtin_fma(q, r, 0x1.6c16c16c16c17p-10);
    q = __builtin_fma(q, r, 0x1.1111111111111p-7);
    q = __builtin_fma(q, r, 0x1.5555555555555p-5);
    q = __builtin_fma(q, r, 0x1.5555555555555p-3);
    q = __builtin_fma(q, r, 0x1.0000000000000p-1);
    const double r2 = r * r;
    const double p = q * r2;
    const double s = 1.0 + rh;
    const double e = (1.0 - s) + rh;
    const double tail = e + (rl + p);
    const double y = s + tail;
    return __builtin_ldexp(y, (int)kd);
}

// calculateReactionProbability, byte_reaction_dev.cu:130-137.
__device__ inline float reaction_probability(float a)
{
    return (float)(1.0 - exp_proj(-(double)a));
}

// lookupTransitionProbability, byte_diffusion_1d_dev.cu:70-79, in table Tt
// (p.T or a copy of it, e.g. in LDS).  The three range tests are combined
// into one condition, so that one branch guards the load.
// With en false the result is 0 (a neighbour that does not exist).  The
// index products use 24-bit multiplies (full rate; 32-bit integer multiplies
// are quarter rate).  Both are exact: nST <= 256 gives src * nST + dst <
// 2^16, and the model check nST * nS^2 < 2^32 gives nS < 2^16 and a table of
// nST^2 * nS < 2^32 entries.
__device__ inline float lookup_T_if(const Params &p, const float *Tt, bool en, uint32_t particle, uint32_t src,
                                    uint32_t dst)
{
    const uint32_t sp = particle - 1u;
    const bool ok = en & (sp < p.nS) & (dst < p.nST) & (src < p.nST);
    return ok ? Tt[__umul24(__umul24(src, p.nST) + dst, p.nS) + sp] : 0.0f;
}

__device__ inline float lookup_T_in(const Params &p, const float *Tt, uint32_t particle, uint32_t src, uint32_t dst)
{
    return lookup_T_if(p, Tt, true, particle, src, dst);
}

// The same lookup in p.T.
__device__ inline float lookup_T(const Params &p, uint32_t particle, uint32_t src, uint32_t dst)
{
    return lookup_T_in(p, p.T, particle, src, dst);
}

// ---------------------------------------------------------------------------
// Slot packing.  A site's particles are W 32-bit words (slot 4w+b = byte b of
// word w).  SlotBuf appends particles in order and keeps the first P.
// ---------------------------------------------------------------------------
template <int W>
struct SlotBuf {
    uint32_t w[W];
    uint32_t n;
    __device__ inline void clear()
    {
#pragma unroll
        for (int k = 0; k < W; k++) w[k] = 0;
        n = 0;
    }
    // Returns true if the particle fits; otherwise the caller records an overflow.
    __device__ inline bool push(uint32_t particle)
    {
        const uint32_t k = n;
        n = k + 1;
        if (k >= 4u * W) return false;
#pragma unroll
        for (int q = 0; q < W; q++)
            if ((k >> 2) == (uint32_t)q) w[q] |= particle << (8 * (k & 3));
        return true;
    }
};

__device__ inline uint32_t slot_of(const uint32_t *w, int j)
{
    return (w[j >> 2] >> (8 * (j & 3))) & 0xFFu;
}

// Appends an overflow entry (phase, site, emission, particle) to the step's list.
__device__ inline void record_overflow(Control *ctl, uint32_t phase, uint32_t site, uint32_t emission,
                                       uint32_t particle)
{
    const uint32_t k = atomicAdd(&ctl->ovf_count, 1u);
    if (k < kMaxOverflows) {
        ctl->ovf_part[k] = particle;
        ctl->ovf_keys[k] = overflow_key(phase, site, emission, k);
    }
}

// Whether word k of a site (k >= 1) can hold a particle, given word k-1: on a
// packed lattice only if word k-1's last slot is occupied.  Every kernel
// writes packed sites, so only a lattice written from outside can be
// unpacked (Params::packed is 0 for the first step after that).
__device__ inline bool need_next_byte_word(const Params &p, uint32_t prev)
{
    return !p.packed || (prev >> 24) != 0u;
}

__device__ inline bool need_next_particle_word(const Params &p, uint32_t prev)
{
    return !p.packed || prev != 0u;
}

// Chunk flags of the byte lattice (Params::plane_flags): after a buffer's W
// planes, for each plane k >= 1, one word per chunk of kChunkGroups groups
// (256 sites, aligned), 1 while that chunk of plane k is known to be all
// zero.  Every kernel that writes a nonzero word into a chunk of plane k
// clears its flag; a wave that writes a whole chunk and writes only zeros sets
// it; a zero word may be left unwritten if its chunk's flag was 1 when the
// launch began (each word has one writer per launch).  The host zeroes the
// flags (nothing known) whenever it writes a buffer or runs kernels that do
// not keep them.
constexpr uint32_t kChunkGroups = 64;

__host__ __device__ inline uint32_t chunk_count(uint32_t N) { return ((N >> 2) + kChunkGroups - 1) / kChunkGroups; }

__device__ inline uint32_t *chunk_flag(const Params &p, uint32_t *lat, uint32_t k, uint32_t g)
{
    return lat + (size_t)p.W * p.N + (size_t)(k - 1) * chunk_count(p.N) + g / kChunkGroups;
}

// Clears the flag of group g's chunk in plane k >= 1 after a nonzero write.
__device__ inline void clear_chunk_flag(const Params &p, uint32_t *lat, uint32_t k, uint32_t g)
{
    if (!p.plane_flags) return;
    uint32_t *f = chunk_flag(p, lat, k, g);
    if (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
        __hip_atomic_store(f, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One particle per 32-bit word (the uint32 lattice of IntMpdRdmeSolver):
// WordBuf appends particles in order and keeps the first P.
template <int P>
struct WordBuf {
    uint32_t w[P];
    uint32_t n;
    __device__ inline void clear()
    {
#pragma unroll
        for (int k = 0; k < P; k++) w[k] = 0;
        n = 0;
    }
    __device__ inline bool push(uint32_t particle)
    {
        const uint32_t k = n;
        n = k + 1;
        if (k >= (uint32_t)P) return false;
#pragma unroll
        for (int q = 0; q < P; q++)
            if (k == (uint32_t)q) w[q] = particle;
        return true;
    }
};

// Slot layouts of a site's W words.  ByteSites<W>: slot j is byte j % 4 of
// word j / 4 (P = 4W, byte_diffusion_1d_dev.cu).  WordSites<P>: slot j is
// word j (W = P, word_diffusion_1d_dev.cu).
template <int W_>
struct ByteSites {
    static constexpr int W = W_, P = 4 * W_, kPerWord = 4;
    using Buf = SlotBuf<W_>;
    __device__ static inline bool need_next(const Params &p, uint32_t prev) { return need_next_byte_word(p, prev); }
    __device__ static inline uint32_t part(uint32_t word, int q) { return (word >> (8 * q)) & 0xFFu; }
    __device__ static inline uint32_t slot(const uint32_t *w, int j) { return slot_of(w, j); }
};

template <int P_>
struct WordSites {
    static constexpr int W = P_, P = P_, kPerWord = 1;
    using Buf = WordBuf<P_>;
    __device__ static inline bool need_next(const Params &p, uint32_t prev) { return need_next_particle_word(p, prev); }
    __device__ static inline uint32_t part(uint32_t word, int) { return word; }
    __device__ static inline uint32_t slot(const uint32_t *w, int j) { return w[j]; }
};

}  // namespace lmx
