/*
 * Serial CPU restatement of the MPD-RDME timestep.  TEST INFRASTRUCTURE ONLY:
 * the checker for the HIP product path, never part of it.  Compile with
 * -ffp-contract=off so that every float/double operation is the single IEEE
 * operation written here.
 *
 * Reference paths are relative to the Lattice Microbes 2.5.0 source tree.
 */
#include "mpd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* RNG                                                                      */
/* ------------------------------------------------------------------------ */

/* Counter-based hash, src/rdme/dev/xor_random_dev.cu:46-53 (also the CPU
 * XORShift::getRandom, src/rng/XORShift.cpp:54-62). */
uint32_t orc_hash(uint64_t v, uint64_t seed)
{
    v ^= seed;
    v = v * 3935559000370003845ULL + 2691343689449507681ULL;
    v ^= v >> 21;
    v ^= v << 37;
    v ^= v >> 4;
    v *= 2685821657736338717ULL;
    return (uint32_t)(v >> 32);
}

/* MpdRdmeSolver::getTimestepSeed, src/rdme/MpdRdmeSolver.cu:987-994. */
uint64_t orc_timestep_hash(uint32_t seed32, uint32_t timestep, uint32_t substep)
{
    uint64_t h = (((((uint64_t)seed32) << 30) + timestep) << 2) + substep;
    h = h * 3202034522624059733ULL + 4354685564936845319ULL;
    h ^= h >> 20;
    h ^= h << 41;
    h ^= h >> 5;
    h *= 7664345821815920749ULL;
    return h;
}

/* getRandomHashFloat(v1, shift, v2, seed), xor_random_dev.cu:66-81: the
 * u32 hash converted to float (round to nearest) times 2^-32. */
float orc_uniform(uint32_t v1, uint32_t shift, uint32_t v2, uint64_t seed)
{
    uint64_t v = ((uint64_t)v1 << shift) | (uint64_t)v2;
    return (float)orc_hash(v, seed) * 2.32830643654e-10f;
}

/* MPD_PARTICLE_COUNT_BITS, src/rdme/dev/byte_diffusion_1d_dev.cu:51-58. */
uint32_t orc_particle_count_bits(uint32_t P)
{
    if (P <= 4) return 2;
    if (P <= 8) return 3;
    return 4;
}

/* ------------------------------------------------------------------------ */
/* Project double exp.                                                      */
/*                                                                          */
/* The reference evaluates (float)(1.0-exp(-(double)a)) with the platform   */
/* double exp (byte_reaction_dev.cu:130-137).  Here the project's own exp,  */
/* identical on host and device: Cody-Waite argument reduction              */
/* (x = k*ln2 + r, ln2 split into a 32-bit head with trailing zeros and a   */
/* tail) followed by the degree-13 Taylor polynomial of exp(r) for          */
/* |r| <= ln2/2, the leading 1+r_hi summed exactly (Fast2Sum), the result   */
/* scaled by 2^k.                                                           */
/* ------------------------------------------------------------------------ */
static const double ORC_LOG2E = 1.4426950408889634;
static const double ORC_LN2_HI = 6.93147180369123816490e-01; /* 0x3fe62e42fee00000 */
static const double ORC_LN2_LO = 1.90821492927058770002e-10; /* ln2 - LN2_HI */

double orc_exp(double x)
{
    if (x != x) return x + x;
    if (x > 709.782712893384) return HUGE_VAL;
    if (x < -745.1332191019412) return 0.0;
    double kd = rint(x * ORC_LOG2E);
    double rh = fma(-kd, ORC_LN2_HI, x);
    double rl = -kd * ORC_LN2_LO;
    double r = rh + rl;
    double q = 1.0 / 6227020800.0;           /* 1/13! */
    q = fma(q, r, 1.0 / 479001600.0);        /* 1/12! */
    q = fma(q, r, 1.0 / 39916800.0);         /* 1/11! */
    q = fma(q, r, 1.0 / 3628800.0);          /* 1/10! */
    q = fma(q, r, 1.0 / 362880.0);           /* 1/9!  */
    q = fma(q, r, 1.0 / 40320.0);            /* 1/8!  */
    q = fma(q, r, 1.0 / 5040.0);             /* 1/7!  */
    q = fma(q, r, 1.0 / 720.0);              /* 1/6!  */
    q = fma(q, r, 1.0 / 120.0);              /* 1/5!  */
    q = fma(q, r, 1.0 / 24.0);               /* 1/4!  */
    q = fma(q, r, 1.0 / 6.0);                /* 1/3!  */
    q = fma(q, r, 0.5);                      /* 1/2!  */
    double r2 = r * r;
    double p = q * r2;                       /* exp(r) - 1 - r */
    double s = 1.0 + rh;
    double e = (1.0 - s) + rh;               /* exact: |rh| <= 1 */
    double tail = e + (rl + p);
    double y = s + tail;
    return ldexp(y, (int)kd);
}

/* calculateReactionProbability, byte_reaction_dev.cu:130-137 (default
 * RNG_CUDA_DOUBLE_PRECISION branch). */
float orc_reaction_probability(float a)
{
    return (float)(1.0 - orc_exp(-(double)a));
}

/* ------------------------------------------------------------------------ */
/* Model tables                                                             */
/* ------------------------------------------------------------------------ */

/* CMESolver::buildModel (src/cme/CMESolver.cpp:307-420) restricted to the
 * four reaction types MPD accepts, then MpdRdmeSolver::buildModel
 * (src/rdme/MpdRdmeSolver.cu:264-325): order codes, D1/D2 = reactant+1,
 * S transposed to reaction-major int8. */
int orc_build_reaction_tables(uint32_t nS, uint32_t R, const uint32_t *types,
                              const double *K, uint32_t kCols, const int32_t *S,
                              const uint32_t *D,
                              uint32_t *order_out, uint32_t *D1_out, uint32_t *D2_out,
                              int8_t *S_out, uint32_t *si_out, uint32_t *s2i_out, double *k_out)
{
    if (R > 0 && kCols == 0) return ORC_ERR_ARG;
    for (uint32_t i = 0; i < R; i++) {
        uint32_t deps[3] = {0, 0, 0}, nd = 0;
        for (uint32_t j = 0; j < nS; j++) {
            if (D[j * R + i] == 1) {
                if (nd < 3) deps[nd] = j;
                nd++;
            }
        }
        k_out[i] = K[i * kCols];
        switch (types[i]) {
        case 0:
            if (nd != 0) return ORC_ERR_ARG;
            order_out[i] = ORC_ZERO_ORDER; D1_out[i] = 0; D2_out[i] = 0;
            si_out[i] = 0; s2i_out[i] = 0;
            break;
        case 1:
            if (nd != 1) return ORC_ERR_ARG;
            order_out[i] = ORC_FIRST_ORDER; D1_out[i] = deps[0] + 1; D2_out[i] = 0;
            si_out[i] = deps[0]; s2i_out[i] = 0;
            break;
        case 2:
            if (nd != 2) return ORC_ERR_ARG;
            order_out[i] = ORC_SECOND_ORDER; D1_out[i] = deps[0] + 1; D2_out[i] = deps[1] + 1;
            si_out[i] = deps[0]; s2i_out[i] = deps[1];
            break;
        case 3:
            if (nd != 1) return ORC_ERR_ARG;
            order_out[i] = ORC_SECOND_ORDER_SELF; D1_out[i] = deps[0] + 1; D2_out[i] = 0;
            si_out[i] = deps[0]; s2i_out[i] = 0;
            break;
        default:
            return ORC_ERR_ARG;
        }
    }
    for (uint32_t rx = 0; rx < R; rx++)
        for (uint32_t p = 0; p < nS; p++)
            S_out[rx * nS + p] = (int8_t)S[R * p + rx];
    return ORC_OK;
}

/* buildDiffusionModel transition table, MpdRdmeSolver.cu:408-421. */
int orc_build_T(uint32_t nS, uint32_t nST, const double *DF, double tau, double lambda, float *T_out)
{
    size_t n = (size_t)nS * nST * nST;
    for (size_t i = 0; i < n; i++) {
        float q = (float)(DF[i] * tau / pow(lambda, 2));
        if (q > 0.50f) return ORC_ERR_ARG;
        T_out[i] = q;
    }
    return ORC_OK;
}

/* MpdRdmeSolver::computePropensities, MpdRdmeSolver.cu:697-774.  The float
 * members are updated with double arithmetic exactly as the compound
 * assignments there evaluate. */
int orc_compute_propensities(uint32_t nS, uint32_t nST, uint32_t R, const uint32_t *types,
                             const double *k, const uint32_t *si, const uint32_t *s2i,
                             const uint32_t *RL, double tau, uint32_t X, uint32_t Y, uint32_t Z,
                             float *rate_out, float *qp0, float *qp1, float *qp2)
{
    uint32_t nsites32 = X * Y * Z; /* unsigned product, as in the reference */
    for (uint32_t i = 0; i < R; i++) {
        switch (types[i]) {
        case 0: rate_out[i] = (float)(k[i] * tau / (double)nsites32); break;
        case 1: rate_out[i] = (float)(k[i] * tau); break;
        case 2:
        case 3: rate_out[i] = (float)(k[i] * tau * (double)X * (double)Y * (double)Z); break;
        default: return ORC_ERR_ARG;
        }
    }
    float scale = (float)nsites32;
    for (uint32_t site = 0; site < nST; site++) {
        uint32_t o1 = site * nS, o2 = site * nS * nS;
        qp0[site] = 0.0f;
        for (uint32_t i = 0; i < nS; i++) {
            qp1[o1 + i] = 0.0f;
            for (uint32_t j = 0; j < nS; j++) qp2[o2 + i * nS + j] = 0.0f;
        }
        for (uint32_t i = 0; i < R; i++) {
            if (!RL[i * nST + site]) continue;
            switch (types[i]) {
            case 0: qp0[site] = (float)((double)qp0[site] + k[i] * tau / (double)scale); break;
            case 1: qp1[o1 + si[i]] = (float)((double)qp1[o1 + si[i]] + k[i] * tau); break;
            case 2: {
                float v = (float)(k[i] * tau * (double)scale);
                qp2[o2 + si[i] * nS + s2i[i]] = v;
                qp2[o2 + s2i[i] * nS + si[i]] = v;
            } break;
            case 3:
                qp2[o2 + si[i] * nS + si[i]] =
                    (float)((double)qp2[o2 + si[i] * nS + si[i]] + k[i] * tau * (double)scale * 2);
                break;
            }
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Lattice layout                                                           */
/* ------------------------------------------------------------------------ */

/* Byte lattice: slot j of site i is byte j%4 of word i of plane j/4
 * (ByteLattice.cpp:371-435).  uint32 lattice: slot j is word i of plane j
 * (IntLattice.cpp, one particle per word). */
static inline uint32_t byte_get(const uint32_t *lat, size_t N, size_t i, uint32_t j)
{
    return (lat[(size_t)(j >> 2) * N + i] >> (8 * (j & 3))) & 0xFFu;
}

static inline void byte_set(uint32_t *lat, size_t N, size_t i, uint32_t j, uint32_t v)
{
    uint32_t *w = &lat[(size_t)(j >> 2) * N + i];
    uint32_t sh = 8 * (j & 3);
    *w = (*w & ~(0xFFu << sh)) | ((v & 0xFFu) << sh);
}

static inline int word_mode(const orc_model *m) { return m->bpp == 4; }

static inline uint32_t slot_get(const orc_model *m, const uint32_t *lat, size_t N, size_t i, uint32_t j)
{
    return word_mode(m) ? lat[(size_t)j * N + i] : byte_get(lat, N, i, j);
}

static inline void slot_set(const orc_model *m, uint32_t *lat, size_t N, size_t i, uint32_t j, uint32_t v)
{
    if (word_mode(m)) lat[(size_t)j * N + i] = v;
    else byte_set(lat, N, i, j, v);
}

/* ByteLattice::getOccupancy, ByteLattice.cpp:206-222 (IntLattice.cpp:172-190):
 * slots up to the first 0. */
static uint32_t occupancy(const orc_model *m, const uint32_t *lat, size_t N, size_t i)
{
    for (uint32_t j = 0; j < m->P; j++)
        if (slot_get(m, lat, N, i, j) == 0) return j;
    return m->P;
}

/* ByteLattice::addParticle, ByteLattice.cpp:264-285: first empty slot. */
static int add_particle(const orc_model *m, uint32_t *lat, size_t N, size_t i, uint32_t p)
{
    for (uint32_t j = 0; j < m->P; j++) {
        if (slot_get(m, lat, N, i, j) == 0) {
            slot_set(m, lat, N, i, j, p);
            return 1;
        }
    }
    return 0;
}

/* RDMESolver::setLatticeData, RDMESolver.cpp:144-167. */
static int from_row_major(uint32_t X, uint32_t Y, uint32_t Z, uint32_t P, uint32_t nS, uint32_t bpp,
                          const void *rm, uint32_t *native)
{
    orc_model m;
    memset(&m, 0, sizeof(m));
    m.P = P;
    m.bpp = bpp;
    size_t N = (size_t)X * Y * Z;
    memset(native, 0, N * (bpp == 4 ? P : P / 4) * sizeof(uint32_t));
    size_t idx = 0;
    for (uint32_t x = 0; x < X; x++)
        for (uint32_t y = 0; y < Y; y++)
            for (uint32_t z = 0; z < Z; z++)
                for (uint32_t l = 0; l < P; l++, idx++) {
                    uint32_t v = bpp == 4 ? ((const uint32_t *)rm)[idx] : ((const uint8_t *)rm)[idx];
                    if (v != 0) {
                        if (v > nS) return ORC_ERR_ARG;
                        size_t i = (size_t)x + (size_t)X * y + (size_t)X * Y * z;
                        if (!add_particle(&m, native, N, i, v)) return ORC_ERR_ARG;
                    }
                }
    return ORC_OK;
}

int orc_lattice_from_row_major(uint32_t X, uint32_t Y, uint32_t Z, uint32_t P, uint32_t nS,
                               const uint8_t *rm, uint32_t *native)
{
    return from_row_major(X, Y, Z, P, nS, 1, rm, native);
}

/* IntLattice::setFromRowMajorData: the same insertion, 32-bit particles. */
int orc_lattice_from_row_major_words(uint32_t X, uint32_t Y, uint32_t Z, uint32_t P, uint32_t nS,
                                     const uint32_t *rm, uint32_t *native)
{
    return from_row_major(X, Y, Z, P, nS, 4, rm, native);
}

/* ByteLattice::copyNativeToRowMajorByte, ByteLattice.cpp:371-403. */
void orc_lattice_to_row_major(uint32_t X, uint32_t Y, uint32_t Z, uint32_t P,
                              const uint32_t *native, uint8_t *rm)
{
    size_t N = (size_t)X * Y * Z;
    for (uint32_t x = 0; x < X; x++)
        for (uint32_t y = 0; y < Y; y++)
            for (uint32_t z = 0; z < Z; z++) {
                size_t i = (size_t)x + (size_t)X * y + (size_t)X * Y * z;
                size_t d = (((size_t)x * Y + y) * Z + z) * P;
                for (uint32_t j = 0; j < P; j++) rm[d + j] = (uint8_t)byte_get(native, N, i, j);
            }
}

/* IntLattice::copyNativeToRowMajor: word j of site i to (x, y, z, j). */
void orc_lattice_to_row_major_words(uint32_t X, uint32_t Y, uint32_t Z, uint32_t P,
                                    const uint32_t *native, uint32_t *rm)
{
    size_t N = (size_t)X * Y * Z;
    for (uint32_t x = 0; x < X; x++)
        for (uint32_t y = 0; y < Y; y++)
            for (uint32_t z = 0; z < Z; z++) {
                size_t i = (size_t)x + (size_t)X * y + (size_t)X * Y * z;
                size_t d = (((size_t)x * Y + y) * Z + z) * P;
                for (uint32_t j = 0; j < P; j++) rm[d + j] = native[(size_t)j * N + i];
            }
}

/* RDMESolver::setLatticeSitesData, RDMESolver.cpp:169-188. */
int orc_sites_from_row_major(uint32_t X, uint32_t Y, uint32_t Z, uint32_t nST,
                             const uint8_t *rm, uint8_t *native)
{
    size_t idx = 0;
    memset(native, 0, (size_t)X * Y * Z);
    for (uint32_t x = 0; x < X; x++)
        for (uint32_t y = 0; y < Y; y++)
            for (uint32_t z = 0; z < Z; z++, idx++) {
                uint8_t v = rm[idx];
                if (v != 0) {
                    if (v >= nST) return ORC_ERR_ARG;
                    native[(size_t)x + (size_t)X * y + (size_t)X * Y * z] = v;
                }
            }
    return ORC_OK;
}

/* ByteLattice::getParticleCounts, ByteLattice.cpp:495-520: each word is
 * counted up to its first empty byte. */
void orc_species_counts(uint32_t N, uint32_t P, const uint32_t *native, uint32_t nS, uint32_t *counts)
{
    memset(counts, 0, nS * sizeof(uint32_t));
    size_t nw = (size_t)N * (P / 4);
    for (size_t w = 0; w < nw; w++) {
        uint32_t word = native[w];
        for (uint32_t b = 0; b < 4; b++) {
            uint32_t p = (word >> (8 * b)) & 0xFF;
            if (p == 0) break;
            if (p <= nS) counts[p - 1]++;
        }
    }
}

/* IntLattice::getParticleCounts, IntLattice.cpp:419-435: every nonzero word. */
void orc_species_counts_words(uint32_t N, uint32_t P, const uint32_t *native, uint32_t nS, uint32_t *counts)
{
    memset(counts, 0, nS * sizeof(uint32_t));
    size_t nw = (size_t)N * P;
    for (size_t w = 0; w < nw; w++) {
        uint32_t p = native[w];
        if (p >= 1 && p <= nS) counts[p - 1]++;
    }
}

/* ------------------------------------------------------------------------ */
/* Diffusion                                                                */
/* ------------------------------------------------------------------------ */

#define MOVE_NONE 0
#define MOVE_STAY 1
#define MOVE_PLUS 2
#define MOVE_MINUS 3
#define NO_SITE ((size_t)-1)

/* lookupTransitionProbability, byte_diffusion_1d_dev.cu:70-79.  A missing
 * neighbour (reflecting boundary) has no type and gives 0. */
static inline float lookup_T(const orc_model *m, uint32_t p, uint32_t src, size_t dst_site,
                             const uint8_t *sites)
{
    if (dst_site == NO_SITE) return 0.0f;
    uint32_t dst = sites[dst_site];
    if ((uint32_t)(p - 1) >= m->nS || dst >= m->nST || src >= m->nST) return 0.0f;
    return m->T[(size_t)src * m->nST * m->nS + (size_t)dst * m->nS + (p - 1)];
}

typedef struct {
    size_t stride;
    uint32_t len;
    size_t N;
} axis_t;

static axis_t make_axis(const orc_model *m, uint32_t axis)
{
    axis_t a;
    a.N = (size_t)m->X * m->Y * m->Z;
    if (axis == 0) { a.stride = 1; a.len = m->X; }
    else if (axis == 1) { a.stride = m->X; a.len = m->Y; }
    else { a.stride = (size_t)m->X * m->Y; a.len = m->Z; }
    return a;
}

static inline uint32_t coord(const axis_t *a, size_t i)
{
    return (uint32_t)((i / a->stride) % a->len);
}

/* Periodic neighbours as ByteLattice::getNeighboringSites
 * (ByteLattice.cpp:81-98) and the periodic window loaders
 * (lattice_sim_1d_dev.cu:125-435); none at a reflecting edge. */
static inline size_t nbr_minus(const orc_model *m, const axis_t *a, size_t i)
{
    uint32_t c = coord(a, i);
    if (c > 0) return i - a->stride;
    return m->boundary ? NO_SITE : i + (size_t)(a->len - 1) * a->stride;
}

static inline size_t nbr_plus(const orc_model *m, const axis_t *a, size_t i)
{
    uint32_t c = coord(a, i);
    if (c + 1 < a->len) return i + a->stride;
    return m->boundary ? NO_SITE : i - (size_t)(a->len - 1) * a->stride;
}

static int push_overflow(uint32_t *ovf_site, uint32_t *ovf_particle, uint32_t *ovf_count,
                         size_t site, uint32_t particle)
{
    uint32_t k = (*ovf_count)++;
    if (k < ORC_MAX_OVERFLOWS) {
        ovf_site[k] = (uint32_t)site;
        ovf_particle[k] = particle;
    }
    return 0;
}

/* Formulation 0: every particle's choice (makeDiffusionChoices,
 * byte_diffusion_1d_dev.cu:86-121), then each site gathers its stayers, the
 * minus neighbour's PLUS movers and the plus neighbour's MINUS movers in slot
 * order, keeps the first P and sends the rest to the overflow list
 * (performPropagation, :199-270). */
static int sweep_choices(const orc_model *m, const uint32_t *src, const uint8_t *sites, uint32_t *dst,
                         uint32_t axis, uint64_t hash,
                         uint32_t *ovf_site, uint32_t *ovf_particle, uint32_t *ovf_count)
{
    axis_t a = make_axis(m, axis);
    size_t N = a.N;
    uint32_t P = m->P, pcb = orc_particle_count_bits(P);
    uint8_t *choice = (uint8_t *)calloc(N * P, 1);
    if (!choice) return ORC_ERR_ARG;
    for (size_t i = 0; i < N; i++) {
        size_t mi = nbr_minus(m, &a, i), pi = nbr_plus(m, &a, i);
        uint32_t st = sites[i];
        for (uint32_t j = 0; j < P; j++) {
            uint32_t p = slot_get(m, src, N, i, j);
            if (p == 0) continue;
            float qm = lookup_T(m, p, st, mi, sites);
            float qp = lookup_T(m, p, st, pi, sites);
            float u = orc_uniform((uint32_t)i, pcb, j, hash);
            uint8_t c = (u < qm) ? MOVE_MINUS : MOVE_STAY;
            c = (u >= 0.5f && u < (qp + 0.5f)) ? MOVE_PLUS : c;
            choice[i * P + j] = c;
        }
    }
    uint32_t nw[48];
    for (size_t i = 0; i < N; i++) {
        uint32_t n = 0;
        size_t mi = nbr_minus(m, &a, i), pi = nbr_plus(m, &a, i);
        for (uint32_t j = 0; j < P; j++)
            if (choice[i * P + j] == MOVE_STAY) nw[n++] = slot_get(m, src, N, i, j);
        if (mi != NO_SITE)
            for (uint32_t j = 0; j < P; j++)
                if (choice[mi * P + j] == MOVE_PLUS) nw[n++] = slot_get(m, src, N, mi, j);
        if (pi != NO_SITE)
            for (uint32_t j = 0; j < P; j++)
                if (choice[pi * P + j] == MOVE_MINUS) nw[n++] = slot_get(m, src, N, pi, j);
        for (uint32_t j = 0; j < P; j++) slot_set(m, dst, N, i, j, j < n ? nw[j] : 0);
        for (uint32_t j = P; j < n; j++) push_overflow(ovf_site, ovf_particle, ovf_count, i, nw[j]);
    }
    free(choice);
    return ORC_OK;
}

/* Formulation 1: each site pulls only the decisions it depends on.  A
 * particle leaves towards minus iff u < qm and towards plus iff
 * 0.5 <= u < qp+0.5; with q <= 0.5 (enforced when T is built) the two
 * are exclusive, so this agrees with formulation 0 bit for bit. */
static int sweep_pull(const orc_model *m, const uint32_t *src, const uint8_t *sites, uint32_t *dst,
                      uint32_t axis, uint64_t hash,
                      uint32_t *ovf_site, uint32_t *ovf_particle, uint32_t *ovf_count)
{
    axis_t a = make_axis(m, axis);
    size_t N = a.N;
    uint32_t P = m->P, pcb = orc_particle_count_bits(P);
    uint32_t nw[48];
    for (size_t i = 0; i < N; i++) {
        uint32_t n = 0;
        size_t mi = nbr_minus(m, &a, i), pi = nbr_plus(m, &a, i);
        uint32_t st = sites[i];
        for (uint32_t j = 0; j < P; j++) {
            uint32_t p = slot_get(m, src, N, i, j);
            if (p == 0) continue;
            float u = orc_uniform((uint32_t)i, pcb, j, hash);
            float qm = lookup_T(m, p, st, mi, sites);
            float qp = lookup_T(m, p, st, pi, sites);
            int leaves = (u < qm) || (u >= 0.5f && u < (qp + 0.5f));
            if (!leaves) nw[n++] = p;
        }
        if (mi != NO_SITE) {
            uint32_t stm = sites[mi];
            for (uint32_t j = 0; j < P; j++) {
                uint32_t p = slot_get(m, src, N, mi, j);
                if (p == 0) continue;
                float u = orc_uniform((uint32_t)mi, pcb, j, hash);
                float qp = lookup_T(m, p, stm, i, sites);
                if (u >= 0.5f && u < (qp + 0.5f)) nw[n++] = p;
            }
        }
        if (pi != NO_SITE) {
            uint32_t stp = sites[pi];
            for (uint32_t j = 0; j < P; j++) {
                uint32_t p = slot_get(m, src, N, pi, j);
                if (p == 0) continue;
                float u = orc_uniform((uint32_t)pi, pcb, j, hash);
                float qm = lookup_T(m, p, stp, i, sites);
                if (u < qm) nw[n++] = p;
            }
        }
        for (uint32_t j = 0; j < P; j++) slot_set(m, dst, N, i, j, j < n ? nw[j] : 0);
        for (uint32_t j = P; j < n; j++) push_overflow(ovf_site, ovf_particle, ovf_count, i, nw[j]);
    }
    return ORC_OK;
}

int orc_sweep(const orc_model *m, const uint32_t *src, const uint8_t *sites, uint32_t *dst,
              uint32_t axis, uint64_t hash, int formulation,
              uint32_t *ovf_site, uint32_t *ovf_particle, uint32_t *ovf_count)
{
    if (axis > 2 || (m->P != 4 && m->P != 8 && m->P != 16)) return ORC_ERR_ARG;
    if (formulation == 0)
        return sweep_choices(m, src, sites, dst, axis, hash, ovf_site, ovf_particle, ovf_count);
    return sweep_pull(m, src, sites, dst, axis, hash, ovf_site, ovf_particle, ovf_count);
}

/* ------------------------------------------------------------------------ */
/* Reaction                                                                 */
/* ------------------------------------------------------------------------ */

/* calculateReactionPropensity, global-R branch, byte_reaction_dev.cu:77-100. */
static float reaction_propensity(const orc_model *m, uint32_t st, const uint32_t *slots, uint32_t r)
{
    if (!m->RL[(size_t)r * m->nST + st]) return 0.0f;
    float n1 = 0.0f, n2 = 0.0f;
    for (uint32_t i = 0; i < m->P; i++) {
        if (slots[i] > 0) {
            n1 += (slots[i] == m->D1[r]) ? 1.0f : 0.0f;
            n2 += (slots[i] == m->D2[r]) ? 1.0f : 0.0f;
        }
    }
    switch (m->order[r]) {
    case ORC_ZERO_ORDER: return m->rate[r];
    case ORC_FIRST_ORDER: return m->rate[r] * n1;
    case ORC_SECOND_ORDER: return m->rate[r] * n1 * n2;
    case ORC_SECOND_ORDER_SELF: return m->rate[r] * n1 * (n1 - 1.0f);
    }
    return 0.0f;
}

/* precomp_reaction_kernel's total propensity, MpdRdmeSolver.cu:1617-1639. */
static float total_propensity(const orc_model *m, uint32_t st, const uint32_t *slots)
{
    uint32_t nS = m->nS;
    float a = m->qp0[st];
    for (uint32_t i = 0; i < m->P; i++) {
        uint32_t p1 = slots[i];
        if (p1 > 0) {
            a += m->qp1[(size_t)st * nS + (p1 - 1)];
            for (uint32_t j = i + 1; j < m->P; j++) {
                uint32_t p2 = slots[j];
                if (p2 > 0)
                    a += m->qp2[(size_t)st * nS * nS + (size_t)(p1 - 1) * nS + (p2 - 1)];
            }
        }
    }
    return a;
}

int orc_react(const orc_model *m, uint32_t *lat, const uint8_t *sites, uint64_t hash,
              uint32_t *ovf_site, uint32_t *ovf_particle, uint32_t *ovf_count, uint64_t *reactions)
{
    size_t N = (size_t)m->X * m->Y * m->Z;
    uint32_t P = m->P, nS = m->nS;
    uint32_t slots[16], nw[16];
    int8_t *Sloc = (int8_t *)malloc(nS ? nS : 1);
    if (!Sloc) return ORC_ERR_ARG;
    int err = ORC_OK;
    for (size_t i = 0; i < N; i++) {
        uint32_t st = sites[i];
        for (uint32_t j = 0; j < P; j++) slots[j] = slot_get(m, lat, N, i, j);
        float a = total_propensity(m, st, slots);
        if (a == 0.0f) continue;
        float prob = orc_reaction_probability(a);
        /* checkForReaction, byte_reaction_dev.cu:139-144. */
        if (!(prob > 0.0f && orc_uniform((uint32_t)i, 1, 0, hash) <= prob)) continue;
        /* determineReactionIndex, byte_reaction_dev.cu:146-177. */
        float rp = orc_uniform((uint32_t)i, 1, 1, hash) * a;
        uint32_t idx = 0;
        for (uint32_t r = 0; r < m->R; r++) {
            float pr = reaction_propensity(m, st, slots, r);
            if (pr > 0.0f) {
                if (rp > 0.0f) idx = r;
                rp -= pr;
            }
        }
        /* evaluateReaction, byte_reaction_dev.cu:179-257. */
        memcpy(Sloc, &m->S[(size_t)idx * nS], nS);
        uint32_t n = 0;
        for (uint32_t j = 0; j < P; j++) {
            uint32_t p = slots[j];
            if (p > 0) {
                if (Sloc[p - 1] >= 0) nw[n++] = p;
                else Sloc[p - 1]++;
            }
        }
        for (uint32_t s = 0; s < nS; s++) {
            if (Sloc[s] < 0) {
                /* The reference loops (unsigned)S times here and floods the
                 * overflow list; it ends in its too-many-overflows error. */
                err = ORC_ERR_REACTION_UNDERFLOW;
                continue;
            }
            for (int j = 0; j < Sloc[s]; j++) {
                if (n < P) nw[n++] = s + 1;
                else push_overflow(ovf_site, ovf_particle, ovf_count, i, s + 1);
            }
        }
        for (uint32_t j = 0; j < P; j++) slot_set(m, lat, N, i, j, j < n ? nw[j] : 0);
        if (reactions) (*reactions)++;
    }
    free(Sloc);
    return err;
}

/* ------------------------------------------------------------------------ */
/* Overflow resolution                                                      */
/* ------------------------------------------------------------------------ */

/* Classic handler (MpdRdmeSolver.cu:1103-1167) with the candidate order of
 * Lattice::getNearbySites (Lattice.cpp:106-128): radius 0..4, shell
 * max(R-1,0)^2 < d^2 <= R^2 (d=0 at R=0), x outer, z inner, bounds clamped
 * with the reference's unsigned arithmetic.  A candidate outside the lattice
 * makes the reference's getOccupancy throw: reported as an error here. */
int orc_resolve_overflows(const orc_model *m, uint32_t *lat, const uint8_t *sites,
                          const uint32_t *ovf_site, const uint32_t *ovf_particle, uint32_t count)
{
    if (count > ORC_MAX_OVERFLOWS) return ORC_ERR_OVERFLOW_CAP;
    size_t N = (size_t)m->X * m->Y * m->Z;
    uint32_t X = m->X, Y = m->Y, Z = m->Z;
    for (uint32_t e = 0; e < count; e++) {
        uint32_t li = ovf_site[e];
        uint32_t particle = ovf_particle[e];
        uint32_t xc = li % X, yc = (li / X) % Y, zc = li / (X * Y);
        uint8_t st0 = sites[li];
        int placed = 0;
        for (uint32_t R = 0; !placed && R <= ORC_MAX_REPLACEMENT_DIST; R++) {
            uint32_t minD = R > 0 ? R - 1 : 0, maxD = R;
            double minD2 = (double)(minD * minD), maxD2 = (double)(maxD * maxD);
            uint32_t x0 = xc >= maxD ? xc - maxD : 0, x1 = (xc <= X - 1 - maxD) ? xc + maxD : X - 1;
            uint32_t y0 = yc >= maxD ? yc - maxD : 0, y1 = (yc <= Y - 1 - maxD) ? yc + maxD : Y - 1;
            uint32_t z0 = zc >= maxD ? zc - maxD : 0, z1 = (zc <= Z - 1 - maxD) ? zc + maxD : Z - 1;
            for (uint32_t x = x0; !placed && x <= x1; x++)
                for (uint32_t y = y0; !placed && y <= y1; y++)
                    for (uint32_t z = z0; !placed && z <= z1; z++) {
                        double dx = (double)x - (double)xc, dy = (double)y - (double)yc,
                               dz = (double)z - (double)zc;
                        double r2 = dx * dx + dy * dy + dz * dz;
                        if (!(r2 <= maxD2 && (r2 > minD2 || minD == maxD))) continue;
                        if (x >= X || y >= Y || z >= Z) return ORC_ERR_OVERFLOW_OUT_OF_RANGE;
                        size_t s = (size_t)x + (size_t)X * y + (size_t)X * Y * z;
                        if (occupancy(m, lat, N, s) < m->P && sites[s] == st0) {
                            add_particle(m, lat, N, s, particle);
                            placed = 1;
                        }
                    }
        }
        if (!placed) return ORC_ERR_OVERFLOW_UNPLACEABLE;
    }
    return ORC_OK;
}

/* Relaxed handler (correct_overflows, MpdRdmeSolver.cu:1677-1761): the list
 * as the kernel sees it holds the entries left unresolved by earlier steps
 * first (it rewrites them from index 0, :1745-1749, and the next step's
 * sweeps append after them), then this step's.  Entries of one site go in
 * list order (its rounds, :1717-1729), each into the site's first empty slot
 * (:1733-1742); a full site keeps the entry in the list.  Entries of
 * different sites do not interact, so list order is the whole schedule.  The
 * kernel asserts total < TUNE_MPD_MAX_PARTICLE_OVERFLOWS (:1697); here more
 * than ORC_MAX_OVERFLOWS entries is the list-capacity error. */
int orc_resolve_relaxed(const orc_model *m, uint32_t *lat, const uint32_t *ovf_site,
                        const uint32_t *ovf_particle, uint32_t count, uint32_t *carry_site,
                        uint32_t *carry_particle, uint32_t *carry_count)
{
    uint32_t carried = *carry_count;
    if (count > ORC_MAX_OVERFLOWS || carried + count > ORC_MAX_OVERFLOWS) return ORC_ERR_OVERFLOW_CAP;
    size_t N = (size_t)m->X * m->Y * m->Z;
    uint32_t kept = 0;
    for (uint32_t e = 0; e < carried + count; e++) {
        uint32_t site = e < carried ? carry_site[e] : ovf_site[e - carried];
        uint32_t particle = e < carried ? carry_particle[e] : ovf_particle[e - carried];
        if (!add_particle(m, lat, N, site, particle)) {
            carry_site[kept] = site;   /* kept <= e: entry e has been read */
            carry_particle[kept] = particle;
            kept++;
        }
    }
    *carry_count = kept;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Timestep                                                                 */
/* ------------------------------------------------------------------------ */

/* runTimestep, MpdRdmeSolver.cu:996-1193: x, y, z sweeps (substeps 0-2),
 * the reaction phase (substep 3) when R > 0, then overflow resolution. */
int orc_run(const orc_model *m, uint32_t *lat, const uint8_t *sites, uint32_t t0,
            uint32_t nsteps, int formulation, orc_stats *stats)
{
    return orc_run_handler(m, lat, sites, t0, nsteps, formulation, 0, NULL, NULL, NULL, stats);
}

int orc_run_handler(const orc_model *m, uint32_t *lat, const uint8_t *sites, uint32_t t0,
                    uint32_t nsteps, int formulation, int relaxed, uint32_t *carry_site,
                    uint32_t *carry_particle, uint32_t *carry_count, orc_stats *stats)
{
    if (relaxed && (!carry_site || !carry_particle || !carry_count)) return ORC_ERR_ARG;
    size_t N = (size_t)m->X * m->Y * m->Z;
    size_t words = N * (word_mode(m) ? m->P : m->P / 4);
    uint32_t *tmp = (uint32_t *)malloc(words * sizeof(uint32_t));
    uint32_t *ovf_site = (uint32_t *)malloc(ORC_MAX_OVERFLOWS * sizeof(uint32_t));
    uint32_t *ovf_particle = (uint32_t *)malloc(ORC_MAX_OVERFLOWS * sizeof(uint32_t));
    if (!tmp || !ovf_site || !ovf_particle) {
        free(tmp); free(ovf_site); free(ovf_particle);
        return ORC_ERR_ARG;
    }
    int err = ORC_OK;
    uint32_t t = t0;
    for (uint32_t s = 0; s < nsteps && err == ORC_OK; s++) {
        t = t0 + s;
        uint32_t count = 0;
        err = orc_sweep(m, lat, sites, tmp, 0, orc_timestep_hash(m->seed32, t, 0), formulation,
                        ovf_site, ovf_particle, &count);
        if (err) break;
        err = orc_sweep(m, tmp, sites, lat, 1, orc_timestep_hash(m->seed32, t, 1), formulation,
                        ovf_site, ovf_particle, &count);
        if (err) break;
        err = orc_sweep(m, lat, sites, tmp, 2, orc_timestep_hash(m->seed32, t, 2), formulation,
                        ovf_site, ovf_particle, &count);
        if (err) break;
        memcpy(lat, tmp, words * sizeof(uint32_t));
        if (m->R > 0) {
            err = orc_react(m, lat, sites, orc_timestep_hash(m->seed32, t, 3), ovf_site,
                            ovf_particle, &count, stats ? &stats->reactions : NULL);
            if (err) break;
        }
        if (relaxed && (count > 0 || *carry_count > 0))
            err = orc_resolve_relaxed(m, lat, ovf_site, ovf_particle, count, carry_site, carry_particle,
                                      carry_count);
        else if (count > 0)
            err = orc_resolve_overflows(m, lat, sites, ovf_site, ovf_particle, count);
        if (count > 0) {
            if (stats) {
                stats->overflow_entries += count;
                stats->overflow_steps++;
            }
        }
        if (err) break;
    }
    if (err && stats) stats->last_error_step = t;
    free(tmp);
    free(ovf_site);
    free(ovf_particle);
    return err;
}
