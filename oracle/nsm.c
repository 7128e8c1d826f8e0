/*
 * Next-subvolume method counterpart (TEST INFRASTRUCTURE ONLY: CPU baseline
 * and statistical companion, never part of the product path).
 * Reference paths are relative to the Lattice Microbes 2.5.0 source tree.
 */
#include "nsm.h"
#include "mpd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RNG_CACHE 2048 /* TUNE_LOCAL_RNG_CACHE_SIZE, src/cmake/options.cmake:87 */
#define NSM_EPS 1e-12  /* EPS, src/core/Math.h:52 */

/* ---------------- CPU hash RNG (src/rng/XORShift.cpp:54-135) ---------------- */
typedef struct {
    uint64_t state;
    uint64_t seed;
} xs_rng;

static uint32_t xs_next(xs_rng *r)
{
    return orc_hash(r->state++, r->seed);
}

static void xs_uniform(xs_rng *r, double *out, int n)
{
    for (int i = 0; i < n; i++) out[i] = ((double)xs_next(r)) * (2.328306436539e-10);
}

static void xs_exp(xs_rng *r, double *out, int n)
{
    for (int i = 0; i < n; i++) {
        uint64_t v;
        while ((v = (uint64_t)xs_next(r)) == 0) {}
        out[i] = -log(((double)v) * (2.328306435997e-10));
    }
}

/* ---------------- binary heap (src/reaction/ReactionQueue.h:60-170) ---------------- */
typedef struct {
    uint32_t n;
    double *time;
    double *prop;
    uint32_t *queue; /* 1-based heap of indices */
    uint32_t *pos;
} heap_t;

static int heap_init(heap_t *h, uint32_t n)
{
    h->n = n;
    h->time = (double *)malloc(sizeof(double) * n);
    h->prop = (double *)malloc(sizeof(double) * n);
    h->queue = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    h->pos = (uint32_t *)malloc(sizeof(uint32_t) * n);
    if (!h->time || !h->prop || !h->queue || !h->pos) return 0;
    h->queue[0] = 0;
    for (uint32_t i = 0; i < n; i++) {
        h->time[i] = INFINITY;
        h->prop[i] = 0.0;
        h->queue[i + 1] = i;
        h->pos[i] = i + 1;
    }
    return 1;
}

static void heap_free(heap_t *h)
{
    free(h->time); free(h->prop); free(h->queue); free(h->pos);
}

static void heap_update(heap_t *h, uint32_t ri, double newTime, double newProp)
{
    double oldTime = h->time[ri];
    h->time[ri] = newTime;
    h->prop[ri] = newProp;
    uint32_t cur = h->pos[ri];
    if (newTime <= oldTime) {
        uint32_t par = cur >> 1;
        uint32_t pri = h->queue[par];
        while (par >= 1 && newTime < h->time[pri]) {
            h->queue[par] = ri;
            h->queue[cur] = pri;
            h->pos[ri] = par;
            h->pos[pri] = cur;
            cur = par;
            par = cur >> 1;
            pri = h->queue[par];
        }
    } else {
        uint32_t c1 = cur << 1, c2 = c1 + 1;
        while ((c1 <= h->n && newTime > h->time[h->queue[c1]]) ||
               (c2 <= h->n && newTime > h->time[h->queue[c2]])) {
            uint32_t mc = (c2 > h->n) ? c1
                          : ((h->time[h->queue[c1]] <= h->time[h->queue[c2]]) ? c1 : c2);
            uint32_t mri = h->queue[mc];
            h->queue[mc] = ri;
            h->queue[cur] = mri;
            h->pos[ri] = mc;
            h->pos[mri] = cur;
            cur = mc;
            c1 = cur << 1;
            c2 = c1 + 1;
        }
    }
}

int orc_heap_sequence(uint32_t n, uint32_t nupd, const uint32_t *idx, const double *times,
                      uint32_t *top_out, uint32_t *pos_out)
{
    heap_t h;
    if (!heap_init(&h, n)) { heap_free(&h); return ORC_ERR_ARG; }
    for (uint32_t u = 0; u < nupd; u++) {
        if (idx[u] >= n) { heap_free(&h); return ORC_ERR_ARG; }
        heap_update(&h, idx[u], times[u], 1.0);
        top_out[u] = h.queue[1];
        if (pos_out)
            for (uint32_t i = 0; i < n; i++) pos_out[(size_t)u * n + i] = h.pos[i];
    }
    heap_free(&h);
    return ORC_OK;
}

/* ---------------- solver ---------------- */
struct orc_nsm {
    uint32_t X, Y, Z, P, nS, nST, R;
    size_t N;
    uint32_t *types;
    double *k;       /* second-order rates scaled by N (NextSubvolumeSolver.cpp:96-107) */
    uint32_t *si, *s2i;
    int32_t *S;      /* species-major */
    double *DF;
    uint32_t *RL;
    double lambda2;
    uint8_t *slots;  /* [N][P] compacted particle bytes */
    uint8_t *sites;
    uint32_t *cnt;   /* per-subvolume scratch counts [nS] */
    uint32_t *ndep;  /* per reaction: number of changed species */
    uint32_t *dep;   /* [R][nS] changed species */
    int32_t *dchg;   /* [R][nS] change */
    int has_zero_order; /* an empty subvolume has nonzero propensity only then */
    heap_t heap;
    xs_rng rng;
    double expv[RNG_CACHE], univ[RNG_CACHE];
    int expn, unin;
    double time;
    uint64_t events, rxevents;
};

static void neighbours(const orc_nsm *s, size_t index, size_t *nb)
{
    /* ByteLattice::getNeighboringSites, ByteLattice.cpp:81-98 (periodic). */
    size_t X = s->X, Y = s->Y, Z = s->Z, XY = X * Y;
    size_t z = index / XY, xy = index - z * XY, y = xy / X, x = xy - y * X;
    nb[0] = (x > 0) ? index - 1 : index - 1 + X;
    nb[1] = (x < X - 1) ? index + 1 : index + 1 - X;
    nb[2] = (y > 0) ? index - X : index - X + XY;
    nb[3] = (y < Y - 1) ? index + X : index + X - XY;
    nb[4] = (z > 0) ? index - XY : index - XY + s->N;
    nb[5] = (z < Z - 1) ? index + XY : index + XY - s->N;
}

static void count_subvolume(orc_nsm *s, size_t sv)
{
    /* updateSpeciesCountsForSubvolume, NextSubvolumeSolver.cpp:455-465. */
    for (uint32_t i = 0; i < s->nS; i++) s->cnt[i] = 0;
    const uint8_t *p = &s->slots[sv * s->P];
    for (uint32_t j = 0; j < s->P && p[j]; j++) s->cnt[p[j] - 1]++;
}

static double reaction_prop(const orc_nsm *s, uint32_t r)
{
    /* CMESolver.cpp:797-823 */
    switch (s->types[r]) {
    case 0: return s->k[r];
    case 1: return s->k[r] * (double)s->cnt[s->si[r]];
    case 2: return s->k[r] * ((double)s->cnt[s->si[r]]) * ((double)s->cnt[s->s2i[r]]);
    case 3: {
        uint32_t c = s->cnt[s->si[r]];
        return c >= 2 ? 0.5 * s->k[r] * ((double)c) * ((double)(c - 1)) : 0.0;
    }
    }
    return 0.0;
}

static double subvolume_propensity(orc_nsm *s, size_t sv)
{
    /* calculateSubvolumePropensity, NextSubvolumeSolver.cpp:421-452.  An
     * empty subvolume without zero-order reactions sums only zeros: 0.0. */
    if (!s->has_zero_order && s->slots[sv * s->P] == 0) return 0.0;
    count_subvolume(s, sv);
    uint32_t src = s->sites[sv];
    double a = 0.0;
    for (uint32_t r = 0; r < s->R; r++)
        if (s->RL[r * s->nST + src]) a += reaction_prop(s, r);
    size_t nb[6];
    neighbours(s, sv, nb);
    for (uint32_t i = 0; i < s->nS; i++) {
        if (s->cnt[i] > 0) {
            for (int j = 0; j < 6; j++)
                a += s->cnt[i] * (s->DF[(size_t)src * s->nST * s->nS + (size_t)s->sites[nb[j]] * s->nS + i] / s->lambda2);
        }
    }
    return a;
}

static void update_subvolume(orc_nsm *s, size_t sv)
{
    double a = subvolume_propensity(s, sv);
    double t = INFINITY;
    if (a > 0.0) {
        if (s->expn >= RNG_CACHE) { xs_exp(&s->rng, s->expv, RNG_CACHE); s->expn = 0; }
        t = s->time + s->expv[s->expn++] / a;
    }
    heap_update(&s->heap, (uint32_t)sv, t, a);
}

static int write_counts(orc_nsm *s, size_t sv)
{
    /* updateSubvolumeWithSpeciesCounts: rebuild the site from counts in
     * species order (addParticles). */
    uint8_t *p = &s->slots[sv * s->P];
    uint32_t n = 0;
    for (uint32_t i = 0; i < s->nS; i++)
        for (uint32_t c = 0; c < s->cnt[i]; c++) {
            if (n >= s->P) return ORC_ERR_OVERFLOW_CAP;
            p[n++] = (uint8_t)(i + 1);
        }
    while (n < s->P) p[n++] = 0;
    return ORC_OK;
}

static int add_one(orc_nsm *s, size_t sv, uint8_t particle)
{
    uint8_t *p = &s->slots[sv * s->P];
    for (uint32_t j = 0; j < s->P; j++)
        if (p[j] == 0) { p[j] = particle; return ORC_OK; }
    return ORC_ERR_OVERFLOW_CAP; /* the reference's addParticle throws */
}

orc_nsm *orc_nsm_create(uint32_t X, uint32_t Y, uint32_t Z, uint32_t P, uint32_t nS, uint32_t nST,
                        uint32_t R, const uint32_t *types, const double *k, const uint32_t *si,
                        const uint32_t *s2i, const int32_t *S, const double *DF, const uint32_t *RL,
                        double lambda, const uint32_t *lattice, const uint8_t *sites,
                        uint32_t seedTop, uint32_t seedBottom)
{
    orc_nsm *s = (orc_nsm *)calloc(1, sizeof(orc_nsm));
    if (!s) return NULL;
    s->X = X; s->Y = Y; s->Z = Z; s->P = P; s->nS = nS; s->nST = nST; s->R = R;
    s->N = (size_t)X * Y * Z;
    size_t Rn = R ? R : 1, nSn = nS ? nS : 1;
    s->types = (uint32_t *)malloc(sizeof(uint32_t) * Rn);
    s->k = (double *)malloc(sizeof(double) * Rn);
    s->si = (uint32_t *)malloc(sizeof(uint32_t) * Rn);
    s->s2i = (uint32_t *)malloc(sizeof(uint32_t) * Rn);
    s->S = (int32_t *)malloc(sizeof(int32_t) * nSn * Rn);
    s->DF = (double *)malloc(sizeof(double) * (size_t)nSn * nST * nST);
    s->RL = (uint32_t *)malloc(sizeof(uint32_t) * Rn * nST);
    s->slots = (uint8_t *)malloc(s->N * P);
    s->sites = (uint8_t *)malloc(s->N);
    s->cnt = (uint32_t *)malloc(sizeof(uint32_t) * nSn);
    s->ndep = (uint32_t *)malloc(sizeof(uint32_t) * Rn);
    s->dep = (uint32_t *)malloc(sizeof(uint32_t) * Rn * nSn);
    s->dchg = (int32_t *)malloc(sizeof(int32_t) * Rn * nSn);
    if (!s->types || !s->k || !s->si || !s->s2i || !s->S || !s->DF || !s->RL || !s->slots ||
        !s->sites || !s->cnt || !s->ndep || !s->dep || !s->dchg || !heap_init(&s->heap, (uint32_t)s->N)) {
        orc_nsm_destroy(s);
        return NULL;
    }
    memcpy(s->types, types, sizeof(uint32_t) * R);
    memcpy(s->si, si, sizeof(uint32_t) * R);
    memcpy(s->s2i, s2i, sizeof(uint32_t) * R);
    memcpy(s->S, S, sizeof(int32_t) * nS * R);
    memcpy(s->DF, DF, sizeof(double) * (size_t)nS * nST * nST);
    memcpy(s->RL, RL, sizeof(uint32_t) * R * nST);
    for (uint32_t r = 0; r < R; r++) {
        s->k[r] = k[r];
        if (types[r] == 0) s->has_zero_order = 1;
        if (types[r] == 2 || types[r] == 3) s->k[r] *= (double)s->N;
        uint32_t n = 0;
        for (uint32_t j = 0; j < nS; j++) {
            int32_t v = S[(size_t)j * R + r];
            if (v != 0) { s->dep[r * nSn + n] = j; s->dchg[r * nSn + n] = v; n++; }
        }
        s->ndep[r] = n;
    }
    for (size_t i = 0; i < s->N; i++)
        for (uint32_t j = 0; j < P; j++)
            s->slots[i * P + j] = (uint8_t)(lattice[(size_t)(j >> 2) * s->N + i] >> (8 * (j & 3)));
    memcpy(s->sites, sites, s->N);
    s->lambda2 = lambda * lambda;
    s->rng.state = 1;
    s->rng.seed = (((uint64_t)seedTop) << 32) | (uint64_t)seedBottom;
    xs_exp(&s->rng, s->expv, RNG_CACHE);
    xs_uniform(&s->rng, s->univ, RNG_CACHE);
    s->expn = 0; s->unin = 0;
    s->time = 0.0;
    for (size_t sv = 0; sv < s->N; sv++) update_subvolume(s, sv);
    return s;
}

void orc_nsm_destroy(orc_nsm *s)
{
    if (!s) return;
    free(s->types); free(s->k); free(s->si); free(s->s2i); free(s->S); free(s->DF); free(s->RL);
    free(s->slots); free(s->sites); free(s->cnt); free(s->ndep); free(s->dep); free(s->dchg);
    heap_free(&s->heap);
    free(s);
}

static void total_counts(const orc_nsm *s, uint32_t *counts)
{
    memset(counts, 0, sizeof(uint32_t) * s->nS);
    size_t n = s->N * s->P;
    for (size_t i = 0; i < n; i++)
        if (s->slots[i] && s->slots[i] <= s->nS) counts[s->slots[i] - 1]++;
}

/* performSubvolumeReaction, NextSubvolumeSolver.cpp:478-551. */
static int perform_event(orc_nsm *s, size_t sv, int *neighbour_hit, size_t *neighbour)
{
    *neighbour_hit = 0;
    count_subvolume(s, sv);
    if (s->unin >= RNG_CACHE) { xs_uniform(&s->rng, s->univ, RNG_CACHE); s->unin = 0; }
    double rv = s->univ[s->unin++] * s->heap.prop[sv];
    uint32_t src = s->sites[sv];
    for (uint32_t r = 0; r < s->R; r++) {
        if (!s->RL[r * s->nST + src]) continue;
        double pr = reaction_prop(s, r);
        if (pr > 0.0) {
            if (rv <= pr) {
                uint32_t nS = s->nS ? s->nS : 1;
                for (uint32_t d = 0; d < s->ndep[r]; d++)
                    s->cnt[s->dep[r * nS + d]] += s->dchg[r * nS + d];
                return write_counts(s, sv);
            }
            rv -= pr;
        }
    }
    size_t nb[6];
    neighbours(s, sv, nb);
    for (uint32_t i = 0; i < s->nS; i++) {
        if (s->cnt[i] == 0) continue;
        for (int j = 0; j < 6; j++) {
            double dp = ((double)s->cnt[i]) *
                        (s->DF[(size_t)src * s->nST * s->nS + (size_t)s->sites[nb[j]] * s->nS + i] / s->lambda2);
            if (rv <= dp) {
                s->cnt[i]--;
                int e = write_counts(s, sv);
                if (e) return e;
                e = add_one(s, nb[j], (uint8_t)(i + 1));
                *neighbour_hit = 1;
                *neighbour = nb[j];
                return e;
            }
            rv -= dp;
        }
    }
    return ORC_ERR_ARG; /* "Unable to determine correct reaction or diffusion event" */
}

int orc_nsm_run(orc_nsm *s, double maxTime, uint64_t maxEvents, double record_interval,
                uint32_t nrec, uint32_t *records, uint32_t *nrec_written)
{
    double nextRec = record_interval;
    uint32_t written = 0;
    uint64_t ev = 0;
    int err = ORC_OK;
    while (ev < maxEvents) {
        size_t sv = s->heap.queue[1];
        double t = s->heap.time[sv];
        if (t >= maxTime) { s->time = maxTime; break; }
        while (record_interval > 0.0 && written < nrec && nextRec <= t + NSM_EPS) {
            total_counts(s, &records[(size_t)written * s->nS]);
            written++;
            nextRec += record_interval;
        }
        s->time = t;
        int hit = 0;
        size_t nbr = 0;
        err = perform_event(s, sv, &hit, &nbr);
        if (err) break;
        ev++;
        s->events++;
        if (!hit) s->rxevents++;
        update_subvolume(s, sv);
        if (hit) update_subvolume(s, nbr);
    }
    if (!err && s->time >= maxTime) {
        while (record_interval > 0.0 && written < nrec && nextRec <= maxTime + NSM_EPS) {
            total_counts(s, &records[(size_t)written * s->nS]);
            written++;
            nextRec += record_interval;
        }
    }
    if (nrec_written) *nrec_written = written;
    return err;
}

double orc_nsm_time(const orc_nsm *s) { return s->time; }
uint64_t orc_nsm_events(const orc_nsm *s) { return s->events; }
uint64_t orc_nsm_reaction_events(const orc_nsm *s) { return s->rxevents; }
void orc_nsm_counts(const orc_nsm *s, uint32_t *counts) { total_counts(s, counts); }

void orc_nsm_lattice(const orc_nsm *s, uint32_t *native)
{
    memset(native, 0, sizeof(uint32_t) * s->N * (s->P / 4));
    for (size_t i = 0; i < s->N; i++)
        for (uint32_t j = 0; j < s->P; j++)
            native[(size_t)(j >> 2) * s->N + i] |= ((uint32_t)s->slots[i * s->P + j]) << (8 * (j & 3));
}
