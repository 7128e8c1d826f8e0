/*
 * Next-subvolume method (exact spatial SSA) on the byte lattice: this
 * project's counterpart of lm::rdme::NextSubvolumeSolver.
 *
 * TEST INFRASTRUCTURE ONLY: used as the CPU baseline timed next to the GPU
 * path and as a statistical companion (ensemble means/variances).  It is not
 * part of the product and is never called by it.
 *
 * Follows src/rdme/NextSubvolumeSolver.cpp:123-551, the binary-heap queue of
 * src/reaction/ReactionQueue.h:60-170 and the CPU hash RNG of
 * src/rng/XORShift.cpp:54-135.  Single-threaded, like one reference replicate.
 */
#ifndef LMX_NSM_H
#define LMX_NSM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_nsm orc_nsm;

/* types/k/si/s2i as produced by orc_build_reaction_tables (k in the units
 * the reference's buildModel receives); S and D species-major [nS][R];
 * DF [nST][nST][nS]; RL [R][nST]; lattice native (P/4 planes of N words);
 * sites native [N].  seed = (seedTop<<32)|seedBottom as RandomGenerator. */
orc_nsm *orc_nsm_create(uint32_t X, uint32_t Y, uint32_t Z, uint32_t P, uint32_t nS, uint32_t nST,
                        uint32_t R, const uint32_t *types, const double *k, const uint32_t *si,
                        const uint32_t *s2i, const int32_t *S, const double *DF, const uint32_t *RL,
                        double lambda, const uint32_t *lattice, const uint8_t *sites,
                        uint32_t seedTop, uint32_t seedBottom);
void orc_nsm_destroy(orc_nsm *s);

/* Runs until maxTime or maxEvents events.  When record_interval > 0, species
 * counts are recorded at record_interval, 2*record_interval, ... (at most
 * nrec records, each nS counts) with the reference's write rule
 * (NextSubvolumeSolver.cpp:216-222).  Returns 0, or an error code. */
int orc_nsm_run(orc_nsm *s, double maxTime, uint64_t maxEvents, double record_interval,
                uint32_t nrec, uint32_t *records, uint32_t *nrec_written);

double orc_nsm_time(const orc_nsm *s);
uint64_t orc_nsm_events(const orc_nsm *s);
uint64_t orc_nsm_reaction_events(const orc_nsm *s);
void orc_nsm_counts(const orc_nsm *s, uint32_t *counts);
void orc_nsm_lattice(const orc_nsm *s, uint32_t *native_out);

/* The binary heap (ReactionQueue.h:60-170) exposed for known-answer tests:
 * run a sequence of updates and report the top and, if pos_out is given,
 * every index's heap position (nupd x n) after each update. */
int orc_heap_sequence(uint32_t n, uint32_t nupd, const uint32_t *idx, const double *times,
                      uint32_t *top_out, uint32_t *pos_out);

#ifdef __cplusplus
}
#endif
#endif
