/*
 * lmx_mpd: MI355X-native MPD-RDME timestep behind a C ABI.
 *
 * This is the drop-in boundary for lm::rdme::MpdRdmeSolver.  The reference
 * exposes a C++ virtual-class protocol (src/me/MESolver.h:70-77,
 * src/cme/CMESolver.h:200-201, src/rdme/RDMESolver.h:59-60,
 * src/rdme/MpdRdmeSolver.h:66-156); every entry point below replaces one step
 * of that protocol and cites it.  Plain pointers and sizes only; no C++ types,
 * no exceptions across the ABI.  All per-replicate state lives in the opaque
 * handle (the reference keeps the model in __constant__ symbols,
 * src/cuda/constant.cu:45-89, so two solvers on one device collide; these do
 * not).
 *
 * Status codes: every function returns LMX_OK (0) or an LMX_ERR_* code;
 * lmx_last_error() gives the calling thread's last message.
 *
 * Layouts (the reference's, src/rdme/ByteLattice.cpp:371-435):
 *   row-major particles : uint8 [X][Y][Z][P]      (file/IO layout)
 *   row-major sites     : uint8 [X][Y][Z]
 *   native particles    : uint32 [P/4][Z][Y][X]    (word w holds slots 4w..4w+3,
 *                                                   little-endian bytes)
 *   native sites        : uint8 [Z][Y][X]
 */
#ifndef LMX_MPD_H
#define LMX_MPD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LMX_ABI_VERSION 1

enum lmx_status {
    LMX_OK = 0,
    LMX_ERR_ARG = 1,                    /* InvalidArgException / Exception on bad input */
    LMX_ERR_HIP = 2,                    /* HIP runtime failure (CUDAException analogue) */
    LMX_ERR_OVERFLOW_CAP = 3,           /* "Too many particle overflows" (MpdRdmeSolver.cu:1123) */
    LMX_ERR_OVERFLOW_UNPLACEABLE = 4,   /* "Unable to find an available site" (:1166) */
    LMX_ERR_OVERFLOW_OUT_OF_RANGE = 5,  /* getNearbySites reached outside the lattice (Lattice.cpp:106-128) */
    LMX_ERR_REACTION_UNDERFLOW = 6,     /* reaction applied without its reactants (byte_reaction_dev.cu:226) */
    LMX_ERR_STATE = 7,                  /* call out of protocol order */
    LMX_ERR_NOMEM = 8,                  /* device or host allocation failed */
    LMX_ERR_INVALID_LATTICE = 9         /* a lattice changed on the device holds an invalid species or site type */
};

/* Boundary conditions.  Every reference build is periodic
 * (src/config.h.in:61 always defines MPD_BOUNDARY_PERIODIC); reflecting is the
 * reference's dormant branch (lattice_sim_1d_dev.cu:84-90). */
enum lmx_boundary { LMX_BOUNDARY_PERIODIC = 0, LMX_BOUNDARY_REFLECTING = 1 };

/* Step implementations.  Every one gives bit-identical lattices. */
enum lmx_path {
    LMX_PATH_SWEEPS = 0,   /* x, y, z sweep kernels + reaction kernel + resolver */
    LMX_PATH_FUSED = 1     /* one fused diffuse+react kernel per step + resolver */
};

typedef struct lmx_solver lmx_solver;

/* Message for the last failing call on this thread ("" if none). */
const char *lmx_last_error(void);
int lmx_abi_version(void);

/* Number of HIP devices visible (0 on a host without a GPU). */
int lmx_device_count(int *count);

/* MpdRdmeSolver::MpdRdmeSolver + the stream/overflow-list part of
 * MpdRdmeSolver::initialize (MpdRdmeSolver.cu:95-162).  device selects the
 * GPU (ReplicateRunner.cpp:92 setCurrentDevice). */
int lmx_create(int device, lmx_solver **out);
/* MpdRdmeSolver::~MpdRdmeSolver (MpdRdmeSolver.cu:164-190). */
int lmx_destroy(lmx_solver *s);

/* "rdme.mpd.overflowhandler" (MpdRdmeSolver.cu:126-146). */
enum lmx_overflow_mode {
    /* "classic" or "" (the default): each entry into the first site with room
     * among the nearby sites of the same type (:1103-1189, Lattice.cpp:106-128) */
    LMX_OVERFLOW_CLASSIC = 0,
    /* "relaxed": each entry into its own site's first empty slot after the
     * step; an entry whose site is full stays listed and is retried after the
     * next step (correct_overflows, :1677-1761, launched at :1086-1101) */
    LMX_OVERFLOW_RELAXED = 1
};

/* MpdRdmeSolver::initialize (MpdRdmeSolver.cu:107-162): parameters "seed",
 * "timestep" (tau, seconds) and "rdme.mpd.overflowhandler" (overflow_mode,
 * enum lmx_overflow_mode).  seedTop == 0 draws the seed from the clock as the
 * reference does (:113-122).  seed32 = (seedTop<<16) | (replicate & 0xFFFF)
 * (:123). */
int lmx_initialize(lmx_solver *s, uint32_t replicate, uint32_t seedTop, double timestep,
                   int overflow_mode, int boundary);
/* The seed actually used (after a clock seed). */
int lmx_get_seed(lmx_solver *s, uint32_t *seed32);

/* CMESolver::buildModel + MpdRdmeSolver::buildModel (CMESolver.cpp:307-460,
 * MpdRdmeSolver.cu:232-364).  types[R] in {0,1,2,3}; K[R*kCols]; S and D
 * species-major [nS][R] as in the ReactionModel message. */
int lmx_build_model(lmx_solver *s, uint32_t numberSpecies, uint32_t numberReactions,
                    const uint32_t *initialSpeciesCounts, const uint32_t *reactionTypes,
                    const double *K, const int32_t *S, const uint32_t *D, uint32_t kCols);

/* RDMESolver::buildDiffusionModel + MpdRdmeSolver::buildDiffusionModel
 * (RDMESolver.cpp:110-141, MpdRdmeSolver.cu:366-483): DF [nST][nST][nS]
 * (m^2/s), RL [R][nST], lattice and sites in row-major layout (rowMajor must
 * be 1, as the reference requires).  Any X, Y, Z >= 1 is accepted (the
 * reference's X%32, Y%8, Z%8 limits of :201-217 are launch-shape limits, not
 * semantics); particlesPerSite in {4, 8, 16}.
 * bytesPerParticle 1: the byte lattice (MpdRdmeSolver), at most 255 species;
 * latticeData holds X*Y*Z*P bytes.  bytesPerParticle 4: the uint32 lattice
 * of IntMpdRdmeSolver (IntMpdRdmeSolver.cu:155-170, word_diffusion_1d_dev.cu,
 * word_reaction_dev.cu), one particle per 32-bit word, the same update rules
 * and RNG keys, species beyond 255; latticeData holds X*Y*Z*P uint32 values
 * and the native layout is P planes of N words.  numberSiteTypes *
 * numberSpecies^2 must stay below 2^32. */
int lmx_build_diffusion_model(lmx_solver *s, uint32_t numberSiteTypes, const double *DF,
                              const uint32_t *RL, uint32_t X, uint32_t Y, uint32_t Z,
                              uint32_t particlesPerSite, uint32_t bytesPerParticle,
                              double latticeSpacing, const uint8_t *latticeData,
                              const uint8_t *latticeSitesData, int rowMajorData);

/* MpdRdmeSolver::setReactionRate (MpdRdmeSolver.cu:670-695): tables are
 * recomputed (computePropensities, :697-774) and uploaded before the next step. */
int lmx_set_reaction_rate(lmx_solver *s, uint32_t rxid, float rate);

/* Select the step implementation (enum lmx_path); results do not depend on it.
 * LMX_PATH_FUSED needs X % 4 == 0 and X <= 256 (a lattice row fits one
 * wavefront); other shapes, and LMX_OPT_FORCE_GENERIC, run the sweeps path. */
int lmx_set_path(lmx_solver *s, int path);

/* Launch-shape options (no reference counterpart: the reference fixes its
 * block shapes at compile time, src/cmake/options.cmake:87-95).  Results do
 * not depend on them.
 *   LMX_OPT_MARCH_SEGMENT  rows per thread of the marching y/z sweep: 4, or 8
 *                          (default; any other value selects 8)
 *   LMX_OPT_FORCE_GENERIC  1 = one-site-per-thread kernels even when X%4==0
 *   LMX_OPT_YZ_KERNEL      y/z sweep kernel: 0 = LDS row tiles, one group
 *                          per thread (default), 1 = register march, 2 = row
 *                          tiles with compacted work lists, 3 = register
 *                          columns of LMX_OPT_MARCH_SEGMENT rows (axes whose
 *                          length the segment does not divide use 0)
 *   LMX_OPT_FUSED_ZSEG     z planes per block of the fused step; 0 (default)
 *                          picks one block per compute unit
 *   LMX_OPT_FUSE_REACTION  1 (default) = the tiled z sweep lists its
 *                          output sites that hold particles, with their
 *                          words, and the reaction phase evaluates only the
 *                          listed sites (models without zero-order
 *                          reactions; a full list falls back to the full
 *                          pass); 0 = full reaction kernel */
enum lmx_option {
    LMX_OPT_MARCH_SEGMENT = 1, LMX_OPT_FORCE_GENERIC = 2, LMX_OPT_YZ_KERNEL = 3, LMX_OPT_FUSED_ZSEG = 4,
    LMX_OPT_FUSE_REACTION = 5
};
int lmx_set_launch_option(lmx_solver *s, int option, int value);

/* Launch-shape autotuner (no reference counterpart: the reference fixes its
 * block shapes at build time, options.cmake TUNE_MPD_*).  From the current
 * state it times nsteps steps (after one untimed step) of each y/z sweep
 * kernel and window shape: LDS row tiles with and without the reaction list,
 * register march and register columns of 4 and 8 rows, compacted tiles, and
 * the fused step.  It restores the state after each, times two interleaved
 * passes and keeps the fastest by the minimum of each, but keeps the current
 * configuration unless another is at least 2% faster.  With
 * LMX_OPT_FORCE_GENERIC set only the current configuration is timed and every
 * option stays as it is.  Every option gives bit-identical
 * results, so afterwards the lattice, step counter, overflow statistics and
 * carried entries are as before the call.  best_ms (optional) receives the
 * best time per step; chosen (optional, 4 ints) the path, LMX_OPT_YZ_KERNEL,
 * LMX_OPT_MARCH_SEGMENT and LMX_OPT_FUSE_REACTION kept.  Allocates a copy of
 * both lattice buffers for the duration of the call; not for ensemble
 * members (LMX_ERR_STATE). */
int lmx_autotune(lmx_solver *s, uint32_t nsteps, double *best_ms, int *chosen);

/* runTimestep (MpdRdmeSolver.cu:996-1193) nsteps times, enqueued on the
 * solver's stream without a host synchronization per step (the reference
 * synchronizes every step, :1091).  Device-side errors are sticky and are
 * reported by the next lmx_synchronize. */
int lmx_run_steps(lmx_solver *s, uint32_t nsteps);
/* No counterpart in the reference: lmx_run_steps for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * The same kernels and results, with the z sweep's reaction list in a static
 * device array of the library (1024 shards x 256 entries, 8 MiB per GPU,
 * held until the process ends) instead of the solver's own list; the call
 * allocates nothing.  Lattices of at most 2,097,152 sites (128^3); larger
 * ones are rejected (LMX_ERR_ARG), as is a call after a reaction-rate change
 * whose tables lmx_run_steps has not uploaded yet (LMX_ERR_STATE).  Every
 * call on one GPU shares the array: calls that overlap on one GPU (solvers
 * stepping concurrently on different streams) race. */
int lmx_run_steps_static(lmx_solver *s, uint32_t nsteps);
int lmx_synchronize(lmx_solver *s);
/* Current timestep counter (number of steps taken since the lattice was set). */
int lmx_get_timestep(lmx_solver *s, uint32_t *timestep);
int lmx_set_timestep(lmx_solver *s, uint32_t timestep);

/* CudaByteLattice::copyFromGPU + ByteLattice::copyNativeToRowMajorByte
 * (CudaByteLattice.cu:128-141, ByteLattice.cpp:371-403; CudaIntLattice and
 * IntLattice for bytesPerParticle 4).  size is the buffer's size in bytes,
 * X*Y*Z*P*bytesPerParticle; native buffers hold N*W words (W = P/4, or P for
 * the uint32 lattice). */
int lmx_get_lattice(lmx_solver *s, uint8_t *rowMajor, size_t size);
int lmx_get_lattice_native(lmx_solver *s, uint32_t *native, size_t words);
int lmx_get_sites(lmx_solver *s, uint8_t *rowMajor, size_t size);
/* Hook write-back (hookCheckSimulation return 1/2, MpdRdmeSolver.cu:958-985). */
int lmx_set_lattice(lmx_solver *s, const uint8_t *rowMajor, size_t size);
int lmx_set_lattice_native(lmx_solver *s, const uint32_t *native, size_t words);
int lmx_set_sites(lmx_solver *s, const uint8_t *rowMajor, size_t size);
/* Device-side hooks (the host copies of copyFromGPU/copyToGPU around
 * hookSimulation, MpdRdmeSolver.cu:958-985, without the copies): the current
 * device particle planes (W planes of N uint32, the native layout; the
 * library's own bookkeeping words follow them and are not to be written) and site
 * types (N bytes, index x + X*y + X*Y*z), after the solver's stream has
 * drained.  They stay valid until the next lmx_run_steps or lmx_set_*; the
 * caller finishes its own device work on them before the next step.  The
 * caller may write any slot, so the next step reads every word plane of the
 * lattice (the kernels otherwise skip a plane after a word whose last slot is
 * empty, which the packed sites they write allow).
 * lmx_validate_lattice then checks them on the device (particle types at most
 * numberSpecies, site types below numberSiteTypes; LMX_ERR_INVALID_LATTICE). */
int lmx_device_lattice(lmx_solver *s, void **particles, void **sites);
int lmx_validate_lattice(lmx_solver *s);

/* ByteLattice::getParticleCounts (ByteLattice.cpp:495-520;
 * IntLattice::getParticleCounts, IntLattice.cpp:419-435), computed on the
 * device; counts[nS] for species 1..nS. */
int lmx_get_species_counts(lmx_solver *s, uint32_t *counts, uint32_t n);

/* Overflow statistics (overflowTimesteps/overflowListUses, :1104-1187):
 * total entries resolved and number of steps that had any. */
int lmx_get_overflow_stats(lmx_solver *s, uint64_t *entries, uint64_t *steps);
/* Relaxed handler: entries left unresolved after the last step, whose own
 * site was full (the reference's "%d unresolved overflows" warning,
 * MpdRdmeSolver.cu:1094-1101).  They stay listed and count against the list's
 * 2048 entries: more than that at the end of a step is LMX_ERR_OVERFLOW_CAP.
 * Always 0 with the classic handler. */
int lmx_get_unresolved_overflows(lmx_solver *s, uint32_t *count);

/* Measurement (no reference counterpart; replaces the lptf PROF_CUDA_* events,
 * MpdRdmeSolver.cu:1014-1083).  Runs nsteps steps bracketed by HIP events on
 * the solver's stream.  If kernel_ms != NULL, every launch is also bracketed
 * and the summed milliseconds per kernel class (enum lmx_kernel) are written
 * to kernel_ms[0..nkernels-1] and the launch counts to kernel_launches. */
enum lmx_kernel {
    LMX_K_SWEEP_X = 0, LMX_K_SWEEP_Y = 1, LMX_K_SWEEP_Z = 2, LMX_K_REACTION = 3,
    LMX_K_RESOLVE = 4, LMX_K_FUSED = 5, LMX_K_COUNT = 6
};
int lmx_time_steps(lmx_solver *s, uint32_t nsteps, double *total_ms, double *kernel_ms,
                   uint64_t *kernel_launches, uint32_t nkernels);

/* Debug tables (for parity tests of the host-side table construction):
 * which: 0 T, 1 qp0, 2 qp1, 3 qp2, 4 reaction rates.  n is the capacity. */
int lmx_get_table_f32(lmx_solver *s, int which, float *out, size_t n, size_t *written);

/* Evaluate the device's reaction-probability function (float)(1-exp(-(double)a))
 * and the device exp itself on n inputs (numerics tests). */
/* Measured copy peak: reps copies of a bytes-long device buffer by a 16-byte
 * grid-stride copy kernel, as (read + write) GB/s (the STREAM-copy figure the
 * HBM roofline is also quoted against).  Allocates its own buffers. */
int lmx_copy_bandwidth(int device, size_t bytes, int reps, double *gbps);
int lmx_device_exp(int device, const double *x, double *y, size_t n);
int lmx_device_reaction_probability(int device, const float *a, float *prob, size_t n);

/* ---- Ensembles -----------------------------------------------------------
 * No reference counterpart: the reference runs every replicate in its own
 * thread or process (src/core/ReplicateRunner.cpp, src/lm/main/Main.cpp).
 * A batch steps many independently configured solvers on one device with one
 * launch per kernel class per step (replayed from a captured HIP graph by
 * default); members may differ in lattice size, model, seed and step count.
 * Every member's lattice is bit-identical to running that solver alone.
 * Members need X % 4 == 0 and the same particles per site.  A solver belongs
 * to at most one batch and leaves it when destroyed; its own calls stay
 * valid between batch runs (they are ordered after the batch's work). */
typedef struct lmx_batch lmx_batch;
int lmx_batch_create(int device, lmx_batch **out);
int lmx_batch_destroy(lmx_batch *b);
int lmx_batch_add(lmx_batch *b, lmx_solver *s);
int lmx_batch_remove(lmx_batch *b, lmx_solver *s);
int lmx_batch_size(const lmx_batch *b, uint32_t *n);
/* LMX_BATCH_GRAPH: 1 (default) = replay a captured graph of one step.
 * LMX_BATCH_REACTION_LIST: 1 (default) = a member without zero-order
 * reactions runs its reaction phase over the occupied sites its z sweep lists
 * (as a single solver does); 0 = the full in-place pass for every member.
 * LMX_BATCH_TILE_THREADS: threads per block of the tiled y/z sweeps, which
 * sets every member's tile shape: 0 (default) = 256 (tiles 32 sites wide and
 * 32 rows tall), else a multiple of 64 from 256 up to 1024 (512 with 16
 * slots per site); above 256, tiles 256 sites wide.  Results are identical
 * whatever the options. */
enum lmx_batch_option { LMX_BATCH_GRAPH = 1, LMX_BATCH_REACTION_LIST = 2, LMX_BATCH_TILE_THREADS = 3 };
int lmx_batch_set_option(lmx_batch *b, int option, int value);
/* No reference counterpart (the reference's launch shapes are build-time
 * TUNE_MPD_* constants): times nsteps batched steps (after one untimed step)
 * for each launch configuration -- reaction list on and off, tile blocks of
 * 256, half the maximum and the maximum threads -- in two interleaved passes,
 * restoring every member's lattice buffers, control block and step counter
 * before each and at the end, and keeps the fastest (the current one unless
 * another is at least 2% faster).  chosen[0] = reaction list (0/1),
 * chosen[1] = tile threads.  Allocates a copy of every member's buffers for
 * the duration of the call. */
int lmx_batch_autotune(lmx_batch *b, uint32_t nsteps, double *best_ms, int *chosen);
/* Advances every member by nsteps (asynchronous; errors are reported by
 * lmx_batch_synchronize, or by a member's own calls). */
int lmx_batch_run_steps(lmx_batch *b, uint32_t nsteps);
int lmx_batch_synchronize(lmx_batch *b);
/* lmx_batch_run_steps + synchronize, timed with HIP events (milliseconds). */
int lmx_batch_time_steps(lmx_batch *b, uint32_t nsteps, double *total_ms);


#ifdef __cplusplus
}
#endif
#endif
