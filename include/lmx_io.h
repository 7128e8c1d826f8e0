/*
 * lmx_io: Lattice Microbes SimulationFile (HDF5) input and output.
 *
 * C-ABI replacement for the parts of lm::io::hdf5::SimulationFile that the
 * MPD-RDME path reads and writes (src/io/SimulationFile.cpp,
 * src/io/SimulationFile_create.cpp; layout in docs/HDF5FileFormat.text):
 * parameters, the reaction and diffusion models with the initial lattice, and
 * the per-replicate species-count and lattice trajectories.  Files written
 * here open in the reference and vice versa (format version 4, "LMH5" user
 * block, the same group/dataset names, types, chunking and gzip level).
 *
 * Status codes: 0 on success, otherwise LMXIO_ERR_*; lmxio_last_error() gives
 * the message (thread-local).  A handle is not safe for concurrent use.
 */
#ifndef LMX_IO_H
#define LMX_IO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum lmxio_status { LMXIO_OK = 0, LMXIO_ERR_ARG = 1, LMXIO_ERR_IO = 2, LMXIO_ERR_FORMAT = 3 };

/* Columns of /Model/Reaction/ReactionRateConstants
 * (SimulationFile::MAX_REACTION_RATE_CONSTANTS) and of ReactionRateNoise. */
#define LMXIO_MAX_RATE_CONSTANTS 10
#define LMXIO_NOISE_COLS 2

typedef struct lmxio_file lmxio_file;

const char *lmxio_last_error(void);

/* SimulationFile::create (SimulationFile_create.cpp:56-122): a new file with
 * the 512-byte "LMH5" user block, formatVersion 4 and the /Parameters,
 * /Model and /Simulations groups; with initializeModel also /Model/Reaction
 * holding only numberSpecies.  Fails if the file exists. */
int lmxio_create(const char *path, int initializeModel, uint32_t numberSpecies);

/* SimulationFile::SimulationFile + open + isValidFile (SimulationFile.cpp:82-134,
 * :225-252): rejects paths that are not regular files, lack the "LMH5" magic
 * or are not HDF5, and format versions outside 2..4. */
int lmxio_open(const char *path, int writable, lmxio_file **out);
int lmxio_close(lmxio_file *f);
int lmxio_format_version(const lmxio_file *f, uint32_t *version);
/* SimulationFile::flush (SimulationFile.cpp:148-156). */
int lmxio_flush(lmxio_file *f);

/* getParameters + parseParameter (SimulationFile.cpp:260-310): every attribute of /Parameters
 * as "key=value\n" lines (version 2: doubles, formatted with %g; version 3+:
 * strings; other attribute types are skipped).  *needed receives the size
 * including the terminating NUL; buf may be NULL to query it. */
int lmxio_get_parameters(lmxio_file *f, char *buf, size_t cap, size_t *needed);
/* setParameter (SimulationFile.cpp:322-336): a string attribute (version 3+; a
 * double parsed from the value in version 2 files). */
int lmxio_set_parameter(lmxio_file *f, const char *key, const char *value);

typedef struct lmxio_reaction_info {
    uint32_t present;          /* /Model/Reaction exists */
    uint32_t numberSpecies;
    uint32_t hasTables;        /* the numberReactions attribute exists: initial counts and tables follow */
    uint32_t numberReactions;
    uint32_t hasRateNoise;     /* /Model/Reaction/ReactionRateNoise exists */
} lmxio_reaction_info;
int lmxio_get_reaction_info(lmxio_file *f, lmxio_reaction_info *info);

/* getReactionModel (SimulationFile.cpp:363-470).  Arrays as stored:
 * initialCounts[nS], types[R], K[R][10] (unused columns NaN), noise[R][2]
 * (variance, tau; zeros when the file has no noise table), S and D [nS][R].
 * Any output may be NULL. */
int lmxio_get_reaction_model(lmxio_file *f, uint32_t *initialCounts, uint32_t *types, double *K, double *noise,
                             int32_t *S, uint32_t *D);
/* setReactionModel (SimulationFile.cpp:472-555); noise may be NULL (no
 * noise table is written). */
int lmxio_set_reaction_model(lmxio_file *f, uint32_t nS, uint32_t R, const uint32_t *initialCounts,
                             const uint32_t *types, const double *K, const double *noise, const int32_t *S,
                             const uint32_t *D);

typedef struct lmxio_diffusion_info {
    uint32_t present;          /* /Model/Diffusion exists */
    uint32_t numberSpecies, numberReactions, numberSiteTypes;
    uint32_t X, Y, Z, P;
    uint32_t bytesPerParticle; /* 1 (uint8 lattice) or 4 (uint32 lattice); 1 when the attribute is absent */
    double latticeSpacing;
} lmxio_diffusion_info;
int lmxio_get_diffusion_info(lmxio_file *f, lmxio_diffusion_info *info);

/* getDiffusionModel + getDiffusionModelLattice (SimulationFile.cpp:557-659):
 * DF[nST][nST][nS], RL[R][nST], lattice row-major (x,y,z,p) of
 * bytesPerParticle-byte entries (latticeBytes = X*Y*Z*P*bytesPerParticle) and
 * sites row-major (x,y,z) bytes.  Any output may be NULL. */
int lmxio_get_diffusion_model(lmxio_file *f, double *DF, uint32_t *RL, void *lattice, size_t latticeBytes,
                              uint8_t *sites, size_t sitesBytes);
/* setDiffusionModel (SimulationFile.cpp:696-761), and with lattice and sites
 * (either may be NULL: not written) setDiffusionModelLattice. */
int lmxio_set_diffusion_model(lmxio_file *f, uint32_t nS, uint32_t R, uint32_t nST, const double *DF,
                              const uint32_t *RL, double latticeSpacing, uint32_t X, uint32_t Y, uint32_t Z,
                              uint32_t P, uint32_t bytesPerParticle, const void *lattice, const uint8_t *sites);
/* setDiffusionModelLattice (SimulationFile.cpp:763-870) into an existing
 * /Model/Diffusion, replacing Lattice and LatticeSites. */
int lmxio_set_diffusion_model_lattice(lmxio_file *f, uint32_t X, uint32_t Y, uint32_t Z, uint32_t P,
                                      uint32_t bytesPerParticle, const void *lattice, const uint8_t *sites);
/* setSpatialModel (SimulationFile.cpp:970-1060) for a model without regions
 * or obstacles: /Model/Spatial/Regions (numberRegions 0) and
 * /Model/Spatial/Obstacles (numberObstacles 0). */
int lmxio_set_empty_spatial_model(lmxio_file *f);

/* Trajectories in /Simulations/%07d: appendSpeciesCounts
 * (SimulationFile.cpp:1090-1164, SpeciesCounts I32 [n][nS] and
 * SpeciesCountTimes F64, chunks of 100 rows), appendLattice (:1166-1265,
 * Lattice/%010d gzip frames and LatticeTimes), appendSites (:1347-1444,
 * Sites/%010d and SiteTimes). */
int lmxio_append_species_counts(lmxio_file *f, uint32_t replicate, uint32_t n, uint32_t nS, const double *times,
                                const uint32_t *counts);
int lmxio_append_lattice(lmxio_file *f, uint32_t replicate, double time, uint32_t X, uint32_t Y, uint32_t Z,
                         uint32_t P, uint32_t bytesPerParticle, const void *lattice);
int lmxio_append_sites(lmxio_file *f, uint32_t replicate, double time, uint32_t X, uint32_t Y, uint32_t Z,
                       const uint8_t *sites);
/* replicateExists (SimulationFile.cpp:1077-1083). */
int lmxio_replicate_exists(lmxio_file *f, uint32_t replicate, int *exists);
/* getSpeciesCountTimes / getSpeciesCounts (SimulationFile.cpp:1951-2008); counts
 * [n][nS] row-major. */
int lmxio_get_species_counts_size(lmxio_file *f, uint32_t replicate, uint32_t *n, uint32_t *nS);
int lmxio_get_species_counts(lmxio_file *f, uint32_t replicate, double *times, uint32_t *counts);
/* getLatticeTimes / getLattice (SimulationFile.cpp:1929-1949, :2010-2056):
 * the number of frames, then frame `index` (latticeBytes must equal the
 * stored frame's size; *bytesPerParticle receives 1 or 4). */
int lmxio_get_lattice_count(lmxio_file *f, uint32_t replicate, uint32_t *n);
int lmxio_get_lattice(lmxio_file *f, uint32_t replicate, uint32_t index, double *time, void *lattice,
                      size_t latticeBytes, uint32_t *bytesPerParticle);

/* A new rank-1 uint32 dataset in the replicate's group, gzip-compressed in
 * chunks of min(16, n) (makeNewDataset, SimulationFile.cpp:2368-2393; the
 * solver writes MaxParticleCounts and MaxSiteCounts this way,
 * MpdRdmeSolver.cu:919-931).  Fails if the dataset exists. */
int lmxio_new_replicate_dataset_u32(lmxio_file *f, uint32_t replicate, const char *name, uint32_t n,
                                    const uint32_t *data);
/* Reads such a dataset: *n receives its length; data (capacity cap
 * entries) may be NULL to query it. */
int lmxio_get_replicate_dataset_u32(lmxio_file *f, uint32_t replicate, const char *name, uint32_t *data,
                                    uint32_t cap, uint32_t *n);

#ifdef __cplusplus
}
#endif
#endif
