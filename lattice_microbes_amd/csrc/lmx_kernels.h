// Host-side launchers of the MPD-RDME kernels (implemented in mpd_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include "lmx_device.h"

namespace lmx {

struct Params;
struct Member;

hipError_t launch_sweep(const Params &p, const uint32_t *src, uint32_t *dst, int axis, hipStream_t st);
hipError_t launch_reaction(const Params &p, uint32_t *lat, hipStream_t st);
hipError_t launch_resolve(const Params &p, uint32_t *lat, hipStream_t st);
bool fast_path_ok(const Params &p);
hipError_t launch_sweep_fast(const Params &p, const uint32_t *src, uint32_t *dst, int axis, uint32_t seg,
                             hipStream_t st);
bool col_ok(const Params &p, int axis, uint32_t seg);
hipError_t launch_sweep_col(const Params &p, const uint32_t *src, uint32_t *dst, int axis, uint32_t seg,
                            hipStream_t st);
hipError_t launch_sweep_tile(const Params &p, const uint32_t *src, uint32_t *dst, int axis, bool react,
                             const ReactionList &rl, hipStream_t st);
hipError_t launch_sweep_sparse(const Params &p, const uint32_t *src, uint32_t *dst, int axis, bool react,
                               const ReactionList &rl, hipStream_t st);
hipError_t launch_reaction_list(const Params &p, uint32_t *lat, const ReactionList &rl, int num_cu, hipStream_t st);
// The static reaction list (lmx_run_steps_static): kListShards shards of
// kStaticListCap entries in a device array of the library; `device` must be
// current.  Lattices of up to kStaticListMaxSites sites use it.
constexpr uint32_t kStaticListCap = 256;
constexpr size_t kStaticListMaxSites = (size_t)kListShards * kStaticListCap * 8;
constexpr int kMaxDevices = 64;
hipError_t static_reaction_list(int device, uint32_t W, ReactionList *out);
bool fused_ok(const Params &p);
uint32_t fused_tile_rows(uint32_t X, uint32_t W);
hipError_t launch_fused(const Params &p, const uint32_t *src, uint32_t *dst, uint32_t zseg, hipStream_t st);
hipError_t launch_reaction_fast(const Params &p, uint32_t *lat, hipStream_t st);
hipError_t launch_prob_tables(const Params &p, float *prob1, float *prob0, hipStream_t st);
uint32_t x4_blocks(const Params &p);
uint32_t reaction_blocks(const Params &p);
struct TileGeom;
uint32_t batch_list_shards(const TileGeom &zgeo, uint32_t threads);
constexpr uint32_t kBatchListBlockSites = 1u << 22;
bool batch_list_by_block(uint32_t sites);
uint32_t batch_list_blocks(uint32_t nshards, uint32_t sites);
// One batched step without the resolver: x, y, z sweeps and (if total[3] > 0)
// the reaction phase.  For launch k (x, y, z, reaction), start[k] is a device
// prefix array of n+1 block offsets and map[k] a device array of total[k]
// member indices (block b belongs to member map[k][b]); total[] holds the
// four grid sizes.
// Tiled sweeps in blocks of tile_threads (a multiple of 64, at most
// tile_threads_for(W)), the size the members' tile geometry was made for.
hipError_t launch_batch_sweeps(const Member *m, const uint32_t *const start[4], const uint32_t *const map[4],
                               uint32_t n, const uint32_t *total, uint32_t W, uint32_t tile_threads, hipStream_t st);
hipError_t launch_batch_resolve(Member *m, uint32_t n, hipStream_t st);
hipError_t launch_copy(const void *src, void *dst, size_t bytes, int num_cu, hipStream_t st);
hipError_t launch_fill(uint32_t *ptr, size_t n, uint32_t value, hipStream_t st);
hipError_t launch_validate(const Params &p, const uint32_t *lat, hipStream_t st);
hipError_t launch_counts(const uint32_t *lat, size_t nwords, uint32_t nS, bool words, uint32_t *counts,
                         hipStream_t st);
hipError_t launch_exp(const double *x, double *y, size_t n, hipStream_t st);
hipError_t launch_prob(const float *a, float *y, size_t n, hipStream_t st);

}  // namespace lmx
