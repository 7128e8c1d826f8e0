// Host-side launchers of the MPD-RDME kernels (implemented in mpd_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

namespace lmx {

struct Params;

hipError_t launch_sweep(const Params &p, const uint32_t *src, uint32_t *dst, int axis, hipStream_t st);
hipError_t launch_reaction(const Params &p, uint32_t *lat, hipStream_t st);
hipError_t launch_resolve(const Params &p, uint32_t *lat, hipStream_t st);
bool fast_path_ok(const Params &p);
hipError_t launch_sweep_fast(const Params &p, const uint32_t *src, uint32_t *dst, int axis, uint32_t seg,
                             hipStream_t st);
hipError_t launch_sweep_tile(const Params &p, const uint32_t *src, uint32_t *dst, int axis, bool react,
                             uint32_t *deferred, uint32_t defer_cap, hipStream_t st);
hipError_t launch_sweep_sparse(const Params &p, const uint32_t *src, uint32_t *dst, int axis, bool react,
                               uint32_t *deferred, uint32_t defer_cap, hipStream_t st);
hipError_t launch_reaction_deferred(const Params &p, uint32_t *lat, const uint32_t *deferred, uint32_t defer_cap,
                                    int num_cu, hipStream_t st);
bool fused_ok(const Params &p);
uint32_t fused_tile_rows(uint32_t X, uint32_t W);
hipError_t launch_fused(const Params &p, const uint32_t *src, uint32_t *dst, uint32_t zseg, hipStream_t st);
hipError_t launch_reaction_fast(const Params &p, uint32_t *lat, hipStream_t st);
hipError_t launch_prob_tables(const Params &p, float *prob1, float *prob0, hipStream_t st);
hipError_t launch_counts(const uint32_t *lat, size_t nwords, uint32_t nS, uint32_t *counts, hipStream_t st);
hipError_t launch_exp(const double *x, double *y, size_t n, hipStream_t st);
hipError_t launch_prob(const float *a, float *y, size_t n, hipStream_t st);

}  // namespace lmx
