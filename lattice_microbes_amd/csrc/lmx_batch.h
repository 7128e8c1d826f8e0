// Device-side description of a batch of replicates stepped by one launch per
// kernel class.  Each member is an independent solver (its own model tables,
// lattice, seed, step counter and overflow list); a batched kernel maps its
// block to a member and runs that member's share with the single-solver body,
// so every member's lattice is bit-identical to its plain run.
#pragma once

#include <stdint.h>

#include "lmx_device.h"

namespace lmx {

// Geometry of a tiled y (axis 1) or z (axis 2) sweep: tiles of Gx groups of
// 4 x-sites by Ty rows along the axis, nxchunk x chunks and nouter rows of
// the remaining axis.
struct TileGeom {
    uint32_t stride, len, ostride, nouter, Gx, Ty, nxchunk, ntiles, tile_first;
};

__host__ __device__ inline uint32_t tile_threads_for(uint32_t W) { return W == 4 ? 512u : 1024u; }

__host__ __device__ inline TileGeom tile_geom(const Params &p, int axis, uint32_t threads)
{
    TileGeom g;
    g.stride = axis == 1 ? p.X : p.XY;
    g.len = axis == 1 ? p.Y : p.Z;
    g.ostride = axis == 1 ? p.XY : p.X;
    g.nouter = axis == 1 ? p.Z : p.Y;
    const uint32_t gxn = p.X >> 2;
    g.Gx = gxn < 64u ? gxn : 64u;
    g.Ty = threads / g.Gx;
    g.nxchunk = (gxn + g.Gx - 1) / g.Gx;
    g.ntiles = (g.len + g.Ty - 1) / g.Ty;
    // Adjacent tiles share apron rows: number them consecutively when they
    // are closer in memory than the outer rows (y sweep).
    g.tile_first = (unsigned long long)g.Ty * g.stride < (unsigned long long)g.ostride ? 1u : 0u;
    return g;
}

// Blocks of one tiled sweep: a multiple of 8 so that block b runs on XCD b % 8.
__host__ __device__ inline uint32_t tile_blocks(const TileGeom &g)
{
    const uint32_t nblocks = g.nxchunk * g.nouter * g.ntiles;
    return (nblocks + 7) / 8 * 8;
}

struct Member {
    Params p;
    uint32_t *lat[2];
    TileGeom geo[2];     // y and z tiled sweeps
    uint32_t cur;        // lat[cur] holds the state at the start of a step; the batched resolver flips it
    uint32_t pad;
};

// Member of global block b: start[] holds, for n members, the first block of
// each member and start[n] the total (members with no blocks repeat a start).
__device__ inline uint32_t find_member(const uint32_t *__restrict__ start, uint32_t n, uint32_t b)
{
    uint32_t lo = 0, hi = n;  // start[lo] <= b < start[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

}  // namespace lmx
