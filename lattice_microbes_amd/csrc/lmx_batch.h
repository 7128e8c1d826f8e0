// Device-side description of a batch of replicates stepped by one launch per
// kernel class.  Each member is an independent solver (its own model tables,
// lattice, seed, step counter and overflow list); a batched kernel maps its
// block to a member and runs that member's share with the single-solver body,
// so every member's lattice is bit-identical to its plain run.
#pragma once

#include <stdint.h>

#include "lmx_device.h"

namespace lmx {

// Division by a divisor fixed per launch, without a division instruction
// sequence: q = floor(n / d) = umulhi(n, mul) >> shr for n < 2^31 (the
// round-up method of Granlund and Montgomery, "Division by invariant
// integers using multiplication", 1994): with l = ceil(log2 d),
// mul = ceil(2^(31+l) / d) < 2^32 and shr = l - 1; d = 1 is the identity.
struct FastDiv {
    uint32_t d, mul, shr;
};

__host__ __device__ inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f{d, 0u, 0u};
    if (d > 1) {
        uint32_t l = 0;
        while ((1ull << l) < d) l++;
        f.mul = (uint32_t)(((1ull << (31 + l)) + d - 1) / d);
        f.shr = l - 1;
    }
    return f;
}

__host__ __device__ inline uint32_t fast_div(uint32_t n, const FastDiv &f)
{
#ifdef __HIP_DEVICE_COMPILE__
    return f.d == 1 ? n : __umulhi(n, f.mul) >> f.shr;
#else
    return f.d == 1 ? n : (uint32_t)(((unsigned long long)n * f.mul) >> 32) >> f.shr;
#endif
}

// Geometry of a tiled y (axis 1) or z (axis 2) sweep: tiles of Gx groups of
// 4 x-sites by Ty rows along the axis, nxchunk x chunks and nouter rows of
// the remaining axis.  A block holds No tiles of consecutive outer rows (more
// than one when a row is shorter than the block), nob blocks cover the outer
// rows; the divisors of the block and thread decompositions as FastDiv.
struct TileGeom {
    uint32_t stride, len, ostride, nouter, Gx, Ty, nxchunk, ntiles, tile_first, No, nob;
    FastDiv div_nob, div_Gx, div_nxchunk, div_ntiles, div_tile;
};

__host__ __device__ inline uint32_t tile_threads_for(uint32_t W) { return W == 4 ? 512u : 1024u; }

// Blocks of at most kNarrowTileThreads threads take tiles kNarrowTileGx
// groups (32 sites) wide and threads / kNarrowTileGx rows tall: eight such
// blocks share a CU, so one block's row loads overlap another's barrier and
// stores.  Larger blocks take tiles one wave (64 groups) wide.  The narrow
// tile is the default of the single-solver sweeps and of batches.
constexpr uint32_t kNarrowTileThreads = 256u, kNarrowTileGx = 8u;

__host__ __device__ inline TileGeom tile_geom(const Params &p, int axis, uint32_t threads)
{
    TileGeom g;
    g.stride = axis == 1 ? p.X : p.XY;
    g.len = axis == 1 ? p.Y : p.Z;
    g.ostride = axis == 1 ? p.XY : p.X;
    g.nouter = axis == 1 ? p.Z : p.Y;
    const uint32_t gxn = p.X >> 2;
    const uint32_t gx_max = threads <= kNarrowTileThreads ? kNarrowTileGx : 64u;
    g.Gx = gxn < gx_max ? gxn : gx_max;
    g.Ty = threads / g.Gx < g.len ? threads / g.Gx : g.len;
    g.nxchunk = (gxn + g.Gx - 1) / g.Gx;
    g.ntiles = (g.len + g.Ty - 1) / g.Ty;
    // Tiles per block: as many as the threads hold, with their two apron
    // rows each within the block's LDS rows (threads + 128 entries: No * Gx
    // <= 64).
    uint32_t no = threads / (g.Gx * g.Ty);
    if (no * g.Gx > 64u) no = 64u / g.Gx;
    if (no > g.nouter) no = g.nouter;
    g.No = no > 0u ? no : 1u;
    g.nob = (g.nouter + g.No - 1) / g.No;
    // Adjacent tiles share apron rows: wide tiles number them consecutively
    // when they are closer in memory than the outer rows (y sweep).  Narrow
    // tiles run over the outer rows first in both sweeps (on the whole-cell
    // lattice 1.5 us faster per y and per z sweep than tiles first).
    g.tile_first = threads > kNarrowTileThreads &&
                   (unsigned long long)g.Ty * g.stride < (unsigned long long)g.ostride ? 1u : 0u;
    g.div_nob = make_fastdiv(g.nob);
    g.div_Gx = make_fastdiv(g.Gx);
    g.div_nxchunk = make_fastdiv(g.nxchunk);
    g.div_ntiles = make_fastdiv(g.ntiles);
    g.div_tile = make_fastdiv(g.Gx * g.Ty);
    return g;
}

// Blocks of one tiled sweep: a multiple of 8 so that block b runs on XCD b % 8.
__host__ __device__ inline uint32_t tile_blocks(const TileGeom &g)
{
    const uint32_t nblocks = g.nxchunk * g.nob * g.ntiles;
    return (nblocks + 7) / 8 * 8;
}

// Reaction phase of a member: none (no reactions), the full in-place pass, or
// the list its z sweep fills (no zero-order reactions).
constexpr uint32_t kReactNone = 0, kReactFull = 1, kReactList = 2;

struct Member {
    Params p;
    // lat[0] holds the state at the start of a step, lat[1] the other
    // buffer; the batched resolver swaps them (reading a parity first would
    // put one more dependent load before every block's first lattice load)
    uint32_t *lat[2];
    TileGeom geo[2];     // y and z tiled sweeps
    ReactionList rl;     // the member's own reaction list (kReactList)
    uint32_t react;      // kReactNone, kReactFull or kReactList
    uint32_t nshards;    // list shards its z sweep fills (kReactList)
    uint32_t list_by_block;  // 1: a block per list shard (large members); 0: a wave per several shards
    // timestep_hash(seed32, timestep, s) of the step about to run for the
    // x, y and z sweeps: set by the host, advanced by the batched resolver,
    // so that a batched block reads it with the descriptor instead of
    // waiting for the control block's timestep and hashing it
    unsigned long long h[3];
};

// Work units (x4 blocks, sweep tiles) per batched block: a block loops over
// its member's units with the member's block count as stride, so that the
// member lookup and descriptor reads at a block's start are paid once per
// kBatchUnits units of a smaller member.  A member of at least
// kBatchOwnUnits units fills the device alone and keeps one unit per block
// (its blocks then overlap each other's latencies).  A member's block count
// is a multiple of 8, so every unit of a block keeps the XCD of the
// single-solver block order.
constexpr uint32_t kBatchUnits = 4, kBatchOwnUnits = 8192;
__host__ __device__ inline uint32_t batch_blocks(uint32_t units)
{
    const uint32_t k = units >= kBatchOwnUnits ? 1u : kBatchUnits;
    const uint32_t b = (units + k - 1) / k;
    return (b + 7) / 8 * 8;
}

// Member of block b of a batched launch: map[b], written by the host with the
// block offsets (one scalar load; a binary search over the offsets was a
// chain of eight dependent loads at the start of every block of a batch of
// 256 members).  A batch of one member reads no map.
__device__ inline uint32_t member_of(const uint32_t *__restrict__ map, uint32_t n, uint32_t b)
{
    return n == 1u ? 0u : map[b];
}

}  // namespace lmx
