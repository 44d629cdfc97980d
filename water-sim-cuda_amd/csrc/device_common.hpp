// device_common.hpp -- device-side math and layout shared by the HIP kernels.
//
// The per-particle and per-node arithmetic restates the reference's device
// helpers (common.cuh:131-378) with the same operation order; the layout and
// work decomposition are this engine's own (chunked Morton-sorted particles,
// per-wave LDS tiles, 4^3 activity flags).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

// Device-side copy of one scene's scalar config (the reference keeps a single
// SimConfig in __constant__ d_config, config.cu:3).  An engine holds one or
// more independent scenes; their grids, flags and SDFs are concatenated, and
// goff/foff locate the scene's part.  Kernels read the scene's DevCfg from a
// device table indexed by the scene id of their work item.
struct DevCfg {
    int sx, sy, sz;  // SIZE_X/Y/Z
    int sxy;         // SIZE_X*SIZE_Y
    uint32_t pad;    // PADDING
    float gravity, dt, gamma, compression;
    float vol, mass;  // per-particle V = RESOLUTION^3 and m = 1 (initFMatrices)
    int fbx, fby, fbz;  // activity-flag blocks (4^3 nodes) per axis
    int goff;           // first node of the scene in the concatenated grid
    int foff;           // first flag block of the scene
};

// Scene lookup for per-particle kernels: particle (caller index) i belongs to
// scene s with pstart[s] <= i < pstart[s + 1].
struct SceneTable {
    const DevCfg* cfgs;
    const int* pstart;  // nscenes + 1 entries
    int nscenes;
};

// A wave-uniform struct read from memory, moved into scalar registers field by
// field (the loads themselves stay vector loads: the kernels also write
// memory, so the compiler cannot use the scalar cache for them).
template <class T> __device__ __forceinline__ T uniform(const T& v) {
    static_assert(sizeof(T) % 4 == 0, "32-bit fields");
    T r;
    const int* src = reinterpret_cast<const int*>(&v);
    int* dst = reinterpret_cast<int*>(&r);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return r;
}

__device__ __forceinline__ int scene_of(const SceneTable& t, int i) {
    int lo = 0, hi = t.nscenes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t.pstart[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Particle planes: attribute-major, attr*stride + slot.  Attributes 0..23 of
// the reference layout (pos, vel, C, F); VOL and MASS are engine scalars.
enum : int {
    A_PX = 0, A_PY, A_PZ, A_VX, A_VY, A_VZ,
    A_C0, A_C1, A_C2, A_C3, A_C4, A_C5, A_C6, A_C7, A_C8,
    A_F0, A_F1, A_F2, A_F3, A_F4, A_F5, A_F6, A_F7, A_F8,
    A_STATE = 24
};

// Chunk = up to CHUNK consecutive particles (in base-cell order) whose base
// cells (trunc(x - 0.5), the lowest node of the quadratic stencil) lie in one
// 8 x 4 x 4 block, processed by one transfer wave.  It occupies the slots
// [start, start + count) in S rounds; round r holds n_r <= 64 particles stored
// contiguously, and no two of them share a base cell, so the lanes of a round
// update distinct arena nodes at each stencil offset.  The layout is
// McNaughton's wrap-around rule: with the chunk's particles listed cell after
// cell and S = max(largest cell, ceil(count/64)), list entry q goes to round
// q % S (as if to lane q / S; the re-sort then orders each round's lanes by
// LDS bank group).  A cell's entries are at most S consecutive list
// positions, so they never meet in one round.  Hence n_r = nl = ceil(count/S)
// for r < full = count - (nl - 1) S and nl - 1 after.
constexpr int CHUNK_CAP_MAX = 2048;  // largest CHUNK (the re-sort's LDS and per-thread items)
struct ChunkDesc {
    int start;       // first slot (= first index of the chunk in sorted order)
    int count;       // particles in the chunk
    int ox, oy, oz;  // node origin of the chunk's LDS tile
    int S;           // rounds
    int scene;       // scene of the chunk (index into the DevCfg table)
    int block;       // its chunk block when the block is one chunk, else -1
    int paired;      // 1: a round may hold two members of a cell, in lanes 2k and 2k + 1
};

// Chunk block: 8 x 4 x 4 base cells, i.e. two x-adjacent 4^3 flag blocks
// (the grid's activity flags keep their 4^3 granularity).  With 128 cells
// per block, the rounds of a chunk can be filled close to 64 lanes even where
// cell occupancies are uneven (see k_reorder).
constexpr int BLOCK_X = 8, BLOCK_Y = 4, BLOCK_Z = 4;
constexpr int BLOCK_BITS = 7;                    // log2(8 * 4 * 4): cell index bits of a bucket key
constexpr int BLOCK_CELLS_N = 1 << BLOCK_BITS;   // 128
// Nodes of a chunk's LDS tile per axis: the block's base cells plus the two
// further nodes a quadratic stencil reaches.  Particles whose base node moved
// out of their block since the last re-sort take the global-memory path.
constexpr int TILE_DX = BLOCK_X + 2, TILE_DY = BLOCK_Y + 2, TILE_DZ = BLOCK_Z + 2;  // 10 x 6 x 6
constexpr int TILE_N = TILE_DX * TILE_DY * TILE_DZ;                                // 360
constexpr int TRANSFER_THREADS = 64;  // one wave per chunk (wave-private LDS arena)

// LDS slot pitches of the tile/arena (16-byte slots), x fastest.  The
// defaults store the 10 x 6 x 6 tile densely.
#ifndef MPM_TILE_PY
#define MPM_TILE_PY 10
#endif
#ifndef MPM_TILE_PZ
#define MPM_TILE_PZ 60
#endif
constexpr int TILE_PY = MPM_TILE_PY;
constexpr int TILE_PZ = MPM_TILE_PZ;
constexpr int TILE_SLOTS = (TILE_DX - 1) + TILE_PY * (TILE_DY - 1) + TILE_PZ * (TILE_DZ - 1) + 1;
static_assert(TILE_PY >= TILE_DX && TILE_PZ >= TILE_PY * TILE_DY, "tile pitches must not overlap");

// Tile coordinates of entry i of a dense x-fastest walk over the tile.
__device__ __forceinline__ void tile_coords(int i, int& lx, int& ly, int& lz) {
    lx = i % TILE_DX;
    ly = (i / TILE_DX) % TILE_DY;
    lz = i / (TILE_DX * TILE_DY);
}

// getGridIdx, common.cuh:131-134, offset to the scene's part of the grid.
__device__ __forceinline__ int gidx(const DevCfg& c, int x, int y, int z) { return c.goff + z * c.sxy + y * c.sx + x; }

// isInBounds, common.cuh:212-218 (negative coordinates compare as large unsigned).
__device__ __forceinline__ bool inb(const DevCfg& c, int x, int y, int z) {
    return (unsigned)x < (unsigned)c.sx && (unsigned)y < (unsigned)c.sy && (unsigned)z < (unsigned)c.sz;
}

__device__ __forceinline__ int flag_block(const DevCfg& c, int x, int y, int z) {
    return c.foff + ((z >> 2) * c.fby + (y >> 2)) * c.fbx + (x >> 2);
}

// spline, common.cuh:205-210.  `0.75 - x*x` rounds the float product first and
// the double difference is exactly the float difference, so float arithmetic
// without contraction reproduces it; the second branch squares an exact
// difference and halves it, which float also reproduces.
__device__ __forceinline__ float spline(float x) {
#pragma clang fp contract(off)
    float ax = fabsf(x);
    if (ax < 0.5f) return 0.75f - x * x;
    if (ax < 1.5f) {
        float t = 1.5f - ax;
        return 0.5f * (t * t);
    }
    return 0.0f;
}

// (int)(x - 0.5f): C truncation, kernels.cu:14, :204.
__device__ __forceinline__ int base_of(float x) { return (int)(x - 0.5f); }

// det3x3, common.cuh:250-253.
__device__ __forceinline__ float det3(const float* A) {
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}

// expandBits / calculateMorton, common.cuh:277-284, :304-307.
__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
__device__ __forceinline__ uint32_t key_coord(float p) {
    float v = fmaxf(0.0f, p);
    return v < 4294967296.0f ? (uint32_t)v : 0xFFFFFFFFu;
}
// setKeys, kernels.cu:361-371.  Returned with the sign bit flipped so that an
// unsigned sort orders the keys as the reference's signed-int sort does.
__device__ __forceinline__ uint32_t sort_key(float x, float y, float z) {
    uint32_t k = (expand_bits(key_coord(z)) << 2) | (expand_bits(key_coord(y)) << 1) | expand_bits(key_coord(x));
    return k ^ 0x80000000u;
}
// Engine-internal ordering key: Morton code of the base cell with each
// coordinate clamped into the grid (particles outside the grid only need some
// block), so keys have 3 * ceil(log2(max SIZE)) bits.
__device__ __forceinline__ uint32_t base_key(const DevCfg& c, float x, float y, float z) {
    const uint32_t bx = (uint32_t)min(max(base_of(x), 0), c.sx - 1);
    const uint32_t by = (uint32_t)min(max(base_of(y), 0), c.sy - 1);
    const uint32_t bz = (uint32_t)min(max(base_of(z), 0), c.sz - 1);
    return (expand_bits(bz) << 2) | (expand_bits(by) << 1) | expand_bits(bx);
}
// Inverse of expand_bits on the 10 low interleaved bits.
__device__ __forceinline__ uint32_t compact_bits(uint32_t v) {
    v &= 0x49249249u;
    v = (v ^ (v >> 2)) & 0xC30C30C3u;
    v = (v ^ (v >> 4)) & 0x0F00F00Fu;
    v = (v ^ (v >> 8)) & 0xFF0000FFu;
    v = (v ^ (v >> 16)) & 0x000003FFu;
    return v;
}

// Trilinear blend of the 8 corner values, in getSDF's order (common.cuh:154-166).
__device__ __forceinline__ float trilerp(float n0, float n1, float n2, float n3, float n4, float n5, float n6,
                                         float n7, float tx, float ty, float tz) {
    float a0 = n0 * (1.0f - tx) + n1 * tx, a1 = n2 * (1.0f - tx) + n3 * tx;
    float a2 = n4 * (1.0f - tx) + n5 * tx, a3 = n6 * (1.0f - tx) + n7 * tx;
    float b0 = a0 * (1.0f - ty) + a1 * ty, b1 = a2 * (1.0f - ty) + a3 * ty;
    return b0 * (1.0f - tz) + b1 * tz;
}

// getSDF, common.cuh:136-167: trilinear on node values, floor cells, clamped.
__device__ __forceinline__ float sdf_at(const DevCfg& c, const float* __restrict__ sdf, float px, float py, float pz) {
    int i = (int)floorf(px), j = (int)floorf(py), k = (int)floorf(pz);
    float tx = px - (float)i, ty = py - (float)j, tz = pz - (float)k;
    int x0 = max(0, min(i, c.sx - 1)), x1 = max(0, min(i + 1, c.sx - 1));
    int y0 = max(0, min(j, c.sy - 1)), y1 = max(0, min(j + 1, c.sy - 1));
    int z0 = max(0, min(k, c.sz - 1)), z1 = max(0, min(k + 1, c.sz - 1));
    return trilerp(sdf[gidx(c, x0, y0, z0)], sdf[gidx(c, x1, y0, z0)], sdf[gidx(c, x0, y1, z0)],
                   sdf[gidx(c, x1, y1, z0)], sdf[gidx(c, x0, y0, z1)], sdf[gidx(c, x1, y0, z1)],
                   sdf[gidx(c, x0, y1, z1)], sdf[gidx(c, x1, y1, z1)], tx, ty, tz);
}

// calculateNormal, common.cuh:169-182.
__device__ __forceinline__ void sdf_normal(const DevCfg& c, const float* __restrict__ sdf, float px, float py, float pz,
                                           float& nx, float& ny, float& nz) {
    const float eps = 0.1f;
    nx = sdf_at(c, sdf, px + eps, py, pz) - sdf_at(c, sdf, px - eps, py, pz);
    ny = sdf_at(c, sdf, px, py + eps, pz) - sdf_at(c, sdf, px, py - eps, pz);
    nz = sdf_at(c, sdf, px, py, pz + eps) - sdf_at(c, sdf, px, py, pz - eps);
    float len = sqrtf(nx * nx + ny * ny + nz * nz) + 1e-6f;
    nx /= len; ny /= len; nz /= len;
}

// getNormalAtNode, common.cuh:195-203.
__device__ __forceinline__ void node_normal(const DevCfg& c, const float* __restrict__ sdf, int x, int y, int z,
                                            float& nx, float& ny, float& nz) {
    nx = sdf[gidx(c, min(x + 1, c.sx - 1), y, z)] - sdf[gidx(c, max(x - 1, 0), y, z)];
    ny = sdf[gidx(c, x, min(y + 1, c.sy - 1), z)] - sdf[gidx(c, x, max(y - 1, 0), z)];
    nz = sdf[gidx(c, x, y, min(z + 1, c.sz - 1))] - sdf[gidx(c, x, y, max(z - 1, 0))];
    float len = sqrtf(nx * nx + ny * ny + nz * nz) + 1e-9f;
    nx /= len; ny /= len; nz /= len;
}

__device__ __forceinline__ void atomic_add_global(float* p, float v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace mpm
