// kernels_grid.hip -- grid update, mesh SDF, state export and small fills.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"

namespace mpm {
namespace {

// gridUpdate, kernels.cu:144-196, per node: velocity from momentum, gravity,
// collider slip, PADDING walls.  Split into the node's loads and the rest, so
// that a wave can have several nodes' loads in flight.
struct GuNode {
    int g;  // < 0: nothing to do
    float mass, mx, my, mz, dist;
};

__device__ __forceinline__ GuNode grid_update_load(GridBuf m, const float* __restrict__ sdf, int g, bool fm) {
    GuNode n;
    n.g = g;
    n.mass = n.mx = n.my = n.mz = n.dist = 0.0f;
    if (fm && g >= 0) {
        n.mass = m.plane[0][g];
        n.mx = m.plane[1][g];
        n.my = m.plane[2][g];
        n.mz = m.plane[3][g];
        n.dist = sdf[g];
    }
    return n;
}

// fm: the node's block was touched by the last P2G (update it); fv: the
// node's block of the previous velocity buffer is to be zeroed.  With a
// velocity copy (planes 1..3 of `vc`), the new velocity is written there too.
template <bool COPY>
__device__ __forceinline__ void grid_update_store(const DevCfg& c, GridBuf m, GridBuf v, const float* __restrict__ sdf,
                                                  const GuNode& n, int x, int y, int z, bool fm, bool fv,
                                                  const GridBuf& vc) {
    if (n.g < 0) return;
    const int g = n.g;
    if (fv) {
        v.plane[0][g] = 0.0f;
        v.plane[1][g] = 0.0f;
        v.plane[2][g] = 0.0f;
        v.plane[3][g] = 0.0f;
    }
    if (!fm) return;
    float vx = 0.0f, vy = 0.0f, vz = 0.0f;
    if (n.mass > 1e-9f) {
        vx = n.mx / n.mass;
        vy = n.my / n.mass - c.gravity * c.dt;
        vz = n.mz / n.mass;
        if (n.dist < 0.0f) {
            float nx, ny, nz;
            node_normal(c, sdf, x, y, z, nx, ny, nz);
            float vn = vx * nx + vy * ny + vz * nz;
            if (vn < 0.0f) {
                vx -= nx * vn; vy -= ny * vn; vz -= nz * vn;
            }
        }
        if ((uint32_t)x < c.pad && vx < 0) vx = 0.0f;
        if ((uint32_t)y < c.pad && vy < 0) vy = 0.0f;
        if ((uint32_t)z < c.pad && vz < 0) vz = 0.0f;
        if ((uint32_t)x > (uint32_t)c.sx - 1u - c.pad && vx > 0) vx = 0.0f;
        if ((uint32_t)y > (uint32_t)c.sy - 1u - c.pad && vy > 0) vy = 0.0f;
        if ((uint32_t)z > (uint32_t)c.sz - 1u - c.pad && vz > 0) vz = 0.0f;
    }
    m.plane[1][g] = vx;
    m.plane[2][g] = vy;
    m.plane[3][g] = vz;
    if (COPY) {
        vc.plane[1][g] = vx;
        vc.plane[2][g] = vy;
        vc.plane[3][g] = vz;
    }
}

// The grid update is restricted to the 4^3 node blocks that the last P2G
// touched; the same pass zeroes the blocks of the previous velocity buffer
// that its own P2G had touched (and clears their flags), so that buffer is
// all-zero again when it next accumulates.  Work goes by row segments of four
// x-adjacent blocks (16 x 4 x 4 nodes), so that a wave's accesses cover 64
// contiguous bytes per grid row.
//
// 1. The segments with a flagged block: entry {first block, touched mask |
// to-clear mask << 4} (bit k: block first + k).  One thread per block; those
// with bx % 4 == 0 read their segment's flags.  One list append (atomic) per
// workgroup.  *count is zero on entry.  single: every block belongs to scene
// 0, whose row length is then read once into scalar registers (no per-block
// scene and config loads before the flags).
__global__ __launch_bounds__(256) void k_grid_list(const DevCfg* __restrict__ cfgs,
                                                   const uint16_t* __restrict__ block_scene,
                                                   const uint8_t* __restrict__ mflags, uint8_t* __restrict__ vflags,
                                                   int clear_v, int nblocks, uint2* __restrict__ list,
                                                   uint32_t* __restrict__ count, int single) {
    __shared__ uint32_t wcount[4];
    __shared__ uint32_t wbase;
    const int blk = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t fm = 0u, fv = 0u;
    int foff = 0, fbx = 1;
    if (single) {
        foff = __builtin_amdgcn_readfirstlane(cfgs[0].foff);
        fbx = __builtin_amdgcn_readfirstlane(cfgs[0].fbx);
    }
    if (blk < nblocks) {
        if (!single) {
            const DevCfg& c = cfgs[block_scene[blk]];
            foff = c.foff;
            fbx = c.fbx;
        }
        const int bx = (blk - foff) % fbx;
        if ((bx & 3) == 0) {
            const int nk = min(4, fbx - bx);
            for (int k = 0; k < nk; ++k) {
                if (mflags[blk + k]) fm |= 1u << k;
                if (clear_v && vflags[blk + k]) {
                    fv |= 1u << k;
                    vflags[blk + k] = 0;
                }
            }
        }
    }
    const bool any = (fm | fv) != 0u;
    const uint64_t m = __ballot(any);
    if (lane == 0) wcount[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t tot = wcount[0] + wcount[1] + wcount[2] + wcount[3];
        wbase = tot ? atomicAdd(count, tot) : 0u;
    }
    __syncthreads();
    if (any) {
        uint32_t base = wbase;
        for (int w = 0; w < wave; ++w) base += wcount[w];
        const uint64_t below = lane == 0 ? 0ull : (m & (~0ull >> (64 - lane)));
        const uint32_t slot = base + (uint32_t)__popcll(below);
        if (slot < (uint32_t)nblocks) list[slot] = make_uint2((uint32_t)blk, fm | (fv << 4));
    }
}

__global__ void k_set_u32_grid(uint32_t* p, uint32_t v) { *p = v; }

// 2. One wave per segment at a time: lane = (x 0..15, y 0..3), the four z
// slices' loads issued together.  single: every block belongs to scene 0.
__global__ __launch_bounds__(256) void k_grid_update(const DevCfg* __restrict__ cfgs,
                                                     const uint16_t* __restrict__ block_scene, GridBuf m, GridBuf v,
                                                     const float* __restrict__ sdf, const uint2* __restrict__ list,
                                                     const uint32_t* __restrict__ count,
                                                     uint32_t* __restrict__ next_count, int single, int cap) {
    const int n = (int)min(*count, (uint32_t)cap);
    if (next_count && blockIdx.x == 0 && threadIdx.x == 0) *next_count = 0u;  // the next update's list starts empty
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6), nw = gridDim.x * 4;
    DevCfg c0{};
    if (single) c0 = uniform(cfgs[0]);
    for (int i = __builtin_amdgcn_readfirstlane(wave); i < n; i += nw) {
        const uint2 e = uniform(list[i]);
        const int blk = (int)e.x;
        const DevCfg c = single ? c0 : uniform(cfgs[block_scene[blk]]);
        const int lb = blk - c.foff;
        const int bx = lb % c.fbx, by = (lb / c.fbx) % c.fby, bz = lb / (c.fbx * c.fby);
        const int k = (lane & 15) >> 2;
        const bool fm = (e.y >> k) & 1u, fv = (e.y >> (4 + k)) & 1u;
        const int x = bx * 4 + (lane & 15), y = by * 4 + (lane >> 4);
        GuNode nd[4];
#pragma unroll
        for (int zz = 0; zz < 4; ++zz) {
            const int z = bz * 4 + zz;
            const int g = (fm || fv) && inb(c, x, y, z) ? gidx(c, x, y, z) : -1;
            nd[zz] = grid_update_load(m, sdf, g, fm);
        }
#pragma unroll
        for (int zz = 0; zz < 4; ++zz) grid_update_store<false>(c, m, v, sdf, nd[zz], x, y, bz * 4 + zz, fm, fv, m);
    }
}

// The grid update of a corrected substep: the same per-segment work, plus the
// velocity copy (planes 1..3 of vc) and the list of nodes with mass > 1e-9
// ({x | y << 10 | z << 20, scene}, at most acap stored) that the corrector
// runs over.  A workgroup's four waves take four segments per iteration in
// step, so that the list takes one atomic per workgroup and iteration.
__global__ __launch_bounds__(256) void k_grid_update_corr(const DevCfg* __restrict__ cfgs,
                                                          const uint16_t* __restrict__ block_scene, GridBuf m,
                                                          GridBuf v, const float* __restrict__ sdf,
                                                          const uint2* __restrict__ list,
                                                          const uint32_t* __restrict__ count,
                                                          uint32_t* __restrict__ next_count, int single, int cap,
                                                          GridBuf vc, int2* __restrict__ active, int acap,
                                                          uint32_t* __restrict__ acount) {
    __shared__ uint32_t wcount[4];
    __shared__ uint32_t wbase;
    const uint32_t n = min(*count, (uint32_t)cap);
    if (next_count && blockIdx.x == 0 && threadIdx.x == 0) *next_count = 0u;  // the next update's list starts empty
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    DevCfg c0{};
    if (single) c0 = uniform(cfgs[0]);
    // the trip count is the workgroup's (barriers inside)
    for (uint32_t b0 = blockIdx.x * 4u; b0 < n; b0 += gridDim.x * 4u) {
        const uint32_t i = b0 + (uint32_t)wave;
        uint64_t am[4] = {0ull, 0ull, 0ull, 0ull};
        int x = 0, y = 0, bz = 0, sc = 0;
        if (i < n) {
            const uint2 e = uniform(list[i]);
            const int blk = (int)e.x;
            sc = single ? 0 : (int)block_scene[blk];
            const DevCfg c = single ? c0 : uniform(cfgs[sc]);
            const int lb = blk - c.foff;
            const int bx = lb % c.fbx, by = (lb / c.fbx) % c.fby;
            bz = lb / (c.fbx * c.fby);
            const int k = (lane & 15) >> 2;
            const bool fm = (e.y >> k) & 1u, fv = (e.y >> (4 + k)) & 1u;
            x = bx * 4 + (lane & 15);
            y = by * 4 + (lane >> 4);
            GuNode nd[4];
#pragma unroll
            for (int zz = 0; zz < 4; ++zz) {
                const int z = bz * 4 + zz;
                const int g = (fm || fv) && inb(c, x, y, z) ? gidx(c, x, y, z) : -1;
                nd[zz] = grid_update_load(m, sdf, g, fm);
            }
#pragma unroll
            for (int zz = 0; zz < 4; ++zz) {
                grid_update_store<true>(c, m, v, sdf, nd[zz], x, y, bz * 4 + zz, fm, fv, vc);
                am[zz] = __ballot(fm && nd[zz].g >= 0 && nd[zz].mass > 1e-9f);
            }
        }
        const uint32_t mine = (uint32_t)(__popcll(am[0]) + __popcll(am[1]) + __popcll(am[2]) + __popcll(am[3]));
        if (lane == 0) wcount[wave] = mine;
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t tot = wcount[0] + wcount[1] + wcount[2] + wcount[3];
            wbase = tot ? atomicAdd(acount, tot) : 0u;
        }
        __syncthreads();
        uint32_t base = wbase;
        for (int w = 0; w < wave; ++w) base += wcount[w];
        const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
#pragma unroll
        for (int zz = 0; zz < 4; ++zz) {
            const uint32_t slot = base + (uint32_t)__popcll(am[zz] & below);
            if (((am[zz] >> lane) & 1ull) && slot < (uint32_t)acap)
                active[slot] = make_int2(x | (y << 10) | ((bz * 4 + zz) << 20), sc);
            base += (uint32_t)__popcll(am[zz]);
        }
        __syncthreads();  // wcount and wbase are reused by the next iteration
    }
}

// makeSDF, kernels.cu:282-300, one thread per node.  The triangle index is
// uniform across the wave, so the 36-byte triangle loads are scalar loads.
// Winding by summed solid angles (calculateSolidAngle, common.cuh:310-323)
// and the minimum point-triangle distance (common.cuh:325-378).
__device__ __forceinline__ float point_tri_dist2(float px, float py, float pz, const float* t) {
    const float ax = t[0], ay = t[1], az = t[2];
    const float abx = t[3] - ax, aby = t[4] - ay, abz = t[5] - az;
    const float acx = t[6] - ax, acy = t[7] - ay, acz = t[8] - az;
    const float apx = px - ax, apy = py - ay, apz = pz - az;
    const float d1 = abx * apx + aby * apy + abz * apz;
    const float d2 = acx * apx + acy * apy + acz * apz;
    if (d1 <= 0.0f && d2 <= 0.0f) return apx * apx + apy * apy + apz * apz;
    const float bpx = px - t[3], bpy = py - t[4], bpz = pz - t[5];
    const float d3 = abx * bpx + aby * bpy + abz * bpz;
    const float d4 = acx * bpx + acy * bpy + acz * bpz;
    if (d3 >= 0.0f && d4 <= d3) return bpx * bpx + bpy * bpy + bpz * bpz;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float v = d1 / (d1 - d3);
        const float qx = px - (ax + v * abx), qy = py - (ay + v * aby), qz = pz - (az + v * abz);
        return qx * qx + qy * qy + qz * qz;
    }
    const float cpx = px - t[6], cpy = py - t[7], cpz = pz - t[8];
    const float d5 = abx * cpx + aby * cpy + abz * cpz;
    const float d6 = acx * cpx + acy * cpy + acz * cpz;
    if (d6 >= 0.0f && d5 <= d6) return cpx * cpx + cpy * cpy + cpz * cpz;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float w = d2 / (d2 - d6);
        const float qx = px - (ax + w * acx), qy = py - (ay + w * acy), qz = pz - (az + w * acz);
        return qx * qx + qy * qy + qz * qz;
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        const float qx = px - (t[3] + w * (t[6] - t[3]));
        const float qy = py - (t[4] + w * (t[7] - t[4]));
        const float qz = pz - (t[5] + w * (t[8] - t[5]));
        return qx * qx + qy * qy + qz * qz;
    }
    const float denom = 1.0f / (va + vb + vc);
    const float v = vb * denom, w = vc * denom;
    const float qx = px - (ax + abx * v + acx * w);
    const float qy = py - (ay + aby * v + acy * w);
    const float qz = pz - (az + abz * v + acz * w);
    return qx * qx + qy * qy + qz * qz;
}

__device__ __forceinline__ float solid_angle(float px, float py, float pz, const float* t) {
    const float ax = t[0] - px, ay = t[1] - py, az = t[2] - pz;
    const float bx = t[3] - px, by = t[4] - py, bz = t[5] - pz;
    const float cx = t[6] - px, cy = t[7] - py, cz = t[8] - pz;
    const float la = norm3df(ax, ay, az), lb = norm3df(bx, by, bz), lc = norm3df(cx, cy, cz);
    const float det = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
    const float div = (la * lb * lc) + (ax * bx + ay * by + az * bz) * lc + (ax * cx + ay * cy + az * cz) * lb +
                      (bx * cx + by * cy + bz * cz) * la;
    return 2.0f * atan2f(det, div);
}

__global__ __launch_bounds__(256) void k_make_sdf(DevCfg c, const float* __restrict__ tris, int nt,
                                                  float* __restrict__ sdf) {
    const int G = c.sxy * c.sz;
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= G) return;
    const int x = idx % c.sx, y = (idx / c.sx) % c.sy, z = idx / c.sxy;
    const float px = (float)x, py = (float)y, pz = (float)z;
    float wn = 0.0f, mind = 1e30f;
    for (int t = 0; t < nt; ++t) {
        const float* tri = tris + 9 * (size_t)t;
        wn += solid_angle(px, py, pz, tri);
        const float d2 = point_tri_dist2(px, py, pz, tri);
        if (d2 < mind) mind = d2;
    }
    wn = wn / (4.0f * 3.1415f);
    sdf[c.goff + idx] = (wn > 0.5f ? -1.0f : 1.0f) * sqrtf(mind);
}

// ---- makeSDF over a bounding-volume hierarchy (same results, less work) ----
// Nodes in depth-first order: an internal node's left child follows it, its
// right child starts at the left child's `skip`, and `skip` is the first node
// after its subtree, so both traversals below are stackless loops.
struct SdfNode {
    float lo[3];
    int skip;
    float hi[3];
    int first;   // leaf: first triangle of the reordered array
    float an[3]; // sum over the node's triangles of area x unit normal (= cross / 2)
    float rad;   // radius of a ball around `cen` holding every vertex
    float cen[3];
    int count;   // leaf: triangle count; 0 for internal nodes
};

// Far-field rule of the winding sum: a node whose ball is seen from farther
// than SDF_BETA radii contributes its dipole term an . (cen - q) / |cen - q|^3.
// A node whose approximate winding number lies within SDF_TAU of 1/2 gets the
// exact sum over all triangles in input order, as makeSDF computes it.
constexpr float SDF_BETA = 3.0f;
constexpr float SDF_TAU = 0.25f;

__device__ __forceinline__ float box_dist2(const SdfNode& n, float px, float py, float pz) {
    const float dx = fmaxf(fmaxf(n.lo[0] - px, px - n.hi[0]), 0.0f);
    const float dy = fmaxf(fmaxf(n.lo[1] - py, py - n.hi[1]), 0.0f);
    const float dz = fmaxf(fmaxf(n.lo[2] - pz, pz - n.hi[2]), 0.0f);
    return dx * dx + dy * dy + dz * dz;
}

// One thread per node.  The distance is the minimum of point_tri_dist2 over
// every triangle that a box test cannot exclude (with a relative margin that
// covers rounding), hence the brute force's minimum bit for bit; the sign
// follows the exact winding sum wherever the approximation is not clearly
// away from 1/2.
__global__ __launch_bounds__(256) void k_make_sdf_bvh(DevCfg c, const SdfNode* __restrict__ nodes, int nn,
                                                      const float* __restrict__ ptris,
                                                      const float* __restrict__ tris, int nt,
                                                      float* __restrict__ sdf, unsigned int* __restrict__ exact) {
    const int G = c.sxy * c.sz;
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= G) return;
    const int x = idx % c.sx, y = (idx / c.sx) % c.sy, z = idx / c.sxy;
    const float px = (float)x, py = (float)y, pz = (float)z;
    // an upper bound from the leaf reached by always taking the nearer box
    float best = 1e30f;
    int i = 0;
    while (nodes[i].count == 0) {
        const int l = i + 1, r = nodes[l].skip;
        i = box_dist2(nodes[l], px, py, pz) <= box_dist2(nodes[r], px, py, pz) ? l : r;
    }
    for (int t = 0; t < nodes[i].count; ++t)
        best = fminf(best, point_tri_dist2(px, py, pz, ptris + 9 * (size_t)(nodes[i].first + t)));
    i = 0;
    while (i < nn) {
        const SdfNode& n = nodes[i];
        if (box_dist2(n, px, py, pz) > best * 1.0001f + 1e-6f) {
            i = n.skip;
            continue;
        }
        if (n.count) {
            for (int t = 0; t < n.count; ++t)
                best = fminf(best, point_tri_dist2(px, py, pz, ptris + 9 * (size_t)(n.first + t)));
            i = n.skip;
        } else {
            ++i;
        }
    }
    float w = 0.0f;
    i = 0;
    while (i < nn) {
        const SdfNode& n = nodes[i];
        const float dx = n.cen[0] - px, dy = n.cen[1] - py, dz = n.cen[2] - pz;
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > (SDF_BETA * n.rad) * (SDF_BETA * n.rad)) {
            w += (n.an[0] * dx + n.an[1] * dy + n.an[2] * dz) / (d2 * sqrtf(d2));
            i = n.skip;
        } else if (n.count) {
            for (int t = 0; t < n.count; ++t) w += solid_angle(px, py, pz, ptris + 9 * (size_t)(n.first + t));
            i = n.skip;
        } else {
            ++i;
        }
    }
    float wn = w / (4.0f * 3.1415f);
    if (fabsf(wn - 0.5f) < SDF_TAU) {
        float we = 0.0f;
        for (int t = 0; t < nt; ++t) we += solid_angle(px, py, pz, tris + 9 * (size_t)t);
        wn = we / (4.0f * 3.1415f);
        atomicAdd(exact, 1u);
    }
    sdf[c.goff + idx] = (wn > 0.5f ? -1.0f : 1.0f) * sqrtf(best);
}

__global__ void k_fill_f32(float* p, size_t n, float v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u8(uint8_t* p, size_t n, uint8_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// initFMatrices, kernels.cu:349-358 (diagonal of F; V and m are engine scalars).
__global__ void k_init_F(float* soa26, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    soa26[(size_t)A_F0 * n + i] = 1.0f;
    soa26[(size_t)A_F4 * n + i] = 1.0f;
    soa26[(size_t)A_F8 * n + i] = 1.0f;
}

// VOL and MASS come from the particle's scene (they are per-scene constants).
__global__ __launch_bounds__(256) void k_export_soa26(const float* __restrict__ part, size_t stride,
                                                      const int* __restrict__ orig, int n, SceneTable t,
                                                      float* __restrict__ out) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const int id = orig[s];
        for (int a = 0; a < A_STATE; ++a) out[(size_t)a * n + id] = part[(size_t)a * stride + s];
        const DevCfg& c = t.cfgs[scene_of(t, id)];
        out[(size_t)24 * n + id] = c.vol;
        out[(size_t)25 * n + id] = c.mass;
    }
}

__global__ __launch_bounds__(256) void k_export_positions(const float* __restrict__ part, size_t stride,
                                                          const int* __restrict__ orig, int n,
                                                          float* __restrict__ out) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const int id = orig[s];
        out[id] = part[A_PX * stride + s];
        out[(size_t)n + id] = part[A_PY * stride + s];
        out[(size_t)2 * n + id] = part[A_PZ * stride + s];
    }
}

// dst plane a, slot i = src plane a, slot i for a < planes, and the caller
// indices (one launch for a whole particle state).
__global__ __launch_bounds__(256) void k_copy_state(const float* __restrict__ src, size_t src_stride,
                                                    const int* __restrict__ src_orig, float* __restrict__ dst,
                                                    size_t dst_stride, int* __restrict__ dst_orig, int n,
                                                    int planes) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int a = 0; a < planes; ++a) dst[(size_t)a * dst_stride + i] = src[(size_t)a * src_stride + i];
        dst_orig[i] = src_orig[i];
    }
}

// Positions in caller order, AoS (x, y, z per particle): changeFormat's
// layout (kernels.cu:341-347).
__global__ __launch_bounds__(256) void k_export_xyz(const float* __restrict__ part, size_t stride,
                                                    const int* __restrict__ orig, int n, float* __restrict__ out) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const size_t id = (size_t)orig[s];
        out[3 * id] = part[A_PX * stride + s];
        out[3 * id + 1] = part[A_PY * stride + s];
        out[3 * id + 2] = part[A_PZ * stride + s];
    }
}

// Streaming copy (measurement aid: calibrates HBM byte counters and gives the
// achievable copy bandwidth).
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// The same copy with four float4 per thread in flight (all loads issued
// before the stores) and nontemporal stores: a stronger estimate of the
// achievable streaming bandwidth.
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy_wide(const f32x4* __restrict__ src, f32x4* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const f32x4 v0 = src[i], v1 = src[i + stride], v2 = src[i + 2 * stride], v3 = src[i + 3 * stride];
        __builtin_nontemporal_store(v0, dst + i);
        __builtin_nontemporal_store(v1, dst + i + stride);
        __builtin_nontemporal_store(v2, dst + i + 2 * stride);
        __builtin_nontemporal_store(v3, dst + i + 3 * stride);
    }
    for (; i < n; i += stride) __builtin_nontemporal_store(src[i], dst + i);
}

}  // namespace

void launch_export_xyz(const float* part, size_t stride, const int* orig, int n, float* out_aos, hipStream_t s) {
    if (n <= 0) return;
    const int blocks = std::min((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_export_xyz, dim3(blocks), dim3(256), 0, s, part, stride, orig, n, out_aos);
}

void launch_copy_state(const float* src, size_t src_stride, const int* src_orig, float* dst, size_t dst_stride,
                       int* dst_orig, int n, int planes, hipStream_t s) {
    if (n <= 0) return;
    const int blocks = std::min((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_copy_state, dim3(blocks), dim3(256), 0, s, src, src_stride, src_orig, dst, dst_stride,
                       dst_orig, n, planes);
}

void launch_copy(const float4* src, float4* dst, size_t n, hipStream_t s, bool wide) {
    if (!n) return;
    if (wide)
        hipLaunchKernelGGL(k_copy_wide, dim3(256 * 8 * 4), dim3(256), 0, s, reinterpret_cast<const f32x4*>(src),
                           reinterpret_cast<f32x4*>(dst), n);
    else
        hipLaunchKernelGGL(k_copy, dim3(256 * 8 * 4), dim3(256), 0, s, src, dst, n);
}

void launch_grid_update(const DevCfg* cfgs, const uint16_t* block_scene, int nscenes, GridBuf m, GridBuf v_clear,
                        bool clear_v, const float* sdf, int nblocks, uint2* list, uint32_t* count,
                        uint32_t* next_count, hipStream_t s, const CorrGridOut* corr) {
    if (nblocks <= 0) return;
    if (!next_count) hipLaunchKernelGGL(k_set_u32_grid, dim3(1), dim3(1), 0, s, count, 0u);
    const int single = nscenes == 1 ? 1 : 0;
    hipLaunchKernelGGL(k_grid_list, dim3((nblocks + 255) / 256), dim3(256), 0, s, cfgs, block_scene, m.flags,
                       v_clear.flags, clear_v ? 1 : 0, nblocks, list, count, single);
    const int waves = std::min((nblocks + 3) / 4, 8192);
    if (corr)
        hipLaunchKernelGGL(k_grid_update_corr, dim3((waves + 3) / 4), dim3(256), 0, s, cfgs, block_scene, m, v_clear,
                           sdf, list, count, next_count, single, nblocks, corr->vcopy, corr->active, corr->acap,
                           corr->acount);
    else
        hipLaunchKernelGGL(k_grid_update, dim3((waves + 3) / 4), dim3(256), 0, s, cfgs, block_scene, m, v_clear, sdf,
                           list, count, next_count, single, nblocks);
}

void launch_make_sdf(const DevCfg& c, const float* tris, int nt, float* sdf, hipStream_t s) {
    const int G = c.sxy * c.sz;
    hipLaunchKernelGGL(k_make_sdf, dim3((G + 255) / 256), dim3(256), 0, s, c, tris, nt, sdf);
}

namespace {

// Median-split hierarchy over the triangles (host).  perm receives the
// triangle order of the leaves.
struct SdfBuilder {
    const float* tris;
    std::vector<int> perm;
    std::vector<SdfNode> nodes;
    std::vector<float> cx;  // centroids, 3 per triangle

    int build(int begin, int end) {
        const int id = (int)nodes.size();
        nodes.push_back(SdfNode{});
        SdfNode n{};
        for (int a = 0; a < 3; ++a) {
            n.lo[a] = 3.4e38f;
            n.hi[a] = -3.4e38f;
        }
        double an[3] = {0, 0, 0}, ac[3] = {0, 0, 0}, area = 0.0, mc[3] = {0, 0, 0};
        float clo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, chi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
        for (int k = begin; k < end; ++k) {
            const float* t = tris + 9 * (size_t)perm[k];
            for (int v = 0; v < 3; ++v)
                for (int a = 0; a < 3; ++a) {
                    n.lo[a] = std::min(n.lo[a], t[3 * v + a]);
                    n.hi[a] = std::max(n.hi[a], t[3 * v + a]);
                }
            const double e1[3] = {(double)t[3] - t[0], (double)t[4] - t[1], (double)t[5] - t[2]};
            const double e2[3] = {(double)t[6] - t[0], (double)t[7] - t[1], (double)t[8] - t[2]};
            const double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                  e1[0] * e2[1] - e1[1] * e2[0]};
            const double A = 0.5 * std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
            for (int a = 0; a < 3; ++a) {
                an[a] += 0.5 * cr[a];
                const float cc = cx[3 * (size_t)perm[k] + a];
                ac[a] += A * cc;
                mc[a] += cc;
                clo[a] = std::min(clo[a], cc);
                chi[a] = std::max(chi[a], cc);
            }
            area += A;
        }
        const int cnt = end - begin;
        for (int a = 0; a < 3; ++a) {
            n.an[a] = (float)an[a];
            n.cen[a] = (float)(area > 0.0 ? ac[a] / area : mc[a] / cnt);
        }
        double r2 = 0.0;
        for (int k = begin; k < end; ++k) {
            const float* t = tris + 9 * (size_t)perm[k];
            for (int v = 0; v < 3; ++v) {
                double d2 = 0.0;
                for (int a = 0; a < 3; ++a) {
                    const double d = (double)t[3 * v + a] - n.cen[a];
                    d2 += d * d;
                }
                r2 = std::max(r2, d2);
            }
        }
        // rounded up so that the float ball still holds every vertex
        n.rad = (float)(std::sqrt(r2) * (1.0 + 1e-6) + 1e-6);
        if (cnt <= 4) {
            n.first = begin;
            n.count = cnt;
            n.skip = id + 1;
            nodes[id] = n;
            return id;
        }
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
        const int mid = begin + cnt / 2;
        std::nth_element(perm.begin() + begin, perm.begin() + mid, perm.begin() + end, [&](int p, int q) {
            const float a = cx[3 * (size_t)p + axis], b = cx[3 * (size_t)q + axis];
            return a < b || (a == b && p < q);
        });
        n.first = -1;
        n.count = 0;
        nodes[id] = n;
        build(begin, mid);
        build(mid, end);
        nodes[id].skip = (int)nodes.size();
        return id;
    }
};

}  // namespace

int launch_make_sdf_bvh(const DevCfg& c, const float* tris_dev, const float* tris_host, int nt, float* sdf,
                        unsigned int* exact_count, hipStream_t s) {
    SdfBuilder b;
    b.tris = tris_host;
    b.perm.resize((size_t)nt);
    b.cx.resize((size_t)3 * nt);
    for (int t = 0; t < nt; ++t) {
        b.perm[t] = t;
        for (int a = 0; a < 3; ++a)
            b.cx[3 * (size_t)t + a] = (tris_host[9 * (size_t)t + a] + tris_host[9 * (size_t)t + 3 + a] +
                                       tris_host[9 * (size_t)t + 6 + a]) / 3.0f;
    }
    b.nodes.reserve((size_t)2 * nt);
    b.build(0, nt);
    std::vector<float> pt((size_t)9 * nt);
    for (int k = 0; k < nt; ++k)
        std::copy(tris_host + 9 * (size_t)b.perm[k], tris_host + 9 * (size_t)b.perm[k] + 9, pt.begin() + 9 * (size_t)k);
    SdfNode* dn = nullptr;
    float* dp = nullptr;
    unsigned int* dc = nullptr;
    if (hipMalloc((void**)&dn, b.nodes.size() * sizeof(SdfNode)) != hipSuccess) return -1;
    if (hipMalloc((void**)&dp, pt.size() * sizeof(float)) != hipSuccess) {
        (void)hipFree(dn);
        return -1;
    }
    if (hipMalloc((void**)&dc, sizeof(unsigned int)) != hipSuccess) {
        (void)hipFree(dn);
        (void)hipFree(dp);
        return -1;
    }
    hipError_t e = hipMemcpyAsync(dn, b.nodes.data(), b.nodes.size() * sizeof(SdfNode), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dp, pt.data(), pt.size() * sizeof(float), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(dc, 0, sizeof(unsigned int), s);
    const int G = c.sxy * c.sz;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_make_sdf_bvh, dim3((G + 255) / 256), dim3(256), 0, s, c, dn, (int)b.nodes.size(), dp,
                           tris_dev, nt, sdf, dc);
        e = hipGetLastError();
    }
    unsigned int cnt = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&cnt, dc, sizeof(unsigned int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dn);
    (void)hipFree(dp);
    (void)hipFree(dc);
    if (exact_count) *exact_count = cnt;
    return e == hipSuccess ? 0 : -1;
}

void launch_fill_f32(float* p, size_t n, float v, hipStream_t s) {
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

void launch_fill_u8(uint8_t* p, size_t n, uint8_t v, hipStream_t s) {
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

void launch_init_F(float* soa26, int n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_init_F, dim3((n + 255) / 256), dim3(256), 0, s, soa26, n);
}

void launch_export_soa26(const float* part, size_t stride, const int* orig, int n, const SceneTable& t, float* out26,
                         hipStream_t s) {
    if (n <= 0) return;
    const int blocks = std::min((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_export_soa26, dim3(blocks), dim3(256), 0, s, part, stride, orig, n, t, out26);
}

void launch_export_positions(const float* part, size_t stride, const int* orig, int n, float* out_planes3,
                             hipStream_t s) {
    if (n <= 0) return;
    const int blocks = std::min((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_export_positions, dim3(blocks), dim3(256), 0, s, part, stride, orig, n, out_planes3);
}

}  // namespace mpm
