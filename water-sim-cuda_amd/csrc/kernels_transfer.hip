// kernels_transfer.hip -- particle<->grid transfer kernels.
//
// One single-wave workgroup per chunk (<= CHUNK particles of one 8 x 4 x 4
// block of base cells).  The wave stages the grid velocity of its block's
// neighbourhood in an LDS tile (G2P reads) and accumulates its P2G
// contributions in an LDS arena of the same footprint, which it flushes with
// one global float atomic per node and channel.  Particles whose stencil
// leaves the tile (their base cell moved out of their block since the last
// re-sort, or they sit far outside the grid) fall back to global
// loads/atomics.
//
// Arena accumulation.  FP32 LDS atomics (ds_add_f32) run at about a third of
// a lane per clock per CU on gfx950, some 40x below integer LDS atomics, so
// the arena is updated with plain float4 read-modify-writes instead.  The
// arena belongs to one wave; the 27 stencil nodes are visited in program
// order and LDS executes one wave's operations in order, so a later offset
// always sees an earlier offset's write.  Within one offset, lanes must hit
// distinct nodes: of the lanes whose particles share a base node (after
// drifting since the last re-sort) one wins a claim byte and the others use
// LDS float atomics.  The sums are the same float additions the reference
// performs, in a different order.
//
// Three modes share the code:
//   XFER_P2G   : p2GTransferScatter, kernels.cu:6-142
//   XFER_G2P   : g2PTransfer, kernels.cu:199-280
//   XFER_G2P2G : G2P of substep s followed, in registers, by P2G of substep
//                s+1 into the other grid buffer.  Between substeps only
//                position and F then travel through HBM; v and C stay in
//                registers (they are written out by the last transfer of a
//                steps() call).
//   XFER_G2P2G_FULL : the same, also writing v and C: the last transfer of a
//                steps() call, after which the particles are complete and the
//                next call starts from the P2G already in the grid.
#include "internal.hpp"

namespace mpm {
namespace {

struct Particle {
    float x, y, z;
    float vx, vy, vz;
    float C[9];
    float F[9];
};

__device__ __forceinline__ void weights(float p, int b, float w[3], float d[3]) {
    // fr = p - b is exact, and for fr in [0.5, 1.5) each of the three spline
    // branches is fixed, so the closed forms below equal spline(p - (b + i))
    // bit for bit (every difference involved is exact).  Otherwise (particles
    // below 0.5 or far outside the grid) the generic spline is evaluated.
    const float fr = p - (float)b;
    d[0] = -fr;
    d[1] = -(fr - 1.0f);
    d[2] = -(fr - 2.0f);
    if (fr >= 0.5f && fr < 1.5f) {
#pragma clang fp contract(off)
        const float t0 = 1.5f - fr, t1 = fr - 1.0f, t2 = fr - 0.5f;
        w[0] = 0.5f * (t0 * t0);
        w[1] = 0.75f - t1 * t1;
        w[2] = 0.5f * (t2 * t2);
    } else {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float r = p - (float)(b + i);  // pPos - (pos + i), kernels.cu:54, :211
            w[i] = spline(r);
            d[i] = -r;                           // dist = node - pPos
        }
    }
}

// ---- G2P (kernels.cu:199-280) ----
// vel = sum w v and C = 4 sum w v (x_i - x_p)^T over the 27 nodes.  The
// sums are grouped by axis (z innermost, then y, then x).  Per axis the node
// offset is x_i - x_p = o - f with o = 0, 1, 2 and f = -d[0], so a sum
// sum_o w_o (o - f) v_o is w_1 v_1 + 2 w_2 v_2 - f sum_o w_o v_o: the o terms
// take two products instead of three, and the -f vel terms are added once at
// the end (about 250 multiply-adds; 650 node by node).  The additions happen
// in a different order than the reference's.
template <class Fetch>
__device__ __forceinline__ void g2p_gather(Fetch fetch, const float wx[3], const float wy[3], const float wz[3],
                                           const float dx[3], const float dy[3], const float dz[3], float vel[3],
                                           float C[9]) {
    const float wz2 = 2.0f * wz[2], wy2 = 2.0f * wy[2], wx2 = 2.0f * wx[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float S[3] = {0.f, 0.f, 0.f}, Y[3] = {0.f, 0.f, 0.f}, T[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float3 v0 = fetch(i, j, 0), v1 = fetch(i, j, 1), v2 = fetch(i, j, 2);
            // s = sum_k wz v, t = sum_k wz k v
            const float s0 = wz[0] * v0.x + wz[1] * v1.x + wz[2] * v2.x;
            const float s1 = wz[0] * v0.y + wz[1] * v1.y + wz[2] * v2.y;
            const float s2 = wz[0] * v0.z + wz[1] * v1.z + wz[2] * v2.z;
            const float t0 = wz[1] * v1.x + wz2 * v2.x;
            const float t1 = wz[1] * v1.y + wz2 * v2.y;
            const float t2 = wz[1] * v1.z + wz2 * v2.z;
            S[0] += wy[j] * s0; S[1] += wy[j] * s1; S[2] += wy[j] * s2;
            T[0] += wy[j] * t0; T[1] += wy[j] * t1; T[2] += wy[j] * t2;
            if (j > 0) {
                const float wyo = j == 1 ? wy[1] : wy2;  // wy_j j
                Y[0] += wyo * s0; Y[1] += wyo * s1; Y[2] += wyo * s2;
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            vel[a] += wx[i] * S[a];
            if (i > 0) C[3 * a + 0] += (i == 1 ? wx[1] : wx2) * S[a];
            C[3 * a + 1] += wx[i] * Y[a];
            C[3 * a + 2] += wx[i] * T[a];
        }
    }
    const float fx = -dx[0], fy = -dy[0], fz = -dz[0];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        C[3 * a + 0] = 4.0f * (C[3 * a + 0] - fx * vel[a]);
        C[3 * a + 1] = 4.0f * (C[3 * a + 1] - fy * vel[a]);
        C[3 * a + 2] = 4.0f * (C[3 * a + 2] - fz * vel[a]);
    }
}

// The same sums with the 27 node velocities delivered nine at a time, one
// x slice (j, k = 0..2) per call of `slice(i, vx, vy, vz)`; used where some
// nodes come from global memory, so that a slice's loads are all issued
// before the first of them is used (one memory latency per slice, not one per
// node).
template <class Slice>
__device__ __forceinline__ void g2p_gather_slices(Slice slice, const float wx[3], const float wy[3],
                                                  const float wz[3], const float dx[3], const float dy[3],
                                                  const float dz[3], float vel[3], float C[9]) {
    float wxd[3], wyd[3], wzd[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        wxd[i] = wx[i] * dx[i];
        wyd[i] = wy[i] * dy[i];
        wzd[i] = wz[i] * dz[i];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float vx[9], vy[9], vz[9];
        slice(i, vx, vy, vz);
        float S[3] = {0.f, 0.f, 0.f}, Y[3] = {0.f, 0.f, 0.f}, T[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float s[3] = {0.f, 0.f, 0.f}, t[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int n = 3 * j + k;
                s[0] += wz[k] * vx[n]; s[1] += wz[k] * vy[n]; s[2] += wz[k] * vz[n];
                t[0] += wzd[k] * vx[n]; t[1] += wzd[k] * vy[n]; t[2] += wzd[k] * vz[n];
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                S[a] += wy[j] * s[a];
                Y[a] += wyd[j] * s[a];
                T[a] += wy[j] * t[a];
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            vel[a] += wx[i] * S[a];
            C[3 * a + 0] += wxd[i] * S[a];
            C[3 * a + 1] += wx[i] * Y[a];
            C[3 * a + 2] += wx[i] * T[a];
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) C[i] *= 4.0f;
}

// getSDF at p from the .w lanes of the LDS tile when the sample's 2x2x2 cell
// lies inside the tile and needs no clamping (then both read the same node
// values); otherwise from the global SDF.
__device__ __forceinline__ float sdf_sample(const DevCfg& c, const float* __restrict__ sdf, const float4* vt, int ox,
                                            int oy, int oz, float px, float py, float pz) {
    const int i = (int)floorf(px), j = (int)floorf(py), k = (int)floorf(pz);
    const int lx = i - ox, ly = j - oy, lz = k - oz;
    if ((unsigned)lx <= (unsigned)(TILE_DX - 2) && (unsigned)ly <= (unsigned)(TILE_DY - 2) &&
        (unsigned)lz <= (unsigned)(TILE_DZ - 2) && i >= 0 && j >= 0 && k >= 0 && i + 1 < c.sx && j + 1 < c.sy &&
        k + 1 < c.sz) {
        const float4* b = vt + (lx + TILE_PY * ly + TILE_PZ * lz);
        return trilerp(b[0].w, b[1].w, b[TILE_PY].w, b[TILE_PY + 1].w, b[TILE_PZ].w, b[TILE_PZ + 1].w,
                       b[TILE_PZ + TILE_PY].w, b[TILE_PZ + TILE_PY + 1].w, px - (float)i, py - (float)j,
                       pz - (float)k);
    }
    return sdf_at(c, sdf, px, py, pz);
}

// calculateNormal (common.cuh:169-182) with each of the six samples taken
// from the LDS tile when its cell lies there (same node values, same
// arithmetic as the global path).
__device__ __forceinline__ void sdf_normal_tile(const DevCfg& c, const float* __restrict__ sdf, const float4* vt,
                                                int ox, int oy, int oz, float px, float py, float pz, float& nx,
                                                float& ny, float& nz) {
    const float eps = 0.1f;
    nx = sdf_sample(c, sdf, vt, ox, oy, oz, px + eps, py, pz) - sdf_sample(c, sdf, vt, ox, oy, oz, px - eps, py, pz);
    ny = sdf_sample(c, sdf, vt, ox, oy, oz, px, py + eps, pz) - sdf_sample(c, sdf, vt, ox, oy, oz, px, py - eps, pz);
    nz = sdf_sample(c, sdf, vt, ox, oy, oz, px, py, pz + eps) - sdf_sample(c, sdf, vt, ox, oy, oz, px, py, pz - eps);
    const float len = sqrtf(nx * nx + ny * ny + nz * nz) + 1e-6f;
    nx /= len; ny /= len; nz /= len;
}

// F update, J floor, advection and collider projection (kernels.cu:237-279).
__device__ __forceinline__ void g2p_finish(const DevCfg& c, const float* __restrict__ sdf, const float4* vt, int ox,
                                           int oy, int oz, Particle& p, const float vel_in[3], const float C[9]) {
    float A[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) A[i] = c.dt * C[i];
    A[0] = 1.0f + A[0]; A[4] = 1.0f + A[4]; A[8] = 1.0f + A[8];
    float nF[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q)
            nF[3 * r + q] = A[3 * r] * p.F[q] + A[3 * r + 1] * p.F[3 + q] + A[3 * r + 2] * p.F[6 + q];
    const float J = det3(nF);
    if (J < 0.3f) {
        const float scale = powf(0.3f / J, 1.0f / 3.0f);
#pragma unroll
        for (int i = 0; i < 9; ++i) nF[i] *= scale;
    }
    float vx = vel_in[0], vy = vel_in[1], vz = vel_in[2];
    float nx_ = p.x + vx * c.dt, ny_ = p.y + vy * c.dt, nz_ = p.z + vz * c.dt;
    float dist = sdf_sample(c, sdf, vt, ox, oy, oz, nx_, ny_, nz_);
    if (dist < 0.0f) {
        float nx, ny, nz;
        sdf_normal_tile(c, sdf, vt, ox, oy, oz, nx_, ny_, nz_, nx, ny, nz);
        nx_ -= nx * dist; ny_ -= ny * dist; nz_ -= nz * dist;
        float vn = vx * nx + vy * ny + vz * nz;
        if (vn < 0.0f) {
            vx -= nx * vn; vy -= ny * vn; vz -= nz * vn;
        }
    }
    p.x = nx_; p.y = ny_; p.z = nz_;
    p.vx = vx; p.vy = vy; p.vz = vz;
#pragma unroll
    for (int i = 0; i < 9; ++i) p.C[i] = C[i];
#pragma unroll
    for (int i = 0; i < 9; ++i) p.F[i] = nF[i];
}

// ---- P2G (kernels.cu:14-119) ----
// Per-particle quantities: pressure-augmented affine matrix (kernels.cu:26-35),
// weights and node offsets.
struct P2GPrep {
    float C[9];
    float wx[3], wy[3], wz[3], dx[3], dy[3], dz[3];
};

__device__ __forceinline__ void p2g_prepare(const DevCfg& c, const Particle& p, int bx, int by, int bz, P2GPrep& q) {
    float J = det3(p.F);
    J = fmaxf(0.9f, fminf(J, 1.1f));
    // J^GAMMA for J in [0.9, 1.1] as 2^(GAMMA log2 J) on the hardware log and
    // exp units (4 instructions instead of a general powf's ~150); the
    // relative error is a few 1e-7, far inside the step's tolerance.
    float pressure = c.compression * (__builtin_amdgcn_exp2f(c.gamma * __builtin_amdgcn_logf(J)) - 1.0f);
    if (pressure < 0.0f) pressure = 0.0f;
    float volume = c.vol * J;
    float multi = 4.0f * c.dt * volume * pressure / c.mass;
#pragma unroll
    for (int i = 0; i < 9; ++i) q.C[i] = p.C[i];
    q.C[0] += multi; q.C[4] += multi; q.C[8] += multi;
    weights(p.x, bx, q.wx, q.dx);
    weights(p.y, by, q.wy, q.dy);
    weights(p.z, bz, q.wz, q.dz);
}

// Node contributions m w (v + C' (x_i - x_p)) in the reference's loop order
// (x outer, z inner).  With w = wx wy and e = v + C' (dx, dy, 0) per (x, y)
// column, mz = m wz and g = m wz dz C'[:, 2] per z offset, a node receives
// mass w mz and momentum w (mz e + g).  `scatter(i, j, k, n)` gets a Node
// that either returns that contribution (value(), for atomics) or adds it to
// a float4 with fused multiply-adds (add_to(), 7 instructions for the 4
// components).
struct P2GNode {
    float w, mz, e0, e1, e2, g0, g1, g2;
    __device__ __forceinline__ float4 value() const {
        return make_float4(w * mz, w * fmaf(mz, e0, g0), w * fmaf(mz, e1, g1), w * fmaf(mz, e2, g2));
    }
    __device__ __forceinline__ void add_to(float4& v) const {
        v.x = fmaf(w, mz, v.x);
        v.y = fmaf(w, fmaf(mz, e0, g0), v.y);
        v.z = fmaf(w, fmaf(mz, e1, g1), v.z);
        v.w = fmaf(w, fmaf(mz, e2, g2), v.w);
    }
};

template <class Scatter>
__device__ __forceinline__ void p2g_scatter(const DevCfg& c, const Particle& p, const P2GPrep& q, Scatter scatter) {
    float mz[3], g0[3], g1[3], g2[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        mz[k] = c.mass * q.wz[k];
        const float mzd = mz[k] * q.dz[k];
        g0[k] = mzd * q.C[2];
        g1[k] = mzd * q.C[5];
        g2[k] = mzd * q.C[8];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float f0 = p.vx + q.C[0] * q.dx[i], f1 = p.vy + q.C[3] * q.dx[i], f2 = p.vz + q.C[6] * q.dx[i];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float w = q.wx[i] * q.wy[j];
            const float e0 = f0 + q.C[1] * q.dy[j], e1 = f1 + q.C[4] * q.dy[j], e2 = f2 + q.C[7] * q.dy[j];
#pragma unroll
            for (int k = 0; k < 3; ++k) scatter(i, j, k, P2GNode{w, mz[k], e0, e1, e2, g0[k], g1[k], g2[k]});
        }
    }
}

__device__ __forceinline__ void load_particle(const float* __restrict__ part, size_t stride, uint32_t slot, Particle& p,
                                              bool full) {
    p.x = (part + (size_t)A_PX * stride)[slot];
    p.y = (part + (size_t)A_PY * stride)[slot];
    p.z = (part + (size_t)A_PZ * stride)[slot];
#pragma unroll
    for (int i = 0; i < 9; ++i) p.F[i] = (part + (size_t)(A_F0 + i) * stride)[slot];
    if (full) {
        p.vx = (part + (size_t)A_VX * stride)[slot];
        p.vy = (part + (size_t)A_VY * stride)[slot];
        p.vz = (part + (size_t)A_VZ * stride)[slot];
#pragma unroll
        for (int i = 0; i < 9; ++i) p.C[i] = (part + (size_t)(A_C0 + i) * stride)[slot];
    }
}

// A node contribution added to an arena slot with LDS float atomics
// (ds_add_f32), for the few lanes that cannot use the read-modify-write chain.
__device__ __forceinline__ void lds_add4(float4* slot, const float4 v) {
    float* n = reinterpret_cast<float*>(slot);
    atomicAdd(n + 0, v.x);
    atomicAdd(n + 1, v.y);
    atomicAdd(n + 2, v.z);
    atomicAdd(n + 3, v.w);
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }

// The value of lane l ^ 1 (a DPP quad permutation [1, 0, 3, 2]); every lane
// must be active.
__device__ __forceinline__ int swap_adjacent(int v) { return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false); }
__device__ __forceinline__ float swap_adjacent(float v) {
    return __int_as_float(swap_adjacent(__float_as_int(v)));
}

__device__ __forceinline__ bool node_in_tile(int tx, int ty, int tz) {
    return (unsigned)tx < (unsigned)TILE_DX && (unsigned)ty < (unsigned)TILE_DY && (unsigned)tz < (unsigned)TILE_DZ;
}

__device__ __forceinline__ bool stencil_in_tile(int lx, int ly, int lz) {
    return (unsigned)lx <= (unsigned)(TILE_DX - 3) && (unsigned)ly <= (unsigned)(TILE_DY - 3) &&
           (unsigned)lz <= (unsigned)(TILE_DZ - 3);
}

// GATHER: the particles are read through a.map from a.src (the transfers
// after a re-sort that left the planes in place, up to the first G2P).
template <int MODE, bool GATHER>
// At most 168 VGPRs: three waves per SIMD (LDS allows about 13 chunks per CU).
__global__ __launch_bounds__(TRANSFER_THREADS, 3) void k_transfer(TransferArgs a) {
    constexpr bool DO_G2P = (MODE & XFER_G2P) != 0;
    constexpr bool DO_P2G = (MODE & XFER_P2G) != 0;
    constexpr bool WRITE_VC = !DO_P2G || MODE == XFER_G2P2G_FULL;  // G2P also stores v and C
    // Workgroups are dispatched to the 8 XCDs round-robin; chunk ranges are
    // split into 8 contiguous parts so that neighbouring chunks (which share
    // tile nodes) run on the same XCD and share its L2.
    const uint32_t nch = *a.nchunks;
    const int per = (int)((nch + 7u) >> 3);
    // Diagnostic counts (particles outside their tile, lost claims), kept
    // wave-uniform and added to the device counters once per wave: one global
    // atomic per lane on a single address serialises at the memory side.
    uint32_t n_outside = 0u, n_lost = 0u;
    const int xcd = blockIdx.x & 7, nl = gridDim.x >> 3;
    for (int t = blockIdx.x >> 3; t < per; t += nl) {
    const int ci = __builtin_amdgcn_readfirstlane(xcd * per + t);
    if (ci >= (int)nch) break;
    // The chunk's descriptor and its scene's config are wave-uniform: moved
    // to scalar registers once per chunk, they take no vector registers.
    const ChunkDesc d = uniform(a.chunks[ci]);
    const DevCfg c = uniform(a.cfgs[d.scene]);

    __shared__ float4 vt[DO_G2P ? TILE_SLOTS : 1];
    __shared__ float4 ar[DO_P2G ? TILE_SLOTS : 1];  // mass, momentum x, y, z
    __shared__ uint8_t claim[DO_P2G ? TILE_SLOTS : 1];

    const int lane = threadIdx.x;
    if constexpr (DO_G2P) {
        // All of the tile's loads are issued before any node is stored (one
        // memory latency per chunk, not one per node of a lane): each node is
        // read unconditionally at coordinates clamped into the grid (a guarded
        // load would be waited for before the next one is issued), and nodes
        // outside the grid are set to zero afterwards.  The node indices pass
        // through an empty asm so that the compiler does not hoist their
        // (lane-only) coordinates out of the chunk loop, which would hold them
        // in registers through the rounds.
        constexpr int TL_ITEMS = (TILE_N + TRANSFER_THREADS - 1) / TRANSFER_THREADS;
        {
            float4 tv[TL_ITEMS];
#pragma unroll
            for (int k = 0; k < TL_ITEMS; ++k) {
                int i = min(lane + k * TRANSFER_THREADS, TILE_N - 1);
                asm volatile("" : "+v"(i));
                int lx, ly, lz;
                tile_coords(i, lx, ly, lz);
                const int x = d.ox + lx, y = d.oy + ly, z = d.oz + lz;
                const int g =
                    gidx(c, min(max(x, 0), c.sx - 1), min(max(y, 0), c.sy - 1), min(max(z, 0), c.sz - 1));
                tv[k] = make_float4(a.vel[0][g], a.vel[1][g], a.vel[2][g], a.sdf[g]);
            }
#pragma unroll
            for (int k = 0; k < TL_ITEMS; ++k) {
                int i = lane + k * TRANSFER_THREADS;
                asm volatile("" : "+v"(i));
                if (i < TILE_N) {
                    int lx, ly, lz;
                    tile_coords(i, lx, ly, lz);
                    const bool in = inb(c, d.ox + lx, d.oy + ly, d.oz + lz);
                    vt[lx + TILE_PY * ly + TILE_PZ * lz] = in ? tv[k] : make_float4(0.f, 0.f, 0.f, 0.f);
                }
            }
        }
    }
    if constexpr (DO_P2G) {
        for (int i = lane; i < TILE_SLOTS; i += TRANSFER_THREADS) ar[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();

    const int S = d.S;
    // Round j is the n_j particles in slots [start + off_j, start + off_j + n_j):
    // n_j = nl for j < full and nl - 1 after (the wrap-around layout).  The
    // next round's particle is loaded while the current one is processed,
    // hiding HBM latency; with GATHER the map entries are loaded one round
    // earlier still, so the particle loads never wait for them.
    const int nl = (d.count + S - 1) / S;
    const int full = d.count - (nl - 1) * S;
    auto round_n = [&](int j) { return j < S ? (j < full ? nl : nl - 1) : 0; };
    int nxt_n = round_n(0);
    int nxt_off = 0;
    Particle nxt;
    bool nxt_valid = lane < nxt_n;
    uint32_t nxt_slot = (uint32_t)(d.start + lane);
    uint32_t nn_src = 0u;  // GATHER: old slot of this lane's particle in the round after the next
    if constexpr (GATHER) {
        if (lane < round_n(1)) nn_src = a.map[d.start + nxt_n + lane];
        if (nxt_valid) load_particle(a.src, a.src_stride, a.map[nxt_slot], nxt, !DO_G2P);
    } else {
        if (nxt_valid) load_particle(a.part, a.stride, nxt_slot, nxt, !DO_G2P);
    }
    for (int j = 0; j < S; ++j) {
        const bool valid = nxt_valid;
        const uint32_t slot = nxt_slot;
        Particle p = nxt;
        nxt_off += nxt_n;
        nxt_n = round_n(j + 1);
        nxt_valid = lane < nxt_n;
        nxt_slot = (uint32_t)(d.start + nxt_off + lane);
        if constexpr (GATHER) {
            const uint32_t si = nn_src;
            if (lane < round_n(j + 2)) nn_src = a.map[d.start + nxt_off + nxt_n + lane];
            if (nxt_valid) load_particle(a.src, a.src_stride, si, nxt, !DO_G2P);
        } else {
            if (nxt_valid) load_particle(a.part, a.stride, nxt_slot, nxt, !DO_G2P);
        }
        if (valid) {
            if constexpr (DO_G2P) {
                const int bx = base_of(p.x), by = base_of(p.y), bz = base_of(p.z);
                float wx[3], wy[3], wz[3], dx[3], dy[3], dz[3];
                weights(p.x, bx, wx, dx);
                weights(p.y, by, wy, dy);
                weights(p.z, bz, wz, dz);
                float vel[3] = {0.f, 0.f, 0.f};
                float C[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                const int lx = bx - d.ox, ly = by - d.oy, lz = bz - d.oz;
                if (stencil_in_tile(lx, ly, lz)) {
                    const float4* base = vt + (lx + TILE_PY * ly + TILE_PZ * lz);
                    g2p_gather(
                        [&](int i, int jj, int k) {
                            float4 v = base[i + TILE_PY * jj + TILE_PZ * k];
                            return make_float3(v.x, v.y, v.z);
                        },
                        wx, wy, wz, dx, dy, dz, vel, C);
                } else {
                    // stencil partly outside the tile: nodes inside it from the
                    // tile (the same values), the others from global memory.
                    // A slice's nine nodes are loaded unconditionally at
                    // coordinates clamped into the grid, so the 27 loads are in
                    // flight together; nodes in the tile or outside the grid
                    // are then taken from the tile or set to zero.
                    g2p_gather_slices(
                        [&](int i, float vx[9], float vy[9], float vz[9]) {
                            const int x = bx + i;
                            const int xc = min(max(x, 0), c.sx - 1);
#pragma unroll
                            for (int n = 0; n < 9; ++n) {
                                const int y = by + n / 3, z = bz + n % 3;
                                const int g = gidx(c, xc, min(max(y, 0), c.sy - 1), min(max(z, 0), c.sz - 1));
                                vx[n] = a.vel[0][g];
                                vy[n] = a.vel[1][g];
                                vz[n] = a.vel[2][g];
                            }
#pragma unroll
                            for (int n = 0; n < 9; ++n) {
                                const int y = by + n / 3, z = bz + n % 3;
                                const int tx = x - d.ox, ty = y - d.oy, tz = z - d.oz;
                                if (node_in_tile(tx, ty, tz)) {
                                    const float4 v = vt[tx + TILE_PY * ty + TILE_PZ * tz];
                                    vx[n] = v.x; vy[n] = v.y; vz[n] = v.z;
                                } else if (!inb(c, x, y, z)) {
                                    vx[n] = 0.f; vy[n] = 0.f; vz[n] = 0.f;
                                }
                            }
                        },
                        wx, wy, wz, dx, dy, dz, vel, C);
                }
                g2p_finish(c, a.sdf, vt, d.ox, d.oy, d.oz, p, vel, C);
                (a.part + (size_t)A_PX * a.stride)[slot] = p.x;
                (a.part + (size_t)A_PY * a.stride)[slot] = p.y;
                (a.part + (size_t)A_PZ * a.stride)[slot] = p.z;
#pragma unroll
                for (int i = 0; i < 9; ++i) (a.part + (size_t)(A_F0 + i) * a.stride)[slot] = p.F[i];
                if constexpr (WRITE_VC) {
                    (a.part + (size_t)A_VX * a.stride)[slot] = p.vx;
                    (a.part + (size_t)A_VY * a.stride)[slot] = p.vy;
                    (a.part + (size_t)A_VZ * a.stride)[slot] = p.vz;
#pragma unroll
                    for (int i = 0; i < 9; ++i) (a.part + (size_t)(A_C0 + i) * a.stride)[slot] = p.C[i];
                }
            }
        }

        if constexpr (DO_P2G) {
            int bx = 0, by = 0, bz = 0;
            bool in_tile = false;
            uint32_t li = 0;
            P2GPrep pq;
            if (valid) {
                bx = base_of(p.x); by = base_of(p.y); bz = base_of(p.z);
                const int lx = bx - d.ox, ly = by - d.oy, lz = bz - d.oz;
                in_tile = stencil_in_tile(lx, ly, lz);
                li = in_tile ? (uint32_t)(lx + TILE_PY * ly + TILE_PZ * lz) : 0u;
                p2g_prepare(c, p, bx, by, bz, pq);
                if (!in_tile) {
                    // stencil partly outside the tile: LDS float atomics for
                    // the nodes inside it (before any lane's read-modify-write
                    // chain below, in program order), global atomics for the
                    // others
                    p2g_scatter(c, p, pq, [&](int i, int jj, int k, const P2GNode& nd) {
                        const int x = bx + i, y = by + jj, z = bz + k;
                        const int tx = x - d.ox, ty = y - d.oy, tz = z - d.oz;
                        const float4 v = nd.value();
                        if (node_in_tile(tx, ty, tz)) {
                            lds_add4(ar + (tx + TILE_PY * ty + TILE_PZ * tz), v);
                            return;
                        }
                        if (!inb(c, x, y, z)) return;
                        const int g = gidx(c, x, y, z);
                        atomic_add_global(a.acc[0] + g, v.x);
                        atomic_add_global(a.acc[1] + g, v.y);
                        atomic_add_global(a.acc[2] + g, v.z);
                        atomic_add_global(a.acc[3] + g, v.w);
                        a.acc_flags[flag_block(c, x, y, z)] = 1;
                    });
                }
            }
            // Lanes with distinct base nodes update the arena together: each
            // lane writes its id into the claim byte of its base slot and
            // reads it back; one writer per slot wins (LDS executes the
            // wave's write before its read) and runs the read-modify-write
            // chain.  A round holds distinct base cells when the chunk is
            // laid out, so only particles whose base cell changed since the
            // last re-sort lose; they add their contributions with LDS float
            // atomics afterwards (slow per lane, but one or two lanes need no
            // 27-step dependent chain).  Their inputs are laundered so that
            // the compiler does not share the two paths' contributions (it
            // would compute them all above the claim test).
            auto arena_chain = [&](const Particle& pp, const P2GPrep& qq) {
                float4* base = ar + li;
                p2g_scatter(c, pp, qq, [&](int i, int jj, int k, const P2GNode& nd) {
                    float4* n = base + (i + TILE_PY * jj + TILE_PZ * k);
                    float4 v = *n;
                    nd.add_to(v);
                    *n = v;
                    asm volatile("" ::: "memory");  // keep the 27 updates in program order
                });
            };
            n_outside += (uint32_t)__popcll(__ballot(valid && !in_tile));
            // Same-cell pairs: in a paired chunk lanes 2k and 2k + 1 can hold
            // two members of one base cell (k_reorder); the high lane's
            // contributions reach the low lane through a cross-lane move and
            // the low lane's chain adds both, so the high lane takes no part
            // in the claims.
            bool lo = false, hi = false;
            if (d.paired) {
                // the cross-lane moves run with every lane active: no
                // short-circuit puts them under a branch
                const int key = in_tile ? (int)li : -1 - lane;
                const int other = swap_adjacent(key);
                const bool same = in_tile & (other == key);
                lo = same & ((lane & 1) == 0);
                hi = same & ((lane & 1) != 0);
            }
            const bool any_pair = __ballot(lo) != 0ull;
            const bool rmw = in_tile && !hi;
            if (rmw) claim[li] = (uint8_t)lane;
            asm volatile("" ::: "memory");
            const bool mine = rmw && claim[li] == (uint8_t)lane;
            asm volatile("" ::: "memory");
            bool covered = mine;
            if (any_pair) {
                // a pair's high lane is covered when its low lane won the claim
                const int partner_mine = swap_adjacent(mine ? 1 : 0);
                covered = mine | (hi & (partner_mine != 0));
                float4* base = ar + li;
                p2g_scatter(c, p, pq, [&](int i, int jj, int k, const P2GNode& nd) {
                    float4 v = nd.value();
                    const float4 o = make_float4(swap_adjacent(v.x), swap_adjacent(v.y), swap_adjacent(v.z),
                                                 swap_adjacent(v.w));
                    if (mine) {
                        if (lo) {
                            v.x += o.x; v.y += o.y; v.z += o.z; v.w += o.w;
                        }
                        float4* n = base + (i + TILE_PY * jj + TILE_PZ * k);
                        float4 u = *n;
                        u.x += v.x; u.y += v.y; u.z += v.z; u.w += v.w;
                        *n = u;
                    }
                    asm volatile("" ::: "memory");  // keep the 27 updates in program order
                });
            } else if (mine) {
                arena_chain(p, pq);
            }
            const bool lost = in_tile && !covered;
            n_lost += (uint32_t)__popcll(__ballot(lost));
            if (lost) {
                Particle p2 = p;
                P2GPrep q2 = pq;
                asm volatile("" : "+v"(p2.vx), "+v"(p2.vy), "+v"(p2.vz));
#pragma unroll
                for (int t = 0; t < 9; ++t) asm volatile("" : "+v"(q2.C[t]));
#pragma unroll
                for (int t = 0; t < 3; ++t)
                    asm volatile("" : "+v"(q2.wx[t]), "+v"(q2.wy[t]), "+v"(q2.wz[t]), "+v"(q2.dx[t]),
                                 "+v"(q2.dy[t]), "+v"(q2.dz[t]));
                float4* base = ar + li;
                p2g_scatter(c, p2, q2, [&](int i, int jj, int k, const P2GNode& nd) {
                    lds_add4(base + (i + TILE_PY * jj + TILE_PZ * k), nd.value());
                });
            }
        }
    }

    if constexpr (DO_P2G) {
        __syncthreads();
        for (int i = lane; i < TILE_N; i += TRANSFER_THREADS) {
            int lx, ly, lz;
            tile_coords(i, lx, ly, lz);
            const float4 v = ar[lx + TILE_PY * ly + TILE_PZ * lz];
            if (v.x > 1e-9f) {  // the reference's tile flush threshold, kernels.cu:132
                int x = d.ox + lx, y = d.oy + ly, z = d.oz + lz;
                if (inb(c, x, y, z)) {
                    int g = gidx(c, x, y, z);
                    atomic_add_global(a.acc[0] + g, v.x);
                    atomic_add_global(a.acc[1] + g, v.y);
                    atomic_add_global(a.acc[2] + g, v.z);
                    atomic_add_global(a.acc[3] + g, v.w);
                    a.acc_flags[flag_block(c, x, y, z)] = 1;
                }
            }
        }
        __syncthreads();  // the next chunk reuses the tile and the arena
    }
    }
    if constexpr (DO_P2G) {
        if (lane_id() == 0) {
            if (n_outside) atomicAdd(a.overflow, n_outside);
            if (n_lost) atomicAdd(a.overflow + 1, n_lost);
        }
    }
}

}  // namespace

void launch_transfer(int mode, const TransferArgs& a, int grid_chunks, hipStream_t s) {
    const int per = (grid_chunks + 7) / 8;
    dim3 grid(8 * (per < 1 ? 1 : per)), block(TRANSFER_THREADS);
    const bool g = a.map != nullptr;  // particles read through the map from a.src
    switch (mode) {
        case XFER_P2G:
            if (g) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_transfer<XFER_P2G, true>), grid, block, 0, s, a);
            else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_transfer<XFER_P2G, false>), grid, block, 0, s, a);
            break;
        case XFER_G2P:
            if (g) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_transfer<XFER_G2P, true>), grid, block, 0, s, a);
            else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_transfer<XFER_G2P, false>), grid, block, 0, s, a);
            break;
        case XFER_G2P2G_FULL:
            if (g) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_transfer<XFER_G2P2G_FULL, true>), grid, block, 0, s, a);
            else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_transfer<XFER_G2P2G_FULL, false>), grid, block, 0, s, a);
            break;
        default:
            if (g) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_transfer<XFER_G2P2G, true>), grid, block, 0, s, a);
            else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_transfer<XFER_G2P2G, false>), grid, block, 0, s, a);
            break;
    }
}

}  // namespace mpm
