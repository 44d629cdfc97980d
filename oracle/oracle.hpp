// oracle.hpp -- CPU restatement of the reference MLS-MPM substep.
//
// TEST INFRASTRUCTURE ONLY.  This file is the reference the HIP engine's
// results are compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load the library built from it; the product path
// (water-sim-cuda_amd/, include/) never links, calls or falls back to it.
//
// Every function cites the reference source it restates (paths relative to the
// reference repository root).  The arithmetic follows the reference's operation
// order; T = float reproduces the reference's FP32 semantics (including the
// double-precision literals of `spline`), T = double is the FP64 oracle whose
// distance from the FP32 one is reported next to the GPU parity error.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace oracle {

// Same field order as SimConfig, common.cuh:65-79.
struct Config {
    uint32_t SIZE_X, SIZE_Y, SIZE_Z;
    uint32_t PADDING;
    float GRAVITY;
    float DT;
    float GAMMA;
    float COMPRESSION;
    float RESOLUTION;
    uint32_t SUBSTEPS;
    uint32_t SDF_RESOLUTION;
    uint32_t GRID_SIZE;
    uint32_t GRID_NUMBER;
    uint32_t GRID_BLOCKS;
};

// ParticleAttr, common.cuh:12-30: attribute-major SoA, element (attr, i) at attr*n + i.
enum Attr {
    POS_X = 0, POS_Y, POS_Z, VEL_X, VEL_Y, VEL_Z,
    C00, C01, C02, C10, C11, C12, C20, C21, C22,
    F00, F01, F02, F10, F11, F12, F20, F21, F22,
    VOL, MASS, ATTR_COUNT
};

template <class T> struct V3 { T x, y, z; };

// getGridIdx, common.cuh:131-134
inline int grid_idx(const Config& c, int x, int y, int z) {
    return z * (int)(c.SIZE_X * c.SIZE_Y) + y * (int)c.SIZE_X + x;
}

// isInBounds, common.cuh:212-218 (the int is compared against uint32 there,
// so negative coordinates are out of bounds).
inline bool in_bounds(const Config& c, int x, int y, int z) {
    if (x < 0 || (uint32_t)x >= c.SIZE_X) return false;
    if (y < 0 || (uint32_t)y >= c.SIZE_Y) return false;
    if (z < 0 || (uint32_t)z >= c.SIZE_Z) return false;
    return true;
}

// spline, common.cuh:205-210.  The reference evaluates `0.75-x*x` with a float
// product and a double subtraction, and `(1.5-|x|)^2*0.5` entirely in double,
// both rounded to float on return.
inline float spline(float x) {
    float ax = std::fabs(x);
    if (ax < 0.5f) {
        float xx = x * x;
        return (float)(0.75 - (double)xx);
    }
    if (ax >= 0.5f && ax < 1.5f) {
        double t = 1.5 - (double)ax;
        return (float)(t * t * 0.5);
    }
    return 0.0f;
}
inline double spline(double x) {
    double ax = std::fabs(x);
    if (ax < 0.5) return 0.75 - x * x;
    if (ax >= 0.5 && ax < 1.5) return (1.5 - ax) * (1.5 - ax) * 0.5;
    return 0.0;
}

// C truncation toward zero of (x - 0.5), kernels.cu:14 and :204.
inline int base_of(float x) { return (int)(x - 0.5f); }
inline int base_of(double x) { return (int)(x - 0.5); }

// det3x3, common.cuh:250-253 (row-major).
template <class T> inline T det3(const T* A) {
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}

// multiply3x3, common.cuh:220-233: C = A * B.
template <class T> inline void mul3(const T* A, const T* B, T* C) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            C[3 * r + c] = A[3 * r + 0] * B[0 + c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
}

// calculateNewF, common.cuh:255-261: newF = (I + DT*C) * oldF.
template <class T> inline void new_F(T dt, const T* C, const T* oldF, T* newF) {
    const T I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    T A[9];
    for (int i = 0; i < 9; ++i) A[i] = dt * C[i];
    for (int i = 0; i < 9; ++i) A[i] = I[i] + A[i];
    mul3(A, oldF, newF);
}

// Grid node buffers: SoA planes mass, mom_x, mom_y, mom_z (Grid, common.cuh:110-121).
template <class T> struct GridView {
    T* mass;
    T* mom[3];
    const T* sdf;
};

// getSDF, common.cuh:136-167: trilinear interpolation of node SDF values with
// floor() cells and indices clamped to the grid.
template <class T> inline T get_sdf(const Config& c, V3<T> p, const T* sdf) {
    int i = (int)std::floor(p.x), j = (int)std::floor(p.y), k = (int)std::floor(p.z);
    T tx = p.x - (T)i, ty = p.y - (T)j, tz = p.z - (T)k;
    auto sample = [&](int x, int y, int z) {
        x = std::max(0, std::min(x, (int)c.SIZE_X - 1));
        y = std::max(0, std::min(y, (int)c.SIZE_Y - 1));
        z = std::max(0, std::min(z, (int)c.SIZE_Z - 1));
        return sdf[grid_idx(c, x, y, z)];
    };
    T N[8] = {sample(i, j, k),     sample(i + 1, j, k),     sample(i, j + 1, k),     sample(i + 1, j + 1, k),
              sample(i, j, k + 1), sample(i + 1, j, k + 1), sample(i, j + 1, k + 1), sample(i + 1, j + 1, k + 1)};
    const T one = 1;
    T inX[4] = {N[0] * (one - tx) + N[1] * tx, N[2] * (one - tx) + N[3] * tx, N[4] * (one - tx) + N[5] * tx,
                N[6] * (one - tx) + N[7] * tx};
    T inY[2] = {inX[0] * (one - ty) + inX[1] * ty, inX[2] * (one - ty) + inX[3] * ty};
    return inY[0] * (one - tz) + inY[1] * tz;
}

// calculateNormal, common.cuh:169-182: central difference (eps 0.1) of six
// trilinear samples, normalised with len + 1e-6.
template <class T> inline V3<T> sdf_normal(const Config& c, V3<T> p, const T* sdf) {
    const T eps = (T)0.1f;
    V3<T> n;
    n.x = get_sdf<T>(c, {p.x + eps, p.y, p.z}, sdf) - get_sdf<T>(c, {p.x - eps, p.y, p.z}, sdf);
    n.y = get_sdf<T>(c, {p.x, p.y + eps, p.z}, sdf) - get_sdf<T>(c, {p.x, p.y - eps, p.z}, sdf);
    n.z = get_sdf<T>(c, {p.x, p.y, p.z + eps}, sdf) - get_sdf<T>(c, {p.x, p.y, p.z - eps}, sdf);
    T len = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z) + (T)1e-6f;
    n.x /= len; n.y /= len; n.z /= len;
    return n;
}

// getNormalAtNode, common.cuh:195-203: clamped central difference of node SDF
// values, normalised with len + 1e-9.
template <class T> inline V3<T> node_normal(const Config& c, int x, int y, int z, const T* sdf) {
    T nx = sdf[grid_idx(c, std::min(x + 1, (int)c.SIZE_X - 1), y, z)] - sdf[grid_idx(c, std::max(x - 1, 0), y, z)];
    T ny = sdf[grid_idx(c, x, std::min(y + 1, (int)c.SIZE_Y - 1), z)] - sdf[grid_idx(c, x, std::max(y - 1, 0), z)];
    T nz = sdf[grid_idx(c, x, y, std::min(z + 1, (int)c.SIZE_Z - 1))] - sdf[grid_idx(c, x, y, std::max(z - 1, 0))];
    T len = std::sqrt(nx * nx + ny * ny + nz * nz) + (T)1e-9f;
    return {nx / len, ny / len, nz / len};
}

// expandBits, common.cuh:277-284 (32-bit wrapping multiplies).
inline uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// calculateMorton, common.cuh:304-307.
inline uint32_t morton(uint32_t x, uint32_t y, uint32_t z) {
    return (expand_bits(z) << 2) | (expand_bits(y) << 1) | expand_bits(x);
}

// (unsigned int)max(0.0f, p), kernels.cu:366-368.  fmaxf(0, NaN) is 0; values at
// or beyond 2^32 saturate, as the GPU conversion does.
inline uint32_t key_coord(float p) {
    float v = std::fmax(0.0f, p);
    if (!(v < 4294967296.0f)) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

// setKeys, kernels.cu:361-371: key stored into an int.
inline int32_t morton_key(float x, float y, float z) {
    return (int32_t)morton(key_coord(x), key_coord(y), key_coord(z));
}

// Engine::sortParticles, engine.cu:92-102: values = iota, then sort_by_key on
// the int keys.  Defined here as the stable ascending (key, original index)
// order; Thrust's tie order is not part of its contract.
inline void sort_perm(int n, const int32_t* keys, int32_t* perm) {
    for (int i = 0; i < n; ++i) perm[i] = i;
    std::stable_sort(perm, perm + n, [&](int32_t a, int32_t b) { return keys[a] < keys[b]; });
}

// Particle state view over a 26-attribute SoA buffer.
template <class T> struct ParticleView {
    T* b;
    int n;
    T& at(int attr, int i) const { return b[(size_t)attr * n + i]; }
};

// initFMatrices, kernels.cu:349-358 (F diagonal = 1, V = RES^3, m = 1; the
// off-diagonals keep the caller's values).
template <class T> inline void init_state(const Config& c, ParticleView<T> p) {
    for (int i = 0; i < p.n; ++i) {
        p.at(F00, i) = 1; p.at(F11, i) = 1; p.at(F22, i) = 1;
        p.at(VOL, i) = (T)(c.RESOLUTION * c.RESOLUTION * c.RESOLUTION);
        p.at(MASS, i) = 1;
    }
}

// Per-particle P2G quantities (kernels.cu:14-35): base cell, weights, affine
// matrix with the pressure term added, mass.
template <class T> struct P2GParticle {
    int base[3];
    T w[3][3];
    V3<T> pos, vel;
    T C[9];
    T m;
};

template <class T> inline P2GParticle<T> p2g_prepare(const Config& c, ParticleView<T> p, int idx) {
    P2GParticle<T> q;
    q.pos = {p.at(POS_X, idx), p.at(POS_Y, idx), p.at(POS_Z, idx)};
    q.vel = {p.at(VEL_X, idx), p.at(VEL_Y, idx), p.at(VEL_Z, idx)};
    q.base[0] = base_of(q.pos.x); q.base[1] = base_of(q.pos.y); q.base[2] = base_of(q.pos.z);
    T F[9];
    for (int i = 0; i < 9; ++i) q.C[i] = p.at(C00 + i, idx);
    for (int i = 0; i < 9; ++i) F[i] = p.at(F00 + i, idx);
    q.m = p.at(MASS, idx);
    T J = det3(F);
    J = std::max((T)0.9f, std::min(J, (T)1.1f));
    T pressure = (T)c.COMPRESSION * (std::pow(J, (T)c.GAMMA) - (T)1);
    if (pressure < 0) pressure = 0;
    T volume = p.at(VOL, idx) * J;
    T multi = (T)4 * (T)c.DT * volume * pressure / q.m;
    const T I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    T stress[9];
    for (int i = 0; i < 9; ++i) stress[i] = multi * I[i];
    for (int i = 0; i < 9; ++i) q.C[i] = q.C[i] + stress[i];
    for (int a = 0; a < 3; ++a) {
        T pa = a == 0 ? q.pos.x : (a == 1 ? q.pos.y : q.pos.z);
        for (int i = 0; i < 3; ++i) q.w[a][i] = spline(pa - (T)(q.base[a] + i));
    }
    return q;
}

// Scatter of one particle onto its 27 nodes, kernels.cu:83-119 (the tile and
// its > 1e-9 flush threshold, :105-140, are a block-level staging detail of the
// reference; the oracle adds directly into the grid).
template <class T> inline void p2g_scatter(const Config& c, const P2GParticle<T>& q, GridView<T> g) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                int gx = q.base[0] + i, gy = q.base[1] + j, gz = q.base[2] + k;
                if (!in_bounds(c, gx, gy, gz)) continue;
                int cell = grid_idx(c, gx, gy, gz);
                V3<T> d = {q.pos.x - (T)gx, q.pos.y - (T)gy, q.pos.z - (T)gz};
                V3<T> dist = {-d.x, -d.y, -d.z};
                T weight = q.w[0][i] * q.w[1][j] * q.w[2][k];
                T wm = q.m * weight * (T)1;
                T cx = q.C[0] * dist.x + q.C[1] * dist.y + q.C[2] * dist.z;
                T cy = q.C[3] * dist.x + q.C[4] * dist.y + q.C[5] * dist.z;
                T cz = q.C[6] * dist.x + q.C[7] * dist.y + q.C[8] * dist.z;
                g.mass[cell] += wm;
                g.mom[0][cell] += wm * (q.vel.x + cx);
                g.mom[1][cell] += wm * (q.vel.y + cy);
                g.mom[2][cell] += wm * (q.vel.z + cz);
            }
}

// gridUpdate, kernels.cu:144-196, one node.
template <class T> inline void grid_update_node(const Config& c, GridView<T> g, int idx) {
    int x = idx % (int)c.SIZE_X, y = (idx / (int)c.SIZE_X) % (int)c.SIZE_Y, z = idx / (int)(c.SIZE_X * c.SIZE_Y);
    T mass = g.mass[idx];
    T dist = g.sdf[idx];
    V3<T> mom = {g.mom[0][idx], g.mom[1][idx], g.mom[2][idx]};
    V3<T> v = {0, 0, 0};
    if (mass > (T)1e-9f) {
        v.x = mom.x / mass;
        v.y = mom.y / mass - (T)c.GRAVITY * (T)c.DT;
        v.z = mom.z / mass;
        if (dist < 0) {
            V3<T> n = node_normal<T>(c, x, y, z, g.sdf);
            T vn = v.x * n.x + v.y * n.y + v.z * n.z;
            if (vn < 0) {
                v.x -= n.x * vn; v.y -= n.y * vn; v.z -= n.z * vn;
            }
        }
        if ((uint32_t)x < c.PADDING && v.x < 0) v.x = 0;
        if ((uint32_t)y < c.PADDING && v.y < 0) v.y = 0;
        if ((uint32_t)z < c.PADDING && v.z < 0) v.z = 0;
        if ((uint32_t)x > c.SIZE_X - 1 - c.PADDING && v.x > 0) v.x = 0;
        if ((uint32_t)y > c.SIZE_Y - 1 - c.PADDING && v.y > 0) v.y = 0;
        if ((uint32_t)z > c.SIZE_Z - 1 - c.PADDING && v.z > 0) v.z = 0;
    }
    g.mom[0][idx] = v.x; g.mom[1][idx] = v.y; g.mom[2][idx] = v.z;
}

// g2PTransfer, kernels.cu:199-280, one particle.  The grid's momentum planes
// hold node velocities after the grid update.
template <class T> inline void g2p_particle(const Config& c, ParticleView<T> p, GridView<T> g, int idx) {
    V3<T> pos = {p.at(POS_X, idx), p.at(POS_Y, idx), p.at(POS_Z, idx)};
    int b[3] = {base_of(pos.x), base_of(pos.y), base_of(pos.z)};
    V3<T> vel = {0, 0, 0};
    T C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    T w[3][3];
    for (int i = 0; i < 3; ++i) w[0][i] = spline(pos.x - (T)(b[0] + i));
    for (int i = 0; i < 3; ++i) w[1][i] = spline(pos.y - (T)(b[1] + i));
    for (int i = 0; i < 3; ++i) w[2][i] = spline(pos.z - (T)(b[2] + i));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                int gx = b[0] + i, gy = b[1] + j, gz = b[2] + k;
                if (!in_bounds(c, gx, gy, gz)) continue;
                int cell = grid_idx(c, gx, gy, gz);
                V3<T> d = {pos.x - (T)gx, pos.y - (T)gy, pos.z - (T)gz};
                T weight = w[0][i] * w[1][j] * w[2][k];
                V3<T> dist = {-d.x, -d.y, -d.z};
                T vx = g.mom[0][cell], vy = g.mom[1][cell], vz = g.mom[2][cell];
                vel.x += weight * vx; vel.y += weight * vy; vel.z += weight * vz;
                T term = (T)4 * weight;
                C[0] += term * vx * dist.x; C[1] += term * vx * dist.y; C[2] += term * vx * dist.z;
                C[3] += term * vy * dist.x; C[4] += term * vy * dist.y; C[5] += term * vy * dist.z;
                C[6] += term * vz * dist.x; C[7] += term * vz * dist.y; C[8] += term * vz * dist.z;
            }
    T oldF[9], newF[9];
    for (int i = 0; i < 9; ++i) oldF[i] = p.at(F00 + i, idx);
    new_F<T>((T)c.DT, C, oldF, newF);
    T J = det3(newF);
    if (J < (T)0.3f) {
        T scale = std::pow((T)0.3f / J, (T)1 / (T)3);
        for (int i = 0; i < 9; ++i) newF[i] *= scale;
    }
    const T dt = (T)c.DT;
    V3<T> next = {pos.x + vel.x * dt, pos.y + vel.y * dt, pos.z + vel.z * dt};
    T dist = get_sdf<T>(c, next, g.sdf);
    if (dist < 0) {
        V3<T> n = sdf_normal<T>(c, next, g.sdf);
        next.x -= n.x * dist; next.y -= n.y * dist; next.z -= n.z * dist;
        T vn = vel.x * n.x + vel.y * n.y + vel.z * n.z;
        if (vn < 0) {
            vel.x -= n.x * vn; vel.y -= n.y * vn; vel.z -= n.z * vn;
        }
    }
    p.at(VEL_X, idx) = vel.x; p.at(VEL_Y, idx) = vel.y; p.at(VEL_Z, idx) = vel.z;
    p.at(POS_X, idx) = next.x; p.at(POS_Y, idx) = next.y; p.at(POS_Z, idx) = next.z;
    for (int i = 0; i < 9; ++i) p.at(C00 + i, idx) = C[i];
    for (int i = 0; i < 9; ++i) p.at(F00 + i, idx) = newF[i];
}

// ---- mesh SDF (makeSDF, kernels.cu:282-300) ----
struct Tri { float v[9]; };  // Triangle, common.cuh:123-126: v0, v1, v2 (36 B)

template <class T> inline T dot3(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> sub3(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> add3(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> scl3(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }

// norm3df: the Euclidean norm, evaluated in double and rounded.
inline float norm3(float x, float y, float z) {
    return (float)std::sqrt((double)x * x + (double)y * y + (double)z * z);
}
inline double norm3(double x, double y, double z) { return std::sqrt(x * x + y * y + z * z); }

// calculateSolidAngle, common.cuh:310-323 (Van Oosterom & Strackee 1983).
template <class T> inline T solid_angle(V3<T> P, V3<T> v0, V3<T> v1, V3<T> v2) {
    V3<T> a = sub3(v0, P), b = sub3(v1, P), cc = sub3(v2, P);
    T la = norm3(a.x, a.y, a.z), lb = norm3(b.x, b.y, b.z), lc = norm3(cc.x, cc.y, cc.z);
    T det = a.x * (b.y * cc.z - b.z * cc.y) + a.y * (b.z * cc.x - b.x * cc.z) + a.z * (b.x * cc.y - b.y * cc.x);
    T div = (la * lb * lc) + dot3(a, b) * lc + dot3(a, cc) * lb + dot3(b, cc) * la;
    return (T)2 * std::atan2(det, div);
}

// pointTriangleDistanceSq, common.cuh:325-378 (closest point by Voronoi region,
// Ericson, Real-Time Collision Detection, 5.1.5).
template <class T> inline T point_tri_dist2(V3<T> p, V3<T> v0, V3<T> v1, V3<T> v2) {
    V3<T> ab = sub3(v1, v0), ac = sub3(v2, v0), ap = sub3(p, v0);
    T d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0 && d2 <= 0) return dot3(ap, ap);
    V3<T> bp = sub3(p, v1);
    T d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0 && d4 <= d3) return dot3(bp, bp);
    T vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        T v = d1 / (d1 - d3);
        V3<T> diff = sub3(p, add3(v0, scl3(v, ab)));
        return dot3(diff, diff);
    }
    V3<T> cp = sub3(p, v2);
    T d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0 && d5 <= d6) return dot3(cp, cp);
    T vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        T w = d2 / (d2 - d6);
        V3<T> diff = sub3(p, add3(v0, scl3(w, ac)));
        return dot3(diff, diff);
    }
    T va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        T w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        V3<T> diff = sub3(p, add3(v1, scl3(w, sub3(v2, v1))));
        return dot3(diff, diff);
    }
    T denom = (T)1 / (va + vb + vc);
    T v = vb * denom, w = vc * denom;
    V3<T> closest = add3(add3(v0, scl3(v, ab)), scl3(w, ac));
    V3<T> diff = sub3(p, closest);
    return dot3(diff, diff);
}

// makeSDF, kernels.cu:282-300, one node: winding number by summed solid angles
// (literal 3.1415f) and the minimum point-triangle distance.
template <class T> inline T make_sdf_node(const Config& c, int idx, const Tri* tris, int nt) {
    int x = idx % (int)c.SIZE_X, y = (idx / (int)c.SIZE_X) % (int)c.SIZE_Y, z = idx / (int)(c.SIZE_X * c.SIZE_Y);
    V3<T> P = {(T)x, (T)y, (T)z};
    T wn_sum = 0, min_d2 = (T)1e30f;
    for (int t = 0; t < nt; ++t) {
        const float* v = tris[t].v;
        V3<T> a = {(T)v[0], (T)v[1], (T)v[2]}, b = {(T)v[3], (T)v[4], (T)v[5]}, cc = {(T)v[6], (T)v[7], (T)v[8]};
        wn_sum += solid_angle<T>(P, a, b, cc);
        T d2 = point_tri_dist2<T>(P, a, b, cc);
        if (d2 < min_d2) min_d2 = d2;
    }
    T wn = wn_sum / ((T)4 * (T)3.1415f);
    return (wn > (T)0.5f ? (T)-1 : (T)1) * std::sqrt(min_d2);
}

}  // namespace oracle
