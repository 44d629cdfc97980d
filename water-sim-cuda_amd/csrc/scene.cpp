// scene.cpp -- seeded scene generators for the benchmark and test scenes.
//
// No reference counterpart: the reference seeds water by voxelising OBJ
// meshes (VoxelEngine.cpp:120-494) and packs the 26-attribute buffer with
// positions only, attributes 3..23 zero and VOL = MASS = 1
// (VoxelPipeline::process, VoxelPipeline.cpp:9-30).  These generators produce
// the same buffer layout from analytic shapes, 8 particles per cell (one per
// 2x2x2 sub-cell, jittered uniformly inside it), from SplitMix64 (Steele, Lea
// and Flood 2014) with an explicit integer-to-float conversion so every
// platform produces the same positions.
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "mpm.h"

namespace mpm {
void set_last_error(const std::string& msg);
}

namespace {

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Uniform float in [0, 1) from the top 24 bits (exactly representable).
inline float u01(uint64_t& s) { return (float)(splitmix64(s) >> 40) * (1.0f / 16777216.0f); }

struct Emitter {
    float* out;
    int64_t cap;
    int64_t count;
    bool overflow = false;
    void emit(float x, float y, float z) {
        if (out) {
            if (count >= cap) {
                overflow = true;
                return;
            }
            out[0 * cap + count] = x;
            out[1 * cap + count] = y;
            out[2 * cap + count] = z;
            for (int a = 3; a < 24; ++a) out[a * cap + count] = 0.0f;
            out[24 * cap + count] = 1.0f;
            out[25 * cap + count] = 1.0f;
        }
        ++count;
    }
};

struct Tris {
    mpm_triangle* out;
    int count = 0;
    void add(const float a[3], const float b[3], const float c[3]) {
        if (out) {
            mpm_triangle& t = out[count];
            for (int i = 0; i < 3; ++i) {
                t.v0[i] = a[i];
                t.v1[i] = b[i];
                t.v2[i] = c[i];
            }
        }
        ++count;
    }
    // Adds triangle (a, b, c) wound so that its normal points along `n`.
    void add_oriented(const float a[3], const float b[3], const float c[3], const float n[3]) {
        const float ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const float ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const float cr[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
        if (cr[0] * n[0] + cr[1] * n[1] + cr[2] * n[2] >= 0.0f) add(a, b, c);
        else add(a, c, b);
    }
    // Planar quad p0 + s*u + t*v, s,t in [0,1], split into nu x nv squares.
    void add_quad(const float p0[3], const float u[3], const float v[3], int nu, int nv, const float n[3]) {
        for (int i = 0; i < nu; ++i)
            for (int j = 0; j < nv; ++j) {
                float s0 = (float)i / nu, s1 = (float)(i + 1) / nu, t0 = (float)j / nv, t1 = (float)(j + 1) / nv;
                float a[3], b[3], c[3], d[3];
                for (int k = 0; k < 3; ++k) {
                    a[k] = p0[k] + s0 * u[k] + t0 * v[k];
                    b[k] = p0[k] + s1 * u[k] + t0 * v[k];
                    c[k] = p0[k] + s1 * u[k] + t1 * v[k];
                    d[k] = p0[k] + s0 * u[k] + t1 * v[k];
                }
                add_oriented(a, b, c, n);
                add_oriented(a, c, d, n);
            }
    }
};

int segs(float len, float seg) {
    int n = (int)std::ceil(std::fabs(len) / seg);
    return n < 1 ? 1 : n;
}

}  // namespace

extern "C" {

int mpm_scene_fill(int shape, const float* lo, const float* hi, uint64_t seed, float* soa26_out, int64_t cap,
                   int64_t* count) {
    if (!lo || !hi || !count) {
        mpm::set_last_error("null argument");
        return MPM_EINVAL;
    }
    if (soa26_out && *count < 0) {
        mpm::set_last_error("negative start count");
        return MPM_EINVAL;
    }
    Emitter e{soa26_out, cap, soa26_out ? *count : 0};
    uint64_t s = seed;
    int x0, x1, y0, y1, z0, z1;
    float cx = 0, cy = 0, cz = 0, r2 = 0;
    if (shape == 0) {
        x0 = (int)std::floor(lo[0]); x1 = (int)std::floor(hi[0]);
        y0 = (int)std::floor(lo[1]); y1 = (int)std::floor(hi[1]);
        z0 = (int)std::floor(lo[2]); z1 = (int)std::floor(hi[2]);
    } else if (shape == 1) {
        cx = lo[0]; cy = lo[1]; cz = lo[2];
        const float r = hi[0];
        r2 = r * r;
        x0 = (int)std::floor(cx - r); x1 = (int)std::ceil(cx + r);
        y0 = (int)std::floor(cy - r); y1 = (int)std::ceil(cy + r);
        z0 = (int)std::floor(cz - r); z1 = (int)std::ceil(cz + r);
    } else {
        mpm::set_last_error("unknown shape (0 = box, 1 = sphere)");
        return MPM_EINVAL;
    }
    for (int z = z0; z < z1; ++z)
        for (int y = y0; y < y1; ++y)
            for (int x = x0; x < x1; ++x)
                for (int sub = 0; sub < 8; ++sub) {
                    const int sx = sub & 1, sy = (sub >> 1) & 1, sz = sub >> 2;
                    const float px = (float)x + 0.5f * ((float)sx + u01(s));
                    const float py = (float)y + 0.5f * ((float)sy + u01(s));
                    const float pz = (float)z + 0.5f * ((float)sz + u01(s));
                    if (shape == 1) {
                        const double dx = (double)px - cx, dy = (double)py - cy, dz = (double)pz - cz;
                        if (dx * dx + dy * dy + dz * dz >= (double)r2) continue;
                    }
                    e.emit(px, py, pz);
                }
    if (e.overflow) {
        mpm::set_last_error("particle buffer capacity exceeded");
        return MPM_ELIMIT;
    }
    *count = e.count;
    return MPM_OK;
}

int mpm_mesh_uv_sphere(const float* c, float radius, int slices, int stacks, mpm_triangle* tris_out, int* ntris) {
    if (!c || !ntris || slices < 3 || stacks < 2) {
        mpm::set_last_error("bad sphere arguments");
        return MPM_EINVAL;
    }
    const double PI = 3.14159265358979323846;
    auto vert = [&](int st, int sl, float v[3]) {
        const double th = PI * st / stacks, ph = 2.0 * PI * sl / slices;
        v[0] = (float)(c[0] + radius * std::sin(th) * std::cos(ph));
        v[1] = (float)(c[1] + radius * std::cos(th));
        v[2] = (float)(c[2] + radius * std::sin(th) * std::sin(ph));
    };
    Tris t{tris_out};
    for (int st = 0; st < stacks; ++st)
        for (int sl = 0; sl < slices; ++sl) {
            float a[3], b[3], cc[3], d[3];
            vert(st, sl, a);
            vert(st + 1, sl, b);
            vert(st + 1, sl + 1, cc);
            vert(st, sl + 1, d);
            // outward normal: direction from the centre through the patch
            float m[3];
            for (int k = 0; k < 3; ++k) m[k] = 0.25f * (a[k] + b[k] + cc[k] + d[k]) - c[k];
            if (st != stacks - 1) t.add_oriented(a, b, cc, m);
            if (st != 0) t.add_oriented(a, cc, d, m);
        }
    *ntris = t.count;
    return MPM_OK;
}

int mpm_mesh_box(const float* lo, const float* hi, mpm_triangle* tris_out, int* ntris) {
    if (!lo || !hi || !ntris) {
        mpm::set_last_error("null argument");
        return MPM_EINVAL;
    }
    Tris t{tris_out};
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    const float ux[3] = {dx, 0, 0}, uy[3] = {0, dy, 0}, uz[3] = {0, 0, dz};
    const float nxm[3] = {-1, 0, 0}, nxp[3] = {1, 0, 0}, nym[3] = {0, -1, 0}, nyp[3] = {0, 1, 0}, nzm[3] = {0, 0, -1},
                nzp[3] = {0, 0, 1};
    const float px[3] = {hi[0], lo[1], lo[2]}, py[3] = {lo[0], hi[1], lo[2]}, pz[3] = {lo[0], lo[1], hi[2]};
    t.add_quad(lo, uy, uz, 1, 1, nxm);
    t.add_quad(px, uy, uz, 1, 1, nxp);
    t.add_quad(lo, ux, uz, 1, 1, nym);
    t.add_quad(py, ux, uz, 1, 1, nyp);
    t.add_quad(lo, ux, uy, 1, 1, nzm);
    t.add_quad(pz, ux, uy, 1, 1, nzp);
    *ntris = t.count;
    return MPM_OK;
}

// A U-shaped channel running along x: the solid is the outer box minus the
// open slot x in [lo.x, hi.x], y in [lo.y + wall, hi.y], z in
// [lo.z + wall, hi.z - wall].  Its boundary is the U outline in the (z, y)
// plane extruded along x, plus the two end caps.
int mpm_mesh_u_channel(const float* lo, const float* hi, float wall, float seg, mpm_triangle* tris_out, int* ntris) {
    if (!lo || !hi || !ntris || wall <= 0 || seg <= 0 || hi[2] - lo[2] <= 2 * wall || hi[1] - lo[1] <= wall) {
        mpm::set_last_error("bad U-channel arguments");
        return MPM_EINVAL;
    }
    Tris t{tris_out};
    const float z0 = lo[2], z1 = lo[2] + wall, z2 = hi[2] - wall, z3 = hi[2];
    const float y0 = lo[1], y1 = lo[1] + wall, y2 = hi[1];
    // outline (z, y), counter-clockwise seen with z right, y up
    const float oz[8] = {z0, z3, z3, z2, z2, z1, z1, z0};
    const float oy[8] = {y0, y0, y2, y2, y1, y1, y2, y2};
    const float lx = hi[0] - lo[0];
    const int nx = segs(lx, seg);
    for (int e = 0; e < 8; ++e) {
        const int f = (e + 1) % 8;
        const float dz = oz[f] - oz[e], dy = oy[f] - oy[e];
        // outward normal of a CCW polygon edge (dz, dy) in the (z, y) plane: (dy, -dz)
        const float n[3] = {0.0f, -dz, dy};
        const float p0[3] = {lo[0], oy[e], oz[e]};
        const float u[3] = {lx, 0.0f, 0.0f};
        const float v[3] = {0.0f, dy, dz};
        t.add_quad(p0, u, v, nx, segs(std::sqrt(dz * dz + dy * dy), seg), n);
    }
    // end caps: left wall, floor middle, right wall rectangles in the (z, y) plane
    const float rz0[3] = {z0, z1, z2}, rz1[3] = {z1, z2, z3};
    const float ry1[3] = {y2, y1, y2};
    for (int cap = 0; cap < 2; ++cap) {
        const float x = cap == 0 ? lo[0] : hi[0];
        const float n[3] = {cap == 0 ? -1.0f : 1.0f, 0.0f, 0.0f};
        for (int r = 0; r < 3; ++r) {
            const float p0[3] = {x, y0, rz0[r]};
            const float u[3] = {0.0f, 0.0f, rz1[r] - rz0[r]};
            const float v[3] = {0.0f, ry1[r] - y0, 0.0f};
            t.add_quad(p0, u, v, segs(u[2], seg), segs(v[1], seg), n);
        }
    }
    *ntris = t.count;
    return MPM_OK;
}

}  // extern "C"
