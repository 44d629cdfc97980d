// voxel.cpp -- the reference's scene front-end, host side: OBJ meshes to
// water particles (surface voxelisation, normalisation, 5x5x5 dilation,
// packing) and to collider triangles (normalisation, culling), and the whole
// config.txt -> particle buffer + triangle list path of the reference's
// Simulation constructor (Simulation.cpp:4-38).
//
// Restated, not translated: the arithmetic follows the reference step by step
// in FP32 (the results are particle positions, which must match), the code
// and data structures are this file's own.
//  * OBJ loading: the reference calls tinyobjloader (>= 2.0.0rc13, vcpkg
//    baseline 55fab67a, ObjLoader.cpp:16-17) with its default triangulation.
//    This parser reads `v` and `f` records (`i`, `i/t`, `i//n`, `i/t/n`,
//    negative = relative indices) and triangulates as that default does for
//    quads: split along the shorter diagonal (0-2 vs 1-3), giving (0,1,2),
//    (0,2,3) when |v2-v0|^2 < |v3-v1|^2 and (0,1,3), (1,2,3) otherwise.
//    Faces with more than four vertices are fanned from their first vertex
//    (tinyobjloader's choice for those is not pinned here).
//  * VoxelEngine::voxelize (VoxelEngine.cpp:120-243): cell size = longest
//    bounding-box side / resolution, grid ceil(extent / cell) per axis, every
//    cell overlapping a triangle's bounding box tested with the separating-
//    axis triangle/box test (triangleAABBIntersection, :902-961; Akenine-
//    Moller 2001), centres of hit cells in z, y, x order.
//  * VoxelEngine::normalize(VoxelData&) (:245-494): centre on normalizeSize/2
//    and scale the longest side to normalizeSize, then to `scale`, displace,
//    snap to the RESOLUTION lattice with round(), clamp to [0, normalizeSize /
//    RESOLUTION], dilate by +-2 lattice steps per axis, keep unique lattice
//    points in (x, y, z) lexicographic order.
//  * VoxelEngine::normalize(std::vector<Triangle>&) (:601-827): the same
//    affine map, then drop triangles with no vertex inside [0, size]^3.
//  * VoxelPipeline::process (VoxelPipeline.cpp:5-33): 26 x n buffer with
//    positions, zeros, VOL = MASS = 1.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "mpm.h"

namespace mpm {
void set_last_error(const std::string& msg);
}

namespace {

struct V3 {
    float x, y, z;
};

inline V3 sub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float mn3(float a, float b, float c) { return std::min(std::min(a, b), c); }
inline float mx3(float a, float b, float c) { return std::max(std::max(a, b), c); }

struct Tri {
    V3 v[3];
};

struct Mesh {
    std::vector<V3> verts;
    std::vector<Tri> tris;
};

int err(int code, const std::string& m) {
    mpm::set_last_error(m);
    return code;
}

// Parses one face corner "i", "i/t", "i//n" or "i/t/n" into a 0-based vertex
// index (negative indices count back from the last vertex read).
bool corner_index(const std::string& tok, int nverts, int& out) {
    const size_t slash = tok.find('/');
    const std::string head = tok.substr(0, slash);
    if (head.empty()) return false;
    char* end = nullptr;
    const long v = std::strtol(head.c_str(), &end, 10);
    if (*end != '\0' || v == 0) return false;
    out = v > 0 ? (int)(v - 1) : nverts + (int)v;
    return out >= 0 && out < nverts;
}

int load_obj(const std::string& path, Mesh& m) {
    std::ifstream f(path);
    if (!f) return err(MPM_EIO, "cannot open " + path);
    std::string line;
    int lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ss(line);
        std::string tag;
        if (!(ss >> tag)) continue;
        if (tag == "v") {
            V3 p;
            if (!(ss >> p.x >> p.y >> p.z)) return err(MPM_EIO, path + ":" + std::to_string(lineno) + ": bad vertex");
            m.verts.push_back(p);
        } else if (tag == "f") {
            std::vector<int> idx;
            std::string tok;
            while (ss >> tok) {
                int i;
                if (!corner_index(tok, (int)m.verts.size(), i))
                    return err(MPM_EIO, path + ":" + std::to_string(lineno) + ": bad face index " + tok);
                idx.push_back(i);
            }
            if (idx.size() < 3) continue;
            auto add = [&](int a, int b, int c) { m.tris.push_back({{m.verts[idx[a]], m.verts[idx[b]], m.verts[idx[c]]}}); };
            if (idx.size() == 3) {
                add(0, 1, 2);
            } else if (idx.size() == 4) {
                const V3 e02 = sub(m.verts[idx[2]], m.verts[idx[0]]);
                const V3 e13 = sub(m.verts[idx[3]], m.verts[idx[1]]);
                if (dot(e02, e02) < dot(e13, e13)) {
                    add(0, 1, 2);
                    add(0, 2, 3);
                } else {
                    add(0, 1, 3);
                    add(1, 2, 3);
                }
            } else {
                for (size_t k = 1; k + 1 < idx.size(); ++k) add(0, (int)k, (int)k + 1);
            }
        }
        // other records (vn, vt, o, g, s, usemtl, mtllib, comments) carry nothing used here
    }
    return MPM_OK;
}

// Separating-axis triangle/box overlap (triangleAABBIntersection,
// VoxelEngine.cpp:902-961): 9 edge-cross-axis tests, the 3 box face normals,
// the triangle's plane.
bool tri_box(const Tri& t, const V3& bmin, const V3& bmax) {
    const V3 c = {(bmin.x + bmax.x) * 0.5f, (bmin.y + bmax.y) * 0.5f, (bmin.z + bmax.z) * 0.5f};
    const V3 h = {(bmax.x - bmin.x) * 0.5f, (bmax.y - bmin.y) * 0.5f, (bmax.z - bmin.z) * 0.5f};
    const V3 v0 = sub(t.v[0], c), v1 = sub(t.v[1], c), v2 = sub(t.v[2], c);
    const V3 e[3] = {sub(v1, v0), sub(v2, v1), sub(v0, v2)};
    for (int k = 0; k < 3; ++k) {
        const V3 ax[3] = {{0.0f, -e[k].z, e[k].y}, {e[k].z, 0.0f, -e[k].x}, {-e[k].y, e[k].x, 0.0f}};
        for (int j = 0; j < 3; ++j) {
            const float p0 = dot(v0, ax[j]), p1 = dot(v1, ax[j]), p2 = dot(v2, ax[j]);
            const float r = h.x * std::fabs(ax[j].x) + h.y * std::fabs(ax[j].y) + h.z * std::fabs(ax[j].z);
            if (std::max(-mx3(p0, p1, p2), mn3(p0, p1, p2)) > r) return false;
        }
    }
    if (mx3(v0.x, v1.x, v2.x) < -h.x || mn3(v0.x, v1.x, v2.x) > h.x) return false;
    if (mx3(v0.y, v1.y, v2.y) < -h.y || mn3(v0.y, v1.y, v2.y) > h.y) return false;
    if (mx3(v0.z, v1.z, v2.z) < -h.z || mn3(v0.z, v1.z, v2.z) > h.z) return false;
    const V3 n = cross(e[0], e[1]);
    const float d = dot(n, v0);
    const float r = h.x * std::fabs(n.x) + h.y * std::fabs(n.y) + h.z * std::fabs(n.z);
    return !(std::fabs(d) > r);
}

// Surface voxelisation: centres of the cells of a (longest side / resolution)
// grid over the vertices' bounding box that a triangle overlaps.
void voxelize(const Mesh& m, float resolution, std::vector<V3>& out) {
    out.clear();
    if (m.verts.empty()) return;
    V3 lo = m.verts[0], hi = m.verts[0];
    for (const V3& p : m.verts) {
        lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
        hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
    }
    const float maxsize = std::max(hi.x - lo.x, std::max(hi.y - lo.y, hi.z - lo.z));
    const float cell = maxsize / resolution;
    const int gx = (int)std::ceil((hi.x - lo.x) / cell);
    const int gy = (int)std::ceil((hi.y - lo.y) / cell);
    const int gz = (int)std::ceil((hi.z - lo.z) / cell);
    if (gx <= 0 || gy <= 0 || gz <= 0) return;
    std::vector<uint8_t> hit((size_t)gx * gy * gz, 0);
    for (const Tri& t : m.tris) {
        const float tminx = mn3(t.v[0].x, t.v[1].x, t.v[2].x), tmaxx = mx3(t.v[0].x, t.v[1].x, t.v[2].x);
        const float tminy = mn3(t.v[0].y, t.v[1].y, t.v[2].y), tmaxy = mx3(t.v[0].y, t.v[1].y, t.v[2].y);
        const float tminz = mn3(t.v[0].z, t.v[1].z, t.v[2].z), tmaxz = mx3(t.v[0].z, t.v[1].z, t.v[2].z);
        const int x0 = std::max(0, (int)std::floor((tminx - lo.x) / cell));
        const int y0 = std::max(0, (int)std::floor((tminy - lo.y) / cell));
        const int z0 = std::max(0, (int)std::floor((tminz - lo.z) / cell));
        const int x1 = std::min(gx - 1, (int)std::ceil((tmaxx - lo.x) / cell));
        const int y1 = std::min(gy - 1, (int)std::ceil((tmaxy - lo.y) / cell));
        const int z1 = std::min(gz - 1, (int)std::ceil((tmaxz - lo.z) / cell));
        for (int z = z0; z <= z1; ++z)
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    const V3 bmin = {lo.x + x * cell, lo.y + y * cell, lo.z + z * cell};
                    const V3 bmax = {bmin.x + cell, bmin.y + cell, bmin.z + cell};
                    if (tri_box(t, bmin, bmax)) hit[((size_t)z * gy + y) * gx + x] = 1;
                }
    }
    for (int z = 0; z < gz; ++z)
        for (int y = 0; y < gy; ++y)
            for (int x = 0; x < gx; ++x)
                if (hit[((size_t)z * gy + y) * gx + x])
                    out.push_back({lo.x + (x + 0.5f) * cell, lo.y + (y + 0.5f) * cell, lo.z + (z + 0.5f) * cell});
}

// The affine map of both normalize() variants, as applied point by point:
// centre, scale to normalizeSize, then to `scale`, then displace.  `bounds`
// gives the point set's bounding box before each scaling stage.
struct Affine {
    V3 centre;
    float s1;
    float half;
    float s2;
    V3 disp;
};

template <class ForEach> bool make_affine(ForEach for_each, float size, float scale, const V3& disp, Affine& A) {
    bool first = true;
    V3 lo{}, hi{};
    for_each([&](const V3& p) {
        if (first) { lo = hi = p; first = false; }
        lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
        hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
    });
    const float maxdim = mx3(hi.x - lo.x, hi.y - lo.y, hi.z - lo.z);
    if (first || maxdim < 1e-6f) return false;
    A.s1 = size / maxdim;
    A.centre = {(lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f};
    A.half = size * 0.5f;
    A.disp = disp;
    // bounds after the first stage
    first = true;
    for_each([&](const V3& p) {
        const V3 q = {(p.x - A.centre.x) * A.s1 + A.half, (p.y - A.centre.y) * A.s1 + A.half,
                      (p.z - A.centre.z) * A.s1 + A.half};
        if (first) { lo = hi = q; first = false; }
        lo = {std::min(lo.x, q.x), std::min(lo.y, q.y), std::min(lo.z, q.z)};
        hi = {std::max(hi.x, q.x), std::max(hi.y, q.y), std::max(hi.z, q.z)};
    });
    const float cur = mx3(hi.x - lo.x, hi.y - lo.y, hi.z - lo.z);
    A.s2 = cur > 1e-6f ? scale / cur : 1.0f;
    return true;
}

inline V3 apply(const Affine& A, V3 p) {
    // same operation order as the reference: subtract, multiply, add, per stage
    p.x -= A.centre.x; p.y -= A.centre.y; p.z -= A.centre.z;
    p.x *= A.s1; p.y *= A.s1; p.z *= A.s1;
    p.x += A.half; p.y += A.half; p.z += A.half;
    p.x -= A.half; p.y -= A.half; p.z -= A.half;
    p.x *= A.s2; p.y *= A.s2; p.z *= A.s2;
    p.x += A.half; p.y += A.half; p.z += A.half;
    p.x += A.disp.x; p.y += A.disp.y; p.z += A.disp.z;
    return p;
}

struct Lattice {
    int x, y, z;
    bool operator<(const Lattice& o) const {
        if (x != o.x) return x < o.x;
        if (y != o.y) return y < o.y;
        return z < o.z;
    }
    bool operator==(const Lattice& o) const { return x == o.x && y == o.y && z == o.z; }
};

// Water from an OBJ: voxelize(mesh, 100), normalize(normalizeSize, scale,
// displacement) with lattice step `res`; lattice points to out.
int water_points(const std::string& path, float size, float scale, const V3& disp, int res, std::vector<V3>& out) {
    if (res <= 0) return err(MPM_EINVAL, "RESOLUTION must be at least 1 (it is read as an integer lattice step)");
    Mesh m;
    int rc = load_obj(path, m);
    if (rc) return rc;
    std::vector<V3> vox;
    voxelize(m, 100.0f, vox);  // Simulation::Prepare_object, Simulation.cpp:84
    if (vox.empty()) return err(MPM_EIO, path + ": the mesh gives no voxels");
    Affine A;
    if (!make_affine([&](auto f) { for (const V3& p : vox) f(p); }, size, scale, disp, A))
        return err(MPM_EIO, path + ": voxels have zero extent");
    const float gres = (float)res;
    const int maxi = (int)(size / gres);
    std::vector<Lattice> pts;
    pts.reserve(vox.size() * 125);
    for (const V3& p0 : vox) {
        const V3 p = apply(A, p0);
        int gx = (int)std::round(p.x / gres), gy = (int)std::round(p.y / gres), gz = (int)std::round(p.z / gres);
        gx = std::max(0, std::min(maxi, gx));
        gy = std::max(0, std::min(maxi, gy));
        gz = std::max(0, std::min(maxi, gz));
        for (int dx = -2; dx <= 2; ++dx)
            for (int dy = -2; dy <= 2; ++dy)
                for (int dz = -2; dz <= 2; ++dz) {
                    const int x = gx + dx, y = gy + dy, z = gz + dz;
                    if (x >= 0 && x <= maxi && y >= 0 && y <= maxi && z >= 0 && z <= maxi) pts.push_back({x, y, z});
                }
    }
    std::sort(pts.begin(), pts.end());
    pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
    out.clear();
    out.reserve(pts.size());
    for (const Lattice& q : pts) out.push_back({q.x * gres, q.y * gres, q.z * gres});
    return MPM_OK;
}

// Collider triangles from an OBJ: all triangles, the affine map, then the
// triangles with at least one vertex inside [0, size]^3.
int collider_tris(const std::string& path, float size, float scale, const V3& disp, std::vector<Tri>& out) {
    Mesh m;
    int rc = load_obj(path, m);
    if (rc) return rc;
    out.clear();
    if (m.tris.empty()) return MPM_OK;  // the reference warns and continues with none
    Affine A;
    if (!make_affine([&](auto f) { for (const Tri& t : m.tris) for (const V3& p : t.v) f(p); }, size, scale, disp, A))
        return MPM_OK;  // zero extent: the reference leaves them unchanged; nothing sensible to collide with
    auto inside = [&](const V3& p) { return p.x >= 0.0f && p.x <= size && p.y >= 0.0f && p.y <= size && p.z >= 0.0f && p.z <= size; };
    for (const Tri& t0 : m.tris) {
        Tri t;
        for (int k = 0; k < 3; ++k) t.v[k] = apply(A, t0.v[k]);
        if (inside(t.v[0]) || inside(t.v[1]) || inside(t.v[2])) out.push_back(t);
    }
    return MPM_OK;
}

std::string resolve(const std::string& p, const std::string& base_dir) {
    if (p.empty() || p[0] == '/' || base_dir.empty()) return p;
    std::ifstream f(p);
    if (f) return p;  // relative to the working directory, as the reference opens it
    return base_dir + "/" + p;
}

}  // namespace

extern "C" {

int mpm_obj_triangles(const char* path, mpm_triangle* out, int cap, int* count) {
    if (!path || !count || cap < 0 || (cap > 0 && !out)) return err(MPM_EINVAL, "bad arguments");
    Mesh m;
    int rc = load_obj(path, m);
    if (rc) return rc;
    *count = (int)m.tris.size();
    for (int i = 0; i < (int)m.tris.size() && i < cap; ++i) {
        float* d = &out[i].v0[0];
        for (int k = 0; k < 3; ++k) {
            d[3 * k] = m.tris[i].v[k].x;
            d[3 * k + 1] = m.tris[i].v[k].y;
            d[3 * k + 2] = m.tris[i].v[k].z;
        }
    }
    return MPM_OK;
}

int mpm_voxelize_obj(const char* path, float resolution, float* xyz, int64_t cap, int64_t* count) {
    if (!path || !count || cap < 0 || (cap > 0 && !xyz) || !(resolution > 0.0f)) return err(MPM_EINVAL, "bad arguments");
    Mesh m;
    int rc = load_obj(path, m);
    if (rc) return rc;
    std::vector<V3> vox;
    voxelize(m, resolution, vox);
    *count = (int64_t)vox.size();
    for (int64_t i = 0; i < (int64_t)vox.size() && i < cap; ++i) {
        xyz[i] = vox[i].x;
        xyz[cap + i] = vox[i].y;
        xyz[2 * cap + i] = vox[i].z;
    }
    return MPM_OK;
}

int mpm_obj_water(const char* path, float normalize_size, float scale, const float* displacement, int lattice,
                  float* xyz, int64_t cap, int64_t* count) {
    if (!path || !displacement || !count || cap < 0 || (cap > 0 && !xyz)) return err(MPM_EINVAL, "bad arguments");
    std::vector<V3> pts;
    int rc = water_points(path, normalize_size, scale, {displacement[0], displacement[1], displacement[2]}, lattice, pts);
    if (rc) return rc;
    *count = (int64_t)pts.size();
    for (int64_t i = 0; i < (int64_t)pts.size() && i < cap; ++i) {
        xyz[i] = pts[i].x;
        xyz[cap + i] = pts[i].y;
        xyz[2 * cap + i] = pts[i].z;
    }
    return MPM_OK;
}

int mpm_obj_collider(const char* path, float normalize_size, float scale, const float* displacement,
                     mpm_triangle* out, int cap, int* count) {
    if (!path || !displacement || !count || cap < 0 || (cap > 0 && !out)) return err(MPM_EINVAL, "bad arguments");
    std::vector<Tri> tris;
    int rc = collider_tris(path, normalize_size, scale, {displacement[0], displacement[1], displacement[2]}, tris);
    if (rc) return rc;
    *count = (int)tris.size();
    for (int i = 0; i < (int)tris.size() && i < cap; ++i) {
        float* d = &out[i].v0[0];
        for (int k = 0; k < 3; ++k) {
            d[3 * k] = tris[i].v[k].x;
            d[3 * k + 1] = tris[i].v[k].y;
            d[3 * k + 2] = tris[i].v[k].z;
        }
    }
    return MPM_OK;
}

int mpm_scene_from_config(const char* config_path, mpm_sim_config* cfg, float* soa26, int64_t cap, int64_t* n,
                          mpm_triangle* tris, int tri_cap, int* ntris) {
    if (!config_path || !cfg || !n || !ntris || cap < 0 || tri_cap < 0 || (cap > 0 && !soa26) ||
        (tri_cap > 0 && !tris))
        return err(MPM_EINVAL, "bad arguments");
    std::vector<mpm_config_object> waters(64), objects(64);
    int nw = 0, no = 0;
    int rc = mpm_config_parse_file(config_path, cfg, waters.data(), &nw, objects.data(), &no, 64);
    if (rc) return rc;
    if (nw > 64 || no > 64) return err(MPM_ELIMIT, "more than 64 WATER or OBJECT entries");
    const std::string cp(config_path);
    const size_t slash = cp.find_last_of('/');
    const std::string dir = slash == std::string::npos ? std::string() : cp.substr(0, slash);
    const float size = (float)cfg->SIZE_X;  // normalize(..., getInt("SIZE_X"), ...), Simulation.cpp:88, :105
    const int lattice = (int)cfg->RESOLUTION;  // getInt("RESOLUTION"): the integer part, VoxelEngine.cpp:369
    std::vector<V3> all;
    for (int i = 0; i < nw; ++i) {
        std::vector<V3> pts;
        const mpm_config_object& w = waters[i];
        if ((rc = water_points(resolve(w.model_path, dir), size, w.scale, {w.x, w.y, w.z}, lattice, pts))) return rc;
        all.insert(all.end(), pts.begin(), pts.end());
    }
    std::vector<Tri> coll;
    for (int i = 0; i < no; ++i) {
        std::vector<Tri> t;
        const mpm_config_object& o = objects[i];
        if ((rc = collider_tris(resolve(o.model_path, dir), size, o.scale, {o.x, o.y, o.z}, t))) return rc;
        coll.insert(coll.end(), t.begin(), t.end());
    }
    *n = (int64_t)all.size();
    *ntris = (int)coll.size();
    if (cap >= (int64_t)all.size()) {  // VoxelPipeline::process: positions, zeros, VOL = MASS = 1
        const int64_t N = (int64_t)all.size();
        for (int64_t i = 0; i < N; ++i) {
            soa26[i] = all[i].x;
            soa26[N + i] = all[i].y;
            soa26[2 * N + i] = all[i].z;
        }
        std::fill(soa26 + 3 * N, soa26 + 24 * N, 0.0f);
        std::fill(soa26 + 24 * N, soa26 + 26 * N, 1.0f);
    }
    for (int i = 0; i < (int)coll.size() && i < tri_cap; ++i) {
        float* d = &tris[i].v0[0];
        for (int k = 0; k < 3; ++k) {
            d[3 * k] = coll[i].v[k].x;
            d[3 * k + 1] = coll[i].v[k].y;
            d[3 * k + 2] = coll[i].v[k].z;
        }
    }
    return MPM_OK;
}

}  // extern "C"
