// oracle.cpp -- C ABI over the CPU restatement in oracle.hpp.
//
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).  Built by oracle/build.sh into
// oracle/liboracle.so and loaded with ctypes by the tests, the smoke check and
// the CPU-baseline leg of the benchmark.
//
// Two step drivers:
//  * serial (nthreads <= 1): Engine::step, engine.cu:54-70 -- keys, stable sort,
//    clear, P2G in sorted order, grid update, G2P.
//  * OpenMP (nthreads > 1): the same arithmetic per particle and per node; P2G
//    runs over 8 colours of 4x4x4-cell blocks so no two threads write the same
//    node, which keeps it race-free and deterministic.  It is the timed CPU
//    baseline.
#include "oracle.hpp"

#include <chrono>
#include <cstdio>
#include <memory>
#include <numeric>
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#endif

using namespace oracle;

namespace {

template <class T> struct GridStore {
    std::vector<T> planes;  // mass, mom_x, mom_y, mom_z (4 * G)
    size_t G;
    explicit GridStore(size_t g) : planes(4 * g, (T)0), G(g) {}
    GridView<T> view(const T* sdf) {
        return GridView<T>{planes.data(), {planes.data() + G, planes.data() + 2 * G, planes.data() + 3 * G}, sdf};
    }
};

template <class T> void compute_keys(ParticleView<T> p, int32_t* keys, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static) if (nthreads > 1)
    for (int i = 0; i < p.n; ++i)
        keys[i] = morton_key((float)p.at(POS_X, i), (float)p.at(POS_Y, i), (float)p.at(POS_Z, i));
}

void stable_perm(int n, const int32_t* keys, int32_t* perm, int nthreads) {
    if (nthreads <= 1) {
        sort_perm(n, keys, perm);
        return;
    }
#ifdef _OPENMP
    std::vector<int64_t> kv(n);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int i = 0; i < n; ++i) kv[i] = ((int64_t)keys[i] << 32) | (uint32_t)i;
    __gnu_parallel::sort(kv.begin(), kv.end());
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int i = 0; i < n; ++i) perm[i] = (int32_t)(kv[i] & 0xFFFFFFFF);
#else
    sort_perm(n, keys, perm);
#endif
}

// P2G over the sorted order, serial.
template <class T> void p2g_serial(const Config& c, ParticleView<T> p, const int32_t* perm, GridView<T> g) {
    for (int t = 0; t < p.n; ++t) {
        P2GParticle<T> q = p2g_prepare<T>(c, p, perm[t]);
        p2g_scatter<T>(c, q, g);
    }
}

// P2G over 8 colours of 4^3-cell blocks of the particles' base cells.  A block
// writes nodes [4b, 4b+6) per axis; same-colour blocks start 8 cells apart, so
// their node ranges never overlap.  Block coordinates are clamped into the
// grid's block range, which only widens the first/last block's write range
// towards the outside of the grid.
template <class T>
void p2g_colored(const Config& c, ParticleView<T> p, const int32_t* perm, GridView<T> g, int nthreads) {
    const int nbx = ((int)c.SIZE_X + 3) / 4, nby = ((int)c.SIZE_Y + 3) / 4, nbz = ((int)c.SIZE_Z + 3) / 4;
    const int nb = nbx * nby * nbz;
    std::vector<int32_t> blk(p.n);
    auto fdiv4 = [](int v) { return v >= 0 ? v / 4 : -((-v + 3) / 4); };
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int t = 0; t < p.n; ++t) {
        int i = perm[t];
        int bx = fdiv4(base_of(p.at(POS_X, i))), by = fdiv4(base_of(p.at(POS_Y, i))), bz = fdiv4(base_of(p.at(POS_Z, i)));
        bx = std::max(0, std::min(bx, nbx - 1));
        by = std::max(0, std::min(by, nby - 1));
        bz = std::max(0, std::min(bz, nbz - 1));
        blk[t] = (bz * nby + by) * nbx + bx;
    }
    // counting sort of sorted positions by block (stable)
    std::vector<int32_t> start(nb + 1, 0), order(p.n);
    for (int t = 0; t < p.n; ++t) start[blk[t] + 1]++;
    for (int b = 0; b < nb; ++b) start[b + 1] += start[b];
    {
        std::vector<int32_t> fill(start.begin(), start.end() - 1);
        for (int t = 0; t < p.n; ++t) order[fill[blk[t]]++] = perm[t];
    }
    for (int color = 0; color < 8; ++color) {
        const int cx = color & 1, cy = (color >> 1) & 1, cz = (color >> 2) & 1;
        const int hx = (nbx - cx + 1) / 2, hy = (nby - cy + 1) / 2, hz = (nbz - cz + 1) / 2;
        const int count = hx * hy * hz;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
        for (int s = 0; s < count; ++s) {
            int bx = cx + 2 * (s % hx), by = cy + 2 * ((s / hx) % hy), bz = cz + 2 * (s / (hx * hy));
            int b = (bz * nby + by) * nbx + bx;
            for (int t = start[b]; t < start[b + 1]; ++t) {
                P2GParticle<T> q = p2g_prepare<T>(c, p, order[t]);
                p2g_scatter<T>(c, q, g);
            }
        }
    }
}

template <class T> void grid_update_all(const Config& c, GridView<T> g, int nthreads) {
    const int G = (int)(c.SIZE_X * c.SIZE_Y * c.SIZE_Z);
#pragma omp parallel for num_threads(nthreads) schedule(static) if (nthreads > 1)
    for (int i = 0; i < G; ++i) grid_update_node<T>(c, g, i);
}

template <class T> void g2p_all(const Config& c, ParticleView<T> p, GridView<T> g, int nthreads) {
#pragma omp parallel for num_threads(nthreads) schedule(static) if (nthreads > 1)
    for (int i = 0; i < p.n; ++i) g2p_particle<T>(c, p, g, i);
}

// Engine::step, engine.cu:54-70, repeated nsteps times.
template <class T> void run_steps(const Config& c, int n, T* soa, const T* sdf, int nsteps, int nthreads) {
    ParticleView<T> p{soa, n};
    const size_t G = (size_t)c.SIZE_X * c.SIZE_Y * c.SIZE_Z;
    GridStore<T> grid(G);
    std::vector<int32_t> keys(n), perm(n);
    for (int s = 0; s < nsteps; ++s) {
        compute_keys<T>(p, keys.data(), nthreads);
        stable_perm(n, keys.data(), perm.data(), nthreads);
        std::fill(grid.planes.begin(), grid.planes.end(), (T)0);
        GridView<T> g = grid.view(sdf);
        if (nthreads > 1) p2g_colored<T>(c, p, perm.data(), g, nthreads);
        else p2g_serial<T>(c, p, perm.data(), g);
        grid_update_all<T>(c, g, nthreads);
        g2p_all<T>(c, p, g, nthreads);
    }
}

}  // namespace

extern "C" {

int oracle_version(void) { return 1; }

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

float oracle_spline_f32(float x) { return spline(x); }
double oracle_spline_f64(double x) { return spline(x); }
uint32_t oracle_expand_bits(uint32_t v) { return expand_bits(v); }
uint32_t oracle_morton(uint32_t x, uint32_t y, uint32_t z) { return morton(x, y, z); }

// setKeys over caller-ordered positions (three planes of n floats).
void oracle_morton_keys(int n, const float* px, const float* py, const float* pz, int32_t* keys) {
    for (int i = 0; i < n; ++i) keys[i] = morton_key(px[i], py[i], pz[i]);
}

// Stable ascending permutation of int keys.
void oracle_sort_perm(int n, const int32_t* keys, int32_t* perm) { sort_perm(n, keys, perm); }

// initFMatrices semantics on a caller-owned 26*n buffer.
void oracle_init_state_f32(const Config* c, int n, float* soa) { init_state<float>(*c, ParticleView<float>{soa, n}); }
void oracle_init_state_f64(const Config* c, int n, double* soa) { init_state<double>(*c, ParticleView<double>{soa, n}); }

void oracle_steps_f32(const Config* c, int n, float* soa, const float* sdf, int nsteps, int nthreads) {
    run_steps<float>(*c, n, soa, sdf, nsteps, nthreads);
}
void oracle_steps_f64(const Config* c, int n, double* soa, const double* sdf, int nsteps, int nthreads) {
    run_steps<double>(*c, n, soa, sdf, nsteps, nthreads);
}

// Stage entry points.  grid4 = 4*G values (mass, mom_x, mom_y, mom_z planes).
void oracle_p2g_f32(const Config* c, int n, float* soa, const int32_t* perm, float* grid4) {
    const size_t G = (size_t)c->SIZE_X * c->SIZE_Y * c->SIZE_Z;
    GridView<float> g{grid4, {grid4 + G, grid4 + 2 * G, grid4 + 3 * G}, nullptr};
    std::vector<int32_t> iota;
    if (!perm) { iota.resize(n); std::iota(iota.begin(), iota.end(), 0); perm = iota.data(); }
    p2g_serial<float>(*c, ParticleView<float>{soa, n}, perm, g);
}
void oracle_p2g_f64(const Config* c, int n, double* soa, const int32_t* perm, double* grid4) {
    const size_t G = (size_t)c->SIZE_X * c->SIZE_Y * c->SIZE_Z;
    GridView<double> g{grid4, {grid4 + G, grid4 + 2 * G, grid4 + 3 * G}, nullptr};
    std::vector<int32_t> iota;
    if (!perm) { iota.resize(n); std::iota(iota.begin(), iota.end(), 0); perm = iota.data(); }
    p2g_serial<double>(*c, ParticleView<double>{soa, n}, perm, g);
}
void oracle_grid_update_f32(const Config* c, float* grid4, const float* sdf) {
    const size_t G = (size_t)c->SIZE_X * c->SIZE_Y * c->SIZE_Z;
    grid_update_all<float>(*c, GridView<float>{grid4, {grid4 + G, grid4 + 2 * G, grid4 + 3 * G}, sdf}, 1);
}
void oracle_grid_update_f64(const Config* c, double* grid4, const double* sdf) {
    const size_t G = (size_t)c->SIZE_X * c->SIZE_Y * c->SIZE_Z;
    grid_update_all<double>(*c, GridView<double>{grid4, {grid4 + G, grid4 + 2 * G, grid4 + 3 * G}, sdf}, 1);
}
void oracle_g2p_f32(const Config* c, int n, float* soa, const float* grid4, const float* sdf) {
    const size_t G = (size_t)c->SIZE_X * c->SIZE_Y * c->SIZE_Z;
    float* g4 = const_cast<float*>(grid4);
    g2p_all<float>(*c, ParticleView<float>{soa, n}, GridView<float>{g4, {g4 + G, g4 + 2 * G, g4 + 3 * G}, sdf}, 1);
}
void oracle_g2p_f64(const Config* c, int n, double* soa, const double* grid4, const double* sdf) {
    const size_t G = (size_t)c->SIZE_X * c->SIZE_Y * c->SIZE_Z;
    double* g4 = const_cast<double*>(grid4);
    g2p_all<double>(*c, ParticleView<double>{soa, n}, GridView<double>{g4, {g4 + G, g4 + 2 * G, g4 + 3 * G}, sdf}, 1);
}

// Single-node helpers exposed for closed-form tests.
float oracle_get_sdf_f32(const Config* c, float x, float y, float z, const float* sdf) {
    return get_sdf<float>(*c, {x, y, z}, sdf);
}
float oracle_solid_angle_f32(const float* P, const float* tri) {
    return solid_angle<float>({P[0], P[1], P[2]}, {tri[0], tri[1], tri[2]}, {tri[3], tri[4], tri[5]},
                              {tri[6], tri[7], tri[8]});
}
float oracle_point_tri_dist2_f32(const float* P, const float* tri) {
    return point_tri_dist2<float>({P[0], P[1], P[2]}, {tri[0], tri[1], tri[2]}, {tri[3], tri[4], tri[5]},
                                  {tri[6], tri[7], tri[8]});
}
double oracle_point_tri_dist2_f64(const double* P, const double* tri) {
    return point_tri_dist2<double>({P[0], P[1], P[2]}, {tri[0], tri[1], tri[2]}, {tri[3], tri[4], tri[5]},
                                   {tri[6], tri[7], tri[8]});
}

// makeSDF over the whole grid (nthreads > 1 parallelises over nodes; each
// node's triangle loop stays in triangle order).
void oracle_make_sdf_f32(const Config* c, const float* tris, int nt, float* sdf, int nthreads) {
    const int G = (int)(c->SIZE_X * c->SIZE_Y * c->SIZE_Z);
    const Tri* t = reinterpret_cast<const Tri*>(tris);
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(dynamic, 256) if (nthreads > 1)
    for (int i = 0; i < G; ++i) sdf[i] = make_sdf_node<float>(*c, i, t, nt);
}

// Time `nsteps` OpenMP substeps of the float oracle (the CPU baseline); returns
// seconds of wall time for the stepping only.
double oracle_time_steps_f32(const Config* c, int n, float* soa, const float* sdf, int nsteps, int nthreads) {
    auto t0 = std::chrono::steady_clock::now();
    run_steps<float>(*c, n, soa, sdf, nsteps, nthreads);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
