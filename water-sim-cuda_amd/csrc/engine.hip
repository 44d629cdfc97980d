// engine.hip -- host side of the engine and the C ABI declared in include/mpm.h.
//
// Replaces the reference's Engine class (engine.h:14-43, engine.cu:36-114).
// Differences of design (not of results):
//  * the particle state lives in a chunked, Morton-sorted slot layout and is
//    physically reordered at each re-sort (the reference only permutes an
//    index array, engine.cu:98-99); the caller's order is kept in `orig`;
//  * a substep is P2G + grid update + G2P as in engine.cu:54-70, but within a
//    k-substep call the G2P of substep s and the P2G of substep s+1 run as one
//    fused kernel on two grid buffers;
//  * the grid is cleared and updated only where the previous P2G wrote
//    (4^3-node activity flags), instead of a dense pass over all nodes;
//  * errors are returned as codes (the reference calls exit(1)).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"
#include "mpm.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(MPM_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

enum Stage { ST_P2G = 0, ST_G2P = 1, ST_G2P2G = 2, ST_GRID = 3, ST_SORT = 4, ST_COUNT = 5 };

}  // namespace

namespace mpm {
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace mpm

using namespace mpm;

struct mpm_engine {
    mpm_sim_config cfg{};
    DevCfg dc{};
    int device = 0;
    hipStream_t stream = nullptr;

    int n = 0;
    bool have_particles = false;
    int chunk_cap = 512;
    int resort_every = 1;
    bool timing = false;

    size_t G = 0;
    int nfb = 0;

    float* planes[2] = {nullptr, nullptr};
    int* orig[2] = {nullptr, nullptr};
    size_t plane_cap[2] = {0, 0};
    int pcur = 0;

    ChunkDesc* chunks = nullptr;
    uint8_t* rounds = nullptr;  // particles per round of each chunk, at chunk.start + round
    int chunk_limit = 0;  // bound on the chunk count (min(n, 4^3 blocks of the grid) + n / CHUNK + 1)
    int chunk_est = 0;    // chunk count at the last host synchronisation (sizes launches)
    int key_bits = 0;     // bits of the engine's base-cell keys

    GridBuf grid[2] = {};
    int vcur = 0;  // grid[vcur]: velocities of the last update; grid[1 - vcur]: accumulates
    float* sdf = nullptr;

    SortScratch sc{};
    uint32_t* overflow = nullptr;  // = sc.counters + 8: [0] tile-overflow particles, [1] extra claim passes
    uint32_t* pinned = nullptr;

    int64_t substeps = 0;
    int64_t sorted_at = 0;
    int64_t last_overflow = 0;
    int64_t last_claim_losers = 0;
    int64_t last_rounds = 0;

    struct Ev {
        hipEvent_t a, b;
        int stage;
    };
    std::vector<Ev> evs;
    std::vector<hipEvent_t> pool;
    double stage_ms[ST_COUNT] = {};
    int64_t stage_cnt[ST_COUNT] = {};
};

namespace {

hipEvent_t take_event(mpm_engine* h) {
    if (!h->pool.empty()) {
        hipEvent_t e = h->pool.back();
        h->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

struct StageTimer {
    mpm_engine* h;
    int stage;
    hipEvent_t a = nullptr;
    StageTimer(mpm_engine* h_, int s) : h(h_), stage(s) {
        if (h->timing) {
            a = take_event(h);
            hipEventRecord(a, h->stream);
        }
    }
    ~StageTimer() {
        if (h->timing && a) {
            hipEvent_t b = take_event(h);
            hipEventRecord(b, h->stream);
            h->evs.push_back({a, b, stage});
        }
    }
};

template <class T> int dev_alloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) return fail(MPM_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return MPM_OK;
}

template <class T> void dev_free(T*& p) {
    if (p) hipFree(p);
    p = nullptr;
}

void free_particles(mpm_engine* h) {
    for (int i = 0; i < 2; ++i) {
        dev_free(h->planes[i]);
        dev_free(h->orig[i]);
        h->plane_cap[i] = 0;
    }
    dev_free(h->chunks);
    dev_free(h->rounds);
    dev_free(h->sc.keys[0]);
    dev_free(h->sc.keys[1]);
    dev_free(h->sc.vals[0]);
    dev_free(h->sc.vals[1]);
    dev_free(h->sc.hist);
    dev_free(h->sc.partial);
    dev_free(h->sc.heads);
    dev_free(h->sc.seg_start);
    dev_free(h->sc.seg_nch);
    h->sc.cap = h->sc.hist_cap = h->sc.partial_cap = 0;
    h->chunk_limit = 0;
    h->chunk_est = 0;
    h->have_particles = false;
    h->n = 0;
}

int ensure_sort_scratch(mpm_engine* h, size_t n) {
    SortScratch& sc = h->sc;
    if (sc.cap >= n + 1 && sc.keys[0]) return MPM_OK;
    dev_free(sc.keys[0]); dev_free(sc.keys[1]); dev_free(sc.vals[0]); dev_free(sc.vals[1]);
    dev_free(sc.hist); dev_free(sc.partial); dev_free(sc.heads); dev_free(sc.seg_start); dev_free(sc.seg_nch);
    const size_t cap = n + 1;
    int rc;
    if ((rc = dev_alloc(&sc.keys[0], cap)) || (rc = dev_alloc(&sc.keys[1], cap)) || (rc = dev_alloc(&sc.vals[0], cap)) ||
        (rc = dev_alloc(&sc.vals[1], cap)) || (rc = dev_alloc(&sc.heads, cap)) || (rc = dev_alloc(&sc.seg_start, cap + 1)) ||
        (rc = dev_alloc(&sc.seg_nch, cap)))
        return rc;
    sc.hist_cap = radix_hist_elems(cap);
    sc.partial_cap = scan_partial_elems(std::max(cap, sc.hist_cap)) + 16;
    if ((rc = dev_alloc(&sc.hist, sc.hist_cap)) || (rc = dev_alloc(&sc.partial, sc.partial_cap))) return rc;
    sc.cap = cap;
    return MPM_OK;
}

int ensure_planes(mpm_engine* h, int which, size_t slots) {
    if (h->plane_cap[which] >= slots && h->planes[which]) return MPM_OK;
    dev_free(h->planes[which]);
    dev_free(h->orig[which]);
    int rc;
    if ((rc = dev_alloc(&h->planes[which], (size_t)A_STATE * slots)) || (rc = dev_alloc(&h->orig[which], slots)))
        return rc;
    h->plane_cap[which] = slots;
    return MPM_OK;
}

// Copies the 16 device counters (SortScratch::counters) into h->pinned and
// waits for the stream.
int read_counters(mpm_engine* h) {
    HIPCHK(hipMemcpyAsync(h->pinned, h->sc.counters, 16 * sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

// Sorts sc.keys[0]/vals[0] (n base-cell keys), builds the chunk table and
// gathers the particles from (src, src_stride, src_orig) into the alternate
// plane buffer.  Nothing here waits for the device: the chunk count stays in
// device memory and is bounded by chunk_limit by construction.
int sort_and_reorder(mpm_engine* h, const float* src, size_t src_stride, const int* src_orig, uint32_t mask) {
    const int n = h->n;
    const int r = radix_sort_pairs(h->sc, n, h->key_bits, h->stream);
    build_chunks(h->sc, r, n, h->chunk_cap, h->chunks, h->chunk_limit, h->stream);
    const int dst = 1 - h->pcur;
    launch_reorder(h->chunks, h->sc.counters + 3, h->chunk_est, h->sc.keys[r], h->sc.vals[r], src, src_stride,
                   src_orig, h->planes[dst], h->plane_cap[dst], h->orig[dst], h->rounds, h->sc.counters + 4, mask,
                   h->stream);
    HIPCHK(hipGetLastError());
    h->pcur = dst;
    return MPM_OK;
}

constexpr uint32_t MASK_ALL = (1u << A_STATE) - 1u;
constexpr uint32_t MASK_POS_F = 0x7u | (0x1FFu << A_F0);

int resort(mpm_engine* h, bool full) {
    StageTimer t(h, ST_SORT);
    const float* pl = h->planes[h->pcur];
    const size_t st = h->plane_cap[h->pcur];
    launch_keys_linear(pl + A_PX * st, pl + A_PY * st, pl + A_PZ * st, h->n, true, h->dc, h->sc.keys[0],
                       h->sc.vals[0], h->stream);
    return sort_and_reorder(h, h->planes[h->pcur], h->plane_cap[h->pcur], h->orig[h->pcur], full ? MASK_ALL : MASK_POS_F);
}

TransferArgs make_args(mpm_engine* h) {
    TransferArgs a{};
    a.c = h->dc;
    a.part = h->planes[h->pcur];
    a.stride = h->plane_cap[h->pcur];
    a.chunks = h->chunks;
    a.nchunks = h->sc.counters + 3;
    a.rounds = h->rounds;
    const GridBuf& V = h->grid[h->vcur];
    const GridBuf& M = h->grid[1 - h->vcur];
    a.vel[0] = V.plane[1];
    a.vel[1] = V.plane[2];
    a.vel[2] = V.plane[3];
    for (int i = 0; i < 4; ++i) a.acc[i] = M.plane[i];
    a.acc_flags = M.flags;
    a.sdf = h->sdf;
    a.overflow = h->overflow;
    return a;
}

void run_transfer(mpm_engine* h, int mode) {
    const int stage = mode == XFER_P2G ? ST_P2G : (mode == XFER_G2P ? ST_G2P : ST_G2P2G);
    StageTimer t(h, stage);
    launch_transfer(mode, make_args(h), h->chunk_est, h->stream);
}

void run_grid_update(mpm_engine* h) {
    StageTimer t(h, ST_GRID);
    launch_grid_update(h->dc, h->grid[1 - h->vcur], h->grid[h->vcur], true, h->sdf, h->nfb, h->stream);
    h->vcur = 1 - h->vcur;
}

int zero_grid(mpm_engine* h, int which) {
    for (int p = 0; p < 4; ++p) HIPCHK(hipMemsetAsync(h->grid[which].plane[p], 0, h->G * sizeof(float), h->stream));
    HIPCHK(hipMemsetAsync(h->grid[which].flags, 0, (size_t)h->nfb, h->stream));
    return MPM_OK;
}

int check_engine(mpm_engine* h, bool need_particles) {
    if (!h) return fail(MPM_EINVAL, "null engine handle");
    if (need_particles && !h->have_particles) return fail(MPM_ESTATE, "particles not initialised (call mpm_engine_init_particles)");
    HIPCHK(hipSetDevice(h->device));
    return MPM_OK;
}

}  // namespace

extern "C" {

const char* mpm_last_error(void) { return g_err.c_str(); }

int mpm_version(void) { return 100; }

int mpm_config_derive(mpm_sim_config* cfg) {
    if (!cfg) return fail(MPM_EINVAL, "null config");
    const uint64_t g = (uint64_t)cfg->SIZE_X * cfg->SIZE_Y * cfg->SIZE_Z;
    cfg->GRID_NUMBER = (uint32_t)g;
    cfg->GRID_SIZE = (uint32_t)(g * 16u);
    cfg->GRID_BLOCKS = (uint32_t)((g + 255u) / 256u);
    return MPM_OK;
}

int mpm_engine_create(const mpm_sim_config* cfg_in, mpm_engine** out) {
    if (!cfg_in || !out) return fail(MPM_EINVAL, "null argument");
    *out = nullptr;
    mpm_sim_config cfg = *cfg_in;
    if (cfg.SIZE_X == 0 || cfg.SIZE_Y == 0 || cfg.SIZE_Z == 0) return fail(MPM_EINVAL, "SIZE_X/Y/Z must be positive");
    if (cfg.SIZE_X > 1024 || cfg.SIZE_Y > 1024 || cfg.SIZE_Z > 1024)
        return fail(MPM_ELIMIT, "SIZE_X/Y/Z above 1024 (Morton keys cover 10 bits per axis)");
    const uint64_t G = (uint64_t)cfg.SIZE_X * cfg.SIZE_Y * cfg.SIZE_Z;
    if (G >= (1ull << 31)) return fail(MPM_ELIMIT, "grid has 2^31 or more nodes");
    mpm_config_derive(&cfg);
    mpm_engine* h = new mpm_engine();
    h->cfg = cfg;
    DevCfg& d = h->dc;
    d.sx = (int)cfg.SIZE_X; d.sy = (int)cfg.SIZE_Y; d.sz = (int)cfg.SIZE_Z;
    d.sxy = d.sx * d.sy;
    d.pad = cfg.PADDING;
    d.gravity = cfg.GRAVITY; d.dt = cfg.DT; d.gamma = cfg.GAMMA; d.compression = cfg.COMPRESSION;
    d.vol = cfg.RESOLUTION * cfg.RESOLUTION * cfg.RESOLUTION;
    d.mass = 1.0f;
    d.fbx = (d.sx + 3) / 4; d.fby = (d.sy + 3) / 4; d.fbz = (d.sz + 3) / 4;
    h->G = (size_t)G;
    h->nfb = d.fbx * d.fby * d.fbz;
    {
        int maxdim = std::max(d.sx, std::max(d.sy, d.sz)), bits = 0;
        while ((1 << bits) < maxdim) ++bits;
        h->key_bits = 3 * std::max(bits, 1);
    }
    auto cleanup = [&](int rc) {
        mpm_engine_destroy(h);
        return rc;
    };
    if (hipGetDevice(&h->device) != hipSuccess) return cleanup(fail(MPM_EHIP, "no HIP device"));
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(MPM_EHIP, "hipStreamCreate failed"));
    int rc;
    for (int b = 0; b < 2; ++b) {
        float* base = nullptr;
        if ((rc = dev_alloc(&base, 4 * h->G))) return cleanup(rc);
        for (int p = 0; p < 4; ++p) h->grid[b].plane[p] = base + p * h->G;
        if ((rc = dev_alloc(&h->grid[b].flags, (size_t)h->nfb))) return cleanup(rc);
        if ((rc = zero_grid(h, b))) return cleanup(rc);
    }
    if ((rc = dev_alloc(&h->sdf, h->G))) return cleanup(rc);
    launch_fill_f32(h->sdf, h->G, 1e15f, h->stream);  // makeSDF with no triangles
    if ((rc = dev_alloc(&h->sc.counters, 64))) return cleanup(rc);
    if (hipMemsetAsync(h->sc.counters, 0, 64 * sizeof(uint32_t), h->stream) != hipSuccess)
        return cleanup(fail(MPM_EHIP, "memset"));
    h->overflow = h->sc.counters + 8;
    if (hipHostMalloc((void**)&h->pinned, 64 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(MPM_ENOMEM, "hipHostMalloc failed"));
    if (hipStreamSynchronize(h->stream) != hipSuccess) return cleanup(fail(MPM_EHIP, "stream sync"));
    *out = h;
    return MPM_OK;
}

int mpm_engine_destroy(mpm_engine* h) {
    if (!h) return MPM_OK;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    free_particles(h);
    for (int b = 0; b < 2; ++b) {
        dev_free(h->grid[b].plane[0]);
        dev_free(h->grid[b].flags);
    }
    dev_free(h->sdf);
    dev_free(h->sc.counters);
    h->overflow = nullptr;
    if (h->pinned) hipHostFree(h->pinned);
    for (auto& e : h->evs) {
        hipEventDestroy(e.a);
        hipEventDestroy(e.b);
    }
    for (auto e : h->pool) hipEventDestroy(e);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return MPM_OK;
}

int mpm_engine_set_option(mpm_engine* h, const char* key, int64_t value) {
    if (!h || !key) return fail(MPM_EINVAL, "null argument");
    std::string k(key);
    if (k == "RESORT_EVERY") {
        if (value < 1) return fail(MPM_EINVAL, "RESORT_EVERY must be >= 1");
        h->resort_every = (int)std::min<int64_t>(value, 1 << 30);
    } else if (k == "CHUNK") {
        if (value < 64 || value > 4096 || (value % 64) != 0)
            return fail(MPM_EINVAL, "CHUNK must be a multiple of 64 in [64, 4096]");
        if (h->have_particles && value != h->chunk_cap)
            return fail(MPM_ESTATE, "CHUNK must be set before mpm_engine_init_particles");
        h->chunk_cap = (int)value;
    } else if (k == "TIMING") {
        h->timing = value != 0;
    } else {
        return fail(MPM_EINVAL, "unknown option " + k);
    }
    return MPM_OK;
}

int mpm_engine_get_option(mpm_engine* h, const char* key, int64_t* value) {
    if (!h || !key || !value) return fail(MPM_EINVAL, "null argument");
    std::string k(key);
    if (k == "RESORT_EVERY") *value = h->resort_every;
    else if (k == "CHUNK") *value = h->chunk_cap;
    else if (k == "TIMING") *value = h->timing ? 1 : 0;
    else return fail(MPM_EINVAL, "unknown option " + k);
    return MPM_OK;
}

int mpm_engine_init_particles(mpm_engine* h, int n, const float* soa26) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (n < 0) return fail(MPM_EINVAL, "negative particle count");
    if (n > 0 && !soa26) return fail(MPM_EINVAL, "null particle buffer");
    if ((int64_t)n > (1ll << 30)) return fail(MPM_ELIMIT, "more than 2^30 particles");
    free_particles(h);
    for (int b = 0; b < 2; ++b)
        if ((rc = zero_grid(h, b))) return rc;
    h->vcur = 0;
    h->n = n;
    h->substeps = 0;
    h->sorted_at = 0;
    if (n == 0) {
        h->have_particles = true;
        HIPCHK(hipStreamSynchronize(h->stream));
        return MPM_OK;
    }
    float* tmp = nullptr;
    if ((rc = dev_alloc(&tmp, (size_t)MPM_PARTICLE_ATTRS * n))) return rc;
    HIPCHK(hipMemcpyAsync(tmp, soa26, sizeof(float) * MPM_PARTICLE_ATTRS * (size_t)n, hipMemcpyHostToDevice, h->stream));
    launch_init_F(tmp, n, h->stream);
    const int64_t blocks = (int64_t)h->dc.fbx * h->dc.fby * h->dc.fbz;
    h->chunk_limit = (int)(std::min<int64_t>(n, blocks) + (n + h->chunk_cap - 1) / h->chunk_cap + 1);
    h->chunk_est = h->chunk_limit;
    if ((rc = ensure_sort_scratch(h, (size_t)n)) || (rc = dev_alloc(&h->chunks, (size_t)h->chunk_limit)) ||
        (rc = dev_alloc(&h->rounds, (size_t)n)) ||
        (rc = ensure_planes(h, 0, (size_t)n)) || (rc = ensure_planes(h, 1, (size_t)n))) {
        dev_free(tmp);
        return rc;
    }
    launch_keys_linear(tmp, tmp + (size_t)n, tmp + 2 * (size_t)n, n, true, h->dc, h->sc.keys[0], h->sc.vals[0],
                       h->stream);
    h->pcur = 1;  // sort_and_reorder writes into buffer 0
    rc = sort_and_reorder(h, tmp, (size_t)n, nullptr, MASK_ALL);
    hipStreamSynchronize(h->stream);
    dev_free(tmp);
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    if ((rc = read_counters(h))) return rc;
    h->chunk_est = (int)h->pinned[3];
    h->last_rounds = h->pinned[4];
    h->have_particles = true;
    return MPM_OK;
}

int mpm_engine_init_sdf_from_triangles(mpm_engine* h, const mpm_triangle* tris, int nt) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (nt < 0 || (nt > 0 && !tris)) return fail(MPM_EINVAL, "bad triangle buffer");
    float* dt = nullptr;
    if ((rc = dev_alloc(&dt, (size_t)9 * std::max(nt, 1)))) return rc;
    if (nt > 0) HIPCHK(hipMemcpyAsync(dt, tris, sizeof(mpm_triangle) * (size_t)nt, hipMemcpyHostToDevice, h->stream));
    launch_make_sdf(h->dc, dt, nt, h->sdf, h->stream);
    hipError_t e = hipStreamSynchronize(h->stream);
    dev_free(dt);
    if (e != hipSuccess) return fail(MPM_EHIP, std::string("makeSDF: ") + hipGetErrorString(e));
    HIPCHK(hipGetLastError());
    return MPM_OK;
}

int mpm_engine_set_sdf(mpm_engine* h, const float* sdf) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (!sdf) return fail(MPM_EINVAL, "null sdf");
    HIPCHK(hipMemcpyAsync(h->sdf, sdf, h->G * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int mpm_engine_get_sdf(mpm_engine* h, float* out) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (!out) return fail(MPM_EINVAL, "null output");
    HIPCHK(hipMemcpyAsync(out, h->sdf, h->G * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int mpm_engine_steps(mpm_engine* h, int k) {
    int rc = check_engine(h, true);
    if (rc) return rc;
    if (k < 0) return fail(MPM_EINVAL, "negative substep count");
    if (k == 0 || h->n == 0) {
        h->substeps += k;
        return MPM_OK;
    }
    HIPCHK(hipMemsetAsync(h->overflow, 0, 2 * sizeof(uint32_t), h->stream));
    for (int s = 0; s < k; ++s) {
        if (s == 0) {
            if (h->substeps - h->sorted_at >= h->resort_every) {
                if ((rc = resort(h, true))) return rc;
                h->sorted_at = h->substeps;
            }
            run_transfer(h, XFER_P2G);
        }
        run_grid_update(h);
        if (s == k - 1) {
            run_transfer(h, XFER_G2P);
            h->substeps++;
        } else {
            run_transfer(h, XFER_G2P2G);
            h->substeps++;
            if (h->substeps - h->sorted_at >= h->resort_every) {
                if ((rc = resort(h, false))) return rc;
                h->sorted_at = h->substeps;
            }
        }
    }
    HIPCHK(hipGetLastError());
    if ((rc = read_counters(h))) return rc;
    h->last_overflow = h->pinned[8];
    h->last_claim_losers = h->pinned[9];
    h->chunk_est = (int)h->pinned[3];
    h->last_rounds = h->pinned[4];
    return MPM_OK;
}

int mpm_engine_step(mpm_engine* h) { return mpm_engine_steps(h, 1); }

int mpm_engine_synchronize(mpm_engine* h) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int mpm_engine_get_number(mpm_engine* h) { return h ? h->n : 0; }

int mpm_engine_get_particles(mpm_engine* h, float* out) {
    int rc = check_engine(h, true);
    if (rc) return rc;
    if (!out) return fail(MPM_EINVAL, "null output");
    if (h->n == 0) return MPM_OK;
    float* tmp = nullptr;
    if ((rc = dev_alloc(&tmp, (size_t)MPM_PARTICLE_ATTRS * h->n))) return rc;
    launch_export_soa26(h->planes[h->pcur], h->plane_cap[h->pcur], h->orig[h->pcur], h->n, h->dc.vol, h->dc.mass, tmp,
                        h->stream);
    hipError_t e = hipMemcpyAsync(out, tmp, sizeof(float) * MPM_PARTICLE_ATTRS * (size_t)h->n, hipMemcpyDeviceToHost,
                                  h->stream);
    hipError_t e2 = hipStreamSynchronize(h->stream);
    dev_free(tmp);
    if (e != hipSuccess || e2 != hipSuccess) return fail(MPM_EHIP, "get_particles copy failed");
    return MPM_OK;
}

int mpm_engine_get_positions(mpm_engine* h, float* xyz) {
    int rc = check_engine(h, true);
    if (rc) return rc;
    if (!xyz) return fail(MPM_EINVAL, "null output");
    if (h->n == 0) return MPM_OK;
    float* tmp = nullptr;
    if ((rc = dev_alloc(&tmp, (size_t)3 * h->n))) return rc;
    launch_export_positions(h->planes[h->pcur], h->plane_cap[h->pcur], h->orig[h->pcur], h->n, tmp, h->stream);
    std::vector<float> planes((size_t)3 * h->n);
    hipError_t e = hipMemcpyAsync(planes.data(), tmp, sizeof(float) * 3 * (size_t)h->n, hipMemcpyDeviceToHost, h->stream);
    hipError_t e2 = hipStreamSynchronize(h->stream);
    dev_free(tmp);
    if (e != hipSuccess || e2 != hipSuccess) return fail(MPM_EHIP, "get_positions copy failed");
    for (int i = 0; i < h->n; ++i) {
        xyz[3 * (size_t)i] = planes[i];
        xyz[3 * (size_t)i + 1] = planes[(size_t)h->n + i];
        xyz[3 * (size_t)i + 2] = planes[2 * (size_t)h->n + i];
    }
    return MPM_OK;
}

int mpm_engine_get_sort(mpm_engine* h, int32_t* keys_out, int32_t* perm_out) {
    int rc = check_engine(h, true);
    if (rc) return rc;
    if (!keys_out || !perm_out) return fail(MPM_EINVAL, "null output");
    const int n = h->n;
    if (n == 0) return MPM_OK;
    float* tmp = nullptr;
    if ((rc = dev_alloc(&tmp, (size_t)3 * n))) return rc;
    launch_export_positions(h->planes[h->pcur], h->plane_cap[h->pcur], h->orig[h->pcur], n, tmp, h->stream);
    launch_keys_linear(tmp, tmp + n, tmp + 2 * (size_t)n, n, false, h->dc, h->sc.keys[0], h->sc.vals[0], h->stream);
    const int r = radix_sort_pairs(h->sc, n, 0, h->stream);
    hipError_t e1 = hipMemcpyAsync(keys_out, h->sc.keys[r], sizeof(int32_t) * n, hipMemcpyDeviceToHost, h->stream);
    hipError_t e2 = hipMemcpyAsync(perm_out, h->sc.vals[r], sizeof(int32_t) * n, hipMemcpyDeviceToHost, h->stream);
    hipError_t e3 = hipStreamSynchronize(h->stream);
    dev_free(tmp);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(MPM_EHIP, "get_sort failed");
    for (int i = 0; i < n; ++i) keys_out[i] = (int32_t)((uint32_t)keys_out[i] ^ 0x80000000u);
    return MPM_OK;
}

int mpm_engine_get_grid(mpm_engine* h, float* out) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (!out) return fail(MPM_EINVAL, "null output");
    for (int p = 0; p < 4; ++p)
        HIPCHK(hipMemcpyAsync(out + p * h->G, h->grid[h->vcur].plane[p], h->G * sizeof(float), hipMemcpyDeviceToHost,
                              h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int mpm_engine_get_stage_times(mpm_engine* h, double* out, int nstages) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    for (auto& e : h->evs) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
            h->stage_ms[e.stage] += ms;
            h->stage_cnt[e.stage] += 1;
        }
        h->pool.push_back(e.a);
        h->pool.push_back(e.b);
    }
    h->evs.clear();
    for (int s = 0; s < nstages && s < ST_COUNT; ++s) {
        if (out) {
            out[2 * s] = h->stage_ms[s];
            out[2 * s + 1] = (double)h->stage_cnt[s];
        }
    }
    for (int s = 0; s < ST_COUNT; ++s) {
        h->stage_ms[s] = 0.0;
        h->stage_cnt[s] = 0;
    }
    return MPM_OK;
}

int mpm_engine_get_counters(mpm_engine* h, int64_t* out, int n) {
    if (!h || !out) return fail(MPM_EINVAL, "null argument");
    int64_t v[5] = {h->chunk_est, h->last_overflow, 0, h->last_claim_losers, h->last_rounds};
    if (n > 2) {
        std::vector<uint8_t> f((size_t)h->nfb);
        hipSetDevice(h->device);
        HIPCHK(hipMemcpyAsync(f.data(), h->grid[h->vcur].flags, f.size(), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        int64_t cnt = 0;
        for (uint8_t b : f) cnt += b != 0;
        v[2] = cnt;
    }
    for (int i = 0; i < n && i < 5; ++i) out[i] = v[i];
    return MPM_OK;
}

int mpm_engine_debug_p2g(mpm_engine* h, float* grid4_out) {
    int rc = check_engine(h, true);
    if (rc) return rc;
    if (!grid4_out) return fail(MPM_EINVAL, "null output");
    const int m = 1 - h->vcur;
    if (h->n > 0) run_transfer(h, XFER_P2G);
    for (int p = 0; p < 4; ++p)
        HIPCHK(hipMemcpyAsync(grid4_out + p * h->G, h->grid[m].plane[p], h->G * sizeof(float), hipMemcpyDeviceToHost,
                              h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return zero_grid(h, m);
}

int mpm_engine_debug_grid_update(mpm_engine* h, const float* in, float* out) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (!in || !out) return fail(MPM_EINVAL, "null argument");
    const int m = 1 - h->vcur;
    for (int p = 0; p < 4; ++p)
        HIPCHK(hipMemcpyAsync(h->grid[m].plane[p], in + p * h->G, h->G * sizeof(float), hipMemcpyHostToDevice, h->stream));
    launch_fill_u8(h->grid[m].flags, (size_t)h->nfb, 1, h->stream);
    launch_grid_update(h->dc, h->grid[m], h->grid[h->vcur], false, h->sdf, h->nfb, h->stream);
    for (int p = 0; p < 4; ++p)
        HIPCHK(hipMemcpyAsync(out + p * h->G, h->grid[m].plane[p], h->G * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return zero_grid(h, m);
}

int mpm_engine_debug_g2p(mpm_engine* h, const float* in) {
    int rc = check_engine(h, true);
    if (rc) return rc;
    if (!in) return fail(MPM_EINVAL, "null argument");
    const int v = h->vcur;
    for (int p = 0; p < 4; ++p)
        HIPCHK(hipMemcpyAsync(h->grid[v].plane[p], in + p * h->G, h->G * sizeof(float), hipMemcpyHostToDevice, h->stream));
    launch_fill_u8(h->grid[v].flags, (size_t)h->nfb, 1, h->stream);
    if (h->n > 0) run_transfer(h, XFER_G2P);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int mpm_probe_copy(size_t bytes, int iters, double* ms_per_copy) {
    if (bytes < 16 || iters < 1 || !ms_per_copy) return fail(MPM_EINVAL, "bad copy probe arguments");
    const size_t n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    int rc;
    if ((rc = dev_alloc(&a, n)) || (rc = dev_alloc(&b, n))) {
        dev_free(a);
        return rc;
    }
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    if (err == hipSuccess) err = hipMemsetAsync(a, 0, n * 16, s);
    if (err == hipSuccess) {
        launch_copy(a, b, n, s);  // warm-up
        hipEventRecord(e0, s);
        for (int i = 0; i < iters; ++i) launch_copy(a, b, n, s);
        hipEventRecord(e1, s);
        err = hipEventSynchronize(e1);
    }
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (s) hipStreamDestroy(s);
    dev_free(a);
    dev_free(b);
    if (err != hipSuccess) return fail(MPM_EHIP, std::string("copy probe: ") + hipGetErrorString(err));
    *ms_per_copy = ms / iters;
    return MPM_OK;
}

}  // extern "C"
