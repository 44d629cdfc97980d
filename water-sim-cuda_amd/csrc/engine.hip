// engine.hip -- host side of the engine and the C ABI declared in include/mpm.h.
//
// Replaces the reference's Engine class (engine.h:14-43, engine.cu:36-114).
// Differences of design (not of results):
//  * an engine holds one or more independent scenes (each with its own
//    SimConfig, grid, SDF and particles); every stage is one launch over all
//    scenes, with per-scene configs in a device table;
//  * the particle state lives in a chunked, Morton-sorted slot layout and is
//    physically reordered at each re-sort (the reference only permutes an
//    index array, engine.cu:98-99); the caller's order is kept in `orig`;
//  * a substep is P2G + grid update + G2P as in engine.cu:54-70, but within a
//    k-substep call the G2P of substep s and the P2G of substep s+1 run as one
//    fused kernel on two grid buffers;
//  * the grid is cleared and updated only where the previous P2G wrote
//    (4^3-node activity flags), instead of a dense pass over all nodes;
//  * a k-substep call can be captured once as a HIP graph and replayed;
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

enum Stage { ST_P2G = 0, ST_G2P = 1, ST_G2P2G = 2, ST_GRID = 3, ST_SORT = 4, ST_CORR = 5, ST_COUNT = 6 };

constexpr int MAX_SCENES = 4096;

}  // namespace

namespace mpm {
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace mpm

using namespace mpm;

struct mpm_engine {
    // ---- scenes ----
    std::vector<mpm_sim_config> cfgs;  // host copies (derived fields filled)
    std::vector<DevCfg> dcs;           // host copies of the device table
    DevCfg* d_cfgs = nullptr;          // device table
    uint16_t* d_block_scene = nullptr; // scene of each flag block
    int S = 1;
    std::vector<int> pstart;  // S + 1: particle ranges (caller order)
    int* d_pstart = nullptr;

    int device = 0;
    hipStream_t stream = nullptr;

    int n = 0;
    bool have_particles = false;
    int chunk_cap = 2048;
    int launch_pct = 100;  // transfer workgroups per 100 chunks
    int resort_every = 1;
    int timing = 0;  // 0 off, 1 every stage, 2 the transfer stages only
    bool use_graph = false;
    bool capturing = false;
    bool sdf_bvh = true;             // makeSDF through the hierarchy (same field)
    int64_t sdf_exact_nodes = 0;     // nodes of the last makeSDF that took the exact sign sum

    size_t G = 0;  // nodes of all scenes
    int nfb = 0;   // flag blocks of all scenes

    float* planes[2] = {nullptr, nullptr};
    int* orig[2] = {nullptr, nullptr};
    size_t plane_cap[2] = {0, 0};
    int pcur = 0;
    // A re-sort inside steps() moves only the caller indices; the next G2P
    // reads positions and F through sc.map from plane set gather_src.
    bool gather_pending = false;
    int gather_src = 0;
    // The last transfer of a steps() call also ran the next substep's P2G
    // into grid[1 - vcur] (XFER_G2P2G_FULL); the next call starts from it.
    // Anything that changes the particles or DT outside steps() discards it.
    bool p2g_pending = false;

    ChunkDesc* chunks = nullptr;
    int chunk_limit = 0;  // bound on the chunk count (sum over scenes of min(n_s, blocks_s) + n / CHUNK + 1)
    int chunk_alloc = 0;  // descriptors allocated
    int chunk_est = 0;    // chunk count at the last host synchronisation (sizes launches)

    GridBuf grid[2] = {};
    uint2* grid_list = nullptr;  // flagged row segments (four x-adjacent 4^3 blocks) of a grid update
    int vcur = 0;  // grid[vcur]: velocities of the last update; grid[1 - vcur]: accumulates
    float* sdf = nullptr;

    SortScratch sc{};
    BlockTable blocks{};
    uint32_t* overflow = nullptr;  // = sc.counters + 8: [0] tile-overflow particles, [1] extra claim passes
    uint32_t* pinned = nullptr;

    int64_t substeps = 0;
    int64_t sorted_at = 0;
    int64_t last_overflow = 0;
    int64_t last_claim_losers = 0;
    int64_t last_rounds = 0;

    // Cached graphs of steps(k) calls, each valid for the start state it was
    // captured from and the launch size it was captured with.  A few are kept
    // because consecutive calls can start from alternating states (an odd
    // number of re-sorts per call swaps the particle plane sets each call).
    struct GraphEntry {
        hipGraphExec_t exec = nullptr;
        int k = 0;
        int64_t phase = 0;  // substeps - sorted_at at the start
        int vcur = 0, pcur = 0;
        bool pending = false;  // p2g_pending at the start (always set at the end)
        int launch = 0;     // transfer launch size baked in
        // end state
        int64_t end_phase = 0;
        int end_vcur = 0, end_pcur = 0;
        int64_t used = 0;   // last use (for eviction)
    };
    static constexpr int GRAPH_CACHE = 4;
    GraphEntry graphs[GRAPH_CACHE];
    int64_t graph_replays = 0;
    int64_t graph_captures = 0;

    // Learned corrector: applied after every grid update while `correct`.
    const CorrState* corr = nullptr;
    bool correct = false;
    bool pair_mode = false;  // collect_pairs: accumulate fine-step velocities after each grid update
    int2* clist = nullptr;   // active nodes
    float* cX = nullptr;     // features, CORR_STRIDE per node
    float* cY = nullptr;     // coarse velocities of collected pairs, 3 per node
    float* cV = nullptr;     // velocity copy of corrected substeps: 3 planes of G (zeroed once)
    float* cacc = nullptr;   // accumulated fine velocities, 3 per node
    float* chits = nullptr;
    int ccap = 0;
    float dt_scale = 1.0f;
    float* check_buf = nullptr;
    float* xbuf = nullptr;      // device staging of the read-back calls
    size_t xbuf_cap = 0;

    // Saved particle state (mpm_engine_snapshot / mpm_engine_restore).
    float* snap_planes = nullptr;  // A_STATE x n
    int* snap_orig = nullptr;
    int64_t snap_substeps = -1;

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
        if (h->timing && !h->capturing && (h->timing == 1 || s <= ST_G2P2G)) {
            a = take_event(h);
            hipEventRecord(a, h->stream);
        }
    }
    ~StageTimer() {
        if (a) {
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

void drop_graph(mpm_engine* h) {
    for (auto& g : h->graphs) {
        if (g.exec) hipGraphExecDestroy(g.exec);
        g = mpm_engine::GraphEntry{};
    }
}

void free_particles(mpm_engine* h) {
    drop_graph(h);
    dev_free(h->snap_planes);
    dev_free(h->snap_orig);
    h->snap_substeps = -1;
    for (int i = 0; i < 2; ++i) {
        dev_free(h->planes[i]);
        dev_free(h->orig[i]);
        h->plane_cap[i] = 0;
    }
    dev_free(h->chunks);
    dev_free(h->d_pstart);
    dev_free(h->sc.keys[0]);
    dev_free(h->sc.keys[1]);
    dev_free(h->sc.vals[0]);
    dev_free(h->sc.vals[1]);
    dev_free(h->sc.hist);
    dev_free(h->sc.partial);
    dev_free(h->sc.map);
    h->gather_pending = false;
    h->sc.cap = h->sc.hist_cap = h->sc.partial_cap = 0;
    h->chunk_limit = 0;
    h->chunk_alloc = 0;
    h->chunk_est = 0;
    h->have_particles = false;
    h->n = 0;
}

int ensure_sort_scratch(mpm_engine* h, size_t n) {
    SortScratch& sc = h->sc;
    if (sc.cap >= n + 1 && sc.keys[0]) return MPM_OK;
    dev_free(sc.keys[0]); dev_free(sc.keys[1]); dev_free(sc.vals[0]); dev_free(sc.vals[1]);
    dev_free(sc.hist); dev_free(sc.partial); dev_free(sc.map);
    const size_t cap = n + 1;
    int rc;
    if ((rc = dev_alloc(&sc.keys[0], cap)) || (rc = dev_alloc(&sc.keys[1], cap)) || (rc = dev_alloc(&sc.vals[0], cap)) ||
        (rc = dev_alloc(&sc.vals[1], cap)) || (rc = dev_alloc(&sc.map, cap)))
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

// Chunk-count bound for the current particles and chunk_cap: chunks are the
// non-empty chunk blocks of every scene split into pieces of at most chunk_cap.
int chunk_bound(const mpm_engine* h, int cap) {
    int64_t b = 0;
    for (int s = 0; s < h->S; ++s) {
        const int64_t ns = h->pstart[s + 1] - h->pstart[s];
        const DevCfg& d = h->dcs[s];
        b += std::min<int64_t>(ns, (int64_t)((d.fbx + 1) / 2) * d.fby * d.fbz);
    }
    return (int)(b + (h->n + cap - 1) / cap + 1);
}

// Sets chunk_cap and makes sure the chunk table can hold its bound.
int set_chunk_cap(mpm_engine* h, int cap) {
    const int lim = chunk_bound(h, cap);
    if (lim > h->chunk_alloc) {
        dev_free(h->chunks);
        int rc;
        if ((rc = dev_alloc(&h->chunks, (size_t)lim))) return rc;
        h->chunk_alloc = lim;
    }
    h->chunk_cap = cap;
    h->chunk_limit = lim;
    return MPM_OK;
}

// Copies the 16 device counters (SortScratch::counters) into h->pinned and
// waits for the stream.
int read_counters(mpm_engine* h) {
    HIPCHK(hipMemcpyAsync(h->pinned, h->sc.counters, 16 * sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

SceneTable scene_table(const mpm_engine* h) {
    SceneTable t;
    t.cfgs = h->d_cfgs;
    t.pstart = h->d_pstart;
    t.nscenes = h->S;
    return t;
}

// Re-sorts the particles of (src, src_stride, src_orig) -- block buckets, the
// chunk table, the round layout -- into the alternate plane buffer, moving the planes
// in `mask`.  Nothing here waits for the device: the chunk count stays in
// device memory and is bounded by chunk_limit by construction.
int sort_and_reorder(mpm_engine* h, const float* src, size_t src_stride, const int* src_orig, uint32_t mask) {
    ResortArgs a{};
    const int dst = 1 - h->pcur;
    a.src = src;
    a.src_stride = src_stride;
    a.src_orig = src_orig;
    a.dst = h->planes[dst];
    a.dst_stride = h->plane_cap[dst];
    a.dst_orig = h->orig[dst];
    a.n = h->n;
    a.mask = mask;
    a.scenes = scene_table(h);
    a.block_scene = h->d_block_scene;
    a.blocks = h->blocks;
    for (int i = 0; i < 2; ++i) {
        a.keys[i] = h->sc.keys[i];
        a.vals[i] = h->sc.vals[i];
    }
    a.counters = h->sc.counters;
    a.map = h->sc.map;
    a.chunks = h->chunks;
    a.chunk_cap = h->chunk_cap;
    a.chunk_limit = h->chunk_limit;
    a.chunk_est = h->chunk_est;
    launch_resort(a, h->stream);
    HIPCHK(hipGetLastError());
    h->pcur = dst;
    return MPM_OK;
}

constexpr uint32_t MASK_ALL = (1u << A_STATE) - 1u;

// full: every plane moves now.  Otherwise (inside steps(), where a G2P always
// follows) no plane moves: until that G2P has written positions and F to the
// new planes, the transfers read the old planes through the map (a P2G all
// attributes, the G2P positions and F; v and C then come from the G2P).
int resort(mpm_engine* h, bool full) {
    StageTimer t(h, ST_SORT);
    const int src = h->pcur;
    const int rc = sort_and_reorder(h, h->planes[src], h->plane_cap[src], h->orig[src], full ? MASK_ALL : 0u);
    if (rc) return rc;
    h->gather_pending = !full;
    h->gather_src = src;
    return MPM_OK;
}

int transfer_launch_size(const mpm_engine* h) {
    return std::max(1, (int)(((int64_t)h->chunk_est * h->launch_pct + 99) / 100));
}

TransferArgs make_args(mpm_engine* h) {
    TransferArgs a{};
    a.cfgs = h->d_cfgs;
    a.part = h->planes[h->pcur];
    a.stride = h->plane_cap[h->pcur];
    a.chunks = h->chunks;
    a.nchunks = h->sc.counters + 3;
    const GridBuf& V = h->grid[h->vcur];
    const GridBuf& M = h->grid[1 - h->vcur];
    a.vel[0] = V.plane[1];
    a.vel[1] = V.plane[2];
    a.vel[2] = V.plane[3];
    for (int i = 0; i < 4; ++i) a.acc[i] = M.plane[i];
    a.acc_flags = M.flags;
    a.sdf = h->sdf;
    a.overflow = h->overflow;
    if (h->gather_pending) {
        a.map = h->sc.map;
        a.src = h->planes[h->gather_src];
        a.src_stride = h->plane_cap[h->gather_src];
    }
    return a;
}

void run_transfer(mpm_engine* h, int mode, int launch) {
    // the full fused transfer ends a steps() call: timed with the final G2Ps
    const int stage = mode == XFER_P2G ? ST_P2G : (mode == XFER_G2P2G ? ST_G2P2G : ST_G2P);
    StageTimer t(h, stage);
    // a pending gather is always consumed by a G2P (resort() keeps it so); a
    // P2G before it reads through the map too
    launch_transfer(mode, make_args(h), launch, h->stream);
    if (mode & XFER_G2P) h->gather_pending = false;
}

uint32_t* corr_count(mpm_engine* h) { return h->sc.counters + 12; }

void run_grid_update(mpm_engine* h) {
    // a corrected substep: the grid update also lists the active nodes and
    // copies the velocities, which the corrector's features read while it
    // corrects the grid in place
    const bool correct = h->correct && h->corr && h->ccap > 0;
    CorrGridOut co{};
    if (correct) {
        for (int k = 0; k < 3; ++k) co.vcopy.plane[1 + k] = h->cV + (size_t)k * h->G;
        co.active = h->clist;
        co.acap = h->ccap;
        co.acount = corr_count(h);
    }
    {
        StageTimer t(h, ST_GRID);
        if (correct) launch_set_u32(corr_count(h), 0, 0u, h->stream);
        // the list counters alternate with the grid buffers: [14 + vcur]
        // now, [15 - vcur] for the next update (cleared by this one)
        launch_grid_update(h->d_cfgs, h->d_block_scene, h->S, h->grid[1 - h->vcur], h->grid[h->vcur], true, h->sdf,
                           h->nfb, h->grid_list, h->sc.counters + 14 + h->vcur, h->sc.counters + 15 - h->vcur,
                           h->stream, correct ? &co : nullptr);
        h->vcur = 1 - h->vcur;
    }
    const GridBuf& V = h->grid[h->vcur];
    if (correct) {
        StageTimer t(h, ST_CORR);
        launch_corr_correct(h->d_cfgs, V, co.vcopy, h->sdf, h->clist, corr_count(h), h->ccap, *h->corr, h->stream);
    }
    if (h->pair_mode)
        launch_corr_gather_vel(h->d_cfgs, V, h->clist, corr_count(h), h->ccap, h->cacc, h->chits, 2, h->stream);
}

int zero_grid(mpm_engine* h, int which) {
    for (int p = 0; p < 4; ++p) HIPCHK(hipMemsetAsync(h->grid[which].plane[p], 0, h->G * sizeof(float), h->stream));
    HIPCHK(hipMemsetAsync(h->grid[which].flags, 0, (size_t)h->nfb, h->stream));
    return MPM_OK;
}

// Drops the next substep's P2G that the last steps() call left in
// grid[1 - vcur] (the particles or DT changed since): that buffer is zeroed
// again, as a P2G expects it.
int discard_pending_p2g(mpm_engine* h) {
    if (!h->p2g_pending) return MPM_OK;
    h->p2g_pending = false;
    return zero_grid(h, 1 - h->vcur);
}

int check_engine(mpm_engine* h, bool need_particles) {
    if (!h) return fail(MPM_EINVAL, "null engine handle");
    if (need_particles && !h->have_particles) return fail(MPM_ESTATE, "particles not initialised (call mpm_engine_init_particles)");
    HIPCHK(hipSetDevice(h->device));
    return MPM_OK;
}

int check_scene(mpm_engine* h, int scene) {
    if (scene < 0 || scene >= h->S) return fail(MPM_EINVAL, "scene index out of range");
    return MPM_OK;
}

// Enqueues k substeps (engine.cu:54-70 per substep) and advances the host-side
// schedule state.  Launches only: safe under stream capture.  The first P2G
// is skipped when the previous call left it in the grid; the last transfer
// writes all particle attributes and runs the next substep's P2G.
int enqueue_steps(mpm_engine* h, int k, int launch) {
    int rc;
    launch_set_u32(h->overflow, 0, 0u, h->stream);
    launch_set_u32(h->overflow, 1, 0u, h->stream);
    for (int s = 0; s < k; ++s) {
        if (s == 0) {
            // A re-sort due now waits for the first fused substep when there
            // is one (reading all attributes through the map would cost the
            // P2G more than one substep's drift).  The grid a pending P2G left
            // does not depend on the layout.
            if (k == 1 && h->substeps - h->sorted_at >= h->resort_every) {
                if ((rc = resort(h, false))) return rc;
                h->sorted_at = h->substeps;
            }
            if (!h->p2g_pending) run_transfer(h, XFER_P2G, launch);
            h->p2g_pending = false;
        }
        run_grid_update(h);
        if (s == k - 1) {
            run_transfer(h, XFER_G2P2G_FULL, launch);
            h->substeps++;
            h->p2g_pending = true;
        } else {
            run_transfer(h, XFER_G2P2G, launch);
            h->substeps++;
            if (h->substeps - h->sorted_at >= h->resort_every) {
                if ((rc = resort(h, false))) return rc;
                h->sorted_at = h->substeps;
            }
        }
    }
    HIPCHK(hipGetLastError());
    return MPM_OK;
}

// Replays the cached graph when it was captured from the same start state
// (schedule phase, buffer parities, k) and a launch size still fitting the
// chunk count; otherwise captures a new one.
int steps_graph(mpm_engine* h, int k) {
    const int64_t phase = h->substeps - h->sorted_at;
    const int want = transfer_launch_size(h);
    // an entry for this start state (launch size aside), else the least
    // recently used slot
    mpm_engine::GraphEntry* slot = nullptr;
    for (auto& c : h->graphs)
        if (c.exec && c.k == k && c.phase == phase && c.vcur == h->vcur && c.pcur == h->pcur &&
            c.pending == h->p2g_pending)
            slot = &c;
    if (!slot) {
        slot = &h->graphs[0];
        for (auto& c : h->graphs)
            if (!c.exec || (slot->exec && c.used < slot->used)) slot = &c;
    }
    mpm_engine::GraphEntry& g = *slot;
    const bool hit = g.exec && g.k == k && g.phase == phase && g.vcur == h->vcur && g.pcur == h->pcur &&
                     g.pending == h->p2g_pending &&
                     want <= g.launch && want * 2 >= g.launch;
    if (!hit) {
        if (g.exec) hipGraphExecDestroy(g.exec);
        g = mpm_engine::GraphEntry{};
        mpm_engine::GraphEntry e;
        e.k = k;
        e.phase = phase;
        e.vcur = h->vcur;
        e.pcur = h->pcur;
        e.pending = h->p2g_pending;
        e.launch = want + want / 8 + 8;  // headroom for chunk-count growth
        const int64_t sub0 = h->substeps, sorted0 = h->sorted_at;
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        h->capturing = true;
        int rc = enqueue_steps(h, k, e.launch);
        h->capturing = false;
        hipGraph_t graph = nullptr;
        hipError_t ce = hipStreamEndCapture(h->stream, &graph);
        if (rc || ce != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            h->substeps = sub0;
            h->sorted_at = sorted0;
            h->vcur = e.vcur;
            h->pcur = e.pcur;
            h->p2g_pending = e.pending;
            return rc ? rc : fail(MPM_EHIP, std::string("graph capture: ") + hipGetErrorString(ce));
        }
        hipError_t ie = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        e.end_phase = h->substeps - h->sorted_at;
        e.end_vcur = h->vcur;
        e.end_pcur = h->pcur;
        h->substeps = sub0;  // the state advances when the graph runs
        h->sorted_at = sorted0;
        h->vcur = e.vcur;
        h->pcur = e.pcur;
        h->p2g_pending = e.pending;
        if (ie != hipSuccess) return fail(MPM_EHIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
        g = e;
        ++h->graph_captures;
    }
    HIPCHK(hipGraphLaunch(g.exec, h->stream));
    ++h->graph_replays;
    g.used = h->graph_replays;
    h->substeps += k;
    h->sorted_at = h->substeps - g.end_phase;
    h->vcur = g.end_vcur;
    h->pcur = g.end_pcur;
    h->p2g_pending = true;  // every steps(k >= 1) call ends with the next P2G in the grid
    return MPM_OK;
}

int finish_steps(mpm_engine* h) {
    int rc;
    if ((rc = read_counters(h))) return rc;
    h->last_overflow = h->pinned[8];
    h->last_claim_losers = h->pinned[9];
    h->chunk_est = (int)h->pinned[3];
    h->last_rounds = h->pinned[4];
    return MPM_OK;
}

// Numbers the chunk blocks (pairs of x-adjacent 4^3 flag blocks; the last
// one of a row is single when a scene has an odd number of flag blocks in x)
// of every scene in Morton order of their coordinates, scene after scene:
// rank[flag block] is its block's number, block[number] its lower flag block.
// Allocates the re-sort buckets (cleared here; each re-sort leaves them
// cleared).
int build_block_table(mpm_engine* h) {
    auto spread = [](uint64_t v) {  // bit i of v -> bit 3i
        uint64_t r = 0;
        for (int i = 0; i < 21; ++i) r |= ((v >> i) & 1ull) << (3 * i);
        return r;
    };
    std::vector<uint32_t> rank((size_t)h->nfb), block;
    std::vector<std::pair<uint64_t, uint32_t>> order;
    uint32_t next = 0;
    for (const DevCfg& d : h->dcs) {
        const int pbx = (d.fbx + 1) / 2;  // blocks per x row: pairs of flag blocks
        order.clear();
        order.reserve((size_t)pbx * d.fby * d.fbz);
        for (int z = 0; z < d.fbz; ++z)
            for (int y = 0; y < d.fby; ++y)
                for (int x = 0; x < pbx; ++x)
                    order.emplace_back((spread((uint64_t)z) << 2) | (spread((uint64_t)y) << 1) | spread((uint64_t)x),
                                       (uint32_t)(d.foff + 2 * x + d.fbx * (y + d.fby * z)));
        std::sort(order.begin(), order.end());
        for (const auto& o : order) {
            rank[o.second] = next;
            if ((int)(o.second - d.foff) % d.fbx + 1 < d.fbx) rank[o.second + 1] = next;
            block.push_back(o.second);
            ++next;
        }
    }
    const int nb = (int)next;
    // bucket keys hold the block number above 7 cell bits
    if (nb >= (1 << (32 - BLOCK_BITS))) return fail(MPM_ELIMIT, "scenes hold 2^25 or more 8x4x4 chunk blocks in all");
    BlockTable& b = h->blocks;
    b.nb = nb;
    int rc;
    if ((rc = dev_alloc(&b.rank, (size_t)h->nfb)) || (rc = dev_alloc(&b.block, (size_t)nb)) ||
        (rc = dev_alloc(&b.count, (size_t)nb)) || (rc = dev_alloc(&b.scan, (size_t)nb)) ||
        (rc = dev_alloc(&b.partial, scan_partial_elems((size_t)nb) + 16)) || (rc = dev_alloc(&b.prev_rounds, (size_t)nb)))
        return rc;
    HIPCHK(hipMemset(b.prev_rounds, 0, (size_t)nb * sizeof(uint16_t)));
    HIPCHK(hipMemcpy(b.rank, rank.data(), rank.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.block, block.data(), block.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(b.count, 0, (size_t)nb * sizeof(uint32_t)));
    return MPM_OK;
}

int create_engine(const mpm_sim_config* cfgs_in, int S, mpm_engine** out) {
    if (!cfgs_in || !out) return fail(MPM_EINVAL, "null argument");
    *out = nullptr;
    if (S < 1 || S > MAX_SCENES) return fail(MPM_ELIMIT, "scene count must be in [1, 4096]");
    mpm_engine* h = new mpm_engine();
    auto cleanup = [&](int rc) {
        mpm_engine_destroy(h);
        return rc;
    };
    h->S = S;
    int64_t G = 0, nfb = 0;
    for (int s = 0; s < S; ++s) {
        mpm_sim_config cfg = cfgs_in[s];
        if (cfg.SIZE_X == 0 || cfg.SIZE_Y == 0 || cfg.SIZE_Z == 0)
            return cleanup(fail(MPM_EINVAL, "SIZE_X/Y/Z must be positive"));
        if (cfg.SIZE_X > 1024 || cfg.SIZE_Y > 1024 || cfg.SIZE_Z > 1024)
            return cleanup(fail(MPM_ELIMIT, "SIZE_X/Y/Z above 1024 (Morton keys cover 10 bits per axis)"));
        mpm_config_derive(&cfg);
        DevCfg d{};
        d.sx = (int)cfg.SIZE_X; d.sy = (int)cfg.SIZE_Y; d.sz = (int)cfg.SIZE_Z;
        d.sxy = d.sx * d.sy;
        d.pad = cfg.PADDING;
        d.gravity = cfg.GRAVITY; d.dt = cfg.DT; d.gamma = cfg.GAMMA; d.compression = cfg.COMPRESSION;
        d.vol = cfg.RESOLUTION * cfg.RESOLUTION * cfg.RESOLUTION;
        d.mass = 1.0f;
        d.fbx = (d.sx + 3) / 4; d.fby = (d.sy + 3) / 4; d.fbz = (d.sz + 3) / 4;
        const int64_t gs = (int64_t)d.sx * d.sy * d.sz;
        if (G + gs >= (1ll << 31)) return cleanup(fail(MPM_ELIMIT, "scenes hold 2^31 or more grid nodes in all"));
        d.goff = (int)G;
        d.foff = (int)nfb;
        G += gs;
        nfb += (int64_t)d.fbx * d.fby * d.fbz;
        h->cfgs.push_back(cfg);
        h->dcs.push_back(d);
    }
    // bucket keys hold a block rank above 6 cell bits
    if (nfb >= (1ll << 26)) return cleanup(fail(MPM_ELIMIT, "scenes hold 2^26 or more 4^3 blocks in all"));
    h->G = (size_t)G;
    h->nfb = (int)nfb;
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
    if ((rc = dev_alloc(&h->d_cfgs, (size_t)S))) return cleanup(rc);
    if ((rc = dev_alloc(&h->d_block_scene, (size_t)h->nfb)) || (rc = dev_alloc(&h->grid_list, (size_t)h->nfb)))
        return cleanup(rc);
    {
        std::vector<uint16_t> bs((size_t)h->nfb);
        for (int s = 0; s < S; ++s) {
            const DevCfg& d = h->dcs[s];
            std::fill(bs.begin() + d.foff, bs.begin() + d.foff + (size_t)d.fbx * d.fby * d.fbz, (uint16_t)s);
        }
        if (hipMemcpy(h->d_block_scene, bs.data(), bs.size() * sizeof(uint16_t), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(h->d_cfgs, h->dcs.data(), (size_t)S * sizeof(DevCfg), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(MPM_EHIP, "scene table upload failed"));
    }
    if ((rc = build_block_table(h))) return cleanup(rc);
    if ((rc = dev_alloc(&h->sc.counters, 64))) return cleanup(rc);
    if (hipMemsetAsync(h->sc.counters, 0, 64 * sizeof(uint32_t), h->stream) != hipSuccess)
        return cleanup(fail(MPM_EHIP, "memset"));
    h->overflow = h->sc.counters + 8;
    if (hipHostMalloc((void**)&h->pinned, 64 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(MPM_ENOMEM, "hipHostMalloc failed"));
    h->pstart.assign((size_t)S + 1, 0);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return cleanup(fail(MPM_EHIP, "stream sync"));
    *out = h;
    return MPM_OK;
}

// Loads counts[s] particles per scene from a 26 x sum(counts) SoA buffer
// (scenes consecutive) and lays them out.
int init_particles_impl(mpm_engine* h, const int* counts, const float* soa26, bool init_f = true) {
    int64_t total = 0;
    for (int s = 0; s < h->S; ++s) {
        if (counts[s] < 0) return fail(MPM_EINVAL, "negative particle count");
        total += counts[s];
    }
    if (total > (1ll << 30)) return fail(MPM_ELIMIT, "more than 2^30 particles");
    const int n = (int)total;
    if (n > 0 && !soa26) return fail(MPM_EINVAL, "null particle buffer");
    free_particles(h);
    int rc;
    for (int b = 0; b < 2; ++b)
        if ((rc = zero_grid(h, b))) return rc;
    h->vcur = 0;
    h->p2g_pending = false;  // both grids were zeroed above
    // both grid-update list counters start at zero with vcur = 0
    HIPCHK(hipMemsetAsync(h->sc.counters + 14, 0, 2 * sizeof(uint32_t), h->stream));
    h->n = n;
    h->substeps = 0;
    h->sorted_at = 0;
    h->pstart[0] = 0;
    for (int s = 0; s < h->S; ++s) h->pstart[s + 1] = h->pstart[s] + counts[s];
    if ((rc = dev_alloc(&h->d_pstart, (size_t)h->S + 1))) return rc;
    HIPCHK(hipMemcpy(h->d_pstart, h->pstart.data(), sizeof(int) * (h->S + 1), hipMemcpyHostToDevice));
    if (n == 0) {
        h->have_particles = true;
        HIPCHK(hipStreamSynchronize(h->stream));
        return MPM_OK;
    }
    float* tmp = nullptr;
    if ((rc = dev_alloc(&tmp, (size_t)MPM_PARTICLE_ATTRS * n))) return rc;
    HIPCHK(hipMemcpyAsync(tmp, soa26, sizeof(float) * MPM_PARTICLE_ATTRS * (size_t)n, hipMemcpyHostToDevice, h->stream));
    if (init_f) launch_init_F(tmp, n, h->stream);
    if ((rc = ensure_sort_scratch(h, (size_t)n)) || (rc = set_chunk_cap(h, h->chunk_cap)) ||
        (rc = ensure_planes(h, 0, (size_t)n)) ||
        (rc = ensure_planes(h, 1, (size_t)n))) {
        dev_free(tmp);
        return rc;
    }
    h->chunk_est = h->chunk_limit;
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

int sdf_from_triangles(mpm_engine* h, int scene, const mpm_triangle* tris, int nt) {
    if (nt < 0 || (nt > 0 && !tris)) return fail(MPM_EINVAL, "bad triangle buffer");
    const DevCfg& d = h->dcs[scene];
    if (nt == 0) {
        launch_fill_f32(h->sdf + d.goff, (size_t)d.sxy * d.sz, 1e15f, h->stream);
        HIPCHK(hipStreamSynchronize(h->stream));
        return MPM_OK;
    }
    float* dt = nullptr;
    int rc;
    if ((rc = dev_alloc(&dt, (size_t)9 * nt))) return rc;
    HIPCHK(hipMemcpyAsync(dt, tris, sizeof(mpm_triangle) * (size_t)nt, hipMemcpyHostToDevice, h->stream));
    hipError_t e = hipSuccess;
    if (h->sdf_bvh) {
        unsigned int ex = 0;
        if (launch_make_sdf_bvh(d, dt, reinterpret_cast<const float*>(tris), nt, h->sdf, &ex, h->stream) != 0)
            e = hipErrorUnknown;
        h->sdf_exact_nodes = ex;
    } else {
        launch_make_sdf(d, dt, nt, h->sdf, h->stream);
        h->sdf_exact_nodes = 0;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    dev_free(dt);
    if (e != hipSuccess) return fail(MPM_EHIP, std::string("makeSDF: ") + hipGetErrorString(e));
    HIPCHK(hipGetLastError());
    return MPM_OK;
}

int ensure_corr_buffers(mpm_engine* h) {
    if (h->ccap > 0) return MPM_OK;
    const size_t cap = h->G;
    int rc;
    if ((rc = dev_alloc(&h->clist, cap)) || (rc = dev_alloc(&h->cX, cap * CORR_STRIDE)) ||
        (rc = dev_alloc(&h->cY, cap * 3)) || (rc = dev_alloc(&h->cacc, cap * 3)) || (rc = dev_alloc(&h->chits, cap)) ||
        (rc = dev_alloc(&h->cV, cap * 3)))
        return rc;
    HIPCHK(hipMemset(h->cV, 0, cap * 3 * sizeof(float)));
    h->ccap = (int)cap;
    return MPM_OK;
}

int upload_dt(mpm_engine* h, float scale) {
    bool same = true;
    for (int s = 0; s < h->S; ++s) same = same && h->dcs[s].dt == h->cfgs[s].DT * scale;
    int rc;
    if (!same && (rc = discard_pending_p2g(h))) return rc;  // its pressure term used the old DT
    for (int s = 0; s < h->S; ++s) h->dcs[s].dt = h->cfgs[s].DT * scale;
    HIPCHK(hipMemcpyAsync(h->d_cfgs, h->dcs.data(), (size_t)h->S * sizeof(DevCfg), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int copy_grid_out(mpm_engine* h, int which, size_t off, size_t count, float* out, size_t out_stride) {
    for (int p = 0; p < 4; ++p)
        HIPCHK(hipMemcpyAsync(out + p * out_stride, h->grid[which].plane[p] + off, count * sizeof(float),
                              hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

}  // namespace

extern "C" {

const char* mpm_last_error(void) { return g_err.c_str(); }

int mpm_version(void) { return 110; }

int mpm_config_derive(mpm_sim_config* cfg) {
    if (!cfg) return fail(MPM_EINVAL, "null config");
    const uint64_t g = (uint64_t)cfg->SIZE_X * cfg->SIZE_Y * cfg->SIZE_Z;
    cfg->GRID_NUMBER = (uint32_t)g;
    cfg->GRID_SIZE = (uint32_t)(g * 16u);
    cfg->GRID_BLOCKS = (uint32_t)((g + 255u) / 256u);
    return MPM_OK;
}

int mpm_engine_create(const mpm_sim_config* cfg, mpm_engine** out) { return create_engine(cfg, 1, out); }

int mpm_batch_create(const mpm_sim_config* cfgs, int nscenes, mpm_engine** out) {
    return create_engine(cfgs, nscenes, out);
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
    dev_free(h->clist);
    dev_free(h->cX);
    dev_free(h->cY);
    dev_free(h->cV);
    dev_free(h->cacc);
    dev_free(h->chits);
    dev_free(h->check_buf);
    dev_free(h->xbuf);
    h->xbuf_cap = 0;
    dev_free(h->d_cfgs);
    dev_free(h->d_block_scene);
    dev_free(h->grid_list);
    dev_free(h->blocks.rank);
    dev_free(h->blocks.block);
    dev_free(h->blocks.count);
    dev_free(h->blocks.scan);
    dev_free(h->blocks.partial);
    dev_free(h->blocks.prev_rounds);
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
        const int re = (int)std::min<int64_t>(value, 1 << 30);
        // a captured steps(k) graph bakes the re-sort schedule in
        if (re != h->resort_every) drop_graph(h);
        h->resort_every = re;
    } else if (k == "CHUNK") {
        if (value < 64 || value > CHUNK_CAP_MAX || (value % 64) != 0)
            return fail(MPM_EINVAL, "CHUNK must be a multiple of 64 in [64, 2048]");
        if (value == h->chunk_cap) return MPM_OK;
        if (!h->have_particles || h->n == 0) {
            h->chunk_cap = (int)value;
            return MPM_OK;
        }
        // Re-lay the current particles out with the new chunk size.
        int rc = check_engine(h, true);
        if (rc) return rc;
        drop_graph(h);
        if ((rc = set_chunk_cap(h, (int)value)) || (rc = resort(h, true))) return rc;
        h->sorted_at = h->substeps;
        if ((rc = read_counters(h))) return rc;
        h->chunk_est = (int)h->pinned[3];
        h->last_rounds = h->pinned[4];
    } else if (k == "LAUNCH_PCT") {
        if (value < 1 || value > 100) return fail(MPM_EINVAL, "LAUNCH_PCT must be in [1, 100]");
        h->launch_pct = (int)value;
    } else if (k == "SDF_BVH") {
        h->sdf_bvh = value != 0;
    } else if (k == "TIMING") {
        if (value < 0 || value > 2) return fail(MPM_EINVAL, "TIMING must be 0, 1 or 2");
        h->timing = (int)value;
    } else if (k == "GRAPH") {
        h->use_graph = value != 0;
        if (!h->use_graph) drop_graph(h);
    } else if (k == "CORRECT") {
        if (value && !h->corr) return fail(MPM_ESTATE, "no corrector attached (mpm_engine_set_corrector)");
        drop_graph(h);
        h->correct = value != 0;
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
    else if (k == "LAUNCH_PCT") *value = h->launch_pct;
    else if (k == "TIMING") *value = h->timing;
    else if (k == "SDF_BVH") *value = h->sdf_bvh ? 1 : 0;
    else if (k == "SDF_EXACT_NODES") *value = h->sdf_exact_nodes;
    else if (k == "GRAPH") *value = h->use_graph ? 1 : 0;
    else if (k == "CORRECT") *value = h->correct ? 1 : 0;
    else if (k == "GRAPH_REPLAYS") *value = h->graph_replays;
    else if (k == "GRAPH_CAPTURES") *value = h->graph_captures;
    else if (k == "SCENES") *value = h->S;
    else return fail(MPM_EINVAL, "unknown option " + k);
    return MPM_OK;
}

int mpm_engine_init_particles(mpm_engine* h, int n, const float* soa26) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (h->S != 1) return fail(MPM_ESTATE, "engine holds several scenes: use mpm_batch_init_particles");
    if (n < 0) return fail(MPM_EINVAL, "negative particle count");
    return init_particles_impl(h, &n, soa26);
}

int mpm_engine_load_state(mpm_engine* h, int n, const float* soa26) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (h->S != 1) return fail(MPM_ESTATE, "engine holds several scenes: use mpm_batch_init_particles");
    if (n < 0) return fail(MPM_EINVAL, "negative particle count");
    return init_particles_impl(h, &n, soa26, false);
}

int mpm_batch_init_particles(mpm_engine* h, const int* counts, const float* soa26) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (!counts) return fail(MPM_EINVAL, "null counts");
    return init_particles_impl(h, counts, soa26);
}

int mpm_engine_init_sdf_from_triangles(mpm_engine* h, const mpm_triangle* tris, int nt) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (h->S != 1) return fail(MPM_ESTATE, "engine holds several scenes: use mpm_batch_init_sdf_from_triangles");
    return sdf_from_triangles(h, 0, tris, nt);
}

int mpm_batch_init_sdf_from_triangles(mpm_engine* h, int scene, const mpm_triangle* tris, int nt) {
    int rc = check_engine(h, false);
    if (rc || (rc = check_scene(h, scene))) return rc;
    return sdf_from_triangles(h, scene, tris, nt);
}

int mpm_engine_set_sdf(mpm_engine* h, const float* sdf) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (!sdf) return fail(MPM_EINVAL, "null sdf");
    HIPCHK(hipMemcpyAsync(h->sdf, sdf, h->G * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int mpm_batch_set_sdf(mpm_engine* h, int scene, const float* sdf) {
    int rc = check_engine(h, false);
    if (rc || (rc = check_scene(h, scene))) return rc;
    if (!sdf) return fail(MPM_EINVAL, "null sdf");
    const DevCfg& d = h->dcs[scene];
    HIPCHK(hipMemcpyAsync(h->sdf + d.goff, sdf, (size_t)d.sxy * d.sz * sizeof(float), hipMemcpyHostToDevice,
                          h->stream));
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

int mpm_batch_scene_info(mpm_engine* h, int scene, int64_t* out4) {
    if (!h || !out4) return fail(MPM_EINVAL, "null argument");
    int rc = check_scene(h, scene);
    if (rc) return rc;
    const DevCfg& d = h->dcs[scene];
    out4[0] = h->pstart[scene];
    out4[1] = h->pstart[scene + 1] - h->pstart[scene];
    out4[2] = d.goff;
    out4[3] = (int64_t)d.sxy * d.sz;
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
    if (h->use_graph) rc = steps_graph(h, k);
    else rc = enqueue_steps(h, k, transfer_launch_size(h));
    if (rc) return rc;
    return finish_steps(h);
}

int mpm_engine_step(mpm_engine* h) { return mpm_engine_steps(h, 1); }

int mpm_engine_set_dt_scale(mpm_engine* h, float scale) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (!(scale > 0.0f) || scale > 1e6f) return fail(MPM_EINVAL, "DT scale must be positive");
    if ((rc = upload_dt(h, scale))) return rc;
    h->dt_scale = scale;
    return MPM_OK;
}

int mpm_engine_set_corrector(mpm_engine* h, mpm_corrector* c) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (c && (rc = ensure_corr_buffers(h))) return rc;
    drop_graph(h);
    h->corr = corrector_state(c);
    if (!c) h->correct = false;
    return MPM_OK;
}

int mpm_engine_check(mpm_engine* h, float depth, double* out3) {
    int rc = check_engine(h, true);
    if (rc) return rc;
    if (!out3) return fail(MPM_EINVAL, "null output");
    if (!h->check_buf && (rc = dev_alloc(&h->check_buf, 3 + 3 * (size_t)CHECK_BLOCKS))) return rc;
    HIPCHK(hipMemsetAsync(h->check_buf, 0, 3 * sizeof(float), h->stream));
    launch_check(h->planes[h->pcur], h->plane_cap[h->pcur], h->orig[h->pcur], h->n, scene_table(h), h->sdf, depth,
                 h->check_buf, h->stream);
    float v[3] = {0, 0, 0};
    HIPCHK(hipMemcpyAsync(v, h->check_buf, sizeof(v), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    out3[0] = v[0];
    out3[1] = h->n ? v[1] / h->n : 0.0;
    out3[2] = v[2];
    return MPM_OK;
}

int mpm_engine_debug_features(mpm_engine* h, int32_t* nodes_out, float* X_out, int cap, int* count_out) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (cap < 0 || !count_out || (cap > 0 && (!nodes_out || !X_out))) return fail(MPM_EINVAL, "bad arguments");
    if ((rc = ensure_corr_buffers(h))) return rc;
    const GridBuf& V = h->grid[h->vcur];
    launch_set_u32(corr_count(h), 0, 0u, h->stream);
    launch_corr_active(h->d_cfgs, h->d_block_scene, V, h->nfb, h->clist, h->ccap, corr_count(h), h->stream);
    launch_corr_features(h->d_cfgs, V, h->sdf, h->clist, corr_count(h), h->ccap, h->cX, h->stream);
    if ((rc = read_counters(h))) return rc;
    const int rows = (int)std::min<uint32_t>(std::min<uint32_t>(h->pinned[12], (uint32_t)h->ccap), (uint32_t)cap);
    if (rows > 0) {
        std::vector<int2> l((size_t)rows);
        HIPCHK(hipMemcpyAsync(l.data(), h->clist, l.size() * sizeof(int2), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(X_out, h->cX, (size_t)rows * CORR_STRIDE * sizeof(float), hipMemcpyDeviceToHost,
                              h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int i = 0; i < rows; ++i) {
            nodes_out[2 * i] = l[i].x;
            nodes_out[2 * i + 1] = l[i].y;
        }
    }
    *count_out = rows;
    return MPM_OK;
}

int mpm_engine_collect_pairs(mpm_engine* h, int ratio, float* X_out, float* T_out, int cap, int* count_out) {
    int rc = check_engine(h, true);
    if (rc) return rc;
    if (ratio < 2 || cap < 0 || !count_out || (cap > 0 && (!X_out || !T_out)))
        return fail(MPM_EINVAL, "bad collect_pairs arguments");
    *count_out = 0;
    if (h->n == 0) return MPM_OK;
    if ((rc = ensure_corr_buffers(h))) return rc;
    drop_graph(h);
    if ((rc = discard_pending_p2g(h))) return rc;  // its own coarse P2G goes into that buffer
    const bool was_correct = h->correct;
    const float scale0 = h->dt_scale;
    h->correct = false;
    if ((rc = mpm_engine_snapshot(h))) return rc;
    // coarse substep (DT * ratio): P2G + grid update, then the features and
    // velocities of its active nodes
    if ((rc = upload_dt(h, scale0 * ratio))) return rc;
    if ((rc = resort(h, true))) return rc;
    h->sorted_at = h->substeps;
    run_transfer(h, XFER_P2G, transfer_launch_size(h));
    run_grid_update(h);
    const GridBuf& V = h->grid[h->vcur];
    launch_set_u32(corr_count(h), 0, 0u, h->stream);
    launch_corr_active(h->d_cfgs, h->d_block_scene, V, h->nfb, h->clist, h->ccap, corr_count(h), h->stream);
    launch_corr_features(h->d_cfgs, V, h->sdf, h->clist, corr_count(h), h->ccap, h->cX, h->stream);
    launch_corr_gather_vel(h->d_cfgs, V, h->clist, corr_count(h), h->ccap, h->cY, nullptr, 0, h->stream);
    if ((rc = upload_dt(h, scale0))) return rc;
    // back to the saved particles; `ratio` plain substeps record the grid
    // velocities of the same nodes
    if ((rc = mpm_engine_restore(h))) return rc;
    launch_fill_f32(h->cacc, (size_t)h->ccap * 3, 0.0f, h->stream);
    launch_fill_f32(h->chits, (size_t)h->ccap, 0.0f, h->stream);
    h->pair_mode = true;
    rc = enqueue_steps(h, ratio, transfer_launch_size(h));
    h->pair_mode = false;
    h->correct = was_correct;
    if (rc || (rc = finish_steps(h))) return rc;
    const int rows = (int)std::min<uint32_t>(h->pinned[12], (uint32_t)h->ccap);
    std::vector<float> X((size_t)rows * CORR_STRIDE), vc((size_t)rows * 3), acc((size_t)rows * 3), hits((size_t)rows);
    if (rows > 0) {
        HIPCHK(hipMemcpyAsync(X.data(), h->cX, X.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(vc.data(), h->cY, vc.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(acc.data(), h->cacc, acc.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(hits.data(), h->chits, hits.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    // target = grid velocity after the last fine substep's grid update -
    // coarse velocity, for nodes that had mass in every fine substep; no
    // correction (0) for the others (surface nodes the fine steps left or
    // reached late).  G2P carries the corrected grid velocity into the
    // particles' velocities, which persist into the next substep, so the
    // correction aims at the velocity at the end of the interval.
    int w = 0;
    for (int i = 0; i < rows && w < cap; ++i) {
        const bool full = hits[i] >= (float)ratio;
        std::memcpy(X_out + (size_t)w * CORR_STRIDE, X.data() + (size_t)i * CORR_STRIDE, CORR_STRIDE * sizeof(float));
        for (int j = 0; j < 3; ++j)
            T_out[(size_t)w * 3 + j] = full ? acc[(size_t)i * 3 + j] - vc[(size_t)i * 3 + j] : 0.0f;
        ++w;
    }
    *count_out = w;
    return MPM_OK;
}

int mpm_engine_snapshot(mpm_engine* h) {
    int rc = check_engine(h, true);
    if (rc) return rc;
    if (h->n == 0) return MPM_OK;
    if (!h->snap_planes) {
        if ((rc = dev_alloc(&h->snap_planes, (size_t)A_STATE * h->n)) || (rc = dev_alloc(&h->snap_orig, (size_t)h->n)))
            return rc;
    }
    // stream-ordered: whatever runs next on the engine's stream sees the copy
    launch_copy_state(h->planes[h->pcur], h->plane_cap[h->pcur], h->orig[h->pcur], h->snap_planes, (size_t)h->n,
                      h->snap_orig, h->n, A_STATE, h->stream);
    HIPCHK(hipGetLastError());
    h->snap_substeps = h->substeps;
    return MPM_OK;
}

int mpm_engine_restore(mpm_engine* h) {
    int rc = check_engine(h, true);
    if (rc) return rc;
    if (h->n == 0) return MPM_OK;
    if (h->snap_substeps < 0) return fail(MPM_ESTATE, "no snapshot taken");
    drop_graph(h);
    if ((rc = discard_pending_p2g(h))) return rc;
    launch_copy_state(h->snap_planes, (size_t)h->n, h->snap_orig, h->planes[h->pcur], h->plane_cap[h->pcur],
                      h->orig[h->pcur], h->n, A_STATE, h->stream);
    HIPCHK(hipGetLastError());
    h->substeps = h->snap_substeps;
    // The saved slots may belong to another chunk layout: lay them out anew.
    if ((rc = resort(h, true))) return rc;
    h->sorted_at = h->substeps;
    if ((rc = read_counters(h))) return rc;
    h->chunk_est = (int)h->pinned[3];
    h->last_rounds = h->pinned[4];
    return MPM_OK;
}

int mpm_engine_synchronize(mpm_engine* h) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int mpm_engine_get_number(mpm_engine* h) { return h ? h->n : 0; }

// Device staging for read-backs, kept between calls (grows as needed).
int export_buffer(mpm_engine* h, size_t floats, float** out) {
    if (h->xbuf_cap < floats) {
        dev_free(h->xbuf);
        h->xbuf_cap = 0;
        int rc;
        if ((rc = dev_alloc(&h->xbuf, floats))) return rc;
        h->xbuf_cap = floats;
    }
    *out = h->xbuf;
    return MPM_OK;
}

int mpm_engine_get_particles(mpm_engine* h, float* out) {
    int rc = check_engine(h, true);
    if (rc) return rc;
    if (!out) return fail(MPM_EINVAL, "null output");
    if (h->n == 0) return MPM_OK;
    float* tmp = nullptr;
    if ((rc = export_buffer(h, (size_t)MPM_PARTICLE_ATTRS * h->n, &tmp))) return rc;
    launch_export_soa26(h->planes[h->pcur], h->plane_cap[h->pcur], h->orig[h->pcur], h->n, scene_table(h), tmp,
                        h->stream);
    hipError_t e = hipMemcpyAsync(out, tmp, sizeof(float) * MPM_PARTICLE_ATTRS * (size_t)h->n, hipMemcpyDeviceToHost,
                                  h->stream);
    hipError_t e2 = hipStreamSynchronize(h->stream);
    if (e != hipSuccess || e2 != hipSuccess) return fail(MPM_EHIP, "get_particles copy failed");
    return MPM_OK;
}

int mpm_engine_get_positions(mpm_engine* h, float* xyz) {
    int rc = check_engine(h, true);
    if (rc) return rc;
    if (!xyz) return fail(MPM_EINVAL, "null output");
    if (h->n == 0) return MPM_OK;
    float* tmp = nullptr;
    if ((rc = export_buffer(h, (size_t)3 * h->n, &tmp))) return rc;
    launch_export_xyz(h->planes[h->pcur], h->plane_cap[h->pcur], h->orig[h->pcur], h->n, tmp, h->stream);
    hipError_t e = hipMemcpyAsync(xyz, tmp, sizeof(float) * 3 * (size_t)h->n, hipMemcpyDeviceToHost, h->stream);
    hipError_t e2 = hipStreamSynchronize(h->stream);
    if (e != hipSuccess || e2 != hipSuccess) return fail(MPM_EHIP, "get_positions copy failed");
    return MPM_OK;
}

int mpm_engine_get_sort(mpm_engine* h, int32_t* keys_out, int32_t* perm_out) {
    int rc = check_engine(h, true);
    if (rc) return rc;
    if (h->S != 1) return fail(MPM_ESTATE, "get_sort is defined for single-scene engines");
    if (!keys_out || !perm_out) return fail(MPM_EINVAL, "null output");
    const int n = h->n;
    if (n == 0) return MPM_OK;
    float* tmp = nullptr;
    if ((rc = dev_alloc(&tmp, (size_t)3 * n))) return rc;
    launch_export_positions(h->planes[h->pcur], h->plane_cap[h->pcur], h->orig[h->pcur], n, tmp, h->stream);
    launch_keys_reference(tmp, tmp + n, tmp + 2 * (size_t)n, n, h->sc.keys[0], h->sc.vals[0], h->stream);
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
    return copy_grid_out(h, h->vcur, 0, h->G, out, h->G);
}

int mpm_batch_get_grid(mpm_engine* h, int scene, float* out) {
    int rc = check_engine(h, false);
    if (rc || (rc = check_scene(h, scene))) return rc;
    if (!out) return fail(MPM_EINVAL, "null output");
    const DevCfg& d = h->dcs[scene];
    const size_t gs = (size_t)d.sxy * d.sz;
    return copy_grid_out(h, h->vcur, (size_t)d.goff, gs, out, gs);
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
    if ((rc = discard_pending_p2g(h))) return rc;
    const int m = 1 - h->vcur;
    if (h->n > 0) run_transfer(h, XFER_P2G, transfer_launch_size(h));
    if ((rc = copy_grid_out(h, m, 0, h->G, grid4_out, h->G))) return rc;
    return zero_grid(h, m);
}

int mpm_engine_debug_grid_update(mpm_engine* h, const float* in, float* out) {
    int rc = check_engine(h, false);
    if (rc) return rc;
    if (!in || !out) return fail(MPM_EINVAL, "null argument");
    if ((rc = discard_pending_p2g(h))) return rc;
    const int m = 1 - h->vcur;
    for (int p = 0; p < 4; ++p)
        HIPCHK(hipMemcpyAsync(h->grid[m].plane[p], in + p * h->G, h->G * sizeof(float), hipMemcpyHostToDevice, h->stream));
    launch_fill_u8(h->grid[m].flags, (size_t)h->nfb, 1, h->stream);
    launch_grid_update(h->d_cfgs, h->d_block_scene, h->S, h->grid[m], h->grid[h->vcur], false, h->sdf, h->nfb,
                       h->grid_list, h->sc.counters + 13, nullptr, h->stream);
    if ((rc = copy_grid_out(h, m, 0, h->G, out, h->G))) return rc;
    return zero_grid(h, m);
}

int mpm_engine_debug_g2p(mpm_engine* h, const float* in) {
    int rc = check_engine(h, true);
    if (rc) return rc;
    if (!in) return fail(MPM_EINVAL, "null argument");
    if ((rc = discard_pending_p2g(h))) return rc;  // the particles change below
    const int v = h->vcur;
    for (int p = 0; p < 4; ++p)
        HIPCHK(hipMemcpyAsync(h->grid[v].plane[p], in + p * h->G, h->G * sizeof(float), hipMemcpyHostToDevice, h->stream));
    launch_fill_u8(h->grid[v].flags, (size_t)h->nfb, 1, h->stream);
    if (h->n > 0) run_transfer(h, XFER_G2P, transfer_launch_size(h));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int mpm_host_alloc(size_t bytes, void** out) {
    if (!out) return fail(MPM_EINVAL, "null output");
    *out = nullptr;
    if (bytes == 0) return MPM_OK;
    if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        return fail(MPM_ENOMEM, "hipHostMalloc failed");
    }
    return MPM_OK;
}

int mpm_host_free(void* p) {
    if (p && hipHostFree(p) != hipSuccess) return fail(MPM_EHIP, "hipHostFree failed");
    return MPM_OK;
}

int mpm_device_alloc(size_t bytes, void** out) {
    if (!out) return fail(MPM_EINVAL, "null output");
    *out = nullptr;
    if (bytes == 0) return MPM_OK;
    if (hipMalloc(out, bytes) != hipSuccess) {
        *out = nullptr;
        return fail(MPM_ENOMEM, "hipMalloc failed");
    }
    return MPM_OK;
}

int mpm_device_free(void* p) {
    if (p && hipFree(p) != hipSuccess) return fail(MPM_EHIP, "hipFree failed");
    return MPM_OK;
}

int mpm_copy_to_device(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return MPM_OK;
    if (!dst || !src) return fail(MPM_EINVAL, "null argument");
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MPM_OK;
}

int mpm_copy_to_host(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return MPM_OK;
    if (!dst || !src) return fail(MPM_EINVAL, "null argument");
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return MPM_OK;
}

static int probe_copy(size_t bytes, int iters, bool wide, double* ms_per_copy) {
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
        launch_copy(a, b, n, s, wide);  // warm-up
        hipEventRecord(e0, s);
        for (int i = 0; i < iters; ++i) launch_copy(a, b, n, s, wide);
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

int mpm_probe_copy(size_t bytes, int iters, double* ms_per_copy) { return probe_copy(bytes, iters, false, ms_per_copy); }

int mpm_probe_copy_wide(size_t bytes, int iters, double* ms_per_copy) {
    return probe_copy(bytes, iters, true, ms_per_copy);
}

}  // extern "C"
