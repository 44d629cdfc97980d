/*
 * mpm.h -- C ABI of the MI355X MLS-MPM engine (libmpm.so).
 *
 * Drop-in boundary for the reference's step path: the C++ class `Engine`
 * (engine.h:14-43, engine.cu:45-114) and the config plumbing around it
 * (SimConfig common.cuh:65-85, uploadConfigToGPU config.cu:5-8,
 * GameConfigData game_configdata.h:15-32, Simulation::initDeviceParams
 * Simulation.cpp:56-77).  Each entry point names the reference interface it
 * replaces.
 *
 * Conventions
 *  - Every function returns MPM_OK (0) or a negative MPM_E* code; the message
 *    of the last failure on the calling thread is in mpm_last_error().  The
 *    reference calls exit(1) instead (engine.cu:21-28, game_configdata.cpp:56-90).
 *  - Host pointers are plain host memory owned by the caller; the library
 *    copies from and to them and never frees them.  (The reference's
 *    Engine::init takes ownership of its buffer, engine.cu:38.)
 *  - Particle buffers use the reference's 26-attribute SoA layout
 *    (ParticleAttr, common.cuh:12-30): element (attr, i) at attr*n + i, attrs
 *    pos[3], vel[3], C[9] row-major, F[9] row-major, VOL, MASS.
 *  - One engine handle per host thread; a handle's calls are synchronous with
 *    respect to the host unless noted.  Config lives in the handle (no
 *    process-global state), so several engines may coexist.
 */
#ifndef MPM_H
#define MPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_OK 0
#define MPM_EINVAL (-1)   /* bad argument */
#define MPM_EHIP (-2)     /* HIP runtime error */
#define MPM_ESTATE (-3)   /* call order violated (e.g. step before init) */
#define MPM_EIO (-4)      /* file or parse error */
#define MPM_ENOMEM (-5)   /* allocation failed */
#define MPM_ELIMIT (-6)   /* size beyond a documented maximum */

#define MPM_PARTICLE_ATTRS 26

/* SimConfig, common.cuh:65-79 -- same fields, same order, same types.
 * GRID_SIZE = 16*G bytes, GRID_NUMBER = G = SIZE_X*SIZE_Y*SIZE_Z,
 * GRID_BLOCKS = ceil(G/256) are derived (Simulation.cpp:70-72). */
typedef struct mpm_sim_config {
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
} mpm_sim_config;

/* Triangle, common.cuh:123-126: three float3 vertices, 36 bytes. */
typedef struct mpm_triangle {
    float v0[3], v1[3], v2[3];
} mpm_triangle;

/* WATER= / OBJECT= entry, ConfigObject game_configdata.h:6-10. */
typedef struct mpm_config_object {
    char model_path[256];
    float scale;
    float x, y, z;
} mpm_config_object;

typedef struct mpm_engine mpm_engine;

/* Last error message of the calling thread ("" if none). */
const char* mpm_last_error(void);

/* Library version (major*10000 + minor*100 + patch). */
int mpm_version(void);

/* ---- configuration (GameConfigData + initDeviceParams) ---- */

/* Fills GRID_SIZE, GRID_NUMBER, GRID_BLOCKS from SIZE_* (Simulation.cpp:70-72). */
int mpm_config_derive(mpm_sim_config* cfg);

/* GameConfigData::setConfigDataFromFile (game_configdata.cpp:22-54) followed
 * by Simulation::initDeviceParams (Simulation.cpp:56-77): KEY=VALUE lines,
 * spaces removed, WATER=/OBJECT= as path;scale;x;y;z.  Every SimConfig key
 * that the reference reads with getInt/getFloat is required (missing key ->
 * MPM_EIO with the key named, where the reference exits).  `waters`/`objects`
 * may be NULL; *n_waters and *n_objects receive the entry counts, at most
 * `cap` entries are written to each array. */
int mpm_config_parse_file(const char* path, mpm_sim_config* cfg, mpm_config_object* waters,
                          int* n_waters, mpm_config_object* objects, int* n_objects, int cap);
/* Same grammar from an in-memory string. */
int mpm_config_parse_string(const char* text, mpm_sim_config* cfg, mpm_config_object* waters,
                            int* n_waters, mpm_config_object* objects, int* n_objects, int cap);
/* Extra engine keys in the same file (RESORT_EVERY=...), returned as int
 * values; returns MPM_EIO if `key` is absent. */
int mpm_config_get_int(const char* text, const char* key, int64_t* value);

/* ---- engine lifetime (Engine ctor/dtor, engine.cu:36-43) ---- */

/* Creates an engine on the current HIP device with its own stream.  The
 * config is copied into the handle (replaces uploadConfigToGPU, config.cu:5-8,
 * and the __constant__ d_config). */
int mpm_engine_create(const mpm_sim_config* cfg, mpm_engine** out);
int mpm_engine_destroy(mpm_engine* h);

/* Engine options (no reference counterpart; defaults in brackets):
 *   "RESORT_EVERY"  Morton re-sort period in substeps [1 = every substep, as
 *                   engine.cu:57].  The sort changes only the order of the
 *                   float atomics, not the exact-arithmetic result.
 *   "CHUNK"         maximum particles per transfer work item, a multiple of
 *                   64 in [64, 2048] [2048]; changing it re-lays the particles
 *   "LAUNCH_PCT"    transfer workgroups per 100 chunks, 1..100 [100]; below
 *                   100 each workgroup loops over several chunks
 *   "TIMING"        1 = record per-stage HIP events, 2 = around the transfer
 *                   stages only (fewer events in the stream) [0]
 *   "GRAPH"         1 = capture a steps(k) call as a HIP graph once and replay
 *                   it while the schedule state matches [0]
 *   "SDF_BVH"       1 = initSDF through a bounding-volume hierarchy [1]: the
 *                   same field as the brute-force makeSDF (kernels.cu:282-300,
 *                   run with 0): equal distances, equal signs wherever the
 *                   distance is non-zero
 * Read-only (get_option): "GRAPH_REPLAYS" graph launches so far,
 * "GRAPH_CAPTURES" graphs captured so far (a steps(k) call whose schedule
 * state matched no cached graph), "SCENES"
 * scene count, "SDF_EXACT_NODES" nodes of the last hierarchy initSDF whose
 * sign took the exact winding sum. */
int mpm_engine_set_option(mpm_engine* h, const char* key, int64_t value);
int mpm_engine_get_option(mpm_engine* h, const char* key, int64_t* value);

/* ---- batched scenes (no reference counterpart) ----
 * One engine can hold nscenes independent scenes, each with its own config
 * (grid size, DT, GRAVITY, ...), grid, SDF and particles.  Every stage runs as
 * one launch over all scenes.  Particles of all scenes form one caller-order
 * array, scene after scene (get_particles/get_positions return it in that
 * order); grids and SDFs are concatenated scene after scene
 * (mpm_batch_scene_info gives the offsets).  mpm_engine_set_sdf/get_sdf and
 * mpm_engine_get_grid address the concatenation.  Limits: nscenes <= 4096,
 * fewer than 2^31 grid nodes, 2^26 4^3-node blocks and 2^25 8x4x4-cell
 * chunk blocks in all. */
int mpm_batch_create(const mpm_sim_config* cfgs, int nscenes, mpm_engine** out);
/* counts[s] particles per scene from a 26 x sum(counts) SoA buffer. */
int mpm_batch_init_particles(mpm_engine* h, const int* counts, const float* soa26);
int mpm_batch_init_sdf_from_triangles(mpm_engine* h, int scene, const mpm_triangle* tris, int nt);
int mpm_batch_set_sdf(mpm_engine* h, int scene, const float* sdf);
/* Planes mass, v_x, v_y, v_z of one scene after its last grid update. */
int mpm_batch_get_grid(mpm_engine* h, int scene, float* grid4_out);
/* out4 = {first particle, particle count, first grid node, grid nodes}. */
int mpm_batch_scene_info(mpm_engine* h, int scene, int64_t* out4);

/* ---- particles and collider ---- */

/* Engine::init + initParticles + initGrid (engine.cu:45-52, :72-90) and
 * initFMatrices (kernels.cu:349-358): copies n particles from the 26*n host
 * buffer, sets F's diagonal to 1, V = RESOLUTION^3, m = 1, zeroes the grid and
 * lays the particles out for the transfer kernels (the initial sort).  n = 0 is
 * allowed (steps are no-ops). */
int mpm_engine_init_particles(mpm_engine* h, int n, const float* soa26);
/* Loads a saved 26*n state (get_particles output) as it is: like
 * init_particles but without initFMatrices, so F, v and C continue (no
 * reference counterpart: checkpoint/resume). */
int mpm_engine_load_state(mpm_engine* h, int n, const float* soa26);

/* Engine::initSDF + makeSDF (engine.cu:105-114, kernels.cu:282-300): builds the
 * node SDF from nt triangles (nt = 0 gives +1e15 everywhere, as the reference). */
int mpm_engine_init_sdf_from_triangles(mpm_engine* h, const mpm_triangle* tris, int nt);

/* Sets the node SDF directly from G = SIZE_X*SIZE_Y*SIZE_Z floats (index
 * z*X*Y + y*X + x, getGridIdx common.cuh:131-134).  For analytic colliders. */
int mpm_engine_set_sdf(mpm_engine* h, const float* sdf);
int mpm_engine_get_sdf(mpm_engine* h, float* sdf_out);

/* ---- stepping (Engine::step, engine.cu:54-70) ---- */

/* One substep: sort (per RESORT_EVERY), clear, P2G, grid update, G2P. */
int mpm_engine_step(mpm_engine* h);
/* k substeps back to back (Simulation::run's inner loop, Simulation.cpp:46-49).
 * Each call ends with the particles complete and the next substep's P2G
 * already in the grid, which the next call starts from (restore, load_state,
 * a DT change and the debug transfers drop it). */
int mpm_engine_steps(mpm_engine* h, int k);
/* Blocks until the engine's stream is idle. */
int mpm_engine_synchronize(mpm_engine* h);

/* Saves the particle state (positions, velocities, C, F, substep count) in
 * device memory, and restores it, re-laid out for the current options (no
 * reference counterpart: used by the autotuner to rewind its trials).  The
 * grid needs no saving: after a restore the next steps() call starts with a
 * P2G of the restored particles. */
int mpm_engine_snapshot(mpm_engine* h);
int mpm_engine_restore(mpm_engine* h);

/* Multiplies every scene's DT by `scale` (relative to its config) for the
 * following steps (no reference counterpart: the learned corrector runs at
 * 4x DT). */
int mpm_engine_set_dt_scale(mpm_engine* h, float scale);

/* Fallback checks of the current particles (no reference counterpart):
 * out3 = {particles whose trilinear SDF (getSDF, common.cuh:136-167) is below
 * -depth, mean |det F - 1|, particles with a non-finite position or
 * velocity}. */
int mpm_engine_check(mpm_engine* h, float depth, double* out3);

/* ---- read-back ---- */

/* Engine::getNumber (engine.h:31). */
int mpm_engine_get_number(mpm_engine* h);
/* Engine::getBuffer / getParticles (engine.h:22, :32): the full state in the
 * caller's original particle order, 26*n floats. */
int mpm_engine_get_particles(mpm_engine* h, float* soa26_out);
/* Engine::getPositions / changeFormat (engine.h:24, kernels.cu:341-347):
 * positions as x,y,z triples in the caller's order, 3*n floats. */
int mpm_engine_get_positions(mpm_engine* h, float* xyz_out);
/* Engine::sortParticles (engine.cu:92-102): Morton keys of the current
 * positions (setKeys, kernels.cu:361-371) and the stable ascending key sort.
 * keys_out receives the sorted keys, perm_out the caller indices in sorted
 * order (the reference's d_keys and d_values after sort_by_key). */
int mpm_engine_get_sort(mpm_engine* h, int32_t* keys_out, int32_t* perm_out);
/* Engine::getGrid (engine.h:23): the grid after the last grid update, 4*G
 * floats as planes mass, v_x, v_y, v_z (the reference's momentum arrays hold
 * velocity after gridUpdate, kernels.cu:193). */
int mpm_engine_get_grid(mpm_engine* h, float* grid4_out);

/* Per-stage device time accumulated while TIMING is on, in milliseconds, and the
 * number of launches: out[2*s] = ms, out[2*s+1] = count for stage s in
 * 0 = P2G, 1 = the last transfer of a call (fused G2P+P2G also writing v and
 * C), 2 = fused G2P+P2G, 3 = grid update, 4 = re-sort.
 * Resets the accumulators. */
int mpm_engine_get_stage_times(mpm_engine* h, double* out, int nstages);

/* Diagnostic counters of the last mpm_engine_steps call (its transfers, the
 * P2G it leaves for the next call included): out[0] = chunks,
 * out[1] = particle-substeps whose stencil had left their chunk's LDS tile
 * (global-memory path), out[2] = active 4^3 grid blocks, out[3] =
 * particle-substeps that shared a base node with another lane of their round
 * and took an extra pass of the LDS update, out[4] = wave rounds of the
 * current particle layout (one round = up to 64 particles). */
int mpm_engine_get_counters(mpm_engine* h, int64_t* out, int n);

/* ---- stage access for parity tests (no reference counterpart) ---- */

/* P2G of the current state alone (sort + clear + p2GTransferScatter,
 * kernels.cu:6-142) into grid4_out (planes mass, mom_x, mom_y, mom_z).  The
 * engine's state is unchanged. */
int mpm_engine_debug_p2g(mpm_engine* h, float* grid4_out);
/* gridUpdate (kernels.cu:144-196) of an arbitrary grid: grid4_in holds
 * mass/momentum planes, grid4_out receives mass/velocity planes.  The
 * engine's particle state is unchanged. */
int mpm_engine_debug_grid_update(mpm_engine* h, const float* grid4_in, float* grid4_out);
/* G2P (kernels.cu:199-280) of the current particles against the given
 * mass/velocity grid; updates the particle state in place. */
int mpm_engine_debug_g2p(mpm_engine* h, const float* grid4_in);

/* ---- page-locked host memory (no reference counterpart) ----
 * Read-backs (get_positions, get_particles, get_grid) into page-locked host
 * memory run at the full host-link rate; into ordinary (pageable) memory the
 * HIP runtime stages them at a fraction of it.  The reference hands positions
 * to OpenGL through CUDA-GL interop instead (kernels.cu:341-347). */
int mpm_host_alloc(size_t bytes, void** out);
int mpm_host_free(void* p);

/* ---- device memory for callers that stage their own buffers (no reference
 * counterpart; the reference hands out device pointers through
 * Engine::getBuffer/getParticles, engine.h:22, :32) ----
 * Plain device allocations on the current HIP device and synchronous copies,
 * for the entry points below that take device pointers. */
int mpm_device_alloc(size_t bytes, void** out);
int mpm_device_free(void* p);
int mpm_copy_to_device(void* dst, const void* src, size_t bytes);
int mpm_copy_to_host(void* dst, const void* src, size_t bytes);

/* ---- key sort that allocates nothing ----
 * No counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It computes what mpm_engine_get_sort
 * returns -- the Morton keys of setKeys (kernels.cu:361-371) and their stable
 * ascending sort with the particle indices (Engine::sortParticles,
 * engine.cu:92-102, which allocates its keys every substep, engine.cu:95) --
 * for n particles whose positions are three device arrays px, py, pz, into
 * the device arrays keys_out (sorted keys) and perm_out (indices in sorted
 * order).  Work is enqueued on `stream` (a hipStream_t; NULL for the null
 * stream) and the call returns without waiting.  Its scratch is a fixed
 * device array of about 34 MB per GPU, reserved by the library's code object
 * for the life of the process: n above MPM_SORT_NOALLOC_MAX returns
 * MPM_ELIMIT.  Calls share that array, so calls that overlap on one GPU
 * (other streams or host threads) race; serialise them. */
#define MPM_SORT_NOALLOC_MAX (1 << 21)
int mpm_sort_keys_noalloc(const float* px, const float* py, const float* pz, int n, int32_t* keys_out,
                          int32_t* perm_out, void* stream);

/* ---- measurement aid (no reference counterpart) ---- */

/* Copies `bytes` (rounded down to 16) between two fresh device buffers
 * `iters` times with a streaming kernel; *ms_per_copy receives the mean time
 * of one copy.  Used to calibrate HBM byte counters and to report the
 * achievable copy bandwidth next to the engine's figures. */
int mpm_probe_copy(size_t bytes, int iters, double* ms_per_copy);
/* The same with four 16-byte loads in flight per thread before their stores,
 * and nontemporal stores: a stronger estimate of the streaming ceiling. */
int mpm_probe_copy_wide(size_t bytes, int iters, double* ms_per_copy);

/* ---- scene generators (build components; no reference counterpart) ---- */

/* Jittered particles, 8 per cell (one per 2x2x2 sub-cell, uniform jitter from
 * SplitMix64(seed)).  shape 0 = box [lo, hi) in cells, shape 1 = sphere
 * centre lo[0..2], radius hi[0].  Appends to `soa26_out` laid out with stride
 * `cap` (attr*cap + i) starting at index *count; with soa26_out == NULL only
 * counts.  Positions only; all other attributes 0 except VOL = MASS = 1
 * (VoxelPipeline::process, VoxelPipeline.cpp:9-30). */
int mpm_scene_fill(int shape, const float* lo, const float* hi, uint64_t seed, float* soa26_out,
                   int64_t cap, int64_t* count);
/* UV sphere mesh: `slices` around, `stacks` pole to pole, 2*slices*(stacks-1)
 * triangles.  tris_out may be NULL (count only). */
int mpm_mesh_uv_sphere(const float* centre, float radius, int slices, int stacks, mpm_triangle* tris_out,
                       int* ntris);
/* Closed U-shaped channel along x: outer box lo..hi, wall thickness `wall`,
 * floor thickness `wall`, faces split into squares of about `seg` cells. */
int mpm_mesh_u_channel(const float* lo, const float* hi, float wall, float seg, mpm_triangle* tris_out,
                       int* ntris);
/* Axis-aligned box mesh lo..hi (12 triangles, outward winding). */
int mpm_mesh_box(const float* lo, const float* hi, mpm_triangle* tris_out, int* ntris);

/* ---- scene front-end: OBJ meshes and config.txt (Simulation.cpp:4-38) ----
 * Host-only; the results are the particle buffer and triangles that the
 * reference's Simulation constructor hands to Engine::init and initSDF.
 * Two-call pattern: call with cap 0 to get the count. */

/* ObjLoader::loadObj (ObjLoader.cpp:9-33, tinyobjloader with its default
 * triangulation) + VoxelEngine::extractTriangles (VoxelEngine.cpp:858-899):
 * the triangles of an OBJ file; quads are split along the shorter diagonal. */
int mpm_obj_triangles(const char* path, mpm_triangle* out, int cap, int* count);
/* VoxelEngine::voxelize (VoxelEngine.cpp:120-243): centres of the cells of a
 * (longest side / resolution) grid over the mesh's bounding box that a
 * triangle overlaps (separating-axis test); xyz planes of cap floats. */
int mpm_voxelize_obj(const char* path, float resolution, float* xyz, int64_t cap, int64_t* count);
/* Simulation::Prepare_object (Simulation.cpp:80-92): voxelize(mesh, 100) and
 * VoxelEngine::normalize (VoxelEngine.cpp:245-494) to normalize_size, scale,
 * displacement, lattice step `lattice` (RESOLUTION as an integer), 5x5x5
 * dilation, unique points in (x, y, z) order; xyz planes of cap floats. */
int mpm_obj_water(const char* path, float normalize_size, float scale, const float* displacement, int lattice,
                  float* xyz, int64_t cap, int64_t* count);
/* Simulation::Prepare_triangles (Simulation.cpp:94-120): the mesh's
 * triangles through VoxelEngine::normalize (VoxelEngine.cpp:601-827),
 * triangles with no vertex inside [0, normalize_size]^3 dropped. */
int mpm_obj_collider(const char* path, float normalize_size, float scale, const float* displacement,
                     mpm_triangle* out, int cap, int* count);
/* The Simulation constructor's scene (Simulation.cpp:4-38): parses the
 * config file, builds every WATER entry (merged in file order) into the
 * 26 x n buffer of VoxelPipeline::process (positions, zeros, VOL = MASS = 1)
 * and every OBJECT entry into triangles.  Model paths are opened as given
 * (relative to the working directory, as the reference does), else relative
 * to the config file's directory.  The buffer is written when cap >= n. */
int mpm_scene_from_config(const char* config_path, mpm_sim_config* cfg, float* soa26, int64_t cap, int64_t* n,
                          mpm_triangle* tris, int tri_cap, int* ntris);

/* ---- learned grid-velocity corrector (no reference counterpart) ----
 * An MLP 66 -> 64 -> 64 -> 64 -> 3 (ReLU; BF16 weights and activations, FP32
 * accumulation on the matrix cores) maps per-node features after the grid
 * update to a velocity correction that is added before G2P, so that a
 * substep of k*DT approximates k plain substeps.  Parameters: W1[66][64],
 * b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[64][3], b4[3], FP32,
 * W[in][out].  Feature rows have MPM_CORRECTOR_STRIDE floats (66 used, the
 * rest zero); the layout is documented in csrc/corrector.hip. */
#define MPM_CORRECTOR_FEATURES 66
#define MPM_CORRECTOR_STRIDE 72
#define MPM_CORRECTOR_PARAMS 12803
typedef struct mpm_corrector mpm_corrector;

int mpm_corrector_create(mpm_corrector** out);
int mpm_corrector_destroy(mpm_corrector* c);
int mpm_corrector_set_params(mpm_corrector* c, const float* params, int n);
int mpm_corrector_get_params(mpm_corrector* c, float* params, int n);
/* Input normalisation x' = (x - mean) * inv_std (66 each) and the output
 * scale: correction = network output * out_scale. */
int mpm_corrector_set_norm(mpm_corrector* c, const float* mean, const float* inv_std, float out_scale);
/* Limit on |network output| (before out_scale) in the forward pass and in
 * the engine [no limit]: a guard against extrapolation. */
int mpm_corrector_set_clamp(mpm_corrector* c, float clamp);
/* Dead zone of the forward pass and the engine [0]: a node whose network
 * output has a magnitude (before out_scale) below `dead` gets no correction.
 * The network's BF16 activations resolve about 1/256 of its output scale, so
 * outputs that small are rounding, and applying them at every node of a
 * resting fluid only adds noise. */
int mpm_corrector_set_deadzone(mpm_corrector* c, float dead);
/* Y[n][3] = corrections of the n feature rows X[n][72] (host arrays). */
int mpm_corrector_forward(mpm_corrector* c, const float* X, int n, float* Y);
/* Training set on the device: features X[n][72], target corrections T[n][3]. */
int mpm_corrector_load_dataset(mpm_corrector* c, const float* X, const float* T, int n);
/* Gradient (MPM_CORRECTOR_PARAMS floats) and mean loss of rows
 * [row0, row0 + n) (wrapping) of the loaded set, loss = mean over rows and
 * outputs of (output - target / out_scale)^2; no update. */
int mpm_corrector_gradient(mpm_corrector* c, int row0, int n, float* grad_out, float* loss_out);
/* `steps` Adam steps (lr, beta 0.9/0.999) over consecutive batches of the
 * loaded set; loss_out (nullable) gets each step's batch loss. */
int mpm_corrector_train(mpm_corrector* c, int steps, int batch, float lr, float* loss_out);

/* Attaches a corrector (null detaches); option "CORRECT" = 1 applies it after
 * every grid update. */
int mpm_engine_set_corrector(mpm_engine* h, mpm_corrector* c);
/* Training pairs at the current state: features of the active nodes of one
 * substep at ratio*DT (after its grid update) and, as targets, the grid
 * velocity of the same nodes after the last of `ratio` plain substeps minus
 * the coarse velocity, for nodes with mass in every plain substep (0 for the
 * others).  The engine ends
 * `ratio` plain substeps later.  Writes at most cap rows. */
int mpm_engine_collect_pairs(mpm_engine* h, int ratio, float* X_out, float* T_out, int cap, int* count_out);
/* Stage access for parity tests: the active nodes of the current grid
 * (nodes_out[2i] = x | y << 10 | z << 20, nodes_out[2i+1] = scene) and their
 * feature rows, at most cap. */
int mpm_engine_debug_features(mpm_engine* h, int32_t* nodes_out, float* X_out, int cap, int* count_out);

#ifdef __cplusplus
}
#endif

#endif /* MPM_H */
