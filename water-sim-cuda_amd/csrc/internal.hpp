// internal.hpp -- host-side declarations shared by the engine and the kernel
// translation units (launchers only; no kernels are visible across files).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

#include "device_common.hpp"

namespace mpm {

void set_last_error(const std::string& msg);

// Grid buffer: four planes of G floats (mass, x, y, z).  While it accumulates
// P2G the planes hold mass and momentum; after the grid update the x/y/z
// planes hold velocity (as the reference's Grid, common.cuh:110-121).
struct GridBuf {
    float* plane[4];
    uint8_t* flags;  // one byte per 4^3 node block: touched by the last P2G into this buffer
};

// XFER_G2P2G_FULL: the fused transfer that also writes v and C, so that the
// particles are complete while the next substep's P2G is already in the grid.
enum TransferMode : int { XFER_P2G = 2, XFER_G2P = 1, XFER_G2P2G = 3, XFER_G2P2G_FULL = 7 };

struct TransferArgs {
    const DevCfg* cfgs;   // per-scene configs (device table), indexed by ChunkDesc::scene
    float* part;          // particle planes, attr*stride + slot
    size_t stride;
    const ChunkDesc* chunks;
    const uint32_t* nchunks;  // chunk count, on the device
    const float* vel[3];  // grid velocity planes read by G2P
    float* acc[4];        // grid mass/momentum planes written by P2G
    uint8_t* acc_flags;
    const float* sdf;
    unsigned int* overflow;  // [0] particles outside their tile, [1] particles that lost a base-node claim
    // Non-null for the transfers after a re-sort that moved no planes, up to
    // the first one with G2P: slot s reads its particle from `src` at old
    // slot map[s]; a G2P writes it to `part` as usual.
    const uint32_t* map;
    const float* src;
    size_t src_stride;
};

// `grid_chunks` sizes the launch (a multiple of 8 workgroups is used); the
// kernel loops until the device chunk count is covered.
void launch_transfer(int mode, const TransferArgs& a, int grid_chunks, hipStream_t s);

// Grid update over the 4^3 blocks flagged in `m` (blocks of all scenes,
// block_scene[b] = scene of block b), zeroing the blocks flagged in `v_clear`
// (the previous velocity buffer) and clearing those flags.  list (nblocks
// entries) and *count are scratch.  With next_count null *count is cleared
// first; otherwise *count must be zero already and the update clears
// *next_count for the next one (two counters used in turn need no clearing
// launch).  With corr (a corrected substep), the new velocities are also
// written to planes 1..3 of corr->vcopy (a copy the corrector reads while it
// updates the grid in place) and the updated nodes with mass > 1e-9 are
// listed in corr->active (*corr->acount zero on entry, at most acap stored).
struct CorrGridOut {
    GridBuf vcopy;
    int2* active;
    int acap;
    uint32_t* acount;
};
void launch_grid_update(const DevCfg* cfgs, const uint16_t* block_scene, int nscenes, GridBuf m, GridBuf v_clear,
                        bool clear_v, const float* sdf, int nblocks, uint2* list, uint32_t* count,
                        uint32_t* next_count, hipStream_t s, const CorrGridOut* corr = nullptr);

// makeSDF of one scene (c: host copy of its DevCfg) into its part of `sdf`.
void launch_make_sdf(const DevCfg& c, const float* tris, int nt, float* sdf, hipStream_t s);
// The same field through a bounding-volume hierarchy built on the host from
// tris_host (tris_dev: the same triangles on the device, for the exact sign
// fallback).  Synchronises the stream; *exact_count receives the number of
// nodes whose sign needed the exact winding sum.  Returns 0 or -1 (HIP error).
int launch_make_sdf_bvh(const DevCfg& c, const float* tris_dev, const float* tris_host, int nt, float* sdf,
                        unsigned int* exact_count, hipStream_t s);
void launch_fill_f32(float* p, size_t n, float v, hipStream_t s);
void launch_copy(const float4* src, float4* dst, size_t n, hipStream_t s, bool wide);
// Copies `planes` particle planes (n slots) and the caller indices.
void launch_copy_state(const float* src, size_t src_stride, const int* src_orig, float* dst, size_t dst_stride,
                       int* dst_orig, int n, int planes, hipStream_t s);
void launch_fill_u8(uint8_t* p, size_t n, uint8_t v, hipStream_t s);
void launch_init_F(float* soa26, int n, hipStream_t s);

// Export of the slot layout into the caller's order (slot s holds caller
// particle orig[s]).
void launch_export_soa26(const float* part, size_t stride, const int* orig, int n, const SceneTable& t, float* out26,
                         hipStream_t s);
// Caller-order AoS positions (3 floats per particle).
void launch_export_xyz(const float* part, size_t stride, const int* orig, int n, float* out_aos, hipStream_t s);
void launch_export_positions(const float* part, size_t stride, const int* orig, int n, float* out_planes3,
                             hipStream_t s);

// ---- sorting and chunking ----
struct SortScratch {
    uint32_t* keys[2];
    uint32_t* vals[2];
    uint32_t* hist;      // 256 * ntiles
    uint32_t* partial;   // scan partials
    // [0] = or, [1] = and, [3] = chunks, [4] = rounds of the last layout,
    // [6..7] = a 64-bit scan total; the engine keeps its counters at [8..]:
    // [8] tile-overflow particles, [9] extra claim passes, [12] corrector
    // nodes, [13] a stand-alone grid update's list, [14..15] the grid
    // update's list (in turn with the grid buffers).
    uint32_t* counters;
    uint32_t* map;       // old slot of each slot after the last re-sort
    size_t cap;          // elements
    size_t hist_cap;     // elements of hist
    size_t partial_cap;
};

// The chunk blocks (8 x 4 x 4 base cells: pairs of x-adjacent 4^3 flag
// blocks) of all scenes numbered densely, scene by scene and within a scene in
// Morton order of their coordinates, and the per-block buckets of a re-sort.
struct BlockTable {
    uint32_t* rank = nullptr;     // one per flag block: the number of its chunk block
    uint32_t* block = nullptr;    // nb: lower flag block of each chunk block
    uint32_t* count = nullptr;    // nb: bucket sizes (all zero between re-sorts)
    uint64_t* scan = nullptr;     // nb: exclusive prefix of count | chunks << 32
    uint64_t* partial = nullptr;  // scan partials
    // nb: the round count the block's chunk was last laid out with (0 before
    // the first layout); a re-sort keeps it while it is within one of the
    // fewest rounds the block needs, so that a count moving back and forth
    // across a multiple of 64 does not re-deal every particle of the chunk
    uint16_t* prev_rounds = nullptr;
    int nb = 0;
};

// One re-sort: the particles of (src, src_stride, src_orig) are grouped by
// chunk block and laid out in rounds; map[new slot] = old slot, the caller
// indices move to dst_orig and the planes in `mask` to (dst, dst_stride).
struct ResortArgs {
    const float* src;
    size_t src_stride;
    const int* src_orig;  // caller index of each source slot, or null (identity)
    float* dst;
    size_t dst_stride;
    int* dst_orig;
    int n;
    uint32_t mask;        // planes that move (bit a = attribute a)
    SceneTable scenes;
    const uint16_t* block_scene;
    BlockTable blocks;
    uint32_t* keys[2];    // n entries each
    uint32_t* vals[2];
    uint32_t* map;        // n entries: the old slot of each new slot
    uint32_t* counters;   // SortScratch::counters
    ChunkDesc* chunks;
    int chunk_cap;
    int chunk_limit;      // descriptors allocated (bounds the chunk count by construction)
    int chunk_est;        // chunk count estimate (sizes the reorder grid)
};

size_t radix_hist_elems(size_t n);
size_t scan_partial_elems(size_t n);

// p[i] = v as a one-thread kernel (usable under stream capture, unlike a
// memset node).
void launch_set_u32(uint32_t* p, int i, uint32_t v, hipStream_t s);

// The reference's setKeys keys of position planes (vals = index).
void launch_keys_reference(const float* px, const float* py, const float* pz, int n, uint32_t* keys, uint32_t* vals,
                           hipStream_t s);
// Stable LSD radix sort of (keys, vals) pairs; result in keys[r], vals[r] with
// r returned (0 or 1).  With key_bits > 0 the keys are known to fit in that
// many bits and only those digits are sorted; with key_bits = 0 the bits on
// which all keys agree are found first (one host synchronisation) and their
// passes skipped (a stable sort by a constant digit is the identity).
int radix_sort_pairs(SortScratch& sc, int n, int key_bits, hipStream_t s);
// Exclusive scan of n uint32 values; total written to *total_dev.
void scan_exclusive_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* partial, uint32_t* total_dev,
                        hipStream_t s);
// The engine's re-sort (chunk table, round sizes, chunk round counts, their
// total in counters[4], chunk count in counters[3]); nothing waits for the
// device.
void launch_resort(const ResortArgs& a, hipStream_t s);

// ---- learned corrector (corrector.hip) ----
constexpr int CORR_FEATURES = 66;
constexpr int CORR_STRIDE = 72;     // floats per feature row (zero padded)
constexpr int CORR_PARAMS = 12803;  // 66*64 + 64 + 2*(64*64 + 64) + 64*3 + 3

struct CorrState {
    float* params;    // FP32 master weights
    void* frags;      // BF16 MFMA operand fragments derived from params (K = 8, then the K = 16 inference set)
    float* mean;      // CORR_STRIDE
    float* inv_std;   // CORR_STRIDE
    float out_scale;  // correction = output * out_scale
    float out_clamp;  // |output| limit applied by the forward pass
    float out_dead;   // outputs of smaller magnitude give no correction
};

void corr_pack(const float* params, void* frags, hipStream_t s);
// Appends the nodes with mass > 1e-9 of V's flagged blocks to list (count on
// the device, at most cap entries stored).
void launch_corr_active(const DevCfg* cfgs, const uint16_t* block_scene, GridBuf V, int nblocks, int2* list, int cap,
                        uint32_t* count, hipStream_t s);
void launch_corr_features(const DevCfg* cfgs, GridBuf V, const float* sdf, const int2* list, const uint32_t* count,
                          int cap, float* X, hipStream_t s);
void launch_corr_forward(const CorrState& cs, const float* X, int cap, const uint32_t* count, float* Y, hipStream_t s);
// Features + forward pass of the listed nodes in one kernel, the corrected
// velocities written into V; the features read V's masses and the velocity
// copy in planes 1..3 of vcopy (as the grid update left them).
void launch_corr_correct(const DevCfg* cfgs, GridBuf V, GridBuf vcopy, const float* sdf, const int2* list,
                         const uint32_t* count, int cap, const CorrState& cs, hipStream_t s);
// Velocities of the listed nodes: copied (mode 0), accumulated (mode 1) or
// copied (mode 2); modes 1 and 2 count the calls in which the node had mass.
void launch_corr_gather_vel(const DevCfg* cfgs, GridBuf V, const int2* list, const uint32_t* count, int cap, float* out,
                            float* hits, int mode, hipStream_t s);
// out3 += {particles below SDF -depth, sum |det F - 1|, non-finite particles}.
// Fallback checks: out3 must hold 3 + 3 * CHECK_BLOCKS floats (per-workgroup
// partials after the three sums).
constexpr int CHECK_BLOCKS = 1024;
void launch_check(const float* part, size_t stride, const int* orig, int n, const SceneTable& t, const float* sdf,
                  float depth, float* out3, hipStream_t s);

}  // namespace mpm

struct mpm_corrector;
namespace mpm {
const CorrState* corrector_state(const mpm_corrector* c);
}  // namespace mpm
