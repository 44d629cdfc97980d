// kernels_sort.hip -- the re-sort that lays particles out for the transfer
// kernels, the reference's Morton-key sort, and scans.
//
// The reference sorts (key, iota) with thrust::sort_by_key every substep and
// only uses the permutation as an index (engine.cu:92-102).  The engine's
// re-sort instead lays particles out in chunks (at most CHUNK particles of one
// 8 x 4 x 4 block of base cells) whose particles sit in rounds of distinct
// base cells:
//   1. bucket count: each particle's key = (dense Morton rank of its base
//      cell's block) << 7 | (cell within the block); runs of equal blocks in
//      a wave take their ranks in the bucket with one atomic;
//   2. a scan over the blocks of (particles, chunks);
//   3. a scatter into block order;
//   4. the chunk table, one thread per block;
//   5. persistent workgroups over the chunks: a counting sort by cell in LDS,
//      the round layout, and the gather of the planes.
// Particles move little between re-sorts, so consecutive slots mostly share a
// block and the buckets see about one atomic per wave.  The order of the
// particles of a block is not fixed (atomics), which changes only the order
// of floating-point sums downstream.
//
// Radix sort (the reference's key order, for inspection): 8-bit digits,
// 4096-key tiles (4 waves x 16 rounds x 64 lanes).  Ranking inside a tile is
// stable: each wave finds the lanes holding the same digit with eight ballots
// (a wave64 "match"), the lowest such lane updates the wave's running digit
// count in LDS, and tiles/waves/rounds/lanes are visited in index order.
#include <algorithm>

#include "internal.hpp"

namespace mpm {
namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;  // 4096 keys

constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 8;
constexpr int SC_TILE = SC_THREADS * SC_ITEMS;  // 2048 values

__device__ __forceinline__ uint64_t lanemask_lt(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

// A value from another lane over the DPP network (an operand modifier of a
// VALU instruction: no LDS round trip, unlike a shuffle).  Lanes whose source
// lies outside the pattern, or whose row is not in ROW_MASK, read 0.  CTRL:
// row_shr:n = 0x110 + n (within rows of 16 lanes), row_bcast:15 = 0x142 and
// row_bcast:31 = 0x143 (the last lane of the previous row, of the lower half).
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp_read(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xF, false);
}

// Inclusive sum and maximum over a full wave64 (every lane active): shifts by
// 1, 2, 4 and 8 within each row, then the rows' totals carried into the rows
// after them.  Values are at least 0, so a lane reading 0 leaves the maximum
// unchanged.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t x) {
    x += dpp_read<0x111>(x);
    x += dpp_read<0x112>(x);
    x += dpp_read<0x114>(x);
    x += dpp_read<0x118>(x);
    x += dpp_read<0x142, 0xA>(x);
    x += dpp_read<0x143, 0xC>(x);
    return x;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t x) {
    x = max(x, dpp_read<0x111>(x));
    x = max(x, dpp_read<0x112>(x));
    x = max(x, dpp_read<0x114>(x));
    x = max(x, dpp_read<0x118>(x));
    x = max(x, dpp_read<0x142, 0xA>(x));
    x = max(x, dpp_read<0x143, 0xC>(x));
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// Lanes of the wave whose `valid` is set and whose 8-bit digit equals ours.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

__global__ void k_set_u32(uint32_t* p, int i, uint32_t v) { p[i] = v; }

__global__ void k_key_or_and(const uint32_t* __restrict__ keys, int n, uint32_t* __restrict__ out) {
    uint32_t o = 0u, a = 0xFFFFFFFFu;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t k = keys[i];
        o |= k;
        a &= k;
    }
    for (int off = 32; off > 0; off >>= 1) {
        o |= __shfl_xor(o, off);
        a &= __shfl_xor(a, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicOr(out + 0, o);
        atomicAnd(out + 1, a);
    }
}

__global__ __launch_bounds__(RS_THREADS) void k_radix_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                           uint32_t* __restrict__ hist, int ntiles) {
    __shared__ uint32_t wh[RS_WAVES][256];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int i = tid; i < RS_WAVES * 256; i += RS_THREADS) (&wh[0][0])[i] = 0u;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * RS_TILE + (size_t)wave * (RS_TILE / RS_WAVES);
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const size_t idx = base + (size_t)r * 64 + lane;
        const bool valid = idx < (size_t)n;
        const uint32_t dg = valid ? (keys[idx] >> shift) & 255u : 0u;
        const uint64_t peers = match_digit(dg, valid);
        if (valid && (peers & lanemask_lt(lane)) == 0ull) wh[wave][dg] += (uint32_t)__popcll(peers);
    }
    __syncthreads();
    const uint32_t sum = wh[0][tid] + wh[1][tid] + wh[2][tid] + wh[3][tid];
    hist[(size_t)tid * ntiles + blockIdx.x] = sum;
}

__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                              const uint32_t* __restrict__ vin,
                                                              uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                              int n, int shift, const uint32_t* __restrict__ hist,
                                                              int ntiles) {
    __shared__ uint32_t wh[RS_WAVES][256];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int i = tid; i < RS_WAVES * 256; i += RS_THREADS) (&wh[0][0])[i] = 0u;
    const size_t base = (size_t)blockIdx.x * RS_TILE + (size_t)wave * (RS_TILE / RS_WAVES);
    uint32_t key[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const size_t idx = base + (size_t)r * 64 + lane;
        key[r] = idx < (size_t)n ? kin[idx] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const size_t idx = base + (size_t)r * 64 + lane;
        const bool valid = idx < (size_t)n;
        const uint32_t dg = (key[r] >> shift) & 255u;
        const uint64_t peers = match_digit(dg, valid);
        if (valid && (peers & lanemask_lt(lane)) == 0ull) wh[wave][dg] += (uint32_t)__popcll(peers);
    }
    __syncthreads();
    {
        uint32_t run = hist[(size_t)tid * ntiles + blockIdx.x];
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) {
            const uint32_t cnt = wh[w][tid];
            wh[w][tid] = run;
            run += cnt;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const size_t idx = base + (size_t)r * 64 + lane;
        const bool valid = idx < (size_t)n;
        const uint32_t dg = (key[r] >> shift) & 255u;
        const uint64_t peers = match_digit(dg, valid);
        if (valid) {
            const uint32_t pos = wh[wave][dg] + (uint32_t)__popcll(peers & lanemask_lt(lane));
            kout[pos] = key[r];
            vout[pos] = vin[idx];
            if ((peers & lanemask_lt(lane)) == 0ull) wh[wave][dg] = pos + (uint32_t)__popcll(peers);
        }
    }
}

// ---- exclusive scan ----
template <class T> __device__ __forceinline__ T block_exclusive_scan(T v, T& total) {
    __shared__ T wsum[SC_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    T wbase = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SC_THREADS / 64; ++w) {
        const T s = wsum[w];
        if (w < wave) wbase += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wbase + x - v;
}

// Scan inputs: plain uint32 values, or a block bucket's (particles, chunks)
// packed as count | ceil(count / cap) << 32 (no carry: counts total < 2^31).
struct LoadU32 {
    const uint32_t* p;
    __device__ uint32_t operator()(size_t i) const { return p[i]; }
};
struct LoadBucket {
    const uint32_t* count;
    uint32_t cap;
    __device__ uint64_t operator()(size_t i) const {
        const uint32_t c = count[i];
        return (uint64_t)c | ((uint64_t)((c + cap - 1u) / cap) << 32);
    }
};

template <class T, class L>
__global__ __launch_bounds__(SC_THREADS) void k_scan_reduce(L in, size_t n, T* __restrict__ partial) {
    const size_t base = (size_t)blockIdx.x * SC_TILE + (size_t)threadIdx.x * SC_ITEMS;
    T s = 0;
#pragma unroll
    for (int i = 0; i < SC_ITEMS; ++i)
        if (base + i < n) s += in(base + i);
    T total;
    block_exclusive_scan(s, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

template <class T>
__global__ __launch_bounds__(SC_THREADS) void k_scan_partials(T* __restrict__ partial, size_t np,
                                                              T* __restrict__ total_out) {
    T carry = 0;
    for (size_t off = 0; off < np; off += SC_TILE) {
        const size_t base = off + (size_t)threadIdx.x * SC_ITEMS;
        T v[SC_ITEMS];
        T s = 0;
#pragma unroll
        for (int i = 0; i < SC_ITEMS; ++i) {
            v[i] = base + i < np ? partial[base + i] : (T)0;
            s += v[i];
        }
        T total;
        T run = carry + block_exclusive_scan(s, total);
#pragma unroll
        for (int i = 0; i < SC_ITEMS; ++i) {
            if (base + i < np) partial[base + i] = run;
            run += v[i];
        }
        carry += total;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// `in` may read the array that `out` writes (each thread reads its items
// before writing them).
template <class T, class L>
__global__ __launch_bounds__(SC_THREADS) void k_scan_apply(L in, T* out, size_t n, const T* __restrict__ partial) {
    const size_t base = (size_t)blockIdx.x * SC_TILE + (size_t)threadIdx.x * SC_ITEMS;
    T v[SC_ITEMS];
    T s = 0;
#pragma unroll
    for (int i = 0; i < SC_ITEMS; ++i) {
        v[i] = base + i < n ? in(base + i) : (T)0;
        s += v[i];
    }
    T total;
    T run = partial[blockIdx.x] + block_exclusive_scan(s, total);
#pragma unroll
    for (int i = 0; i < SC_ITEMS; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
}

template <class T, class L>
void scan_exclusive(L in, T* out, size_t n, T* partial, T* total_dev, hipStream_t s) {
    const size_t tiles = (n + SC_TILE - 1) / SC_TILE;
    if (tiles == 0) return;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_scan_reduce<T, L>), dim3((unsigned)tiles), dim3(SC_THREADS), 0, s, in, n,
                       partial);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_scan_partials<T>), dim3(1), dim3(SC_THREADS), 0, s, partial, tiles,
                       total_dev);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_scan_apply<T, L>), dim3((unsigned)tiles), dim3(SC_THREADS), 0, s, in, out, n,
                       partial);
}

// ---- the reference's keys ----
// setKeys (kernels.cu:361-371): Morton keys of the positions, vals = index.
__global__ void k_keys_reference(const float* __restrict__ px, const float* __restrict__ py,
                                 const float* __restrict__ pz, int n, uint32_t* __restrict__ keys,
                                 uint32_t* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = sort_key(px[i], py[i], pz[i]);
    vals[i] = (uint32_t)i;
}

// ---- the engine's re-sort ----
__device__ __forceinline__ uint64_t lanemask_le(int lane) { return (2ull << lane) - 1ull; }

// Bucket key of a position: the rank of its base cell's 8 x 4 x 4 block (base
// cells clamped into the grid) above the 7-bit index of the cell within the
// block: the half (x bit 2), then the Morton index within the 4^3 half.  Both
// 4^3 flag blocks of a block carry the block's rank.  With the half lowest,
// the cells two apart in this order -- those a round's neighbouring lanes
// hold when cells have about 8 particles -- form 2x2x2 groups, which sit on
// distinct LDS bank groups in the transfer tile.
__device__ __forceinline__ uint32_t bucket_key(const DevCfg& c, const uint32_t* __restrict__ brank, float x, float y,
                                               float z) {
    const int bx = min(max(base_of(x), 0), c.sx - 1);
    const int by = min(max(base_of(y), 0), c.sy - 1);
    const int bz = min(max(base_of(z), 0), c.sz - 1);
    const int fb = c.foff + (bx >> 2) + c.fbx * ((by >> 2) + c.fby * (bz >> 2));
    const uint32_t cell = (uint32_t)(((bx & 4) >> 2) | ((bx & 1) << 1) | ((by & 1) << 2) | ((bz & 1) << 3) |
                                     ((bx & 2) << 3) | ((by & 2) << 4) | ((bz & 2) << 5));
    return (brank[fb] << BLOCK_BITS) | cell;
}

// 1. keys[i] and the particle's place within its bucket (rank[i]); bcount
// holds zeros on entry and the bucket sizes on exit.  The scene of slot i
// comes from its caller index (orig, or i when orig is null).  A wave takes
// BC_BATCH runs of 64 consecutive slots and issues their atomics together.
// Every load of the wave's runs is issued before any is used (indices clamped
// into [0, n), n > 0; lanes past n are masked at the atomics and stores), so
// the wave waits once per dependent step (positions, block ranks, atomics)
// rather than once per run and step.
constexpr int BC_BATCH = 4;
__global__ __launch_bounds__(256) void k_bucket_count(const float* __restrict__ px, const float* __restrict__ py,
                                                      const float* __restrict__ pz, int n, SceneTable t,
                                                      const int* __restrict__ orig,
                                                      const uint32_t* __restrict__ brank,
                                                      uint32_t* __restrict__ bcount, uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ rank) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    uint32_t key[BC_BATCH], b[BC_BATCH], got[BC_BATCH];
    uint64_t heads[BC_BATCH];
    int nvalid[BC_BATCH], at[BC_BATCH];
    float x[BC_BATCH], y[BC_BATCH], z[BC_BATCH];
#pragma unroll
    for (int j = 0; j < BC_BATCH; ++j) {
        at[j] = min((wave * BC_BATCH + j) * 64 + lane, n - 1);
        x[j] = px[at[j]];
        y[j] = py[at[j]];
        z[j] = pz[at[j]];
    }
    if (t.nscenes > 1) {
#pragma unroll
        for (int j = 0; j < BC_BATCH; ++j)
            key[j] = bucket_key(t.cfgs[scene_of(t, orig ? orig[at[j]] : at[j])], brank, x[j], y[j], z[j]);
    } else {
        const DevCfg c0 = uniform(t.cfgs[0]);
#pragma unroll
        for (int j = 0; j < BC_BATCH; ++j) key[j] = bucket_key(c0, brank, x[j], y[j], z[j]);
    }
#pragma unroll
    for (int j = 0; j < BC_BATCH; ++j) {
        const int i = (wave * BC_BATCH + j) * 64 + lane;
        const bool valid = i < n;  // the valid lanes of a batch are a prefix
        b[j] = valid ? key[j] >> BLOCK_BITS : 0xFFFFFFFFu;
        const uint32_t prev = __shfl_up(b[j], 1);
        const bool head = valid && (lane == 0 || prev != b[j]);
        heads[j] = __ballot(head);
        nvalid[j] = __popcll(__ballot(valid));
    }
#pragma unroll
    for (int j = 0; j < BC_BATCH; ++j) {
        got[j] = 0u;
        if ((heads[j] >> lane) & 1ull) {
            const uint64_t later = heads[j] & ~lanemask_le(lane);
            const int end = later ? __ffsll((unsigned long long)later) - 1 : nvalid[j];
            got[j] = atomicAdd(bcount + b[j], (uint32_t)(end - lane));
        }
    }
#pragma unroll
    for (int j = 0; j < BC_BATCH; ++j) {
        const int i = (wave * BC_BATCH + j) * 64 + lane;
        const bool valid = i < n;
        // the head of a valid lane's run: the highest head at or below it
        const uint64_t mine = heads[j] & lanemask_le(lane);
        const int hl = valid ? 63 - __clzll((long long)mine) : lane;
        const uint32_t base = __shfl(got[j], hl);
        if (valid) {
            keys[i] = key[j];
            rank[i] = base + (uint32_t)(lane - hl);
        }
    }
}

// 3. Block order: (key, old slot) to bucket start + rank.
__global__ __launch_bounds__(256) void k_bucket_scatter(const uint32_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ rank, int n,
                                                        const uint64_t* __restrict__ bscan,
                                                        uint32_t* __restrict__ kout, uint32_t* __restrict__ vout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t key = keys[i];
    const uint32_t pos = (uint32_t)bscan[key >> BLOCK_BITS] + rank[i];
    kout[pos] = key;
    vout[pos] = (uint32_t)i;
}

// 4. Chunks of every non-empty bucket: ceil(size / cap) pieces of equal size
// (within one particle), and the bucket cleared for the next re-sort.
// Thread 0 publishes the chunk count (at most `limit`, which bounds it by
// construction) and clears the round total.  bblock holds the lower (even x)
// flag block of each block.
__global__ __launch_bounds__(256) void k_bucket_chunks(uint32_t* __restrict__ bcount,
                                                       const uint64_t* __restrict__ bscan,
                                                       const uint32_t* __restrict__ bblock,
                                                       const uint16_t* __restrict__ block_scene,
                                                       const DevCfg* __restrict__ cfgs, int nb, int cap,
                                                       uint32_t* __restrict__ counters, ChunkDesc* __restrict__ chunks,
                                                       int limit) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) {
        const uint64_t total = *reinterpret_cast<const uint64_t*>(counters + 6);
        counters[3] = min((uint32_t)(total >> 32), (uint32_t)limit);
        counters[4] = 0u;
    }
    if (b >= nb) return;
    const uint32_t cnt = bcount[b];
    if (cnt == 0u) return;
    bcount[b] = 0u;
    const uint64_t sc = bscan[b];
    const int start = (int)(uint32_t)sc;
    int c = (int)(sc >> 32);
    const int fb = (int)bblock[b];
    const int scene = block_scene[fb];
    const DevCfg& cf = cfgs[scene];
    const int local = fb - cf.foff;
    const int bx = local % cf.fbx, byz = local / cf.fbx;
    const int by = byz % cf.fby, bz = byz / cf.fby;
    const int pieces = (int)((cnt + (uint32_t)cap - 1u) / (uint32_t)cap);
    for (int k = 0; k < pieces; ++k, ++c) {
        if (c >= limit) return;
        const int lo = (int)((uint64_t)k * cnt / (uint32_t)pieces);
        const int hi = (int)((uint64_t)(k + 1) * cnt / (uint32_t)pieces);
        ChunkDesc d;
        d.start = start + lo;
        d.count = hi - lo;
        d.ox = 4 * bx;
        d.oy = 4 * by;
        d.oz = 4 * bz;
        d.S = 0;  // set by the reorder
        d.paired = 0;
        d.scene = scene;
        d.block = pieces == 1 ? b : -1;
        chunks[c] = d;
    }
}

constexpr int RO_THREADS = 256;
constexpr int RO_ITEMS = CHUNK_CAP_MAX / RO_THREADS;  // chunk positions per thread
static_assert(RO_ITEMS * RO_THREADS == CHUNK_CAP_MAX, "every chunk position has a thread");

// 5. Persistent workgroups, each taking every gridDim-th chunk.  In LDS: the
// chunk's particles counted and ranked by cell (at most 128 cells: one
// 8 x 4 x 4 block) and listed cell after cell, the members of a cell in
// caller-index order, and the rounds laid out by the wrap-around rule
// (ChunkDesc): with S = max(largest cell, ceil(count / 64)) rounds, list entry
// q goes to round q % S; rounds 0 .. full-1 hold nl = ceil(count / S)
// particles and the others nl - 1.  Within a round the particles are ordered
// by the 16-byte LDS bank group of their base slot in the transfer tile
// (slot mod 8): all first members of the groups, then all second members, and
// so on, so that runs of 8 lanes fall on distinct bank groups where the round
// allows it.  Then map[new slot] = old slot, and the caller indices and the
// planes in `mask` (bit a = attribute a) move to the new slots (PLANES =
// false: mask is 0, only the map and the caller indices).
//
// Chunk position q belongs to thread q % RO_THREADS, which keeps its old slot
// in a register from the load to the final stores.  The old slots, keys and
// caller indices of a workgroup's next chunk are loaded while it lays out the
// current one (the barriers wait for LDS operations only), so the memory
// latency of the loads overlaps the layout work.
//
// Ordering a cell's members by caller index (not by the order of the LDS
// atomics) makes the layout a function of the particles' cells alone: a
// particle that stayed in its cell keeps its list position up to the shifts
// caused by particles that moved, so the map is close to the identity and the
// next transfer's reads through it stay coalesced.  Cells with more than
// RO_SORT_MAX members (far denser than a fluid) keep the atomics' order.
// Dynamic LDS: 3 x cap words (cap <= 2048: chunk positions take 16 bits in
// the list entries).
constexpr int RO_SORT_MAX = 64;
// A chunk takes the paired layout when it needs this share (in %) fewer rounds.
#ifndef MPM_PAIR_SAVE_PCT
#define MPM_PAIR_SAVE_PCT 30
#endif
constexpr uint32_t PAIR_SAVE_PCT = MPM_PAIR_SAVE_PCT;
template <bool PLANES>
__global__ __launch_bounds__(RO_THREADS) void k_reorder(ChunkDesc* __restrict__ ch,
                                                        const uint32_t* __restrict__ counters, int cap,
                                                        const uint32_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ vals,
                                                        const float* __restrict__ src, size_t src_stride,
                                                        const int* __restrict__ src_orig, float* __restrict__ dst,
                                                        size_t dst_stride, int* __restrict__ dst_orig,
                                                        uint32_t* __restrict__ rounds_total,
                                                        uint32_t* __restrict__ map, uint32_t mask,
                                                        uint16_t* __restrict__ prev_rounds) {
    static_assert(BLOCK_CELLS_N == 128 && RO_THREADS >= 128, "one wave scans the cell counts, two per lane");
    extern __shared__ uint32_t ro_lds[];
    uint32_t* cell_rank = ro_lds;          // [cap] cell | atomic rank << BLOCK_BITS per chunk position;
                                           //       then the caller indices in list order; then the new
                                           //       slot (from start) of each chunk position
    uint32_t* org = ro_lds + cap;          // [cap] caller index per chunk position
    uint32_t* lst = ro_lds + 2 * cap;      // [cap] list in atomic order (position | cell << 16);
                                           //       then in final order
    __shared__ uint32_t cnt[BLOCK_CELLS_N];  // zero between chunks
    __shared__ uint32_t cfirst[BLOCK_CELLS_N];
    __shared__ uint32_t cmax;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nch = counters[3];
    uint32_t rsum = 0u;
    // the next chunk's old slots, keys and caller indices, position tid + RO_THREADS i
    uint32_t nsi[RO_ITEMS] = {}, nkey[RO_ITEMS] = {}, norg[RO_ITEMS] = {};
    uint32_t si[RO_ITEMS];  // the current chunk's old slots
    int nstart = 0, ncount = 0, nblk = -1;
    uint32_t nprev = 0u;  // the next chunk's block's previous round count (used when nblk >= 0)
    // (start, count, block) of chunk c, (0, 0, -1) past the last chunk; read
    // two chunks ahead so that no chunk waits for its descriptor
    auto desc_of = [&](uint32_t c) {
        return c < nch ? make_int3(ch[c].start, ch[c].count, ch[c].block) : make_int3(0, 0, -1);
    };
    auto load_slots = [&](int3 d) {
        nstart = __builtin_amdgcn_readfirstlane(d.x);
        ncount = __builtin_amdgcn_readfirstlane(d.y);
        nblk = __builtin_amdgcn_readfirstlane(d.z);
        // loaded with the slots (unconditionally, block 0 standing in for
        // none), so that no wait for it drains the loads issued after it
        // (memory loads complete in order); the index passes through an
        // empty asm so that the compiler does not move the value to a scalar
        // register (and wait for it) right after the load
        int zero = 0;
        asm volatile("" : "+v"(zero));
        nprev = (uint32_t)prev_rounds[max(nblk, 0) + zero];
        // Unconditional loads (positions past the chunk read entry 0, which
        // exists as n > 0): a load under a branch would be copied into the
        // loop-carried register, and the copy waits for the load.
#pragma unroll
        for (int i = 0; i < RO_ITEMS; ++i) {
            const int q = tid + RO_THREADS * i;
            const int at = q < ncount ? nstart + q : 0;
            nsi[i] = vals[at];
            nkey[i] = keys[at];
        }
    };
    auto load_orig = [&]() {  // old slots are < n
        if (src_orig) {
#pragma unroll
            for (int i = 0; i < RO_ITEMS; ++i) norg[i] = (uint32_t)src_orig[nsi[i]];
        } else {
#pragma unroll
            for (int i = 0; i < RO_ITEMS; ++i) norg[i] = nsi[i];
        }
    };
    if (tid < BLOCK_CELLS_N) cnt[tid] = 0u;
    load_slots(desc_of(blockIdx.x));
    int3 dnext = desc_of(blockIdx.x + gridDim.x);
    load_orig();
    // the first chunk's values waited for here, so that the wait at the top
    // of the loop only concerns the loads made during the previous chunk
#pragma unroll
    for (int i = 0; i < RO_ITEMS; ++i) asm volatile("" ::"v"(nsi[i]), "v"(nkey[i]), "v"(norg[i]));
    asm volatile("" ::"v"(nprev));
    __syncthreads();
    for (uint32_t ci = blockIdx.x; ci < nch; ci += gridDim.x) {
        const int start = nstart, count = ncount, blk = nblk;
        const uint32_t prev = __builtin_amdgcn_readfirstlane(nprev);
        // Wait for the loaded values on every path (an empty asm reading
        // them): otherwise the compiler treats their loads as possibly still
        // in flight after the guarded uses below and waits for the next
        // chunk's loads when it reuses these registers.
#pragma unroll
        for (int i = 0; i < RO_ITEMS; ++i) asm volatile("" ::"v"(nsi[i]), "v"(nkey[i]), "v"(norg[i]));
        asm volatile("" ::"v"(nprev));
#pragma unroll
        for (int i = 0; i < RO_ITEMS; ++i) {
            const int q = tid + RO_THREADS * i;
            si[i] = nsi[i];
            if (q < count) {
                const uint32_t c = nkey[i] & (uint32_t)(BLOCK_CELLS_N - 1);
                org[q] = norg[i];
                cell_rank[q] = c | (atomicAdd(&cnt[c], 1u) << BLOCK_BITS);
            }
        }
        load_slots(dnext);  // the next chunk's, in flight during this chunk's layout
        dnext = desc_of(ci + 2u * gridDim.x);
        __syncthreads();
        if (tid < 64) {  // wave 0: cells 2 lane and 2 lane + 1
            const uint32_t a = cnt[2 * lane], b = cnt[2 * lane + 1];
            const uint32_t x = wave_inclusive_sum(a + b);
            const uint32_t m = wave_max(max(a, b));
            cfirst[2 * lane] = x - a - b;
            cfirst[2 * lane + 1] = x - b;
            if (lane == 0) cmax = m;
        }
        __syncthreads();
        for (int q = tid; q < count; q += RO_THREADS) {
            const uint32_t v = cell_rank[q], c = v & (uint32_t)(BLOCK_CELLS_N - 1);
            lst[cfirst[c] + (v >> BLOCK_BITS)] = (uint32_t)q | (c << 16);
        }
        __syncthreads();
        // each list entry's rank among its cell's members by caller index: the
        // entries are taken into registers and their caller indices copied
        // into list order, so that the comparisons read one array; the
        // entries then go back to the list in their final order
        uint32_t ent[RO_ITEMS];
#pragma unroll
        for (int i = 0; i < RO_ITEMS; ++i) {
            const int p = tid + RO_THREADS * i;
            ent[i] = 0u;
            if (p < count) {
                ent[i] = lst[p];
                cell_rank[p] = org[ent[i] & 0xFFFFu];
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < RO_ITEMS; ++i) {
            const int p = tid + RO_THREADS * i;
            if (p < count) {
                const uint32_t c = ent[i] >> 16;
                const uint32_t first = cfirst[c], size = cnt[c];
                uint32_t rank = (uint32_t)p - first;
                if (size <= (uint32_t)RO_SORT_MAX) {
                    const uint32_t g = cell_rank[p];
                    rank = 0u;
#pragma unroll 4
                    for (uint32_t k = 0; k < size; ++k) rank += cell_rank[first + k] < g ? 1u : 0u;
                }
                lst[first + rank] = ent[i];
            }
        }
        load_orig();  // the next chunk's slots have arrived by now
        __syncthreads();
        const uint32_t n = (uint32_t)count;
        // The fewest rounds (a round holds distinct cells, at most 64), or the
        // block's previous round count when that is at most one more: a
        // different S re-deals every list entry (entry q goes to round q % S),
        // so the next transfer would read the chunk's particles in scattered
        // order, while an unchanged S moves them by whole rounds.  Where the
        // fullest cell sets the count (dense cells in a sparse block), a round
        // may hold two members of a cell instead, which the transfer pairs in
        // adjacent lanes (one arena update for both): taken when it saves at
        // least PAIR_SAVE_PCT % of the rounds, as a paired round costs about a
        // third more.
        const uint32_t s_one = max(cmax, (n + 63u) / 64u);
        const uint32_t s_two = max((cmax + 1u) / 2u, (n + 63u) / 64u);
        const bool paired = 100u * s_two <= (100u - PAIR_SAVE_PCT) * s_one;
        const uint32_t s_min = paired ? s_two : s_one;
        uint32_t S = s_min;
        if (blk >= 0 && prev >= s_min && prev <= s_min + 1u) S = prev;
        const uint32_t nl = (n + S - 1u) / S;     // lanes of the first rounds (<= 64 as S >= n / 64)
        const uint32_t full = n - (nl - 1u) * S;  // rounds with nl lanes, 1 .. S
        // one wave per round: lane L holds list entry L * S + r, so two
        // members of a cell in one round (entries q and q + S) are in lanes L
        // and L + 1
        for (uint32_t r = (uint32_t)wave; r < S; r += RO_THREADS / 64) {
            const uint32_t m = r < full ? nl : nl - 1u;
            const bool act = (uint32_t)lane < m;
            const uint32_t e = act ? lst[(uint32_t)lane * S + r] : 0u;
            const uint32_t c = e >> 16;
            // pairs go first, to lanes 2k and 2k + 1 in their order; then the
            // single members (only paired chunks have pairs)
            bool first = false, second = false;
            uint64_t fm = 0ull;
            uint32_t pairs_before = 0u;
            if (paired) {
                const uint32_t c_next = (uint32_t)__shfl_down((int)c, 1), c_prev = (uint32_t)__shfl_up((int)c, 1);
                first = act && (uint32_t)lane + 1u < m && c_next == c;
                second = act && lane > 0 && c_prev == c;
                fm = __ballot(first);
                pairs_before = (uint32_t)__popcll(fm & lanemask_lt(lane));
            }
            const bool single = act && !first && !second;
            const uint32_t cx = ((c >> 1) & 1u) | (((c >> 4) & 1u) << 1) | ((c & 1u) << 2);
            const uint32_t cy = ((c >> 2) & 1u) | (((c >> 5) & 1u) << 1);
            const uint32_t cz = ((c >> 3) & 1u) | (((c >> 6) & 1u) << 1);
            const uint32_t g = (cx + (uint32_t)TILE_PY * cy + (uint32_t)TILE_PZ * cz) & 7u;
            uint32_t cg[8], rank = 0u;
#pragma unroll
            for (uint32_t h = 0; h < 8; ++h) {
                const uint64_t mk = __ballot(single && g == h);
                cg[h] = (uint32_t)__popcll(mk);
                if (g == h) rank = (uint32_t)__popcll(mk & lanemask_lt(lane));
            }
            // singles: members of lower rank in any group, then lower groups of
            // the same rank
            uint32_t pos = 2u * (uint32_t)__popcll(fm);
#pragma unroll
            for (uint32_t h = 0; h < 8; ++h) pos += min(cg[h], rank) + ((h < g && cg[h] > rank) ? 1u : 0u);
            if (first) pos = 2u * pairs_before;
            if (second) pos = 2u * pairs_before - 1u;  // lane - 1 is the pair's first, counted before
            if (act) cell_rank[e & 0xFFFFu] = (r < full ? r * nl : full * nl + (r - full) * (nl - 1u)) + pos;
        }
        if (tid == 0) {
            ch[ci].S = (int)S;
            ch[ci].paired = paired ? 1 : 0;
            rsum += S;
            if (blk >= 0) prev_rounds[blk] = (uint16_t)min(S, 65535u);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < RO_ITEMS; ++i) {
            const int q = tid + RO_THREADS * i;
            if (q < count) {
                const size_t t = (size_t)start + cell_rank[q];
                map[t] = si[i];
                if constexpr (PLANES) {
#pragma unroll
                    for (int a = 0; a < A_STATE; ++a)
                        if (mask & (1u << a)) dst[(size_t)a * dst_stride + t] = src[(size_t)a * src_stride + si[i]];
                }
                dst_orig[t] = (int)org[q];
            }
        }
        if (tid < BLOCK_CELLS_N) cnt[tid] = 0u;
        __syncthreads();  // the LDS arrays are reused by the next chunk
    }
    if (tid == 0 && rsum) atomicAdd(rounds_total, rsum);
}

}  // namespace

size_t radix_hist_elems(size_t n) { return 256 * ((n + RS_TILE - 1) / RS_TILE + 1); }
size_t scan_partial_elems(size_t n) { return (n + SC_TILE - 1) / SC_TILE + 1; }

void scan_exclusive_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* partial, uint32_t* total_dev,
                        hipStream_t s) {
    if (n == 0) {
        if (total_dev) hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, total_dev, 0, 0u);
        return;
    }
    scan_exclusive<uint32_t>(LoadU32{in}, out, n, partial, total_dev, s);
}

void launch_set_u32(uint32_t* p, int i, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, p, i, v);
}

void launch_keys_reference(const float* px, const float* py, const float* pz, int n, uint32_t* keys, uint32_t* vals,
                           hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_keys_reference, dim3((n + 255) / 256), dim3(256), 0, s, px, py, pz, n, keys, vals);
}

int radix_sort_pairs(SortScratch& sc, int n, int key_bits, hipStream_t s) {
    if (n <= 1) return 0;
    uint32_t diff;
    if (key_bits > 0) {
        diff = key_bits >= 32 ? 0xFFFFFFFFu : ((1u << key_bits) - 1u);
    } else {
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, sc.counters, 0, 0u);
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, sc.counters, 1, 0xFFFFFFFFu);
        int blocks = (n + 255) / 256;
        if (blocks > 1024) blocks = 1024;
        hipLaunchKernelGGL(k_key_or_and, dim3(blocks), dim3(256), 0, s, sc.keys[0], n, sc.counters);
        uint32_t oa[2] = {0u, 0u};
        // On a failed read-back every digit is sorted (always correct); the
        // error stays pending for the caller's next status check.
        if (hipMemcpyAsync(oa, sc.counters, sizeof(oa), hipMemcpyDeviceToHost, s) == hipSuccess &&
            hipStreamSynchronize(s) == hipSuccess)
            diff = oa[0] ^ oa[1];
        else
            diff = 0xFFFFFFFFu;
    }
    const int ntiles = (n + RS_TILE - 1) / RS_TILE;
    int r = 0;
    for (int shift = 0; shift < 32; shift += 8) {
        if (((diff >> shift) & 255u) == 0u) continue;
        hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(RS_THREADS), 0, s, sc.keys[r], n, shift, sc.hist, ntiles);
        scan_exclusive_u32(sc.hist, sc.hist, (size_t)256 * ntiles, sc.partial, nullptr, s);
        hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(RS_THREADS), 0, s, sc.keys[r], sc.vals[r],
                           sc.keys[1 - r], sc.vals[1 - r], n, shift, sc.hist, ntiles);
        r = 1 - r;
    }
    return r;
}

void launch_resort(const ResortArgs& a, hipStream_t s) {
    const int n = a.n;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    const BlockTable& bt = a.blocks;
    if (n > 0)
        hipLaunchKernelGGL(k_bucket_count, dim3((unsigned)((n + 256 * BC_BATCH - 1) / (256 * BC_BATCH))), dim3(256),
                           0, s, a.src + (size_t)A_PX * a.src_stride,
                           a.src + (size_t)A_PY * a.src_stride, a.src + (size_t)A_PZ * a.src_stride, n, a.scenes,
                           a.src_orig, bt.rank, bt.count, a.keys[0], a.vals[0]);
    scan_exclusive<uint64_t>(LoadBucket{bt.count, (uint32_t)a.chunk_cap}, bt.scan, (size_t)bt.nb, bt.partial,
                             reinterpret_cast<uint64_t*>(a.counters + 6), s);
    if (n > 0)
        hipLaunchKernelGGL(k_bucket_scatter, dim3(blocks), dim3(256), 0, s, a.keys[0], a.vals[0], n, bt.scan,
                           a.keys[1], a.vals[1]);
    hipLaunchKernelGGL(k_bucket_chunks, dim3((unsigned)((bt.nb + 255) / 256)), dim3(256), 0, s, bt.count, bt.scan,
                       bt.block, a.block_scene, a.scenes.cfgs, bt.nb, a.chunk_cap, a.counters, a.chunks,
                       a.chunk_limit);
    if (n <= 0) return;
    // persistent workgroups: as many as fit on the device at once
    const size_t lds = (size_t)3 * a.chunk_cap * sizeof(uint32_t);
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, a.mask ? (const void*)k_reorder<true>
                                                                     : (const void*)k_reorder<false>,
                                                     RO_THREADS, lds) != hipSuccess)
        cus = per_cu = 0;
    const int resident = std::max(1, cus * per_cu);
    const int grid = std::max(1, std::min(a.chunk_est, resident));
    if (a.mask)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_reorder<true>), dim3((unsigned)grid), dim3(RO_THREADS), lds, s, a.chunks,
                           a.counters, a.chunk_cap, a.keys[1], a.vals[1], a.src, a.src_stride, a.src_orig, a.dst,
                           a.dst_stride, a.dst_orig, a.counters + 4, a.map, a.mask, bt.prev_rounds);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_reorder<false>), dim3((unsigned)grid), dim3(RO_THREADS), lds, s, a.chunks,
                           a.counters, a.chunk_cap, a.keys[1], a.vals[1], a.src, a.src_stride, a.src_orig, a.dst,
                           a.dst_stride, a.dst_orig, a.counters + 4, a.map, 0u, bt.prev_rounds);
}

}  // namespace mpm
