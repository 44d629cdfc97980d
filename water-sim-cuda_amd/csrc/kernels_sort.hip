// kernels_sort.hip -- Morton keys, stable LSD radix sort, scans, chunking and
// the gather that moves particles into the chunked slot layout.
//
// The reference sorts (key, iota) with thrust::sort_by_key every substep and
// only uses the permutation as an index (engine.cu:92-102).  Here the sort's
// output also drives a physical reorder of the particle planes, so the
// transfer kernels read and write them coalesced.
//
// Radix sort: 8-bit digits, 4096-key tiles (4 waves x 16 rounds x 64 lanes).
// Ranking inside a tile is stable: each wave finds the lanes holding the same
// digit with eight ballots (a wave64 "match"), the lowest such lane updates
// the wave's running digit count in LDS, and tiles/waves/rounds/lanes are
// visited in index order.
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

// ---- exclusive scan of uint32 ----
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t wsum[SC_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t wbase = 0u, tot = 0u;
#pragma unroll
    for (int w = 0; w < SC_THREADS / 64; ++w) {
        const uint32_t s = wsum[w];
        if (w < wave) wbase += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wbase + x - v;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_reduce(const uint32_t* __restrict__ in, size_t n,
                                                            uint32_t* __restrict__ partial) {
    const size_t base = (size_t)blockIdx.x * SC_TILE + (size_t)threadIdx.x * SC_ITEMS;
    uint32_t s = 0u;
#pragma unroll
    for (int i = 0; i < SC_ITEMS; ++i)
        if (base + i < n) s += in[base + i];
    uint32_t total;
    block_exclusive_scan(s, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_partials(uint32_t* __restrict__ partial, size_t np,
                                                              uint32_t* __restrict__ total_out) {
    uint32_t carry = 0u;
    for (size_t off = 0; off < np; off += SC_TILE) {
        const size_t base = off + (size_t)threadIdx.x * SC_ITEMS;
        uint32_t v[SC_ITEMS];
        uint32_t s = 0u;
#pragma unroll
        for (int i = 0; i < SC_ITEMS; ++i) {
            v[i] = base + i < np ? partial[base + i] : 0u;
            s += v[i];
        }
        uint32_t total;
        uint32_t run = carry + block_exclusive_scan(s, total);
#pragma unroll
        for (int i = 0; i < SC_ITEMS; ++i) {
            if (base + i < np) partial[base + i] = run;
            run += v[i];
        }
        carry += total;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_apply(const uint32_t* in, uint32_t* out, size_t n,
                                                           const uint32_t* __restrict__ partial) {
    const size_t base = (size_t)blockIdx.x * SC_TILE + (size_t)threadIdx.x * SC_ITEMS;
    uint32_t v[SC_ITEMS];
    uint32_t s = 0u;
#pragma unroll
    for (int i = 0; i < SC_ITEMS; ++i) {
        v[i] = base + i < n ? in[base + i] : 0u;
        s += v[i];
    }
    uint32_t total;
    uint32_t run = partial[blockIdx.x] + block_exclusive_scan(s, total);
#pragma unroll
    for (int i = 0; i < SC_ITEMS; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
}

// ---- keys ----
// Keys of position planes (vals = index): the engine's base-cell keys
// (internal) or the reference's setKeys keys.
// setKeys (kernels.cu:361-371): the reference's Morton keys of the positions.
__global__ void k_keys_reference(const float* __restrict__ px, const float* __restrict__ py,
                                 const float* __restrict__ pz, int n, uint32_t* __restrict__ keys,
                                 uint32_t* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = sort_key(px[i], py[i], pz[i]);
    vals[i] = (uint32_t)i;
}

// The engine's ordering key: scene id above the scene-local Morton code of the
// base cell.  The scene comes from the caller index (orig, or the slot when
// orig is null).
__global__ void k_keys_internal(const float* __restrict__ px, const float* __restrict__ py,
                                const float* __restrict__ pz, int n, SceneTable t, const int* __restrict__ orig,
                                uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int sc = t.nscenes > 1 ? scene_of(t, orig ? orig[i] : i) : 0;
    keys[i] = ((uint32_t)sc << t.local_bits) | base_key(t.cfgs[sc], px[i], py[i], pz[i]);
    vals[i] = (uint32_t)i;
}

// ---- chunking ----
__global__ void k_seg_heads(const uint32_t* __restrict__ keys, int n, uint32_t* __restrict__ heads) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    heads[i] = (i == 0 || (keys[i] >> BLOCK_SHIFT) != (keys[i - 1] >> BLOCK_SHIFT)) ? 1u : 0u;
}

__global__ void k_seg_starts(const uint32_t* __restrict__ heads, const uint32_t* __restrict__ hscan, int n,
                             const uint32_t* __restrict__ counters, int* __restrict__ seg_start) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) seg_start[counters[2]] = n;
    if (i >= n) return;
    if (heads[i]) seg_start[hscan[i]] = i;
}

__global__ void k_seg_nch(const int* __restrict__ seg_start, const uint32_t* __restrict__ counters, int n, int cap,
                          uint32_t* __restrict__ nch) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int nseg = (int)counters[2];
    if (s >= nseg) {
        nch[s] = 0u;
        return;
    }
    const int len = seg_start[s + 1] - seg_start[s];
    nch[s] = (uint32_t)((len + cap - 1) / cap);
}

__global__ void k_chunk_desc(const int* __restrict__ seg_start, const uint32_t* __restrict__ counters, int n,
                             int cap, const uint32_t* __restrict__ first, const uint32_t* __restrict__ keys,
                             int local_bits, ChunkDesc* __restrict__ chunks, int limit) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n || s >= (int)counters[2]) return;
    const int st = seg_start[s], len = seg_start[s + 1] - st;
    const uint32_t kk = keys[st];
    const int scene = (int)(local_bits >= 32 ? 0u : kk >> local_bits);
    const uint32_t k = local_bits >= 32 ? kk : (kk & ((1u << local_bits) - 1u));  // Morton code of the base cell
    const int cx = (int)(compact_bits(k) & ~3u), cy = (int)(compact_bits(k >> 1) & ~3u),
              cz = (int)(compact_bits(k >> 2) & ~3u);
    int c = (int)first[s];
    for (int off = 0; off < len; off += cap, ++c) {
        if (c >= limit) return;
        ChunkDesc d;
        d.start = st + off;
        d.count = min(cap, len - off);
        d.ox = cx - TILE_MARGIN;
        d.oy = cy - TILE_MARGIN;
        d.oz = cz - TILE_MARGIN;
        d.S = 0;  // set by the reorder
        d.scene = scene;
        d.pad1 = 0;
        chunks[c] = d;
    }
}

// One wave per chunk: finds the chunk's groups of equal base-cell keys (at
// most 64: one 4^3 block), then gathers the particles from their old slots
// (vals) into rounds -- round r takes the r-th member of every group larger
// than r -- and records the round sizes and the round count.
__global__ __launch_bounds__(256) void k_reorder(ChunkDesc* __restrict__ ch, const uint32_t* __restrict__ nchunks,
                                                 const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                 const float* __restrict__ src, size_t src_stride,
                                                 const int* __restrict__ src_orig, float* __restrict__ dst,
                                                 size_t dst_stride, int* __restrict__ dst_orig,
                                                 uint8_t* __restrict__ rounds, uint32_t* __restrict__ rounds_total,
                                                 uint32_t mask) {
    __shared__ int gstart_lds[4][65];
    const uint32_t nch = *nchunks;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int* gs = gstart_lds[wave];
    for (uint32_t ci = blockIdx.x * 4u + wave; ci < nch; ci += gridDim.x * 4u) {
        const ChunkDesc d = ch[ci];
        int ng = 0;
        for (int base = 0; base < d.count; base += 64) {
            const int q = base + lane;
            const bool head = q < d.count && (q == 0 || keys[d.start + q] != keys[d.start + q - 1]);
            const uint64_t m = __ballot(head);
            if (head) gs[ng + __popcll(m & lanemask_lt(lane))] = q;
            ng += __popcll(m);
        }
        if (lane == 0) gs[ng] = d.count;
        // The wave's LDS operations execute in order; the clobber keeps the
        // compiler from moving the reads above the writes.
        asm volatile("" ::: "memory");
        const int g0 = lane < ng ? gs[lane] : 0;
        const int gsize = lane < ng ? gs[lane + 1] - g0 : 0;
        int S = gsize;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) S = max(S, __shfl_xor(S, o));
        int off = 0;
        for (int r = 0; r < S; ++r) {
            const bool has = gsize > r;
            const uint64_t m = __ballot(has);
            if (has) {
                const uint32_t si = vals[d.start + g0 + r];
                const size_t slot = (size_t)d.start + off + __popcll(m & lanemask_lt(lane));
#pragma unroll 4
                for (int a = 0; a < A_STATE; ++a)
                    if (mask & (1u << a)) dst[(size_t)a * dst_stride + slot] = src[(size_t)a * src_stride + si];
                dst_orig[slot] = src_orig ? src_orig[si] : (int)si;
            }
            if (lane == 0) rounds[d.start + r] = (uint8_t)__popcll(m);
            off += __popcll(m);
        }
        if (lane == 0) {
            ch[ci].S = S;
            atomicAdd(rounds_total, (uint32_t)S);
        }
        asm volatile("" ::: "memory");  // gs is reused by the wave's next chunk
    }
}

}  // namespace

size_t radix_hist_elems(size_t n) { return 256 * ((n + RS_TILE - 1) / RS_TILE + 1); }
size_t scan_partial_elems(size_t n) { return (n + SC_TILE - 1) / SC_TILE + 1; }

void scan_exclusive_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* partial, uint32_t* total_dev,
                        hipStream_t s) {
    const size_t tiles = (n + SC_TILE - 1) / SC_TILE;
    if (tiles == 0) {
        if (total_dev) hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, total_dev, 0, 0u);
        return;
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)tiles), dim3(SC_THREADS), 0, s, in, n, partial);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(SC_THREADS), 0, s, partial, tiles, total_dev);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)tiles), dim3(SC_THREADS), 0, s, in, out, n, partial);
}

void launch_set_u32(uint32_t* p, int i, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, p, i, v);
}

void launch_keys_reference(const float* px, const float* py, const float* pz, int n, uint32_t* keys, uint32_t* vals,
                           hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_keys_reference, dim3((n + 255) / 256), dim3(256), 0, s, px, py, pz, n, keys, vals);
}

void launch_keys_internal(const float* px, const float* py, const float* pz, int n, const SceneTable& t,
                          const int* orig, uint32_t* keys, uint32_t* vals, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_keys_internal, dim3((n + 255) / 256), dim3(256), 0, s, px, py, pz, n, t, orig, keys, vals);
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
        hipMemcpyAsync(oa, sc.counters, sizeof(oa), hipMemcpyDeviceToHost, s);
        hipStreamSynchronize(s);
        diff = oa[0] ^ oa[1];
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

void build_chunks(SortScratch& sc, int r, int n, int chunk_cap, int local_bits, ChunkDesc* chunks, int chunk_limit,
                  hipStream_t s) {
    const unsigned blocks = (unsigned)((n + 255) / 256);
    uint32_t* heads = reinterpret_cast<uint32_t*>(sc.heads);
    uint32_t* hscan = sc.keys[1 - r];  // free after the sort
    uint32_t* nch = reinterpret_cast<uint32_t*>(sc.seg_nch);
    uint32_t* first = sc.vals[1 - r];
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, sc.counters, 4, 0u);  // round total of the new layout
    hipLaunchKernelGGL(k_seg_heads, dim3(blocks), dim3(256), 0, s, sc.keys[r], n, heads);
    scan_exclusive_u32(heads, hscan, (size_t)n, sc.partial, sc.counters + 2, s);
    hipLaunchKernelGGL(k_seg_starts, dim3(blocks), dim3(256), 0, s, heads, hscan, n, sc.counters, sc.seg_start);
    hipLaunchKernelGGL(k_seg_nch, dim3(blocks), dim3(256), 0, s, sc.seg_start, sc.counters, n, chunk_cap, nch);
    scan_exclusive_u32(nch, first, (size_t)n, sc.partial, sc.counters + 3, s);
    hipLaunchKernelGGL(k_chunk_desc, dim3(blocks), dim3(256), 0, s, sc.seg_start, sc.counters, n, chunk_cap, first,
                       sc.keys[r], local_bits, chunks, chunk_limit);
}

void launch_reorder(ChunkDesc* ch, const uint32_t* nchunks_dev, int chunk_est, const uint32_t* keys_sorted,
                    const uint32_t* vals_sorted, const float* src, size_t src_stride, const int* src_orig, float* dst,
                    size_t dst_stride, int* dst_orig, uint8_t* rounds, uint32_t* rounds_total, uint32_t attr_mask,
                    hipStream_t s) {
    int blocks = (chunk_est + 3) / 4;
    blocks = blocks < 1 ? 1 : (blocks > 65535 ? 65535 : blocks);
    hipLaunchKernelGGL(k_reorder, dim3(blocks), dim3(256), 0, s, ch, nchunks_dev, keys_sorted, vals_sorted, src,
                       src_stride, src_orig, dst, dst_stride, dst_orig, rounds, rounds_total, attr_mask);
}

}  // namespace mpm
