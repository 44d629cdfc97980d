// corrector.hip -- learned grid-velocity corrector (no reference counterpart).
//
// A small MLP (66 -> 64 -> 64 -> 64 -> 3, ReLU) maps per-node features of the
// grid after the grid update to a velocity correction, so that a substep of
// 4x the DT behaves like four plain substeps.  Weights and activations are
// BF16, products accumulate in FP32 on the matrix cores: training on
// v_mfma_f32_32x32x8_bf16, the engine's inference on gfx950's
// v_mfma_f32_32x32x16_bf16 (half the instructions; see k_pack16).  Forward,
// backward and the Adam update are this file's kernels; the CPU oracle
// (oracle/mlp.py) checks them.
//
// Matrix layout.  Every layer computes Out^T = W^T In^T with the nodes of a
// 32-node tile as the MFMA's N dimension: A = W^T (32 output units x 8 inputs),
// B = In^T (8 inputs x 32 nodes), D = Out^T (32 units x 32 nodes).  For the
// 32x32x8 instruction lane l holds A[l%32][4(l/32)+i], B[4(l/32)+i][l%32] and
// D[8(r/4) + 4(l/32) + r%4][l%32] (i = 0..3, r = 0..15), so the D registers
// of one layer, grouped by four, are exactly the B operands of the next: the
// forward pass moves no data between layers.  The weight gradients
// dW = In Out-grad^T contract over nodes and go through two small LDS
// transposes per layer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "internal.hpp"
#include "mpm.h"

namespace mpm {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int NF = CORR_FEATURES;  // 66 features
constexpr int KIN = CORR_STRIDE;   // 72: features padded to 9 K-blocks of 8
constexpr int NH = 64;
constexpr int NO = 3;

constexpr int P_W1 = 0;                 // W1[66][64] (input-major)
constexpr int P_B1 = P_W1 + NF * NH;
constexpr int P_W2 = P_B1 + NH;         // W2[64][64]
constexpr int P_B2 = P_W2 + NH * NH;
constexpr int P_W3 = P_B2 + NH;         // W3[64][64]
constexpr int P_B3 = P_W3 + NH * NH;
constexpr int P_W4 = P_B3 + NH;         // W4[64][3]
constexpr int P_B4 = P_W4 + NH * NO;
static_assert(P_B4 + NO == CORR_PARAMS, "parameter count");

// Fragments (64 lanes x 4 bf16 each).  Forward: A = W^T, blocks (out mb, in kb).
constexpr int F_L1 = 0;    // 2 x 9
constexpr int F_L2 = 18;   // 2 x 8
constexpr int F_L3 = 34;   // 2 x 8
constexpr int F_L4 = 50;   // 1 x 8 (outputs padded to 32)
// Backward: A = W, blocks (in mb, out kb).
constexpr int B_L4 = 58;   // 2 x 1 (outputs padded to 8)
constexpr int B_L3 = 60;   // 2 x 8
constexpr int B_L2 = 76;   // 2 x 8
constexpr int NFRAG = 92;
// Inference fragments for gfx950's K = 16 instruction (v_mfma_f32_32x32x16_bf16,
// 8 BF16 per lane and operand), stored after the NFRAG fragments above:
// A = W^T blocks (out mb, in kb) of 32 x 16.  Layer 1 reads the features
// padded to 80 (5 K-blocks).
constexpr int G_L1 = 0;    // 2 x 5
constexpr int G_L2 = 10;   // 2 x 4
constexpr int G_L3 = 18;   // 2 x 4
constexpr int G_L4 = 26;   // 1 x 4
constexpr int NFRAG16 = 30;
constexpr int KIN16 = 80;

// FP32 -> BF16 with round to nearest even (NaN stays NaN).
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return 0x7FC0u;  // NaN
    u += 0x7FFFu + ((u >> 16) & 1u);                       // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)f2bf(a);
    r[1] = (short)f2bf(b);
    r[2] = (short)f2bf(c);
    r[3] = (short)f2bf(d);
    return r;
}

// Two FP32 values to a BF16 pair (lo in bits 0..15) on gfx950's conversion
// instruction (v_cvt_pk_bf16_f32: round to nearest even, NaN kept): one
// instruction per pair where f2bf above takes six per value.
__device__ __forceinline__ uint32_t bf_pair(float lo, float hi) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef __bf16 b2 __attribute__((ext_vector_type(2)));
    const f2 v = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, b2));
}
__device__ __forceinline__ bf16x4 pack4_hw(float a, float b, float c, float d) {
    u32x2 r;
    r[0] = bf_pair(a, b);
    r[1] = bf_pair(c, d);
    return __builtin_bit_cast(bf16x4, r);
}
// ReLU on the bit pattern: a sign bit makes the integer negative, so the
// signed median of (bits, 0, +inf) is 0 for negative values and the value for
// non-negative ones (one v_med3_i32; fmaxf(x, 0) takes two, as it quiets NaN
// first).  A NaN becomes +inf (sign bit clear, as the hardware makes them) or
// 0 (sign bit set); either way the state it came from is already non-finite,
// which the fallback checks catch.
__device__ __forceinline__ float relu(float x) {
    return __uint_as_float((uint32_t)min(max((int)__float_as_uint(x), 0), 0x7F800000));
}

__device__ __forceinline__ f32x16 mfma(bf16x4 a, bf16x4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mfma16(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 pack8_hw(const float v[8]) {
    u32x4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = bf_pair(v[2 * i], v[2 * i + 1]);
    return __builtin_bit_cast(bf16x8, r);
}

// Weight W[in][out] of layer L (1..4) with zero padding.
__device__ __forceinline__ float w_at(const float* __restrict__ P, int L, int in, int out) {
    switch (L) {
        case 1: return (in < NF && out < NH) ? P[P_W1 + in * NH + out] : 0.0f;
        case 2: return (in < NH && out < NH) ? P[P_W2 + in * NH + out] : 0.0f;
        case 3: return (in < NH && out < NH) ? P[P_W3 + in * NH + out] : 0.0f;
        default: return (in < NH && out < NO) ? P[P_W4 + in * NO + out] : 0.0f;
    }
}

// One thread per (fragment, lane): converts the FP32 master weights into the
// BF16 fragment order the MFMA operands read.
__global__ void k_pack(const float* __restrict__ P, bf16x4* __restrict__ frags) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NFRAG * 64) return;
    const int f = t >> 6, l = t & 63, hi = l >> 5, lo = l & 31;
    int L, mb, kb;
    bool fwd = true;
    if (f < F_L2) { L = 1; mb = f / 9; kb = f % 9; }
    else if (f < F_L3) { L = 2; mb = (f - F_L2) / 8; kb = (f - F_L2) % 8; }
    else if (f < F_L4) { L = 3; mb = (f - F_L3) / 8; kb = (f - F_L3) % 8; }
    else if (f < B_L4) { L = 4; mb = 0; kb = f - F_L4; }
    else if (f < B_L3) { L = 4; mb = f - B_L4; kb = 0; fwd = false; }
    else if (f < B_L2) { L = 3; mb = (f - B_L3) / 8; kb = (f - B_L3) % 8; fwd = false; }
    else { L = 2; mb = (f - B_L2) / 8; kb = (f - B_L2) % 8; fwd = false; }
    float v[4];
    for (int i = 0; i < 4; ++i) {
        const int k = kb * 8 + 4 * hi + i, m = mb * 32 + lo;
        v[i] = fwd ? w_at(P, L, k, m) : w_at(P, L, m, k);
    }
    frags[t] = pack4(v[0], v[1], v[2], v[3]);
}

// The K = 16 fragments.  For 32x32x16, lane l holds A[l%32][8(l/32)+i] and
// B[8(l/32)+i][l%32] (i = 0..7), with the same D layout as 32x32x8.  Layer 1's
// K index is the feature.  A hidden layer's B operand is taken from the
// previous layer's D registers 8(kb%2) .. 8(kb%2)+7 of M-block kb/2 at the same
// lane, so its K index i of lane half h is the unit
// 32(kb/2) + 16(kb%2) + 8(i/4) + 4h + i%4, and the weights are packed in that
// order.
__device__ __forceinline__ int unit16(int kb, int h, int i) {
    return 32 * (kb >> 1) + 16 * (kb & 1) + 8 * (i >> 2) + 4 * h + (i & 3);
}
__global__ void k_pack16(const float* __restrict__ P, bf16x8* __restrict__ frags) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NFRAG16 * 64) return;
    const int f = t >> 6, l = t & 63, hi = l >> 5, lo = l & 31;
    int L, mb, kb;
    if (f < G_L2) { L = 1; mb = f / 5; kb = f % 5; }
    else if (f < G_L3) { L = 2; mb = (f - G_L2) / 4; kb = (f - G_L2) % 4; }
    else if (f < G_L4) { L = 3; mb = (f - G_L3) / 4; kb = (f - G_L3) % 4; }
    else { L = 4; mb = 0; kb = f - G_L4; }
    float v[8];
    for (int i = 0; i < 8; ++i) v[i] = w_at(P, L, L == 1 ? 16 * kb + 8 * hi + i : unit16(kb, hi, i), mb * 32 + lo);
    u32x4 r;
    for (int i = 0; i < 4; ++i) r[i] = (uint32_t)f2bf(v[2 * i]) | ((uint32_t)f2bf(v[2 * i + 1]) << 16);
    frags[t] = __builtin_bit_cast(bf16x8, r);
}

// Bias of the D-layout element r of M-block mb.
__device__ __forceinline__ f32x16 bias_acc(const float* __restrict__ b, int mb, int hi, int nvalid) {
    f32x16 a;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int u = mb * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
        a[r] = u < nvalid ? b[u] : 0.0f;
    }
    return a;
}

__device__ __forceinline__ void relu_pack(const f32x16 a[2], bf16x4 h[8]) {
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) {
        const int mb = kb >> 2, q = (kb & 3) * 4;
        h[kb] = pack4_hw(relu(a[mb][q]), relu(a[mb][q + 1]), relu(a[mb][q + 2]), relu(a[mb][q + 3]));
    }
}

// Layer L's (0..3) bias as the D-layout accumulator of M-block mb, read from
// the parameters in global memory ...
struct BiasParams {
    const float* __restrict__ P;
    __device__ __forceinline__ f32x16 operator()(int L, int mb, int hi) const {
        const int off = L == 0 ? P_B1 : L == 1 ? P_B2 : L == 2 ? P_B3 : P_B4;
        return bias_acc(P + off, mb, hi, L == 3 ? NO : NH);
    }
};
// ... or from a copy in that layout staged once per workgroup in LDS
// ([layer][M-block][lane half][16], zero padded): four 16-byte reads.
constexpr int BIAS_D = 4 * 2 * 2 * 16;
__device__ __forceinline__ void stage_bias(const float* __restrict__ P, float* s) {
    for (int i = threadIdx.x; i < BIAS_D; i += blockDim.x) {
        const int L = i >> 6, mb = (i >> 5) & 1, hi = (i >> 4) & 1, r = i & 15;
        const int u = mb * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
        const int off = L == 0 ? P_B1 : L == 1 ? P_B2 : L == 2 ? P_B3 : P_B4;
        s[i] = u < (L == 3 ? NO : NH) ? P[off + u] : 0.0f;
    }
}
struct BiasLds {
    const float* s;
    __device__ __forceinline__ f32x16 operator()(int L, int mb, int hi) const {
        const float4* q = reinterpret_cast<const float4*>(s + ((L * 2 + mb) * 2 + hi) * 16);
        f32x16 a;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float4 v = q[r];
            a[4 * r] = v.x; a[4 * r + 1] = v.y; a[4 * r + 2] = v.z; a[4 * r + 3] = v.w;
        }
        return a;
    }
};

// Forward pass of one 32-node tile.  xb: input B fragments (9 K-blocks).
template <class Bias>
__device__ __forceinline__ void mlp_fwd(const bf16x4* __restrict__ fr, Bias bias,
                                        const bf16x4 xb[9], int lane, bf16x4 h1[8], bf16x4 h2[8], bf16x4 h3[8],
                                        f32x16& a4) {
    const int hi = lane >> 5;
    f32x16 a[2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        a[mb] = bias(0, mb, hi);
#pragma unroll
        for (int kb = 0; kb < 9; ++kb) a[mb] = mfma(fr[(F_L1 + mb * 9 + kb) * 64 + lane], xb[kb], a[mb]);
    }
    relu_pack(a, h1);
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        a[mb] = bias(1, mb, hi);
#pragma unroll
        for (int kb = 0; kb < 8; ++kb) a[mb] = mfma(fr[(F_L2 + mb * 8 + kb) * 64 + lane], h1[kb], a[mb]);
    }
    relu_pack(a, h2);
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        a[mb] = bias(2, mb, hi);
#pragma unroll
        for (int kb = 0; kb < 8; ++kb) a[mb] = mfma(fr[(F_L3 + mb * 8 + kb) * 64 + lane], h2[kb], a[mb]);
    }
    relu_pack(a, h3);
    a4 = bias(3, 0, hi);
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) a4 = mfma(fr[(F_L4 + kb) * 64 + lane], h3[kb], a4);
}

// The same forward pass on the K = 16 instruction (inference): xb holds the
// 5 K-blocks of layer 1's B operand.
template <class Bias>
__device__ __forceinline__ void mlp_fwd16(const bf16x8* __restrict__ fr, Bias bias, const bf16x8 xb[5], int lane,
                                          f32x16& a4) {
    const int hi = lane >> 5;
    f32x16 a[2];
    bf16x8 h[4];
    const auto relu_pack16 = [&]() {
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            float v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = relu(a[kb >> 1][8 * (kb & 1) + i]);
            h[kb] = pack8_hw(v);
        }
    };
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        a[mb] = bias(0, mb, hi);
#pragma unroll
        for (int kb = 0; kb < 5; ++kb) a[mb] = mfma16(fr[(G_L1 + mb * 5 + kb) * 64 + lane], xb[kb], a[mb]);
    }
    relu_pack16();
#pragma unroll
    for (int L = 1; L < 3; ++L) {
        const int base = L == 1 ? G_L2 : G_L3;
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
            a[mb] = bias(L, mb, hi);
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) a[mb] = mfma16(fr[(base + mb * 4 + kb) * 64 + lane], h[kb], a[mb]);
        }
        relu_pack16();
    }
    a4 = bias(3, 0, hi);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) a4 = mfma16(fr[(G_L4 + kb) * 64 + lane], h[kb], a4);
}

// Input fragments of the tile's node `row` (lanes l and l+32 hold the two
// halves of each K-block): normalised features, zero beyond n.
__device__ __forceinline__ void load_x(const float* __restrict__ X, int row, bool valid, const float* __restrict__ mean,
                                       const float* __restrict__ inv_std, int hi, bf16x4 xb[9]) {
#pragma unroll
    for (int kb = 0; kb < 9; ++kb) {
        const int f = kb * 8 + 4 * hi;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (valid) v = *reinterpret_cast<const float4*>(X + (size_t)row * KIN + f);
        xb[kb] = pack4((v.x - mean[f]) * inv_std[f], (v.y - mean[f + 1]) * inv_std[f + 1],
                       (v.z - mean[f + 2]) * inv_std[f + 2], (v.w - mean[f + 3]) * inv_std[f + 3]);
    }
}

struct Norm {
    const float* mean;     // KIN
    const float* inv_std;  // KIN (0 for padding)
    float out_scale;       // correction = network output * out_scale
    float clamp;           // |network output| limit of the forward pass
    float dead;            // outputs of smaller magnitude give no correction
};

// One node's correction from the network's three outputs: zero when their
// magnitude is below the dead zone, else each limited to the clamp and scaled.
__device__ __forceinline__ void store_correction(float* __restrict__ Y, int row, float o0, float o1, float o2,
                                                 const Norm& nm) {
    const bool live = o0 * o0 + o1 * o1 + o2 * o2 >= nm.dead * nm.dead;
    Y[(size_t)row * 3 + 0] = live ? fminf(fmaxf(o0, -nm.clamp), nm.clamp) * nm.out_scale : 0.0f;
    Y[(size_t)row * 3 + 1] = live ? fminf(fmaxf(o1, -nm.clamp), nm.clamp) * nm.out_scale : 0.0f;
    Y[(size_t)row * 3 + 2] = live ? fminf(fmaxf(o2, -nm.clamp), nm.clamp) * nm.out_scale : 0.0f;
}

// A node's corrected velocity: v + the correction of store_correction, written
// to the grid only where the output is outside the dead zone (elsewhere the
// grid already holds v).
__device__ __forceinline__ void apply_correction(float* __restrict__ ux, float* __restrict__ uy,
                                                 float* __restrict__ uz, int g, float vx, float vy, float vz, float o0,
                                                 float o1, float o2, const Norm& nm) {
    if (o0 * o0 + o1 * o1 + o2 * o2 < nm.dead * nm.dead) return;
    ux[g] = vx + fminf(fmaxf(o0, -nm.clamp), nm.clamp) * nm.out_scale;
    uy[g] = vy + fminf(fmaxf(o1, -nm.clamp), nm.clamp) * nm.out_scale;
    uz[g] = vz + fminf(fmaxf(o2, -nm.clamp), nm.clamp) * nm.out_scale;
}

// Forward over rows [0, *count) (or n when count is null): Y[row][3] =
// out_scale * clamp(MLP(X[row])).  Four waves per workgroup share the forward
// weight fragments staged in LDS once; each wave loops over 32-row tiles.
constexpr int FWD_FRAGS = B_L4;  // forward fragments F_L1..F_L4
__global__ __launch_bounds__(256) void k_forward(const float* __restrict__ X, int n, const uint32_t* __restrict__ count,
                                                 Norm nm, const bf16x4* __restrict__ fr, const float* __restrict__ P,
                                                 float* __restrict__ Y) {
    __shared__ bf16x4 wfr[FWD_FRAGS * 64];
    __shared__ __align__(16) float bsh[BIAS_D];
    for (int i = threadIdx.x; i < FWD_FRAGS * 64; i += blockDim.x) wfr[i] = fr[i];
    stage_bias(P, bsh);
    __syncthreads();
    const int rows = count ? (int)min((uint32_t)n, *count) : n;
    const int lane = threadIdx.x & 63, hi = lane >> 5, wave = threadIdx.x >> 6;
    for (int tile = blockIdx.x * 4 + wave; tile * 32 < rows; tile += gridDim.x * 4) {
        const int row = tile * 32 + (lane & 31);
        const bool valid = row < rows;
        bf16x4 xb[9], h1[8], h2[8], h3[8];
        f32x16 a4;
        load_x(X, row, valid, nm.mean, nm.inv_std, hi, xb);
        // Re-read the weight fragments from LDS in every tile (each is read
        // right before its MFMA) instead of holding all 58 in registers.
        asm volatile("" ::: "memory");
        mlp_fwd(wfr, BiasLds{bsh}, xb, lane, h1, h2, h3, a4);
        if (valid && hi == 0) {
            // guard: outputs beyond the clamp (in target-RMS units, far
            // outside anything the training set holds) are limited
            store_correction(Y, row, a4[0], a4[1], a4[2], nm);
        }
    }
}

// ---- training ----

// LDS transposes of one tile: T[unit][node] holds an activation (or the
// input), U[unit][node] an output gradient, both BF16.
struct TileLds {
    uint16_t T[96 * 32];
    uint16_t U[64 * 32];
};

__device__ __forceinline__ void store_act(uint16_t* T, const bf16x4* h, int nkb, int lane) {
    const int node = lane & 31, hi = lane >> 5;
    for (int kb = 0; kb < nkb; ++kb)
#pragma unroll
        for (int i = 0; i < 4; ++i) T[(kb * 8 + 4 * hi + i) * 32 + node] = (uint16_t)h[kb][i];
}

__device__ __forceinline__ void store_grad(uint16_t* U, const f32x16* g, int nmb, int lane) {
    const int node = lane & 31, hi = lane >> 5;
    for (int mb = 0; mb < nmb; ++mb)
#pragma unroll
        for (int r = 0; r < 16; ++r) U[(mb * 32 + 8 * (r >> 2) + 4 * hi + (r & 3)) * 32 + node] = f2bf(g[mb][r]);
}

__device__ __forceinline__ bf16x4 read_frag(const uint16_t* S, int unit0, int kbn, int lane) {
    return *reinterpret_cast<const bf16x4*>(S + (unit0 + (lane & 31)) * 32 + kbn * 8 + 4 * (lane >> 5));
}

// dW[in][out] += sum over the tile's nodes of act[node][in] grad[node][out]
// for in < nin, out < nout (atomics into the FP32 gradient).
__device__ __forceinline__ void accumulate_dw(const TileLds& s, int nmi, int nno, int nin, int nout,
                                              float* __restrict__ gw, int lane) {
    const int hi = lane >> 5;
    for (int mi = 0; mi < nmi; ++mi)
        for (int no = 0; no < nno; ++no) {
            f32x16 acc = {};
#pragma unroll
            for (int kbn = 0; kbn < 4; ++kbn) acc = mfma(read_frag(s.T, mi * 32, kbn, lane), read_frag(s.U, no * 32, kbn, lane), acc);
            const int out = no * 32 + (lane & 31);
            if (out < nout) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int in = mi * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
                    if (in < nin) atomicAdd(gw + in * nout + out, acc[r]);
                }
            }
        }
}

// db[unit] += sum over nodes of g (FP32, D layout).
__device__ __forceinline__ void accumulate_db(const f32x16* g, int nmb, int nvalid, float* __restrict__ gb, int lane) {
    const int hi = lane >> 5;
    for (int mb = 0; mb < nmb; ++mb)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            float v = g[mb][r];
#pragma unroll
            for (int o = 1; o < 32; o <<= 1) v += __shfl_xor(v, o);
            const int u = mb * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
            if ((lane & 31) == 0 && u < nvalid) atomicAdd(gb + u, v);
        }
}

// grad := W * g^T through the backward fragments, masked by ReLU' of the
// layer input h (given as packed activations).
__device__ __forceinline__ void backprop(const bf16x4* __restrict__ fr, int base, int nkb, const f32x16* g,
                                         const bf16x4 h[8], int lane, f32x16 out[2]) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        out[mb] = f32x16{};
        for (int kb = 0; kb < nkb; ++kb) {
            const int gm = kb >> 2, q = (kb & 3) * 4;
            out[mb] = mfma(fr[(base + mb * nkb + kb) * 64 + lane], pack4(g[gm][q], g[gm][q + 1], g[gm][q + 2], g[gm][q + 3]),
                           out[mb]);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int kb = mb * 4 + (r >> 2);
            if (!(bf2f((uint16_t)h[kb][r & 3]) > 0.0f)) out[mb][r] = 0.0f;
        }
    }
}

// One optimisation step's gradient over rows [row0, row0 + n) of the dataset
// (targets T in correction units): grad += dL/dP of
// L = mean over rows and the 3 outputs of (MLP(x) - t / out_scale)^2; loss
// sum accumulates into *loss.  One wave per 32-row tile.
__global__ __launch_bounds__(64) void k_train(const float* __restrict__ X, const float* __restrict__ Tg, int row0,
                                              int n, int nrows, Norm nm, const bf16x4* __restrict__ fr,
                                              const float* __restrict__ P, float* __restrict__ grad,
                                              float* __restrict__ loss) {
    __shared__ TileLds s;
    const int lane = threadIdx.x, hi = lane >> 5;
    for (int i = lane; i < 24 * 32; i += 64) s.T[72 * 32 + i] = 0;  // input rows 72..95 stay zero
    const float coef = 2.0f / (3.0f * (float)n);
    const float inv_scale = 1.0f / nm.out_scale;
    for (int tile = blockIdx.x; tile * 32 < n; tile += gridDim.x) {
        const int local = tile * 32 + (lane & 31);
        const bool valid = local < n;
        const int row = (row0 + local) % nrows;
        bf16x4 xb[9], h1[8], h2[8], h3[8];
        f32x16 a4;
        load_x(X, row, valid, nm.mean, nm.inv_std, hi, xb);
        mlp_fwd(fr, BiasParams{P}, xb, lane, h1, h2, h3, a4);
        // output gradient (rows 0..2 live in lanes 0..31, registers 0..2)
        f32x16 g4[2] = {};
        float l2 = 0.0f;
        if (valid && hi == 0) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const float d = a4[j] - Tg[(size_t)row * 3 + j] * inv_scale;
                g4[0][j] = coef * d;
                l2 += d * d;
            }
        }
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) l2 += __shfl_xor(l2, o);
        if (lane == 0) atomicAdd(loss, l2);
        // layer 4
        __syncthreads();
        store_act(s.T, h3, 8, lane);
        store_grad(s.U, g4, 1, lane);
        __syncthreads();
        accumulate_dw(s, 2, 1, NH, NO, grad + P_W4, lane);
        accumulate_db(g4, 1, NO, grad + P_B4, lane);
        f32x16 g3[2];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
            g3[mb] = mfma(fr[(B_L4 + mb) * 64 + lane], pack4(g4[0][0], g4[0][1], g4[0][2], g4[0][3]), f32x16{});
#pragma unroll
            for (int r = 0; r < 16; ++r)
                if (!(bf2f((uint16_t)h3[mb * 4 + (r >> 2)][r & 3]) > 0.0f)) g3[mb][r] = 0.0f;
        }
        // layer 3
        __syncthreads();
        store_act(s.T, h2, 8, lane);
        store_grad(s.U, g3, 2, lane);
        __syncthreads();
        accumulate_dw(s, 2, 2, NH, NH, grad + P_W3, lane);
        accumulate_db(g3, 2, NH, grad + P_B3, lane);
        f32x16 g2[2];
        backprop(fr, B_L3, 8, g3, h2, lane, g2);
        // layer 2
        __syncthreads();
        store_act(s.T, h1, 8, lane);
        store_grad(s.U, g2, 2, lane);
        __syncthreads();
        accumulate_dw(s, 2, 2, NH, NH, grad + P_W2, lane);
        accumulate_db(g2, 2, NH, grad + P_B2, lane);
        f32x16 g1[2];
        backprop(fr, B_L2, 8, g2, h1, lane, g1);
        // layer 1
        __syncthreads();
        store_act(s.T, xb, 9, lane);
        store_grad(s.U, g1, 2, lane);
        __syncthreads();
        accumulate_dw(s, 3, 2, NF, NH, grad + P_W1, lane);
        accumulate_db(g1, 2, NH, grad + P_B1, lane);
    }
}

__global__ void k_adam(float* __restrict__ P, float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    g[i] = 0.0f;
}

__global__ void k_zero_f32(float* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// ---- grid side: active nodes, features, corrections ----

// Appends every node with mass > 1e-9 of the flagged 4^3 blocks (all scenes)
// to `list` (x | y << 10 | z << 20 and the scene).  One wave per block.
__global__ __launch_bounds__(256) void k_active(const DevCfg* __restrict__ cfgs, const uint16_t* __restrict__ block_scene,
                                                GridBuf V, int nblocks, int2* __restrict__ list, int cap,
                                                uint32_t* __restrict__ count) {
    const int blk = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (blk >= nblocks) return;
    if (V.flags[blk] == 0) return;
    const int lane = threadIdx.x & 63;
    const int sc = block_scene[blk];
    const DevCfg c = cfgs[sc];
    const int lb = blk - c.foff;
    const int bx = lb % c.fbx, by = (lb / c.fbx) % c.fby, bz = lb / (c.fbx * c.fby);
    const int x = bx * 4 + (lane & 3), y = by * 4 + ((lane >> 2) & 3), z = bz * 4 + (lane >> 4);
    const bool act = inb(c, x, y, z) && V.plane[0][gidx(c, x, y, z)] > 1e-9f;
    const uint64_t m = __ballot(act);
    if (m == 0) return;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, 0);
    const uint32_t slot = base + (uint32_t)__popcll(m & (lane ? (~0ull >> (64 - lane)) : 0ull));
    if (act && slot < (uint32_t)cap) list[slot] = make_int2(x | (y << 10) | (z << 20), sc);
}

// Per-node features (66 used of a 72-float row), from the grid after the
// grid update (mass plane, velocity planes) and the SDF:
//   0      mass / 8
//   1..3   velocity * DT (cells per substep)
//   4      SDF clamped to [-4, 4] / 4
//   5..7   SDF normal (getNormalAtNode)
//   8..34  masses / 8 of the 3^3 neighbourhood (x fastest; 0 outside the grid)
//   35..61 momentum magnitudes |m v| * DT / 8 of the neighbourhood
//   62     GRAVITY * DT^2,  63  COMPRESSION * DT^2
//   64     distance to the nearest padded wall / 8, clamped to [0, 1]
//   65     y / SIZE_Y
// With gv, the node's grid index and velocity are returned as well.
struct NodeGV {
    int g;
    float v[3];
};
// Features 0..7: the node's own mass, velocity, SDF and normal.
__device__ __forceinline__ void features_own(const DevCfg& c, const GridBuf& V, const float* __restrict__ sdf, int x,
                                             int y, int z, float* __restrict__ f, NodeGV* gv) {
    const int g = gidx(c, x, y, z);
    const float dt = c.dt;
    const float v0 = V.plane[1][g], v1 = V.plane[2][g], v2 = V.plane[3][g];
    if (gv) {
        gv->g = g;
        gv->v[0] = v0; gv->v[1] = v1; gv->v[2] = v2;
    }
    f[0] = V.plane[0][g] * 0.125f;
    f[1] = v0 * dt;
    f[2] = v1 * dt;
    f[3] = v2 * dt;
    f[4] = fminf(fmaxf(sdf[g], -4.0f), 4.0f) * 0.25f;
    // getNormalAtNode's central differences; the hardware square root and
    // reciprocal (about 1 ulp) replace the correctly rounded ones here, as
    // for the momentum magnitudes below: these are network inputs rounded
    // to BF16, where the IEEE sequences' dozen instructions buy nothing.
    {
        const float nx = sdf[gidx(c, min(x + 1, c.sx - 1), y, z)] - sdf[gidx(c, max(x - 1, 0), y, z)];
        const float ny = sdf[gidx(c, x, min(y + 1, c.sy - 1), z)] - sdf[gidx(c, x, max(y - 1, 0), z)];
        const float nz = sdf[gidx(c, x, y, min(z + 1, c.sz - 1))] - sdf[gidx(c, x, y, max(z - 1, 0))];
        const float r = __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(nx * nx + ny * ny + nz * nz) + 1e-9f);
        f[5] = nx * r; f[6] = ny * r; f[7] = nz * r;
    }
}
// A neighbour's mass and momentum magnitude * DT from its loaded values: the
// mass as stored (0 outside the grid), the momentum 0 where the mass is not
// positive (a massless node's velocity may be stale in the corrector's copy).
__device__ __forceinline__ float2 nbr_value(bool in, float m, float vx, float vy, float vz, float dt) {
    const float mi = in ? m : 0.0f;
    return make_float2(mi, mi > 0.0f ? mi * __builtin_amdgcn_sqrtf(vx * vx + vy * vy + vz * vz) * dt : 0.0f);
}
// Features 62..65 and the padding.
__device__ __forceinline__ void features_tail(const DevCfg& c, int x, int y, int z, float* __restrict__ f) {
    const float dt = c.dt;
    f[62] = c.gravity * dt * dt;
    f[63] = c.compression * dt * dt;
    const float p = (float)c.pad;
    float w = fminf(fminf((float)x - p, (float)(c.sx - 1) - p - (float)x),
                    fminf(fminf((float)y - p, (float)(c.sy - 1) - p - (float)y),
                          fminf((float)z - p, (float)(c.sz - 1) - p - (float)z)));
    f[64] = fminf(fmaxf(w * 0.125f, 0.0f), 1.0f);
    f[65] = (float)y / (float)c.sy;
#pragma unroll
    for (int i = NF; i < KIN; ++i) f[i] = 0.0f;
}

__device__ __forceinline__ void node_features(const DevCfg& c, const GridBuf& V, const float* __restrict__ sdf, int x, int y, int z,
                              float* __restrict__ f, NodeGV* gv = nullptr) {
    features_own(c, V, sdf, x, y, z, f, gv);
    // The 27 neighbours one z slice at a time: a slice's 36 loads are issued
    // together at coordinates clamped into the grid (one memory latency per
    // slice rather than per neighbour), and neighbours outside the grid are
    // then set to zero.
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) {
        const int zz = z + sl - 1, zc = min(max(zz, 0), c.sz - 1);
        float mm[9], vx[9], vy[9], vz[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const int xx = x + q % 3 - 1, yy = y + q / 3 - 1;
            const int gq = gidx(c, min(max(xx, 0), c.sx - 1), min(max(yy, 0), c.sy - 1), zc);
            mm[q] = V.plane[0][gq];
            vx[q] = V.plane[1][gq];
            vy[q] = V.plane[2][gq];
            vz[q] = V.plane[3][gq];
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const int o = 9 * sl + q;  // x fastest: o % 3 - 1 = dx, (o / 3) % 3 - 1 = dy, o / 9 - 1 = dz
            const float2 nv = nbr_value(inb(c, x + q % 3 - 1, y + q / 3 - 1, zz), mm[q], vx[q], vy[q], vz[q], c.dt);
            f[8 + o] = nv.x * 0.125f;
            f[35 + o] = nv.y * 0.125f;
        }
    }
    features_tail(c, x, y, z, f);
}

extern "C" __device__ int __ockl_wfred_min_i32(int);
extern "C" __device__ int __ockl_wfred_max_i32(int);

// Neighbourhood staging of k_correct: a wave's 64 nodes come from one or two
// z slices of a row segment (16 x 4 nodes), so their 27-neighbourhoods share
// most nodes.  When the wave's nodes lie in one scene and the box around them
// (one node of margin) holds at most STAGE_N nodes, the wave loads the box
// once, coalesced, as (mass, momentum magnitude * DT) pairs into LDS and each
// lane reads its 27 neighbours there: 8 node loads per lane instead of 108.
constexpr int STAGE_N = 512;

__global__ __launch_bounds__(256) void k_features(const DevCfg* __restrict__ cfgs, GridBuf V,
                                                  const float* __restrict__ sdf, const int2* __restrict__ list,
                                                  const uint32_t* __restrict__ count, int cap, float* __restrict__ X) {
    const int n = (int)min(*count, (uint32_t)cap);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int2 e = list[i];
        const DevCfg c = cfgs[e.y];
        node_features(c, V, sdf, e.x & 1023, (e.x >> 10) & 1023, e.x >> 20, X + (size_t)i * KIN);
    }
}

// Fused features + forward over the active-node list, on the K = 16 MFMA.
// Each wave takes 64 nodes at a time, one per lane: the lane computes its
// node's features and packs them, normalised, into BF16 K-block halves
// (16kb..16kb+7 and 16kb+8..16kb+15).  The first layer's B operand for node r
// of a 32-node tile wants the first half at lane r and the second at lane
// r + 32, so one
// v_permlane32_swap per register exchanges the upper lanes of the first
// halves with the lower lanes of the second halves, which leaves tile 0's
// operand (nodes 0..31) in the one and tile 1's (nodes 32..63) in the other,
// with no LDS round trip.  Four waves per workgroup share the forward weight
// fragments, the biases and the normalisation in LDS.  The features read the
// grid's masses and a copy of its velocities (R: planes 1..3 written by the
// grid update), so the corrected velocities can go straight into the grid's
// velocity planes ux, uy, uz.  They are distinct from every array read here
// and declared __restrict__: stored through the grid's plain plane pointers,
// the compiler allowed for aliasing and spilled about a hundred registers.
__global__ __launch_bounds__(256, 3) void k_correct(const DevCfg* __restrict__ cfgs, GridBuf R,
                                                 float* __restrict__ ux, float* __restrict__ uy,
                                                 float* __restrict__ uz,
                                                 const float* __restrict__ sdf, const int2* __restrict__ list,
                                                 const uint32_t* __restrict__ count, int cap, Norm nm,
                                                 const bf16x8* __restrict__ fr, const float* __restrict__ P) {
    __shared__ bf16x8 wfr[NFRAG16 * 64];
    __shared__ float2 stg[4][STAGE_N];  // per wave: the staged neighbourhood
    __shared__ __align__(16) float bsh[BIAS_D];
    __shared__ __align__(16) float2 nsh[KIN16];  // (1 / std, -mean / std) per feature, 0 past KIN
    for (int i = threadIdx.x; i < NFRAG16 * 64; i += blockDim.x) wfr[i] = fr[i];
    stage_bias(P, bsh);
    for (int i = threadIdx.x; i < KIN16; i += blockDim.x)
        nsh[i] = i < KIN ? make_float2(nm.inv_std[i], -nm.mean[i] * nm.inv_std[i]) : make_float2(0.f, 0.f);
    __syncthreads();
    const int rows = (int)min(*count, (uint32_t)cap);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int base = (blockIdx.x * 4 + wave) * 64; base < rows; base += gridDim.x * 4 * 64) {
        // LDS reads stay in the loop (hoisted, the 144 normalisation terms
        // would occupy registers for the whole kernel)
        asm volatile("" ::: "memory");
        bf16x8 x0[5], x1[5];
        NodeGV gv;  // the lane's node (grid index) and velocity
        {
            // lanes past the end repeat the last node (a node's outputs
            // depend on its own inputs only, and theirs are not stored)
            const int2 e = list[min(base + lane, rows - 1)];
            const int x = e.x & 1023, y = (e.x >> 10) & 1023, z = e.x >> 20;
            float f[KIN16];
            // the wave's box: [x0 - 1, x1 + 1] x ... (one scene only)
            const int sc0 = __builtin_amdgcn_readfirstlane(e.y);
            const int bx0 = __ockl_wfred_min_i32(x) - 1, by0 = __ockl_wfred_min_i32(y) - 1,
                      bz0 = __ockl_wfred_min_i32(z) - 1;
            const int dx = __ockl_wfred_max_i32(x) + 2 - bx0, dy = __ockl_wfred_max_i32(y) + 2 - by0,
                      dz = __ockl_wfred_max_i32(z) + 2 - bz0;
            if (__ballot(e.y != sc0) == 0ull && dx * dy * dz <= STAGE_N) {
                const DevCfg c = uniform(cfgs[sc0]);
                float2* st = stg[wave];
                const int vol = dx * dy * dz;
                // all of a lane's box loads first (clamped coordinates), then
                // the values
                float mm[STAGE_N / 64], vx[STAGE_N / 64], vy[STAGE_N / 64], vz[STAGE_N / 64];
                bool in[STAGE_N / 64];
                // box index i -> (i mod dx, i / dx mod dy, i / (dx dy)) with
                // float reciprocals: i + 0.5 lies at least 0.5 / 512 from a
                // multiple of the divisor and the product errs by < 1e-4
                const float rdx = 1.0f / (float)dx, rdxy = 1.0f / (float)(dx * dy);
                // rounds wholly past the box load nothing (a uniform bound)
                const int rounds = (vol + 63) >> 6;
#pragma unroll
                for (int k = 0; k < STAGE_N / 64; ++k) {
                    if (k >= rounds) break;
                    const int i = min(lane + 64 * k, vol - 1);
                    const int iz = (int)(((float)i + 0.5f) * rdxy), r = i - iz * dx * dy;
                    const int iy = (int)(((float)r + 0.5f) * rdx);
                    const int gx = bx0 + r - iy * dx, gy = by0 + iy, gz = bz0 + iz;
                    in[k] = inb(c, gx, gy, gz);
                    const int g = gidx(c, min(max(gx, 0), c.sx - 1), min(max(gy, 0), c.sy - 1),
                                       min(max(gz, 0), c.sz - 1));
                    mm[k] = R.plane[0][g];
                    vx[k] = R.plane[1][g];
                    vy[k] = R.plane[2][g];
                    vz[k] = R.plane[3][g];
                }
#pragma unroll
                for (int k = 0; k < STAGE_N / 64; ++k)
                    if (k < rounds && lane + 64 * k < vol) st[lane + 64 * k] = nbr_value(in[k], mm[k], vx[k], vy[k], vz[k], c.dt);
                asm volatile("" ::: "memory");  // the wave's LDS writes are read below, in order
                features_own(c, R, sdf, x, y, z, f, &gv);
                const int lo = (x - bx0) + dx * ((y - by0) + dy * (z - bz0));
#pragma unroll
                for (int o = 0; o < 27; ++o) {
                    const float2 nv = st[lo + (o % 3 - 1) + dx * ((o / 3) % 3 - 1) + dx * dy * (o / 9 - 1)];
                    f[8 + o] = nv.x * 0.125f;
                    f[35 + o] = nv.y * 0.125f;
                }
                features_tail(c, x, y, z, f);
            } else {
                node_features(cfgs[e.y], R, sdf, x, y, z, f, &gv);
            }
#pragma unroll
            for (int i = KIN; i < KIN16; ++i) f[i] = 0.0f;
#pragma unroll
            for (int kb = 0; kb < 5; ++kb) {
                float v[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const float2 q = nsh[16 * kb + i];
                    v[i] = fmaf(f[16 * kb + i], q.x, q.y);
                }
                u32x4 lo = __builtin_bit_cast(u32x4, pack8_hw(v));
                u32x4 hi = __builtin_bit_cast(u32x4, pack8_hw(v + 8));
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    const auto sw = __builtin_amdgcn_permlane32_swap(lo[d], hi[d], false, false);
                    lo[d] = sw[0];
                    hi[d] = sw[1];
                }
                x0[kb] = __builtin_bit_cast(bf16x8, lo);
                x1[kb] = __builtin_bit_cast(bf16x8, hi);
            }
        }
        // node r of a tile's outputs sits in lane r: tile 0's nodes are lanes
        // 0..31's own, tile 1's come from lanes 32..63
        const auto tile = [&](const bf16x8 xb[5], int row, int g, float vx, float vy, float vz) {
            f32x16 a4;
            mlp_fwd16(wfr, BiasLds{bsh}, xb, lane, a4);
            if (row < rows && lane < 32) apply_correction(ux, uy, uz, g, vx, vy, vz, a4[0], a4[1], a4[2], nm);
        };
        tile(x0, base + lane, gv.g, gv.v[0], gv.v[1], gv.v[2]);
        if (base + 32 < rows) {
            const auto upper = [](float v) {
                return __uint_as_float(__builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false,
                                                                         false)[1]);
            };
            const int g1 = (int)__builtin_amdgcn_permlane32_swap((uint32_t)gv.g, (uint32_t)gv.g, false, false)[1];
            tile(x1, base + 32 + lane, g1, upper(gv.v[0]), upper(gv.v[1]), upper(gv.v[2]));
        }
    }
}

// Training pairs: the velocities of the listed nodes (after a grid update),
// copied (mode 0), accumulated (mode 1) or copied (mode 2), modes 1 and 2
// with a count of the calls in which the node had mass.
__global__ void k_gather_vel(const DevCfg* __restrict__ cfgs, GridBuf V, const int2* __restrict__ list,
                             const uint32_t* __restrict__ count, int cap, float* __restrict__ out,
                             float* __restrict__ hits, int mode) {
    const int n = (int)min(*count, (uint32_t)cap);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int2 e = list[i];
        const DevCfg& c = cfgs[e.y];
        const int g = gidx(c, e.x & 1023, (e.x >> 10) & 1023, e.x >> 20);
        const bool live = V.plane[0][g] > 1e-9f;
        const float vx = live ? V.plane[1][g] : 0.0f, vy = live ? V.plane[2][g] : 0.0f, vz = live ? V.plane[3][g] : 0.0f;
        if (mode == 1) {
            out[(size_t)i * 3] += vx; out[(size_t)i * 3 + 1] += vy; out[(size_t)i * 3 + 2] += vz;
        } else {
            out[(size_t)i * 3] = vx; out[(size_t)i * 3 + 1] = vy; out[(size_t)i * 3 + 2] = vz;
        }
        if (mode != 0) hits[i] += live ? 1.0f : 0.0f;
    }
}

// Fallback checks over the particles: [0] particles whose position samples
// SDF < -depth (collider penetration), [1] sum of |det F - 1|, [2] particles
// with a non-finite position or velocity.
__global__ __launch_bounds__(256) void k_check(const float* __restrict__ part, size_t stride, const int* __restrict__ orig,
                                               int n, SceneTable t, const float* __restrict__ sdf, float depth,
                                               float* __restrict__ out) {
    float pen = 0.0f, drift = 0.0f, bad = 0.0f;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const DevCfg& c = t.cfgs[scene_of(t, orig[s])];
        const float x = part[A_PX * stride + s], y = part[A_PY * stride + s], z = part[A_PZ * stride + s];
        const float vx = part[A_VX * stride + s], vy = part[A_VY * stride + s], vz = part[A_VZ * stride + s];
        if (!(isfinite(x) && isfinite(y) && isfinite(z) && isfinite(vx) && isfinite(vy) && isfinite(vz))) {
            bad += 1.0f;
            continue;
        }
        if (sdf_at(c, sdf, x, y, z) < -depth) pen += 1.0f;
        float F[9];
        for (int i = 0; i < 9; ++i) F[i] = part[(A_F0 + i) * stride + s];
        drift += fabsf(det3(F) - 1.0f);
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        pen += __shfl_xor(pen, o);
        drift += __shfl_xor(drift, o);
        bad += __shfl_xor(bad, o);
    }
    // one partial per workgroup (no shared atomics); k_sum3 adds them up
    __shared__ float part3[4][3];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part3[wave][0] = pen;
        part3[wave][1] = drift;
        part3[wave][2] = bad;
    }
    __syncthreads();
    if (threadIdx.x < 3)
        out[3 + 3 * blockIdx.x + threadIdx.x] =
            part3[0][threadIdx.x] + part3[1][threadIdx.x] + part3[2][threadIdx.x] + part3[3][threadIdx.x];
}

// out[k] = sum over the np partials out[3 + 3 i + k], k < 3 (one workgroup).
__global__ __launch_bounds__(256) void k_sum3(float* __restrict__ out, int np) {
    __shared__ float acc[3][256];
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    for (int i = threadIdx.x; i < np; i += 256) {
        a0 += out[3 + 3 * i];
        a1 += out[4 + 3 * i];
        a2 += out[5 + 3 * i];
    }
    acc[0][threadIdx.x] = a0;
    acc[1][threadIdx.x] = a1;
    acc[2][threadIdx.x] = a2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < 3; ++k) acc[k][threadIdx.x] += acc[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) out[threadIdx.x] = acc[threadIdx.x][0];
}

int grid_blocks(int n, int per, int cap) { return std::max(1, std::min((n + per - 1) / per, cap)); }

}  // namespace

void corr_pack(const float* P, void* frags, hipStream_t s) {
    hipLaunchKernelGGL(k_pack, dim3((NFRAG * 64 + 255) / 256), dim3(256), 0, s, P, (bf16x4*)frags);
    hipLaunchKernelGGL(k_pack16, dim3((NFRAG16 * 64 + 255) / 256), dim3(256), 0, s, P,
                       reinterpret_cast<bf16x8*>(static_cast<char*>(frags) + (size_t)NFRAG * 64 * 8));
}

void launch_corr_active(const DevCfg* cfgs, const uint16_t* block_scene, GridBuf V, int nblocks, int2* list, int cap,
                        uint32_t* count, hipStream_t s) {
    if (nblocks <= 0) return;
    hipLaunchKernelGGL(k_active, dim3((nblocks + 3) / 4), dim3(256), 0, s, cfgs, block_scene, V, nblocks, list, cap,
                       count);
}

void launch_corr_features(const DevCfg* cfgs, GridBuf V, const float* sdf, const int2* list, const uint32_t* count,
                          int cap, float* X, hipStream_t s) {
    hipLaunchKernelGGL(k_features, dim3(grid_blocks(cap, 256, 4096)), dim3(256), 0, s, cfgs, V, sdf, list, count, cap, X);
}

void launch_corr_forward(const CorrState& cs, const float* X, int cap, const uint32_t* count, float* Y, hipStream_t s) {
    Norm nm{cs.mean, cs.inv_std, cs.out_scale, cs.out_clamp, cs.out_dead};
    hipLaunchKernelGGL(k_forward, dim3(grid_blocks(cap, 128, 1024)), dim3(256), 0, s, X, cap, count, nm,
                       (const bf16x4*)cs.frags, cs.params, Y);
}

void launch_corr_correct(const DevCfg* cfgs, GridBuf V, GridBuf vcopy, const float* sdf, const int2* list,
                         const uint32_t* count, int cap, const CorrState& cs, hipStream_t s) {
    Norm nm{cs.mean, cs.inv_std, cs.out_scale, cs.out_clamp, cs.out_dead};
    GridBuf R = vcopy;
    R.plane[0] = V.plane[0];
    const bf16x8* fr16 = reinterpret_cast<const bf16x8*>(static_cast<const char*>(cs.frags) + (size_t)NFRAG * 64 * 8);
    hipLaunchKernelGGL(k_correct, dim3(grid_blocks(cap, 256, 768)), dim3(256), 0, s, cfgs, R, V.plane[1], V.plane[2],
                       V.plane[3], sdf, list, count, cap, nm, fr16, cs.params);
}

void launch_corr_gather_vel(const DevCfg* cfgs, GridBuf V, const int2* list, const uint32_t* count, int cap, float* out,
                            float* hits, int mode, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_vel, dim3(grid_blocks(cap, 256, 4096)), dim3(256), 0, s, cfgs, V, list, count, cap,
                       out, hits, mode);
}

void launch_check(const float* part, size_t stride, const int* orig, int n, const SceneTable& t, const float* sdf,
                  float depth, float* out3, hipStream_t s) {
    if (n <= 0) return;
    const int blocks = grid_blocks(n, 256, CHECK_BLOCKS);
    hipLaunchKernelGGL(k_check, dim3(blocks), dim3(256), 0, s, part, stride, orig, n, t, sdf, depth, out3);
    hipLaunchKernelGGL(k_sum3, dim3(1), dim3(256), 0, s, out3, blocks);
}

}  // namespace mpm

// ---------------- corrector object and its C ABI ----------------

using namespace mpm;

namespace {

int cfail(int code, const std::string& msg) {
    mpm::set_last_error(msg);
    return code;
}

#define CHK(expr)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return cfail(MPM_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T> int alloc(T** p, size_t n) {
    *p = nullptr;
    if (hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return cfail(MPM_ENOMEM, "hipMalloc failed");
    return MPM_OK;
}

}  // namespace

struct mpm_corrector {
    CorrState st{};
    float* grad = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    int64_t adam_t = 0;
    float* loss = nullptr;  // per training step
    int loss_cap = 0;
    float* dX = nullptr;  // dataset
    float* dT = nullptr;
    int nrows = 0;
    hipStream_t stream = nullptr;
};

namespace mpm {
const CorrState* corrector_state(const mpm_corrector* c) { return c ? &c->st : nullptr; }
}  // namespace mpm

extern "C" {

int mpm_corrector_create(mpm_corrector** out) {
    if (!out) return cfail(MPM_EINVAL, "null argument");
    *out = nullptr;
    mpm_corrector* c = new mpm_corrector();
    int rc;
    if ((rc = alloc(&c->st.params, CORR_PARAMS)) || (rc = alloc((uint16_t**)&c->st.frags, (size_t)NFRAG * 256 + (size_t)NFRAG16 * 512)) ||
        (rc = alloc(&c->st.mean, KIN)) || (rc = alloc(&c->st.inv_std, KIN)) || (rc = alloc(&c->grad, CORR_PARAMS)) ||
        (rc = alloc(&c->adam_m, CORR_PARAMS)) || (rc = alloc(&c->adam_v, CORR_PARAMS))) {
        mpm_corrector_destroy(c);
        return rc;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        mpm_corrector_destroy(c);
        return cfail(MPM_EHIP, "hipStreamCreate failed");
    }
    std::vector<float> zero(CORR_PARAMS, 0.0f), mean(KIN, 0.0f), inv(KIN, 0.0f);
    for (int i = 0; i < NF; ++i) inv[i] = 1.0f;
    c->st.out_scale = 1.0f;
    c->st.out_clamp = 1e30f;
    c->st.out_dead = 0.0f;
    hipError_t e = hipMemcpy(c->st.params, zero.data(), sizeof(float) * CORR_PARAMS, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->grad, zero.data(), sizeof(float) * CORR_PARAMS, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->adam_m, zero.data(), sizeof(float) * CORR_PARAMS, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->adam_v, zero.data(), sizeof(float) * CORR_PARAMS, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->st.mean, mean.data(), sizeof(float) * KIN, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->st.inv_std, inv.data(), sizeof(float) * KIN, hipMemcpyHostToDevice);
    corr_pack(c->st.params, c->st.frags, c->stream);
    if (e != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
        mpm_corrector_destroy(c);
        return cfail(MPM_EHIP, "corrector init failed");
    }
    *out = c;
    return MPM_OK;
}

int mpm_corrector_destroy(mpm_corrector* c) {
    if (!c) return MPM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* q : {(void*)c->st.params, c->st.frags, (void*)c->st.mean, (void*)c->st.inv_std, (void*)c->grad,
                    (void*)c->adam_m, (void*)c->adam_v, (void*)c->loss, (void*)c->dX, (void*)c->dT})
        if (q) (void)hipFree(q);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MPM_OK;
}

int mpm_corrector_set_params(mpm_corrector* c, const float* params, int n) {
    if (!c || !params || n != CORR_PARAMS) return cfail(MPM_EINVAL, "parameter count must be MPM_CORRECTOR_PARAMS");
    CHK(hipMemcpyAsync(c->st.params, params, sizeof(float) * CORR_PARAMS, hipMemcpyHostToDevice, c->stream));
    corr_pack(c->st.params, c->st.frags, c->stream);
    std::vector<float> zero(CORR_PARAMS, 0.0f);
    CHK(hipMemcpyAsync(c->adam_m, zero.data(), sizeof(float) * CORR_PARAMS, hipMemcpyHostToDevice, c->stream));
    CHK(hipMemcpyAsync(c->adam_v, zero.data(), sizeof(float) * CORR_PARAMS, hipMemcpyHostToDevice, c->stream));
    c->adam_t = 0;
    CHK(hipStreamSynchronize(c->stream));
    return MPM_OK;
}

int mpm_corrector_get_params(mpm_corrector* c, float* params, int n) {
    if (!c || !params || n != CORR_PARAMS) return cfail(MPM_EINVAL, "parameter count must be MPM_CORRECTOR_PARAMS");
    CHK(hipMemcpyAsync(params, c->st.params, sizeof(float) * CORR_PARAMS, hipMemcpyDeviceToHost, c->stream));
    CHK(hipStreamSynchronize(c->stream));
    return MPM_OK;
}

int mpm_corrector_set_deadzone(mpm_corrector* c, float dead) {
    if (!c || !(dead >= 0.0f)) return cfail(MPM_EINVAL, "dead zone must be >= 0");
    c->st.out_dead = dead;
    return MPM_OK;
}

int mpm_corrector_set_clamp(mpm_corrector* c, float clamp) {
    if (!c || !(clamp > 0.0f)) return cfail(MPM_EINVAL, "clamp must be positive");
    c->st.out_clamp = clamp;
    return MPM_OK;
}

int mpm_corrector_set_norm(mpm_corrector* c, const float* mean, const float* inv_std, float out_scale) {
    if (!c || !mean || !inv_std || !(out_scale > 0.0f)) return cfail(MPM_EINVAL, "bad normalisation");
    std::vector<float> m(KIN, 0.0f), s(KIN, 0.0f);
    for (int i = 0; i < NF; ++i) {
        m[i] = mean[i];
        s[i] = inv_std[i];
    }
    CHK(hipMemcpyAsync(c->st.mean, m.data(), sizeof(float) * KIN, hipMemcpyHostToDevice, c->stream));
    CHK(hipMemcpyAsync(c->st.inv_std, s.data(), sizeof(float) * KIN, hipMemcpyHostToDevice, c->stream));
    CHK(hipStreamSynchronize(c->stream));
    c->st.out_scale = out_scale;
    return MPM_OK;
}

int mpm_corrector_forward(mpm_corrector* c, const float* X, int n, float* Y) {
    if (!c || n < 0 || (n > 0 && (!X || !Y))) return cfail(MPM_EINVAL, "bad arguments");
    if (n == 0) return MPM_OK;
    float *dx = nullptr, *dy = nullptr;
    int rc;
    if ((rc = alloc(&dx, (size_t)n * KIN)) || (rc = alloc(&dy, (size_t)n * 3))) {
        (void)hipFree(dx);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(dx, X, sizeof(float) * KIN * (size_t)n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) launch_corr_forward(c->st, dx, n, nullptr, dy, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(Y, dy, sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dx);
    (void)hipFree(dy);
    if (e != hipSuccess) return cfail(MPM_EHIP, std::string("corrector forward: ") + hipGetErrorString(e));
    return MPM_OK;
}

int mpm_corrector_load_dataset(mpm_corrector* c, const float* X, const float* T, int n) {
    if (!c || n < 1 || !X || !T) return cfail(MPM_EINVAL, "bad dataset");
    if (c->dX) (void)hipFree(c->dX);
    if (c->dT) (void)hipFree(c->dT);
    c->dX = c->dT = nullptr;
    c->nrows = 0;
    int rc;
    if ((rc = alloc(&c->dX, (size_t)n * KIN)) || (rc = alloc(&c->dT, (size_t)n * 3))) return rc;
    CHK(hipMemcpyAsync(c->dX, X, sizeof(float) * KIN * (size_t)n, hipMemcpyHostToDevice, c->stream));
    CHK(hipMemcpyAsync(c->dT, T, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    CHK(hipStreamSynchronize(c->stream));
    c->nrows = n;
    return MPM_OK;
}

// Gradient of the loss over rows [row0, row0 + n) of the loaded dataset
// (wrapping), without an update: grad_out (CORR_PARAMS) and the loss.
int mpm_corrector_gradient(mpm_corrector* c, int row0, int n, float* grad_out, float* loss_out) {
    if (!c || n < 1 || !grad_out || !loss_out) return cfail(MPM_EINVAL, "bad arguments");
    if (!c->nrows) return cfail(MPM_ESTATE, "no dataset loaded");
    if (!c->loss) {
        int rc;
        if ((rc = alloc(&c->loss, 1))) return rc;
        c->loss_cap = 1;
    }
    Norm nm{c->st.mean, c->st.inv_std, c->st.out_scale, c->st.out_clamp, c->st.out_dead};
    hipLaunchKernelGGL(k_zero_f32, dim3((CORR_PARAMS + 255) / 256), dim3(256), 0, c->stream, c->grad, CORR_PARAMS);
    hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, c->stream, c->loss, 1);
    hipLaunchKernelGGL(k_train, dim3(grid_blocks(n, 32, 1024)), dim3(64), 0, c->stream, c->dX, c->dT, row0, n,
                       c->nrows, nm, (const bf16x4*)c->st.frags, c->st.params, c->grad, c->loss);
    CHK(hipMemcpyAsync(grad_out, c->grad, sizeof(float) * CORR_PARAMS, hipMemcpyDeviceToHost, c->stream));
    float l = 0.0f;
    CHK(hipMemcpyAsync(&l, c->loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    CHK(hipStreamSynchronize(c->stream));
    hipLaunchKernelGGL(k_zero_f32, dim3((CORR_PARAMS + 255) / 256), dim3(256), 0, c->stream, c->grad, CORR_PARAMS);
    CHK(hipStreamSynchronize(c->stream));
    *loss_out = l / (3.0f * (float)n);
    return MPM_OK;
}

// `steps` Adam steps over consecutive batches of the loaded dataset (row
// offset advancing by `batch` each step, wrapping); loss_out[s] = the mean
// loss of step s's batch before its update.
int mpm_corrector_train(mpm_corrector* c, int steps, int batch, float lr, float* loss_out) {
    if (!c || steps < 1 || batch < 32 || !(lr > 0.0f)) return cfail(MPM_EINVAL, "bad training arguments");
    if (!c->nrows) return cfail(MPM_ESTATE, "no dataset loaded");
    if (c->loss_cap < steps) {
        if (c->loss) (void)hipFree(c->loss);
        c->loss = nullptr;
        int rc;
        if ((rc = alloc(&c->loss, (size_t)steps))) return rc;
        c->loss_cap = steps;
    }
    hipLaunchKernelGGL(k_zero_f32, dim3((steps + 255) / 256), dim3(256), 0, c->stream, c->loss, steps);
    Norm nm{c->st.mean, c->st.inv_std, c->st.out_scale, c->st.out_clamp, c->st.out_dead};
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int s = 0; s < steps; ++s) {
        const int row0 = (int)(((int64_t)s * batch + c->adam_t * 7919) % c->nrows);
        hipLaunchKernelGGL(k_train, dim3(grid_blocks(batch, 32, 1024)), dim3(64), 0, c->stream, c->dX, c->dT, row0,
                           batch, c->nrows, nm, (const bf16x4*)c->st.frags, c->st.params, c->grad, c->loss + s);
        ++c->adam_t;
        const float c1 = 1.0f / (1.0f - std::pow(b1, (float)c->adam_t));
        const float c2 = 1.0f / (1.0f - std::pow(b2, (float)c->adam_t));
        hipLaunchKernelGGL(k_adam, dim3((CORR_PARAMS + 255) / 256), dim3(256), 0, c->stream, c->st.params, c->grad,
                           c->adam_m, c->adam_v, CORR_PARAMS, lr, b1, b2, eps, c1, c2);
        corr_pack(c->st.params, c->st.frags, c->stream);
    }
    if (loss_out) {
        std::vector<float> l((size_t)steps);
        CHK(hipMemcpyAsync(l.data(), c->loss, sizeof(float) * steps, hipMemcpyDeviceToHost, c->stream));
        CHK(hipStreamSynchronize(c->stream));
        for (int s = 0; s < steps; ++s) loss_out[s] = l[s] / (3.0f * (float)batch);
    }
    CHK(hipStreamSynchronize(c->stream));
    return MPM_OK;
}

}  // extern "C"
