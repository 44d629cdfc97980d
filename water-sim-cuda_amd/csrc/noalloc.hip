// noalloc.hip -- the reference's key sort without any device allocation.
//
// mpm_sort_keys_noalloc computes what mpm_engine_get_sort returns (setKeys,
// kernels.cu:361-371, and the stable ascending key sort of
// Engine::sortParticles, engine.cu:92-102) for positions the caller already
// holds in device memory, into device outputs, on the caller's stream.  Its
// scratch (the radix sort's key/value ping-pong buffers, digit histograms,
// scan partials) is the file-scope device array `scratch` below, so the call
// allocates nothing: it is usable under HIP graph capture and in loops that
// must not call hipMalloc.  The array's device address is taken once per
// device with hipGetSymbolAddress on the first call there and cached; host
// code uses that address, never the symbol itself.  Every call on a device
// shares that device's array: calls that overlap on one device (two streams,
// two host threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"
#include "mpm.h"

namespace mpm {
namespace {

constexpr size_t NA_MAX = MPM_SORT_NOALLOC_MAX;
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
// four key/value planes of NA_MAX + 1 entries
constexpr size_t NA_PLANE = align256((NA_MAX + 1) * sizeof(uint32_t));
// radix histograms: 256 digits per 4096-key tile (radix_hist_elems)
constexpr size_t NA_HIST_ELEMS = 256 * ((NA_MAX + 4095) / 4096 + 1);
// scan partials: one per 2048 values of the largest scan (scan_partial_elems) and headroom
constexpr size_t NA_PART_ELEMS = (NA_MAX + 2047) / 2048 + 1 + 16;
constexpr size_t NA_OFF_HIST = 4 * NA_PLANE;
constexpr size_t NA_OFF_PART = NA_OFF_HIST + align256(NA_HIST_ELEMS * sizeof(uint32_t));
constexpr size_t NA_OFF_CTR = NA_OFF_PART + align256(NA_PART_ELEMS * sizeof(uint32_t));
constexpr size_t NA_BYTES = NA_OFF_CTR + align256(16 * sizeof(uint32_t));

__device__ __align__(256) uint8_t scratch[NA_BYTES];

std::mutex g_mu;
std::vector<uint8_t*> g_ptr;  // per device: scratch's device address (null until the first call there)

// The reference's int keys from the biased keys the sort orders (sort_key
// adds 2^31 so that unsigned order is the signed order of the int keys).
__global__ void k_unbias(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, int n,
                         int32_t* __restrict__ keys_out, int32_t* __restrict__ perm_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys_out[i] = (int32_t)(keys[i] ^ 0x80000000u);
    perm_out[i] = (int32_t)vals[i];
}

int scratch_of_current_device(uint8_t** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_last_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
        return MPM_EHIP;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if ((size_t)dev >= g_ptr.size()) g_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_ptr[(size_t)dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return MPM_EHIP;
        }
        g_ptr[(size_t)dev] = static_cast<uint8_t*>(p);
    }
    *out = g_ptr[(size_t)dev];
    return MPM_OK;
}

}  // namespace
}  // namespace mpm

extern "C" int mpm_sort_keys_noalloc(const float* px, const float* py, const float* pz, int n, int32_t* keys_out,
                                     int32_t* perm_out, void* stream) {
    using namespace mpm;
    if (n < 0) {
        set_last_error("negative particle count");
        return MPM_EINVAL;
    }
    if ((size_t)n > NA_MAX) {
        set_last_error("mpm_sort_keys_noalloc: " + std::to_string(n) + " particles exceed its fixed scratch (at most " +
                       std::to_string(NA_MAX) + ")");
        return MPM_ELIMIT;
    }
    if (n == 0) return MPM_OK;
    if (!px || !py || !pz || !keys_out || !perm_out) {
        set_last_error("null argument");
        return MPM_EINVAL;
    }
    // the sizes the sort routines ask for must fit the fixed layout
    if (radix_hist_elems((size_t)n) > NA_HIST_ELEMS ||
        scan_partial_elems(radix_hist_elems((size_t)n)) + 16 > NA_PART_ELEMS ||
        scan_partial_elems((size_t)n + 1) + 16 > NA_PART_ELEMS) {
        set_last_error("mpm_sort_keys_noalloc: internal scratch layout too small");
        return MPM_ELIMIT;
    }
    uint8_t* base = nullptr;
    int rc = scratch_of_current_device(&base);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    SortScratch sc{};
    for (int i = 0; i < 2; ++i) {
        sc.keys[i] = reinterpret_cast<uint32_t*>(base + (size_t)i * NA_PLANE);
        sc.vals[i] = reinterpret_cast<uint32_t*>(base + (size_t)(2 + i) * NA_PLANE);
    }
    sc.hist = reinterpret_cast<uint32_t*>(base + NA_OFF_HIST);
    sc.partial = reinterpret_cast<uint32_t*>(base + NA_OFF_PART);
    sc.counters = reinterpret_cast<uint32_t*>(base + NA_OFF_CTR);
    sc.map = nullptr;
    sc.cap = NA_MAX + 1;
    sc.hist_cap = NA_HIST_ELEMS;
    sc.partial_cap = NA_PART_ELEMS;
    launch_keys_reference(px, py, pz, n, sc.keys[0], sc.vals[0], s);
    // all 32 key bits: no host read-back of the bits in use (capturable)
    const int r = radix_sort_pairs(sc, n, 32, s);
    hipLaunchKernelGGL(k_unbias, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sc.keys[r], sc.vals[r], n,
                       keys_out, perm_out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_last_error(std::string("mpm_sort_keys_noalloc: ") + hipGetErrorString(e));
        return MPM_EHIP;
    }
    return MPM_OK;
}
