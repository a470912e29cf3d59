// noalloc.hip -- the closest-hit query of lux_hip_trace_closest_device for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// lux_hip_trace_closest_device has one scratch buffer: the ray-order sort
// (keys and values of two radix passes, the digit histogram, the look-back
// words and the tile tickets), grown with hipMalloc on the BVH handle. Here it
// lives in `scratch`, a fixed device array of kScratchBytes (68 MiB) that each
// GPU holds from module load until the process ends, sized for
// kNoAllocMaxRays rays per call. Every call on a GPU shares that GPU's array:
// the histogram, look-back words and tickets are cleared by fill kernels at
// the start of each call, and calls that overlap on one GPU race.
//
// In host code `scratch` names a host-side stand-in, not device memory; the
// device address is taken once per device with hipGetSymbolAddress at library
// initialisation (noalloc_setup) and only that address is used afterwards.
#include "lux_runtime.hpp"

namespace lux {
namespace {

constexpr int kMaxDevices = 64;
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Scratch layout for up to kNoAllocMaxRays rays (the 2^22-ray traversal
// workload in one call). The sort tiles hold 2048 keys (ray_order_status_words
// is checked against kStatusWords at setup).
constexpr size_t kKeyBytes = align_up(sizeof(uint32_t) * size_t(kNoAllocMaxRays));
constexpr size_t kStatusWords = size_t((kNoAllocMaxRays + 2047) / 2048) * 256;
constexpr size_t kStatusBytes = align_up(sizeof(uint64_t) * kStatusWords);
constexpr size_t kHistBytes = align_up(sizeof(uint32_t) * 512);
constexpr size_t kTicketBytes = align_up(sizeof(uint32_t) * 4);
constexpr size_t kOffKeys0 = 0;
constexpr size_t kOffKeys1 = kOffKeys0 + kKeyBytes;
constexpr size_t kOffVals0 = kOffKeys1 + kKeyBytes;
constexpr size_t kOffVals1 = kOffVals0 + kKeyBytes;
constexpr size_t kOffStatus = kOffVals1 + kKeyBytes;
constexpr size_t kOffHist = kOffStatus + kStatusBytes;
constexpr size_t kOffTickets = kOffHist + kHistBytes;
constexpr size_t kScratchBytes = kOffTickets + kTicketBytes;
static_assert(kScratchBytes < (size_t(69) << 20), "scratch stays near 68 MiB");

}  // namespace
}  // namespace lux

static __device__ __align__(256) uint8_t scratch[lux::kScratchBytes];

namespace lux {
namespace {

// Device address of `scratch` on each device (null until noalloc_setup ran
// with that device current).
uint8_t* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

size_t noalloc_scratch_bytes() { return kScratchBytes; }

void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(LUX_ERR_INVALID_ARGUMENT, "device index beyond 64");
    if (g_scratch_ptr[device]) return;
    if (ray_order_status_words(kNoAllocMaxRays) > kStatusWords)
        throw Error(LUX_ERR_DEVICE_CHECK, "no-alloc scratch: look-back words exceed the reserved space");
    void* ptr = nullptr;
    hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) throw Error(LUX_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[device] = static_cast<uint8_t*>(ptr);
}

void trace_closest_noalloc(int device, const BvhDev& d, bool order_rays, const lux_triangle* tris, int tri_count,
                           const lux_ray* rays, int n, lux_hit* hits, hipStream_t s) {
    if (n > kNoAllocMaxRays)
        throw Error(LUX_ERR_INVALID_ARGUMENT,
                    "lux_hip_trace_closest_device_noalloc: more than 4194304 rays in one call (split the batch)");
    if (device < 0 || device >= kMaxDevices || !g_scratch_ptr[device])
        throw Error(LUX_ERR_NOT_INITIALIZED, "no-alloc scratch not set up on this device (call lux_hip_init)");
    if (n <= 0) return;
    const uint32_t* ord = nullptr;
    if (order_rays && n >= kOrderMinRays && d.n >= 2) {
        uint8_t* base = g_scratch_ptr[device];
        const RayOrderBufs b{{reinterpret_cast<uint32_t*>(base + kOffKeys0), reinterpret_cast<uint32_t*>(base + kOffKeys1)},
                             {reinterpret_cast<uint32_t*>(base + kOffVals0), reinterpret_cast<uint32_t*>(base + kOffVals1)},
                             reinterpret_cast<uint32_t*>(base + kOffHist),
                             reinterpret_cast<uint64_t*>(base + kOffStatus),
                             reinterpret_cast<uint32_t*>(base + kOffTickets)};
        // the last call's look-back words and tickets must not be read as this one's
        zero_words(reinterpret_cast<uint32_t*>(b.status), 2 * ray_order_status_words(n), s);
        zero_words(b.tickets, 4, s);
        ord = ray_order_into(b, 1u, d, rays, n, s);
    }
    trace_closest(d, ord, tris, tri_count, rays, n, hits, s);
}

}  // namespace lux
