// abi.hip -- the extern "C" boundary (include/lux_hip.h) over the gfx950
// kernels: handle ownership, host<->device staging for the drop-in calls,
// status codes + thread-local error text, and the kernel-time probe.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lux_runtime.hpp"
#include "two_level.hpp"

struct lux_hip_two_level {
    lux::TwoLevelDev dev;
    bool built = false;
    hipStream_t last_stream = nullptr;
    lux::RaySortDev order;  // ray-order scratch of the query calls
};

struct lux_hip_bvh {
    lux::BvhDev dev;
    bool valid = false;              // GpuBvh::is_valid
    bool host_ids_known = false;     // host copy of the build ids is current
    std::vector<int32_t> host_ids;   // for refit's id check (src/bvh/gpu_bvh.cu:601-605)
    hipStream_t last_stream = nullptr;
    lux::RaySortDev order;           // ray-order scratch of the traversal calls
    bool order_rays = true;          // trace large any-hit ray sets in origin order
};

namespace lux {
namespace {

thread_local std::string g_last_error;
bool g_initialized = false;
int g_device = 0;
hipStream_t g_stream = nullptr;

struct Probe {
    std::mutex mu;
    bool open = false;
    std::string name;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    std::vector<hipEvent_t> pool;
    hipEvent_t pending = nullptr;
} g_probe;

hipEvent_t probe_event() {
    if (!g_probe.pool.empty()) {
        hipEvent_t e = g_probe.pool.back();
        g_probe.pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
}

int fail(int status, const std::string& msg) {
    g_last_error = msg;
    return status;
}

template <typename F>
int guarded(F&& f) {
    try {
        g_last_error.clear();
        return f();
    } catch (const Error& e) {
        return fail(e.status, e.what());
    } catch (const std::bad_alloc&) {
        return fail(LUX_ERR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(LUX_ERR_HIP, e.what());
    }
}

void require_init() {
    if (!g_initialized) throw Error(LUX_ERR_NOT_INITIALIZED, "lux_hip_init() has not been called");
}

hipStream_t pick(lux_stream s) { return s ? static_cast<hipStream_t>(s) : g_stream; }

// Reads and clears the device-side check bits of a handle (caller synchronised).
void check_device_flags(lux_hip_bvh* h) {
    if (!h->dev.info) return;
    DevInfo info;
    hip_check(hipMemcpy(&info, h->dev.info, sizeof(DevInfo), hipMemcpyDeviceToHost), "read device info");
    if (info.error == 0) return;
    uint32_t zero = 0;
    hip_check(hipMemcpy(reinterpret_cast<char*>(h->dev.info) + offsetof(DevInfo, error), &zero,
                        sizeof(zero), hipMemcpyHostToDevice),
              "clear device info");
    std::string what;
    if (info.error & kErrNegativeId) what += " negative primitive id;";
    if (info.error & kErrPrimRange) what += " leaf primitive id outside the triangle array;";
    if (info.error & kErrSpinLimit) what += " radix-sort look-back spin limit;";
    if (info.error & kErrWalkOrder) what += " stackless walk did not advance;";
    if (info.error & kErrClimbLimit) what += " bottom-up pass exceeded n levels;";
    if (info.error & kErrNegativeId) h->valid = false;
    throw Error(LUX_ERR_DEVICE_CHECK, "device check failed:" + what);
}

void sync_handle(lux_hip_bvh* h) {
    hip_check(hipStreamSynchronize(h->last_stream ? h->last_stream : g_stream), "synchronize");
    check_device_flags(h);
}

void require_handle(lux_hip_bvh* h) {
    if (!h) throw Error(LUX_ERR_INVALID_ARGUMENT, "null lux_hip_bvh handle");
}

template <typename T>
struct DeviceBuffer {
    T* p = nullptr;
    explicit DeviceBuffer(size_t count) {
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
        if (e != hipSuccess) throw Error(LUX_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    }
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

}  // namespace

hipStream_t library_stream() { return g_stream; }

void probe_pre_launch(const char* name, hipStream_t s) {
    if (!g_probe.open || g_probe.name != name) return;
    std::lock_guard<std::mutex> lock(g_probe.mu);
    g_probe.pending = probe_event();
    hip_check(hipEventRecord(g_probe.pending, s), "hipEventRecord");
}

void probe_post_launch(const char* name, hipStream_t s) {
    if (!g_probe.open || g_probe.name != name || !g_probe.pending) return;
    std::lock_guard<std::mutex> lock(g_probe.mu);
    hipEvent_t stop = probe_event();
    hip_check(hipEventRecord(stop, s), "hipEventRecord");
    g_probe.events.emplace_back(g_probe.pending, stop);
    g_probe.pending = nullptr;
}

}  // namespace lux

using namespace lux;

// Frees the path tracer's shared engines (defined with the path tracer entry points).
static void release_render_engines();

extern "C" {

int lux_hip_version(void) { return 0x000100; }

const char* lux_hip_last_error(void) { return g_last_error.c_str(); }

int lux_hip_init(int device) {
    return guarded([&] {
        if (g_initialized) {
            if (device != g_device) throw Error(LUX_ERR_INVALID_ARGUMENT, "already initialised on another device");
            return int(LUX_OK);
        }
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device < 0 || device >= count)
            throw Error(LUX_ERR_INVALID_ARGUMENT, "device index out of range");
        hip_check(hipSetDevice(device), "hipSetDevice");
        noalloc_setup(device);  // device address of the no-alloc scratch array
        hip_check(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking), "hipStreamCreate");
        g_device = device;
        g_initialized = true;
        return int(LUX_OK);
    });
}

int lux_hip_shutdown(void) {
    return guarded([&] {
        if (!g_initialized) return int(LUX_OK);
        hip_check(hipStreamSynchronize(g_stream), "synchronize");
        {
            std::lock_guard<std::mutex> lock(g_probe.mu);
            for (auto& pr : g_probe.events) {
                (void)hipEventDestroy(pr.first);
                (void)hipEventDestroy(pr.second);
            }
            for (auto e : g_probe.pool) (void)hipEventDestroy(e);
            g_probe.events.clear();
            g_probe.pool.clear();
            g_probe.open = false;
        }
        release_render_engines();
        (void)hipStreamDestroy(g_stream);
        g_stream = nullptr;
        g_initialized = false;
        return int(LUX_OK);
    });
}

int lux_hip_device_info(char* name, int32_t name_len, char* arch, int32_t arch_len, int32_t* compute_units,
                        uint64_t* global_mem_bytes) {
    return guarded([&] {
        require_init();
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, g_device), "hipGetDeviceProperties");
        if (name && name_len > 0) {
            std::strncpy(name, prop.name, size_t(name_len) - 1);
            name[name_len - 1] = 0;
        }
        if (arch && arch_len > 0) {
            std::strncpy(arch, prop.gcnArchName, size_t(arch_len) - 1);
            arch[arch_len - 1] = 0;
        }
        if (compute_units) *compute_units = prop.multiProcessorCount;
        if (global_mem_bytes) *global_mem_bytes = uint64_t(prop.totalGlobalMem);
        return int(LUX_OK);
    });
}

int lux_hip_malloc(void** ptr, size_t bytes) {
    return guarded([&] {
        require_init();
        if (!ptr) throw Error(LUX_ERR_INVALID_ARGUMENT, "null output pointer");
        *ptr = nullptr;
        hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
        if (e != hipSuccess) throw Error(LUX_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
        return int(LUX_OK);
    });
}

int lux_hip_free(void* ptr) {
    return guarded([&] {
        require_init();
        if (ptr) hip_check(hipFree(ptr), "hipFree");
        return int(LUX_OK);
    });
}

int lux_hip_memcpy_htod(void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        require_init();
        if (bytes && (!dst || !src)) throw Error(LUX_ERR_INVALID_ARGUMENT, "null pointer");
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream), "hipMemcpy H2D");
        hip_check(hipStreamSynchronize(g_stream), "synchronize");
        return int(LUX_OK);
    });
}

int lux_hip_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    return guarded([&] {
        require_init();
        if (bytes && (!dst || !src)) throw Error(LUX_ERR_INVALID_ARGUMENT, "null pointer");
        hip_check(hipStreamSynchronize(g_stream), "synchronize");
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(g_stream), "synchronize");
        return int(LUX_OK);
    });
}

int lux_hip_synchronize(void) {
    return guarded([&] {
        require_init();
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        return int(LUX_OK);
    });
}

int lux_hip_copy_device(void* dst, const void* src, size_t bytes, lux_stream stream) {
    return guarded([&] {
        require_init();
        if (bytes % 16 != 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "copy size must be a multiple of 16 bytes");
        if (bytes && (!dst || !src)) throw Error(LUX_ERR_INVALID_ARGUMENT, "null pointer");
        if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 16 != 0)
            throw Error(LUX_ERR_INVALID_ARGUMENT, "copy pointers must be 16-byte aligned");
        copy_device(dst, src, bytes, pick(stream));
        return int(LUX_OK);
    });
}

int lux_hip_bvh_create(lux_hip_bvh** out) {
    return guarded([&] {
        require_init();
        if (!out) throw Error(LUX_ERR_INVALID_ARGUMENT, "null output pointer");
        *out = new lux_hip_bvh();
        return int(LUX_OK);
    });
}

int lux_hip_bvh_destroy(lux_hip_bvh* bvh) {
    return guarded([&] {
        if (!bvh) return int(LUX_OK);
        if (g_initialized) (void)hipStreamSynchronize(bvh->last_stream ? bvh->last_stream : g_stream);
        bvh_release(bvh->dev);
        raysort_release(bvh->order);
        delete bvh;
        return int(LUX_OK);
    });
}

int lux_hip_bvh_set_ray_order(lux_hip_bvh* bvh, int32_t enable) {
    return guarded([&] {
        require_handle(bvh);
        bvh->order_rays = enable != 0;
        return int(LUX_OK);
    });
}

int lux_hip_bvh_build(lux_hip_bvh* bvh, const lux_aabb* boxes, const int32_t* ids, int32_t n) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (n < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative primitive count");
        if (n > 0 && (!boxes || !ids)) throw Error(LUX_ERR_INVALID_ARGUMENT, "null input arrays");
        if (n >= (1 << 29)) throw Error(LUX_ERR_INVALID_ARGUMENT, "primitive count exceeds 2^29");
        for (int32_t i = 0; i < n; ++i)
            if (ids[i] < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "primitive ids must be >= 0");
        bvh->valid = false;
        bvh->host_ids_known = false;
        if (n == 0) {  // build_gpu_lbvh of an empty list: invalid GpuBvh, count 0
            bvh->dev.n = 0;
            bvh->host_ids.clear();
            bvh->host_ids_known = true;
            return int(LUX_OK);
        }
        bvh_allocate(bvh->dev, n);
        hip_check(hipMemcpyAsync(bvh->dev.boxes, boxes, sizeof(lux_aabb) * n, hipMemcpyHostToDevice, g_stream),
                  "H2D boxes");
        hip_check(hipMemcpyAsync(bvh->dev.ids, ids, sizeof(int32_t) * n, hipMemcpyHostToDevice, g_stream),
                  "H2D ids");
        bvh->last_stream = g_stream;
        lbvh_build(bvh->dev, bvh->dev.boxes, bvh->dev.ids, n, g_stream);
        sync_handle(bvh);
        bvh->host_ids.assign(ids, ids + n);
        bvh->host_ids_known = true;
        bvh->valid = true;
        return int(LUX_OK);
    });
}

int lux_hip_bvh_build_device(lux_hip_bvh* bvh, const lux_aabb* d_boxes, const int32_t* d_ids, int32_t n,
                             lux_stream stream) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (n < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative primitive count");
        if (n > 0 && (!d_boxes || !d_ids)) throw Error(LUX_ERR_INVALID_ARGUMENT, "null input arrays");
        if (n >= (1 << 29)) throw Error(LUX_ERR_INVALID_ARGUMENT, "primitive count exceeds 2^29");
        bvh->valid = false;
        bvh->host_ids_known = false;
        if (n == 0) {
            bvh->dev.n = 0;
            return int(LUX_OK);
        }
        hipStream_t s = pick(stream);
        if (bvh->last_stream && bvh->last_stream != s) hip_check(hipStreamSynchronize(bvh->last_stream), "synchronize");
        bvh->last_stream = s;
        bvh_allocate(bvh->dev, n);
        lbvh_build(bvh->dev, d_boxes, d_ids, n, s);
        bvh->valid = true;  // negative ids are reported by lux_hip_bvh_info
        return int(LUX_OK);
    });
}

int lux_hip_bvh_build_triangles_device(lux_hip_bvh* bvh, const lux_triangle* d_tris, int32_t n, lux_stream stream) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (n < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative primitive count");
        if (n > 0 && !d_tris) throw Error(LUX_ERR_INVALID_ARGUMENT, "null triangle array");
        if (n >= (1 << 29)) throw Error(LUX_ERR_INVALID_ARGUMENT, "primitive count exceeds 2^29");
        bvh->valid = false;
        bvh->host_ids_known = false;
        if (n == 0) {
            bvh->dev.n = 0;
            return int(LUX_OK);
        }
        hipStream_t s = pick(stream);
        if (bvh->last_stream && bvh->last_stream != s) hip_check(hipStreamSynchronize(bvh->last_stream), "synchronize");
        bvh->last_stream = s;
        lbvh_build_triangles(bvh->dev, d_tris, nullptr, nullptr, 0, n, s);
        bvh->valid = true;
        return int(LUX_OK);
    });
}

int lux_hip_bvh_refit(lux_hip_bvh* bvh, const lux_aabb* boxes, const int32_t* ids, int32_t n) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (n < 0 || (n > 0 && (!boxes || !ids))) throw Error(LUX_ERR_INVALID_ARGUMENT, "bad refit input");
        if (n != (bvh->valid ? bvh->dev.n : 0))
            throw Error(LUX_ERR_MISMATCH, "refit primitive count does not match the built BVH");
        if (!bvh->valid) return int(LUX_OK);  // empty == empty (src/bvh/gpu_bvh.cu:599)
        if (!bvh->host_ids_known) {
            sync_handle(bvh);
            bvh->host_ids.resize(n);
            hip_check(hipMemcpy(bvh->host_ids.data(), bvh->dev.ids, sizeof(int32_t) * n, hipMemcpyDeviceToHost),
                      "D2H ids");
            bvh->host_ids_known = true;
        }
        for (int32_t i = 0; i < n; ++i)
            if (bvh->host_ids[i] != ids[i])
                throw Error(LUX_ERR_MISMATCH, "refit primitive ids differ from the built BVH");
        hipStream_t s = g_stream;
        if (bvh->last_stream && bvh->last_stream != s) hip_check(hipStreamSynchronize(bvh->last_stream), "synchronize");
        bvh->last_stream = s;
        hip_check(hipMemcpyAsync(bvh->dev.boxes, boxes, sizeof(lux_aabb) * n, hipMemcpyHostToDevice, s), "H2D boxes");
        lbvh_refit_boxes(bvh->dev, bvh->dev.boxes, s);
        sync_handle(bvh);
        return int(LUX_OK);
    });
}

int lux_hip_bvh_refit_device(lux_hip_bvh* bvh, const lux_aabb* d_boxes, int32_t n, lux_stream stream) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (n != (bvh->valid ? bvh->dev.n : 0))
            throw Error(LUX_ERR_MISMATCH, "refit primitive count does not match the built BVH");
        if (!bvh->valid) return int(LUX_OK);
        if (!d_boxes) throw Error(LUX_ERR_INVALID_ARGUMENT, "null boxes");
        hipStream_t s = pick(stream);
        if (bvh->last_stream && bvh->last_stream != s) hip_check(hipStreamSynchronize(bvh->last_stream), "synchronize");
        bvh->last_stream = s;
        lbvh_refit_boxes(bvh->dev, d_boxes, s);
        return int(LUX_OK);
    });
}

int lux_hip_bvh_refit_triangles_device(lux_hip_bvh* bvh, const lux_triangle* d_tris, int32_t n,
                                       lux_stream stream) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (n != (bvh->valid ? bvh->dev.n : 0))
            throw Error(LUX_ERR_MISMATCH, "refit primitive count does not match the built BVH");
        if (!bvh->valid) return int(LUX_OK);
        if (!d_tris) throw Error(LUX_ERR_INVALID_ARGUMENT, "null triangles");
        hipStream_t s = pick(stream);
        if (bvh->last_stream && bvh->last_stream != s) hip_check(hipStreamSynchronize(bvh->last_stream), "synchronize");
        bvh->last_stream = s;
        lbvh_refit_triangles(bvh->dev, d_tris, s);
        return int(LUX_OK);
    });
}

int lux_hip_bvh_info(lux_hip_bvh* bvh, int32_t* valid, int32_t* count, lux_aabb* root) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (bvh->valid) sync_handle(bvh);
        if (valid) *valid = bvh->valid ? 1 : 0;
        if (count) *count = bvh->valid ? bvh->dev.n : 0;
        if (root) {
            lux_aabb empty;
            empty.p_min = lux_vec3{INFINITY, INFINITY, INFINITY};
            empty.p_max = lux_vec3{-INFINITY, -INFINITY, -INFINITY};
            *root = empty;
            if (bvh->valid)
                hip_check(hipMemcpy(root, &bvh->dev.info->root, sizeof(lux_aabb), hipMemcpyDeviceToHost), "D2H root");
        }
        return int(LUX_OK);
    });
}

int lux_hip_bvh_view(lux_hip_bvh* bvh, lux_bvh_view* out) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (!out) throw Error(LUX_ERR_INVALID_ARGUMENT, "null output");
        lux_bvh_view v;
        std::memset(&v, 0, sizeof(v));
        if (bvh->valid) {
            sync_handle(bvh);
            v.nodes = bvh->dev.nodes;
            v.node_count = 2 * bvh->dev.n - 1;
            hip_check(hipMemcpy(&v.root_bounds, &bvh->dev.info->root, sizeof(lux_aabb), hipMemcpyDeviceToHost),
                      "D2H root");
        }
        *out = v;
        return int(LUX_OK);
    });
}

int lux_hip_bvh_download_nodes(lux_hip_bvh* bvh, lux_bvh_node* nodes, int32_t capacity) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (!bvh->valid) return int(LUX_OK);
        int total = 2 * bvh->dev.n - 1;
        if (!nodes || capacity < total) throw Error(LUX_ERR_INVALID_ARGUMENT, "node buffer too small");
        sync_handle(bvh);
        hip_check(hipMemcpy(nodes, bvh->dev.nodes, sizeof(lux_bvh_node) * total, hipMemcpyDeviceToHost), "D2H nodes");
        return int(LUX_OK);
    });
}

int lux_hip_bvh_download_topology(lux_hip_bvh* bvh, uint32_t* morton_sorted, uint32_t* sorted_indices,
                                  uint32_t* left, uint32_t* right, uint32_t* parent, uint32_t* leaf_parent,
                                  lux_aabb* node_bounds) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (!bvh->valid) return int(LUX_OK);
        sync_handle(bvh);
        const BvhDev& d = bvh->dev;
        int n = d.n;
        auto d2h = [](void* dst, const void* src, size_t bytes) {
            if (dst && bytes) hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "D2H topology");
        };
        d2h(morton_sorted, d.keys[0], 4 * size_t(n));
        d2h(sorted_indices, d.vals[0], 4 * size_t(n));
        if (n > 1) {
            d2h(left, d.left, 4 * size_t(n - 1));
            d2h(right, d.right, 4 * size_t(n - 1));
            d2h(parent, d.parent, 4 * size_t(n - 1));
            d2h(leaf_parent, d.leaf_parent, 4 * size_t(n));
        } else if (leaf_parent) {
            leaf_parent[0] = 0;
        }
        if (node_bounds) {
            int total = 2 * n - 1;
            std::vector<lux_bvh_node> nodes(total);
            std::vector<int32_t> pre_i(std::max(n - 1, 1)), pre_l(n);
            d2h(nodes.data(), d.nodes, sizeof(lux_bvh_node) * total);
            d2h(pre_l.data(), d.pre_leaf, 4 * size_t(n));
            if (n > 1) d2h(pre_i.data(), d.pre_internal, 4 * size_t(n - 1));
            for (int i = 0; i < n - 1; ++i) node_bounds[i] = nodes[pre_i[i]].bounds;
            for (int j = 0; j < n; ++j) node_bounds[n - 1 + j] = nodes[pre_l[j]].bounds;
        }
        return int(LUX_OK);
    });
}

int lux_hip_bvh_download_input(lux_hip_bvh* bvh, lux_aabb* boxes, int32_t* primitive_ids) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (!bvh->valid) return int(LUX_OK);
        sync_handle(bvh);
        const size_t n = size_t(bvh->dev.n);
        if (boxes) hip_check(hipMemcpy(boxes, bvh->dev.boxes, sizeof(lux_aabb) * n, hipMemcpyDeviceToHost), "D2H boxes");
        if (primitive_ids)
            hip_check(hipMemcpy(primitive_ids, bvh->dev.ids, sizeof(int32_t) * n, hipMemcpyDeviceToHost), "D2H ids");
        return int(LUX_OK);
    });
}

int lux_hip_bvh_device_arrays(lux_hip_bvh* bvh, const lux_aabb** d_input_bounds, const int32_t** d_primitive_ids,
                              const uint32_t** d_sorted_indices) {
    return guarded([&] {
        require_handle(bvh);
        const bool v = bvh->valid;
        if (d_input_bounds) *d_input_bounds = v ? bvh->dev.boxes : nullptr;
        if (d_primitive_ids) *d_primitive_ids = v ? bvh->dev.ids : nullptr;
        if (d_sorted_indices) *d_sorted_indices = v ? bvh->dev.vals[0] : nullptr;
        return int(LUX_OK);
    });
}

int lux_hip_trace_closest(lux_hip_bvh* bvh, const lux_triangle* tris, int32_t tri_count, const lux_ray* rays,
                          int32_t n, lux_hit* hits) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (n < 0 || tri_count < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative count");
        if (n > 0 && (!rays || !hits)) throw Error(LUX_ERR_INVALID_ARGUMENT, "null ray/hit arrays");
        for (int32_t i = 0; i < n; ++i) {
            hits[i].primitive_id = -1;
            hits[i].t = INFINITY;
        }
        if (!bvh->valid || n == 0) return int(LUX_OK);
        if (tri_count > 0 && !tris) throw Error(LUX_ERR_INVALID_ARGUMENT, "null triangle array");
        sync_handle(bvh);
        DeviceBuffer<lux_triangle> dt(tri_count);
        DeviceBuffer<lux_ray> dr(n);
        DeviceBuffer<lux_hit> dh(n);
        hip_check(hipMemcpyAsync(dt.p, tris, sizeof(lux_triangle) * tri_count, hipMemcpyHostToDevice, g_stream), "H2D tris");
        hip_check(hipMemcpyAsync(dr.p, rays, sizeof(lux_ray) * n, hipMemcpyHostToDevice, g_stream), "H2D rays");
        bvh->last_stream = g_stream;
        trace_closest(bvh->dev, closest_order(bvh->order, bvh->order_rays, bvh->dev, dr.p, n, g_stream), dt.p, tri_count,
                      dr.p, n, dh.p, g_stream);
        hip_check(hipMemcpyAsync(hits, dh.p, sizeof(lux_hit) * n, hipMemcpyDeviceToHost, g_stream), "D2H hits");
        sync_handle(bvh);
        return int(LUX_OK);
    });
}

int lux_hip_trace_closest_device(lux_hip_bvh* bvh, const lux_triangle* d_tris, int32_t tri_count,
                                 const lux_ray* d_rays, int32_t n, lux_hit* d_hits, lux_stream stream) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (n < 0 || tri_count < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative count");
        if (n > 0 && (!d_rays || !d_hits)) throw Error(LUX_ERR_INVALID_ARGUMENT, "null ray/hit arrays");
        if (!bvh->valid) throw Error(LUX_ERR_INVALID_ARGUMENT, "BVH is not valid");
        hipStream_t s = pick(stream);
        if (bvh->last_stream && bvh->last_stream != s) hip_check(hipStreamSynchronize(bvh->last_stream), "synchronize");
        bvh->last_stream = s;
        trace_closest(bvh->dev, closest_order(bvh->order, bvh->order_rays, bvh->dev, d_rays, n, s), d_tris, tri_count,
                      d_rays, n, d_hits, s);
        return int(LUX_OK);
    });
}

int lux_hip_trace_closest_device_noalloc(lux_hip_bvh* bvh, const lux_triangle* d_tris, int32_t tri_count,
                                         const lux_ray* d_rays, int32_t n, lux_hit* d_hits, lux_stream stream) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (n < 0 || tri_count < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative count");
        if (n > kNoAllocMaxRays)
            throw Error(LUX_ERR_INVALID_ARGUMENT,
                        "lux_hip_trace_closest_device_noalloc: more than 4194304 rays in one call (split the batch)");
        if (n > 0 && (!d_rays || !d_hits)) throw Error(LUX_ERR_INVALID_ARGUMENT, "null ray/hit arrays");
        if (!bvh->valid) throw Error(LUX_ERR_INVALID_ARGUMENT, "BVH is not valid");
        hipStream_t s = pick(stream);
        if (bvh->last_stream && bvh->last_stream != s) hip_check(hipStreamSynchronize(bvh->last_stream), "synchronize");
        bvh->last_stream = s;
        trace_closest_noalloc(g_device, bvh->dev, bvh->order_rays, d_tris, tri_count, d_rays, n, d_hits, s);
        return int(LUX_OK);
    });
}

int lux_hip_trace_any(lux_hip_bvh* bvh, const lux_triangle* tris, int32_t tri_count, const lux_ray* rays,
                      int32_t n, uint8_t* occluded) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (n < 0 || tri_count < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative count");
        if (n > 0 && (!rays || !occluded)) throw Error(LUX_ERR_INVALID_ARGUMENT, "null ray/output arrays");
        std::memset(occluded, 0, size_t(n));
        if (!bvh->valid || n == 0) return int(LUX_OK);
        if (tri_count > 0 && !tris) throw Error(LUX_ERR_INVALID_ARGUMENT, "null triangle array");
        sync_handle(bvh);
        DeviceBuffer<lux_triangle> dt(tri_count);
        DeviceBuffer<lux_ray> dr(n);
        DeviceBuffer<uint8_t> dh(n);
        hip_check(hipMemcpyAsync(dt.p, tris, sizeof(lux_triangle) * tri_count, hipMemcpyHostToDevice, g_stream), "H2D tris");
        hip_check(hipMemcpyAsync(dr.p, rays, sizeof(lux_ray) * n, hipMemcpyHostToDevice, g_stream), "H2D rays");
        bvh->last_stream = g_stream;
        trace_any(bvh->dev, bvh->order_rays ? &bvh->order : nullptr, dt.p, tri_count, dr.p, n, dh.p, g_stream);
        hip_check(hipMemcpyAsync(occluded, dh.p, size_t(n), hipMemcpyDeviceToHost, g_stream), "D2H occluded");
        sync_handle(bvh);
        return int(LUX_OK);
    });
}

int lux_hip_trace_any_device(lux_hip_bvh* bvh, const lux_triangle* d_tris, int32_t tri_count,
                             const lux_ray* d_rays, int32_t n, uint8_t* d_occluded, lux_stream stream) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (n < 0 || tri_count < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative count");
        if (n > 0 && (!d_rays || !d_occluded)) throw Error(LUX_ERR_INVALID_ARGUMENT, "null ray/output arrays");
        if (!bvh->valid) throw Error(LUX_ERR_INVALID_ARGUMENT, "BVH is not valid");
        hipStream_t s = pick(stream);
        if (bvh->last_stream && bvh->last_stream != s) hip_check(hipStreamSynchronize(bvh->last_stream), "synchronize");
        bvh->last_stream = s;
        trace_any(bvh->dev, bvh->order_rays ? &bvh->order : nullptr, d_tris, tri_count, d_rays, n, d_occluded, s);
        return int(LUX_OK);
    });
}

static int overlap_host(lux_hip_bvh* tree, lux_hip_bvh* queries, bool self_mode, int32_t max_pairs,
                        lux_pair* pairs, int32_t* found) {
    require_init();
    require_handle(tree);
    require_handle(queries);
    if (!found) throw Error(LUX_ERR_INVALID_ARGUMENT, "null found_count");
    *found = 0;
    if (!tree->valid || !queries->valid || max_pairs <= 0) return LUX_OK;
    if (!pairs) throw Error(LUX_ERR_INVALID_ARGUMENT, "null pair buffer");
    sync_handle(tree);
    if (queries != tree) sync_handle(queries);
    DeviceBuffer<lux_pair> dp{size_t(max_pairs)};
    tree->last_stream = g_stream;
    overlap_query(tree->dev, queries->dev, self_mode, max_pairs, dp.p, tree->dev.counter, g_stream);
    int32_t count = 0;
    hip_check(hipMemcpyAsync(&count, tree->dev.counter, sizeof(int32_t), hipMemcpyDeviceToHost, g_stream), "D2H count");
    hip_check(hipStreamSynchronize(g_stream), "synchronize");
    int32_t copy = count < max_pairs ? count : max_pairs;
    if (copy > 0) hip_check(hipMemcpy(pairs, dp.p, sizeof(lux_pair) * copy, hipMemcpyDeviceToHost), "D2H pairs");
    check_device_flags(tree);
    *found = count;
    return LUX_OK;
}

int lux_hip_overlap_self(lux_hip_bvh* bvh, int32_t max_pairs, lux_pair* pairs, int32_t* found) {
    return guarded([&] { return overlap_host(bvh, bvh, true, max_pairs, pairs, found); });
}

int lux_hip_overlap_cross(lux_hip_bvh* left, lux_hip_bvh* right, int32_t max_pairs, lux_pair* pairs,
                          int32_t* found) {
    return guarded([&] { return overlap_host(left, right, false, max_pairs, pairs, found); });
}

static int overlap_device(lux_hip_bvh* tree, lux_hip_bvh* queries, bool self_mode, int32_t max_pairs,
                          lux_pair* d_pairs, int32_t* d_found, lux_stream stream) {
    require_init();
    require_handle(tree);
    require_handle(queries);
    if (!d_found) throw Error(LUX_ERR_INVALID_ARGUMENT, "null found_count");
    if (max_pairs > 0 && !d_pairs) throw Error(LUX_ERR_INVALID_ARGUMENT, "null pair buffer");
    hipStream_t s = pick(stream);
    for (lux_hip_bvh* h : {tree, queries})
        if (h->last_stream && h->last_stream != s) hip_check(hipStreamSynchronize(h->last_stream), "synchronize");
    tree->last_stream = s;
    queries->last_stream = s;
    BvhDev none;
    overlap_query(tree->valid ? tree->dev : none, queries->valid ? queries->dev : none, self_mode, max_pairs,
                  d_pairs, d_found, s);
    return LUX_OK;
}

int lux_hip_overlap_self_device(lux_hip_bvh* bvh, int32_t max_pairs, lux_pair* d_pairs, int32_t* d_found,
                                lux_stream stream) {
    return guarded([&] { return overlap_device(bvh, bvh, true, max_pairs, d_pairs, d_found, stream); });
}

int lux_hip_overlap_cross_device(lux_hip_bvh* left, lux_hip_bvh* right, int32_t max_pairs, lux_pair* d_pairs,
                                 int32_t* d_found, lux_stream stream) {
    return guarded([&] { return overlap_device(left, right, false, max_pairs, d_pairs, d_found, stream); });
}

int lux_hip_trace_stats_device(lux_hip_bvh* bvh, const lux_triangle* d_tris, int32_t tri_count,
                               const lux_ray* d_rays, int32_t n, uint64_t* stats) {
    return guarded([&] {
        require_init();
        require_handle(bvh);
        if (n < 0 || tri_count < 0 || !stats) throw Error(LUX_ERR_INVALID_ARGUMENT, "bad stats arguments");
        if (n > 0 && !d_rays) throw Error(LUX_ERR_INVALID_ARGUMENT, "null rays");
        constexpr int kStats = 6;  // {closest, any-hit} x {node units, visits, triangle tests}
        for (int k = 0; k < kStats; ++k) stats[k] = 0;
        if (!bvh->valid) return int(LUX_OK);
        sync_handle(bvh);
        DeviceBuffer<unsigned long long> ds(kStats);
        bvh->last_stream = g_stream;
        trace_stats(bvh->dev, d_tris, tri_count, d_rays, n, ds.p, g_stream);
        unsigned long long h[kStats];
        hip_check(hipMemcpyAsync(h, ds.p, sizeof(h), hipMemcpyDeviceToHost, g_stream), "D2H stats");
        sync_handle(bvh);
        for (int k = 0; k < kStats; ++k) stats[k] = h[k];
        return int(LUX_OK);
    });
}

// ---- two-level scenes -----------------------------------------------------------
namespace lux {
namespace {
void require_two_level(lux_hip_two_level* s) {
    if (!s) throw Error(LUX_ERR_INVALID_ARGUMENT, "null lux_hip_two_level handle");
}

// Waits for the scene's last stream and turns device check bits into an error.
void sync_two_level(lux_hip_two_level* s) {
    hip_check(hipStreamSynchronize(s->last_stream ? s->last_stream : g_stream), "synchronize");
    DevInfo* info = s->dev.tlas.info;
    if (!info) return;
    DevInfo h;
    hip_check(hipMemcpy(&h, info, sizeof(DevInfo), hipMemcpyDeviceToHost), "read device info");
    if (h.error == 0) return;
    uint32_t zero = 0;
    hip_check(hipMemcpy(reinterpret_cast<char*>(info) + offsetof(DevInfo, error), &zero, sizeof(zero),
                        hipMemcpyHostToDevice),
              "clear device info");
    std::string what;
    if (h.error & kErrPrimRange) what += " leaf primitive id outside its mesh;";
    if (h.error & kErrWalkOrder) what += " two-level walk exceeded its step limit;";
    throw Error(LUX_ERR_DEVICE_CHECK, "device check failed:" + what);
}

hipStream_t use_stream(lux_hip_two_level* s, lux_stream stream) {
    hipStream_t st = pick(stream);
    if (s->last_stream && s->last_stream != st) hip_check(hipStreamSynchronize(s->last_stream), "synchronize");
    s->last_stream = st;
    return st;
}
}  // namespace
}  // namespace lux

int lux_hip_two_level_create(lux_hip_two_level** out) {
    return guarded([&] {
        require_init();
        if (!out) throw Error(LUX_ERR_INVALID_ARGUMENT, "null output pointer");
        *out = new lux_hip_two_level();
        return int(LUX_OK);
    });
}

int lux_hip_two_level_destroy(lux_hip_two_level* s) {
    return guarded([&] {
        if (!s) return int(LUX_OK);
        if (g_initialized) (void)hipStreamSynchronize(s->last_stream ? s->last_stream : g_stream);
        two_level_release(s->dev);
        raysort_release(s->order);
        delete s;
        return int(LUX_OK);
    });
}

int lux_hip_two_level_build(lux_hip_two_level* s, const lux_triangle* tris, const int32_t* mesh_offsets,
                            int32_t mesh_count, const lux_instance* instances, int32_t instance_count) {
    return guarded([&] {
        require_init();
        require_two_level(s);
        if (s->last_stream) hip_check(hipStreamSynchronize(s->last_stream), "synchronize");
        s->built = false;
        s->last_stream = g_stream;
        two_level_build(s->dev, tris, mesh_offsets, mesh_count, instances, instance_count, g_stream);
        s->built = true;
        return int(LUX_OK);
    });
}

int lux_hip_two_level_set_instances(lux_hip_two_level* s, const lux_instance* instances, int32_t instance_count) {
    return guarded([&] {
        require_init();
        require_two_level(s);
        if (!s->built) throw Error(LUX_ERR_INVALID_ARGUMENT, "scene has not been built");
        if (s->last_stream) hip_check(hipStreamSynchronize(s->last_stream), "synchronize");
        s->last_stream = g_stream;
        two_level_set_instances(s->dev, instances, instance_count, g_stream);
        return int(LUX_OK);
    });
}

int lux_hip_two_level_intersect(lux_hip_two_level* s, const lux_ray* rays, const int32_t* ignore, int32_t n,
                                lux_scene_hit* hits) {
    return guarded([&] {
        require_init();
        require_two_level(s);
        if (n < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative ray count");
        if (n > 0 && (!rays || !hits)) throw Error(LUX_ERR_INVALID_ARGUMENT, "null ray/hit arrays");
        if (!s->built) throw Error(LUX_ERR_INVALID_ARGUMENT, "scene has not been built");
        if (n == 0) return int(LUX_OK);
        sync_two_level(s);
        DeviceBuffer<lux_ray> dr(n);
        DeviceBuffer<int32_t> di(ignore ? size_t(n) : 1);
        DeviceBuffer<lux_scene_hit> dh(n);
        hip_check(hipMemcpyAsync(dr.p, rays, sizeof(lux_ray) * n, hipMemcpyHostToDevice, g_stream), "H2D rays");
        if (ignore)
            hip_check(hipMemcpyAsync(di.p, ignore, sizeof(int32_t) * n, hipMemcpyHostToDevice, g_stream),
                      "H2D ignore");
        s->last_stream = g_stream;
        two_level_closest(two_level_view(s->dev), two_level_order(s->order, s->dev, dr.p, n, g_stream), dr.p,
                          ignore ? di.p : nullptr, n, dh.p, g_stream);
        hip_check(hipMemcpyAsync(hits, dh.p, sizeof(lux_scene_hit) * n, hipMemcpyDeviceToHost, g_stream),
                  "D2H hits");
        sync_two_level(s);
        return int(LUX_OK);
    });
}

int lux_hip_two_level_intersect_device(lux_hip_two_level* s, const lux_ray* d_rays, const int32_t* d_ignore,
                                       int32_t n, lux_scene_hit* d_hits, lux_stream stream) {
    return guarded([&] {
        require_init();
        require_two_level(s);
        if (n < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative ray count");
        if (n > 0 && (!d_rays || !d_hits)) throw Error(LUX_ERR_INVALID_ARGUMENT, "null ray/hit arrays");
        if (!s->built) throw Error(LUX_ERR_INVALID_ARGUMENT, "scene has not been built");
        const hipStream_t st = use_stream(s, stream);
        two_level_closest(two_level_view(s->dev), two_level_order(s->order, s->dev, d_rays, n, st), d_rays, d_ignore,
                          n, d_hits, st);
        return int(LUX_OK);
    });
}

int lux_hip_two_level_occluded(lux_hip_two_level* s, const lux_ray* rays, const int32_t* ignore, int32_t n,
                               uint8_t* occluded) {
    return guarded([&] {
        require_init();
        require_two_level(s);
        if (n < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative ray count");
        if (n > 0 && (!rays || !occluded)) throw Error(LUX_ERR_INVALID_ARGUMENT, "null ray/output arrays");
        if (!s->built) throw Error(LUX_ERR_INVALID_ARGUMENT, "scene has not been built");
        if (n == 0) return int(LUX_OK);
        sync_two_level(s);
        DeviceBuffer<lux_ray> dr(n);
        DeviceBuffer<int32_t> di(ignore ? size_t(n) : 1);
        DeviceBuffer<uint8_t> dh(n);
        hip_check(hipMemcpyAsync(dr.p, rays, sizeof(lux_ray) * n, hipMemcpyHostToDevice, g_stream), "H2D rays");
        if (ignore)
            hip_check(hipMemcpyAsync(di.p, ignore, sizeof(int32_t) * n, hipMemcpyHostToDevice, g_stream),
                      "H2D ignore");
        s->last_stream = g_stream;
        two_level_occluded(two_level_view(s->dev), two_level_order(s->order, s->dev, dr.p, n, g_stream), dr.p,
                           ignore ? di.p : nullptr, n, dh.p, g_stream);
        hip_check(hipMemcpyAsync(occluded, dh.p, size_t(n), hipMemcpyDeviceToHost, g_stream), "D2H occluded");
        sync_two_level(s);
        return int(LUX_OK);
    });
}

int lux_hip_two_level_occluded_device(lux_hip_two_level* s, const lux_ray* d_rays, const int32_t* d_ignore,
                                      int32_t n, uint8_t* d_occluded, lux_stream stream) {
    return guarded([&] {
        require_init();
        require_two_level(s);
        if (n < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative ray count");
        if (n > 0 && (!d_rays || !d_occluded)) throw Error(LUX_ERR_INVALID_ARGUMENT, "null ray/output arrays");
        if (!s->built) throw Error(LUX_ERR_INVALID_ARGUMENT, "scene has not been built");
        const hipStream_t st = use_stream(s, stream);
        two_level_occluded(two_level_view(s->dev), two_level_order(s->order, s->dev, d_rays, n, st), d_rays,
                           d_ignore, n, d_occluded, st);
        return int(LUX_OK);
    });
}

int lux_hip_two_level_info(lux_hip_two_level* s, int32_t* tlas_instances, int64_t* scene_primitives,
                           uint64_t* device_bytes) {
    return guarded([&] {
        require_two_level(s);
        if (tlas_instances) *tlas_instances = s->dev.tlas_instances;
        if (scene_primitives) *scene_primitives = s->dev.scene_primitives;
        if (device_bytes) *device_bytes = uint64_t(s->dev.device_bytes);
        return int(LUX_OK);
    });
}

int lux_hip_inverse_affine(const float* m16, float* inv16) {
    return guarded([&] {
        if (!m16 || !inv16) throw Error(LUX_ERR_INVALID_ARGUMENT, "null matrix");
        inverse_affine(m16, inv16);
        return int(LUX_OK);
    });
}

int lux_hip_transform_bounds(const float* m16, const lux_aabb* bounds, lux_aabb* out) {
    return guarded([&] {
        if (!m16 || !bounds || !out) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        *out = transform_bounds(m16, *bounds);
        return int(LUX_OK);
    });
}

// ---- film output ----------------------------------------------------------------
namespace lux {
namespace {
void check_film_args(const float* d_film, int32_t w, int32_t h) {
    if (w < 0 || h < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative film size");
    if (int64_t(w) * h > 0 && !d_film) throw Error(LUX_ERR_INVALID_ARGUMENT, "null film");
}
}  // namespace
}  // namespace lux

int lux_hip_film_ppm_device(const float* d_film, int32_t w, int32_t h, int32_t spp, uint8_t* d_rgb8,
                            lux_stream stream) {
    return guarded([&] {
        require_init();
        check_film_args(d_film, w, h);
        if (int64_t(w) * h > 0 && !d_rgb8) throw Error(LUX_ERR_INVALID_ARGUMENT, "null output");
        film_ppm(d_film, w, h, spp, d_rgb8, pick(stream));
        return int(LUX_OK);
    });
}

int lux_hip_film_pfm_device(const float* d_film, int32_t w, int32_t h, int32_t spp, float* d_rgb, lux_stream stream) {
    return guarded([&] {
        require_init();
        check_film_args(d_film, w, h);
        if (int64_t(w) * h > 0 && !d_rgb) throw Error(LUX_ERR_INVALID_ARGUMENT, "null output");
        film_pfm(d_film, w, h, spp, d_rgb, pick(stream));
        return int(LUX_OK);
    });
}

int lux_hip_write_ppm(const char* path, const float* d_film, int32_t w, int32_t h, int32_t spp) {
    return guarded([&] {
        require_init();
        check_film_args(d_film, w, h);
        if (!path) throw Error(LUX_ERR_INVALID_ARGUMENT, "null path");
        write_film_file(path, true, d_film, w, h, spp, g_stream);
        return int(LUX_OK);
    });
}

int lux_hip_write_pfm(const char* path, const float* d_film, int32_t w, int32_t h, int32_t spp) {
    return guarded([&] {
        require_init();
        check_film_args(d_film, w, h);
        if (!path) throw Error(LUX_ERR_INVALID_ARGUMENT, "null path");
        write_film_file(path, false, d_film, w, h, spp, g_stream);
        return int(LUX_OK);
    });
}

int lux_hip_ppm_thresholds(float* out255) {
    return guarded([&] {
        if (!out255) throw Error(LUX_ERR_INVALID_ARGUMENT, "null output");
        ppm_thresholds(out255);
        return int(LUX_OK);
    });
}

int lux_hip_probe_begin(const char* kernel_name) {
    return guarded([&] {
        require_init();
        if (!kernel_name) throw Error(LUX_ERR_INVALID_ARGUMENT, "null kernel name");
        std::lock_guard<std::mutex> lock(g_probe.mu);
        for (auto& pr : g_probe.events) {
            g_probe.pool.push_back(pr.first);
            g_probe.pool.push_back(pr.second);
        }
        g_probe.events.clear();
        g_probe.name = kernel_name;
        g_probe.open = true;
        return int(LUX_OK);
    });
}

int lux_hip_probe_end(double* total_ms, int64_t* launches) {
    return guarded([&] {
        require_init();
        std::lock_guard<std::mutex> lock(g_probe.mu);
        g_probe.open = false;
        double sum = 0;
        for (auto& pr : g_probe.events) {
            hip_check(hipEventSynchronize(pr.second), "hipEventSynchronize");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, pr.first, pr.second), "hipEventElapsedTime");
            sum += ms;
        }
        if (total_ms) *total_ms = sum;
        if (launches) *launches = int64_t(g_probe.events.size());
        return int(LUX_OK);
    });
}

}  // extern "C"

// ---- wavefront path tracer ------------------------------------------------------
#include <memory>

#include "pt_runtime.hpp"
#include "serve.hpp"

struct lux_hip_server {
    std::unique_ptr<lux::pt::Server> server;
    std::vector<uint64_t> ready;  // completed ids not yet returned by poll
};

struct lux_obj_mesh {
    lux::obj::Mesh mesh;
};

struct lux_hip_scene {
    lux::pt::SceneDev dev;
    std::unique_ptr<lux::pt::Renderer> renderer;  // path buffers of this scene's single renders
    lux_hip_bvh bvh_alias;  // borrowed view of dev.bvh for the traversal entry points
};

namespace {
lux::pt::Tuning g_tuning;
std::unique_ptr<lux::pt::Renderer> g_batch;  // engine of the multi-scene batch entry points

void check_settings(const lux_render_settings* s) {
    if (!s) throw lux::Error(LUX_ERR_INVALID_ARGUMENT, "null render settings");
    if (s->width <= 0 || s->height <= 0 || s->spp <= 0)
        throw lux::Error(LUX_ERR_INVALID_ARGUMENT, "WavefrontPathTracer requires a non-empty Film");
    if (s->max_depth < 0) throw lux::Error(LUX_ERR_INVALID_ARGUMENT, "WavefrontPathTracer requires max_depth >= 0");
    if (int64_t(s->width) * s->height >= (int64_t(1) << 31))
        throw lux::Error(LUX_ERR_INVALID_ARGUMENT, "film too large");
    if (int64_t(s->spp) + s->sample_offset >= (int64_t(1) << 31) || s->sample_offset < 0)
        throw lux::Error(LUX_ERR_INVALID_ARGUMENT, "sample range out of bounds");
    if (s->mode != LUX_RENDER_PLAIN && s->mode != LUX_RENDER_WAVEFRONT)
        throw lux::Error(LUX_ERR_INVALID_ARGUMENT, "unknown render mode");
    if (s->batch_size < 0) throw lux::Error(LUX_ERR_INVALID_ARGUMENT, "negative batch size");
    if (s->light_sampler < LUX_LIGHT_SAMPLER_POWER || s->light_sampler > LUX_LIGHT_SAMPLER_BVH)
        throw lux::Error(LUX_ERR_INVALID_ARGUMENT, "unknown light sampler");
}

// Reads and clears the device check bits of a scene (caller synchronised):
// the LBVH's, or the TLAS's for an instanced scene (the two-level walk
// reports there).
void scene_check(lux_hip_scene* sc) {
    lux::DevInfo* d = sc->dev.instanced ? sc->dev.tl.tlas.info : sc->dev.bvh.info;
    if (!d) return;
    lux::DevInfo info;
    lux::hip_check(hipMemcpy(&info, d, sizeof(info), hipMemcpyDeviceToHost), "read device info");
    if (info.error) {
        uint32_t zero = 0;
        lux::hip_check(hipMemcpy(reinterpret_cast<char*>(d) + offsetof(lux::DevInfo, error), &zero, 4,
                                 hipMemcpyHostToDevice),
                       "clear device info");
        throw lux::Error(LUX_ERR_DEVICE_CHECK, "device check failed during render");
    }
}

// A scene whose device state was dropped (or never built: created deferred)
// is rebuilt from its kept description before any call that reads it.
void ensure_resident(lux_hip_scene* sc) {
    if (sc->dev.resident) return;
    if (!sc->dev.source) throw lux::Error(LUX_ERR_INVALID_ARGUMENT, "scene has no device state");
    if (sc->dev.server_jobs > 0)
        throw lux::Error(LUX_ERR_INVALID_ARGUMENT, "scene is being prepared for server requests: wait for them");
    lux::pt::SceneDev fresh;
    try {
        lux::pt::scene_restore(fresh, *sc->dev.source, g_stream);
    } catch (...) {
        lux::pt::scene_release(fresh);
        throw;
    }
    lux::pt::scene_adopt(sc->dev, std::move(fresh));
}

lux::pt::Renderer& scene_renderer(lux_hip_scene* sc) {
    if (!sc->renderer) sc->renderer = std::make_unique<lux::pt::Renderer>();
    sc->renderer->set_tuning(g_tuning);
    return *sc->renderer;
}

lux::pt::Renderer& batch_renderer() {
    if (!g_batch) g_batch = std::make_unique<lux::pt::Renderer>();
    g_batch->set_tuning(g_tuning);
    return *g_batch;
}

void render_one(lux_hip_scene* sc, const lux_render_settings* s, float* d_film, hipStream_t stream) {
    ensure_resident(sc);
    lux::pt::Renderer& r = scene_renderer(sc);
    lux::pt::Request rq;
    rq.sv = lux::pt::scene_view(sc->dev);
    rq.rs = *s;
    rq.film = d_film;
    std::vector<lux::pt::Request*> reqs{&rq};
    r.reset_ray_counts(stream);
    r.render_all(reqs, s->mode, stream, nullptr);
}

void check_batch_args(lux_hip_scene* const* scenes, const lux_render_settings* settings, int32_t n, int mode) {
    if (n < 0 || (n > 0 && (!scenes || !settings))) throw lux::Error(LUX_ERR_INVALID_ARGUMENT, "bad batch arguments");
    if (mode != LUX_RENDER_PLAIN && mode != LUX_RENDER_WAVEFRONT)
        throw lux::Error(LUX_ERR_INVALID_ARGUMENT, "unknown render mode");
    for (int32_t i = 0; i < n; ++i) {
        if (!scenes[i]) throw lux::Error(LUX_ERR_INVALID_ARGUMENT, "null scene in batch");
        check_settings(&settings[i]);
    }
}

void render_batch(lux_hip_scene* const* scenes, const lux_render_settings* settings, float* const* d_films,
                  int32_t n, int mode, hipStream_t stream) {
    for (int32_t i = 0; i < n; ++i) ensure_resident(scenes[i]);
    lux::pt::Renderer& r = batch_renderer();
    std::vector<lux::pt::Request> reqs(n);
    std::vector<lux::pt::Request*> ptrs(n);
    for (int32_t i = 0; i < n; ++i) {
        reqs[i].sv = lux::pt::scene_view(scenes[i]->dev);
        reqs[i].rs = settings[i];
        reqs[i].film = d_films[i];
        reqs[i].tag = uint64_t(i);
        ptrs[i] = &reqs[i];
    }
    r.reset_ray_counts(stream);
    r.render_all(ptrs, mode, stream, nullptr);
}
}  // namespace

static void release_render_engines() { g_batch.reset(); }

extern "C" {

int lux_hip_scene_create(const lux_scene_desc* desc, lux_hip_scene** out) {
    return guarded([&] {
        require_init();
        if (!desc || !out) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        auto* sc = new lux_hip_scene();
        try {
            pt::scene_create(sc->dev, *desc, true, g_stream);
        } catch (...) {
            pt::scene_release(sc->dev);
            delete sc;
            throw;
        }
        *out = sc;
        return int(LUX_OK);
    });
}

int lux_hip_scene_create_deferred(const lux_scene_desc* desc, lux_hip_scene** out) {
    return guarded([&] {
        require_init();
        if (!desc || !out) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        auto* sc = new lux_hip_scene();
        try {
            pt::scene_create(sc->dev, *desc, false, g_stream);
        } catch (...) {
            pt::scene_release(sc->dev);
            delete sc;
            throw;
        }
        *out = sc;
        return int(LUX_OK);
    });
}

int lux_hip_scene_resident(lux_hip_scene* sc, int32_t* resident) {
    return guarded([&] {
        if (!sc || !resident) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        *resident = sc->dev.resident ? 1 : 0;
        return int(LUX_OK);
    });
}

int lux_hip_scene_create_instanced(const lux_instanced_scene_desc* desc, lux_hip_scene** out) {
    return guarded([&] {
        require_init();
        if (!desc || !out) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        auto* sc = new lux_hip_scene();
        try {
            pt::scene_upload_instanced(sc->dev, *desc, g_stream);
        } catch (...) {
            pt::scene_release(sc->dev);
            delete sc;
            throw;
        }
        *out = sc;
        return int(LUX_OK);
    });
}

int lux_hip_scene_set_instances(lux_hip_scene* sc, const lux_instance* instances, int32_t instance_count) {
    return guarded([&] {
        require_init();
        if (!sc) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        pt::scene_set_instances(sc->dev, instances, instance_count, g_stream);
        scene_check(sc);
        return int(LUX_OK);
    });
}

int lux_hip_scene_triangles(lux_hip_scene* sc, float* corners, float* normals, int32_t* material_ids) {
    return guarded([&] {
        require_init();
        if (!sc) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        ensure_resident(sc);
        const pt::SceneDev& d = sc->dev;
        const size_t T = size_t(std::max(d.tri_count, 0));
        hip_check(hipStreamSynchronize(g_stream), "synchronize");
        if (corners && T)
            hip_check(hipMemcpy(corners, d.soup, sizeof(lux_triangle) * T, hipMemcpyDeviceToHost), "D2H triangles");
        if (material_ids && T)
            hip_check(hipMemcpy(material_ids, d.mat_id, 4 * T, hipMemcpyDeviceToHost), "D2H materials");
        if (normals && T) {
            if (!d.normals) {
                std::memset(normals, 0, sizeof(float) * 9 * T);
            } else if (d.instanced) {  // already per corner
                hip_check(hipMemcpy(normals, d.normals, sizeof(float) * 9 * T, hipMemcpyDeviceToHost), "D2H normals");
            } else {  // per vertex: gather the corners on the host
                std::vector<float> vn(3 * size_t(d.vertex_count));
                std::vector<uint32_t> ix(3 * T);
                hip_check(hipMemcpy(vn.data(), d.normals, 4 * vn.size(), hipMemcpyDeviceToHost), "D2H normals");
                hip_check(hipMemcpy(ix.data(), d.indices, 4 * ix.size(), hipMemcpyDeviceToHost), "D2H indices");
                for (size_t c = 0; c < 3 * T; ++c)
                    for (int k = 0; k < 3; ++k) normals[3 * c + k] = vn[3 * size_t(ix[c]) + k];
            }
        }
        return int(LUX_OK);
    });
}

int lux_obj_default_options(lux_obj_options* o) {
    return guarded([&] {
        if (!o) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        lux::obj::default_options(*o);
        return int(LUX_OK);
    });
}

int lux_obj_load(const char* path, const lux_obj_options* o, lux_obj_mesh** out) {
    return guarded([&] {
        if (!path || !o || !out) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        auto m = std::make_unique<lux_obj_mesh>();
        lux::obj::load(path, *o, m->mesh);
        *out = m.release();
        return int(LUX_OK);
    });
}

int lux_obj_mesh_free(lux_obj_mesh* m) {
    delete m;
    return int(LUX_OK);
}

int lux_obj_mesh_counts(const lux_obj_mesh* m, int32_t* vertex_count, int32_t* triangle_count,
                        int32_t* appended_default_material, int32_t* skipped_invalid, int32_t* skipped_nonfinite,
                        int32_t* skipped_degenerate) {
    return guarded([&] {
        if (!m) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        const lux::obj::Mesh& x = m->mesh;
        if (vertex_count) *vertex_count = int32_t(x.vertices.size() / 3);
        if (triangle_count) *triangle_count = int32_t(x.indices.size() / 3);
        if (appended_default_material) *appended_default_material = x.appended_default_material;
        if (skipped_invalid) *skipped_invalid = x.skipped_invalid;
        if (skipped_nonfinite) *skipped_nonfinite = x.skipped_nonfinite;
        if (skipped_degenerate) *skipped_degenerate = x.skipped_degenerate;
        return int(LUX_OK);
    });
}

int lux_obj_mesh_cap_counts(const lux_obj_mesh* m, int32_t* cap_triangles, int32_t* boundary_edges,
                            int32_t* non_simple_vertices, int32_t* incomplete_loops, int32_t* degenerate_loops,
                            int32_t* degenerate_triangles) {
    return guarded([&] {
        if (!m) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        const lux::obj::Mesh& x = m->mesh;
        if (cap_triangles) *cap_triangles = x.cap_triangles;
        if (boundary_edges) *boundary_edges = x.cap_boundary_edges;
        if (non_simple_vertices) *non_simple_vertices = x.cap_non_simple_vertices;
        if (incomplete_loops) *incomplete_loops = x.cap_incomplete_loops;
        if (degenerate_loops) *degenerate_loops = x.cap_degenerate_loops;
        if (degenerate_triangles) *degenerate_triangles = x.cap_degenerate_triangles;
        return int(LUX_OK);
    });
}

int lux_obj_mesh_arrays(const lux_obj_mesh* m, lux_vec3* vertices, lux_vec3* normals, float* uvs, uint32_t* indices,
                        int32_t* material_ids, uint8_t* boundary_edges, float* object_to_world,
                        float* world_to_object) {
    return guarded([&] {
        if (!m) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        const lux::obj::Mesh& x = m->mesh;
        auto put = [](void* dst, const auto& v) {
            if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
        };
        put(vertices, x.vertices);
        put(normals, x.normals);
        put(uvs, x.uvs);
        put(indices, x.indices);
        put(material_ids, x.material_ids);
        put(boundary_edges, x.boundary_edges);
        if (object_to_world) std::memcpy(object_to_world, x.object_to_world, sizeof(x.object_to_world));
        if (world_to_object) std::memcpy(world_to_object, x.world_to_object, sizeof(x.world_to_object));
        return int(LUX_OK);
    });
}

int lux_hip_scene_destroy(lux_hip_scene* sc) {
    return guarded([&] {
        if (!sc) return int(LUX_OK);
        if (sc->dev.server_jobs > 0)
            throw Error(LUX_ERR_INVALID_ARGUMENT, "scene has unfinished server requests: destroy it after they complete");
        if (g_initialized) (void)hipDeviceSynchronize();
        sc->renderer.reset();
        pt::scene_release(sc->dev);
        sc->bvh_alias.dev = BvhDev{};
        raysort_release(sc->bvh_alias.order);
        delete sc;
        return int(LUX_OK);
    });
}

int lux_hip_scene_update_vertices(lux_hip_scene* sc, const lux_vec3* vertices, int32_t vertex_count) {
    return guarded([&] {
        require_init();
        if (!sc || !vertices) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        if (sc->dev.instanced)
            throw Error(LUX_ERR_INVALID_ARGUMENT, "instanced scene: move instances with lux_hip_scene_set_instances");
        if (sc->dev.server_jobs > 0)
            throw Error(LUX_ERR_INVALID_ARGUMENT,
                        "scene has unfinished server requests: submit the update with lux_hip_server_submit");
        if (!sc->dev.resident) {  // kept for the rebuild: the refit happens there
            if (vertex_count != sc->dev.vertex_count)
                throw Error(LUX_ERR_MISMATCH, "vertex update must keep the vertex count");
            sc->dev.source->current.assign(vertices, vertices + vertex_count);
            return int(LUX_OK);
        }
        pt::scene_update_vertices(sc->dev, vertices, vertex_count, g_stream);
        scene_check(sc);
        return int(LUX_OK);
    });
}

int lux_hip_scene_info(lux_hip_scene* sc, int32_t* triangle_count, int32_t* vertex_count, uint64_t* device_bytes) {
    return guarded([&] {
        if (!sc) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        if (triangle_count) *triangle_count = sc->dev.tri_count;
        if (vertex_count) *vertex_count = sc->dev.vertex_count;
        if (device_bytes) *device_bytes = sc->dev.device_bytes;
        return int(LUX_OK);
    });
}

int lux_hip_scene_camera(lux_hip_scene* sc, float* out) {
    return guarded([&] {
        if (!sc || !out) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        std::memcpy(out, &sc->dev.camera, sizeof(float) * 12);
        return int(LUX_OK);
    });
}

int lux_hip_scene_set_camera(lux_hip_scene* sc, const float* v12) {
    return guarded([&] {
        if (!sc || !v12) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        for (int k = 0; k < 12; ++k)
            if (!std::isfinite(v12[k])) throw Error(LUX_ERR_INVALID_ARGUMENT, "camera vectors must be finite");
        std::memcpy(&sc->dev.camera, v12, sizeof(float) * 12);
        return int(LUX_OK);
    });
}

int lux_hip_scene_lights(lux_hip_scene* sc, int32_t* count, int32_t* prim, float* area, float* prob, int32_t* alias,
                         float* pmf, int32_t* light_of_prim) {
    return guarded([&] {
        if (!sc || !count) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        ensure_resident(sc);
        const auto& L = sc->dev.host_lights;
        *count = int32_t(L.size());
        for (size_t i = 0; i < L.size(); ++i) {
            const bool is_area = L[i].kind == lux::pt::kLightArea;  // scene lights: no triangle, no area
            if (prim) prim[i] = is_area ? L[i].prim : -1;
            if (area) area[i] = is_area ? L[i].area : 0.0f;
            if (prob) prob[i] = L[i].prob;
            if (alias) alias[i] = L[i].alias;
            if (pmf) pmf[i] = L[i].pmf;
        }
        if (light_of_prim)
            std::memcpy(light_of_prim, sc->dev.host_light_of_prim.data(), 4 * sc->dev.host_light_of_prim.size());
        return int(LUX_OK);
    });
}

int lux_hip_scene_light_bvh(lux_hip_scene* sc, int32_t* node_count, float* bounds6, float* power, int32_t* links4,
                            int32_t* leaf_of_light, int32_t* root) {
    return guarded([&] {
        if (!sc || !node_count) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        ensure_resident(sc);
        const pt::SceneDev& d = sc->dev;
        *node_count = int32_t(d.host_light_nodes.size());
        if (root) *root = d.light_node_count > 0 ? d.light_root : -1;
        for (size_t i = 0; i < d.host_light_nodes.size(); ++i) {
            const lux::pt::DevLightNode& n = d.host_light_nodes[i];
            if (bounds6)
                for (int a = 0; a < 3; ++a) {
                    bounds6[6 * i + a] = n.lo[a];
                    bounds6[6 * i + 3 + a] = n.hi[a];
                }
            if (power) power[i] = n.power;
            if (links4) {
                links4[4 * i] = n.left;
                links4[4 * i + 1] = n.right;
                links4[4 * i + 2] = n.parent;
                links4[4 * i + 3] = n.light;
            }
        }
        if (leaf_of_light)
            for (size_t i = 0; i < d.host_light_leaf.size(); ++i) leaf_of_light[i] = d.host_light_leaf[i];
        return int(LUX_OK);
    });
}

int lux_hip_scene_bvh(lux_hip_scene* sc, lux_hip_bvh** out) {
    return guarded([&] {
        if (!sc || !out) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        if (sc->dev.instanced) throw Error(LUX_ERR_INVALID_ARGUMENT, "instanced scene: no single-level LBVH");
        ensure_resident(sc);
        // The alias shares the scene's device buffers; it must not be destroyed
        // (and holds them only while the scene is resident: a server may drop it).
        sc->bvh_alias.dev = sc->dev.bvh;
        sc->bvh_alias.valid = sc->dev.bvh.n > 0;
        sc->bvh_alias.host_ids_known = false;
        sc->bvh_alias.last_stream = g_stream;
        *out = &sc->bvh_alias;
        return int(LUX_OK);
    });
}

int lux_hip_render(lux_hip_scene* sc, const lux_render_settings* s, float* film) {
    return guarded([&] {
        require_init();
        if (!sc || !film) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        check_settings(s);
        size_t bytes = sizeof(float) * 3 * size_t(s->width) * s->height;
        DeviceBuffer<float> d_film{3 * size_t(s->width) * s->height};
        render_one(sc, s, d_film.p, g_stream);
        hip_check(hipMemcpyAsync(film, d_film.p, bytes, hipMemcpyDeviceToHost, g_stream), "D2H film");
        hip_check(hipStreamSynchronize(g_stream), "synchronize");
        scene_check(sc);
        return int(LUX_OK);
    });
}

int lux_hip_render_device(lux_hip_scene* sc, const lux_render_settings* s, float* d_film, lux_stream stream) {
    return guarded([&] {
        require_init();
        if (!sc || !d_film) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        check_settings(s);
        render_one(sc, s, d_film, pick(stream));
        return int(LUX_OK);
    });
}

int lux_hip_render_rays(lux_hip_scene* sc, lux_stream stream, uint64_t* extension_rays, uint64_t* shadow_rays) {
    return guarded([&] {
        require_init();
        if (!sc || !extension_rays || !shadow_rays) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        if (!sc->renderer) {
            *extension_rays = *shadow_rays = 0;
            return int(LUX_OK);
        }
        sc->renderer->ray_counts(pick(stream), extension_rays, shadow_rays);
        return int(LUX_OK);
    });
}

int lux_hip_batch_render(lux_hip_scene* const* scenes, const lux_render_settings* settings, float* const* films,
                         int32_t n, int32_t mode) {
    return guarded([&] {
        require_init();
        check_batch_args(scenes, settings, n, mode);
        if (!films && n > 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "null films");
        std::vector<std::unique_ptr<DeviceBuffer<float>>> bufs;
        std::vector<float*> d_films(n);
        for (int32_t i = 0; i < n; ++i) {
            if (!films[i]) throw Error(LUX_ERR_INVALID_ARGUMENT, "null film");
            bufs.push_back(std::make_unique<DeviceBuffer<float>>(3 * size_t(settings[i].width) * settings[i].height));
            d_films[i] = bufs.back()->p;
        }
        render_batch(scenes, settings, d_films.data(), n, mode, g_stream);
        for (int32_t i = 0; i < n; ++i)
            hip_check(hipMemcpyAsync(films[i], d_films[i], sizeof(float) * 3 * size_t(settings[i].width) *
                                                               settings[i].height,
                                     hipMemcpyDeviceToHost, g_stream),
                      "D2H film");
        hip_check(hipStreamSynchronize(g_stream), "synchronize");
        for (int32_t i = 0; i < n; ++i) scene_check(scenes[i]);
        return int(LUX_OK);
    });
}

int lux_hip_batch_render_device(lux_hip_scene* const* scenes, const lux_render_settings* settings,
                                float* const* d_films, int32_t n, int32_t mode, lux_stream stream) {
    return guarded([&] {
        require_init();
        check_batch_args(scenes, settings, n, mode);
        if (!d_films && n > 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "null films");
        for (int32_t i = 0; i < n; ++i)
            if (!d_films[i]) throw Error(LUX_ERR_INVALID_ARGUMENT, "null film");
        render_batch(scenes, settings, d_films, n, mode, pick(stream));
        return int(LUX_OK);
    });
}

int lux_hip_batch_rays(lux_stream stream, uint64_t* extension_rays, uint64_t* shadow_rays) {
    return guarded([&] {
        require_init();
        if (!extension_rays || !shadow_rays) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        batch_renderer().ray_counts(pick(stream), extension_rays, shadow_rays);
        return int(LUX_OK);
    });
}

int lux_hip_server_create(const lux_serve_config* cfg, lux_hip_server** out) {
    return guarded([&] {
        require_init();
        if (!cfg || !out) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        pt::ServeConfig c;
        c.hbm_budget = cfg->hbm_budget_bytes;
        c.slice_spp = cfg->slice_spp;
        c.max_inflight = cfg->max_inflight;
        auto* sv = new lux_hip_server();
        try {
            sv->server = std::make_unique<pt::Server>(c, g_tuning, g_stream);
        } catch (...) {
            delete sv;
            throw;
        }
        *out = sv;
        return int(LUX_OK);
    });
}

int lux_hip_server_destroy(lux_hip_server* sv) {
    return guarded([&] {
        delete sv;
        return int(LUX_OK);
    });
}

int lux_hip_server_submit(lux_hip_server* sv, lux_hip_scene* scene, const lux_render_settings* settings,
                          const lux_vec3* new_vertices, int32_t vertex_count, uint64_t* request_id) {
    return guarded([&] {
        require_init();
        if (!sv || !scene || !request_id) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        check_settings(settings);
        *request_id = sv->server->submit(&scene->dev, *settings, new_vertices, vertex_count, sv->server->now());
        return int(LUX_OK);
    });
}

int lux_hip_server_poll(lux_hip_server* sv, int32_t block, uint64_t* done_ids, int32_t max_ids, int32_t* n_done) {
    return guarded([&] {
        require_init();
        if (!sv || !n_done || max_ids < 0 || (max_ids > 0 && !done_ids))
            throw Error(LUX_ERR_INVALID_ARGUMENT, "bad poll arguments");
        sv->server->progress(block != 0, &sv->ready);
        int32_t n = std::min<int32_t>(max_ids, int32_t(sv->ready.size()));
        for (int32_t i = 0; i < n; ++i) done_ids[i] = sv->ready[i];
        sv->ready.erase(sv->ready.begin(), sv->ready.begin() + n);
        *n_done = n;
        return int(LUX_OK);
    });
}

int lux_hip_server_result(lux_hip_server* sv, uint64_t request_id, float* film_sum, double* latency_s) {
    return guarded([&] {
        require_init();
        if (!sv) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        double a = 0, ad = 0, f = 0;
        sv->server->take_result(request_id, film_sum, &a, &ad, &f);
        if (latency_s) *latency_s = f - a;
        return int(LUX_OK);
    });
}

int lux_hip_server_stats(lux_hip_server* sv, lux_serve_stats* out) {
    return guarded([&] {
        require_init();
        if (!sv || !out) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        pt::ServeStats st = sv->server->stats();
        out->submitted = st.submitted;
        out->admitted = st.admitted;
        out->completed = st.completed;
        out->refits = st.refits;
        out->batches = st.batches;
        out->paths = st.paths;
        out->resident_bytes = st.resident_bytes;
        out->peak_bytes = st.peak_bytes;
        out->shared_batches = st.shared_batches;
        out->max_batch_requests = st.max_batch_requests;
        out->uploads = st.uploads;
        out->evictions = st.evictions;
        out->prepare_seconds = st.prepare_seconds;
        sv->server->ray_counts(&out->extension_rays, &out->shadow_rays);
        return int(LUX_OK);
    });
}

int lux_hip_server_run(lux_hip_server* sv, int32_t n, lux_hip_scene* const* scenes,
                       const lux_render_settings* settings, const double* arrival_s,
                       const lux_vec3* const* new_vertices, const int32_t* vertex_counts, float* const* films,
                       double* latency_s, double* wall_s) {
    return guarded([&] {
        require_init();
        if (!sv || n < 0 || (n > 0 && (!scenes || !settings || !arrival_s || !latency_s)))
            throw Error(LUX_ERR_INVALID_ARGUMENT, "bad trace arguments");
        if (new_vertices && !vertex_counts) throw Error(LUX_ERR_INVALID_ARGUMENT, "vertex counts missing");
        std::vector<pt::SceneDev*> devs(n);
        std::vector<int> counts(n, 0);
        for (int32_t i = 0; i < n; ++i) {
            if (!scenes[i]) throw Error(LUX_ERR_INVALID_ARGUMENT, "null scene in trace");
            check_settings(&settings[i]);
            devs[i] = &scenes[i]->dev;
            if (new_vertices) counts[i] = vertex_counts[i];
        }
        pt::run_trace(*sv->server, n, devs.data(), settings, arrival_s, new_vertices, counts.data(), films, latency_s,
                      wall_s);
        return int(LUX_OK);
    });
}

int lux_hip_set_render_tuning(const lux_render_tuning* t) {
    return guarded([&] {
        if (!t) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        if (t->capacity <= 0 || t->capacity > (1 << 28) || t->grid <= 0 || t->grid > (1 << 20) ||
            (t->block != 128 && t->block != 256 && t->block != 512))
            throw Error(LUX_ERR_INVALID_ARGUMENT, "tuning out of range");
        g_tuning.capacity = t->capacity;
        g_tuning.grid = t->grid;
        g_tuning.block = t->block;
        g_tuning.use_graph = t->use_graph ? 1 : 0;
        g_tuning.preorder_walk = t->preorder_walk ? 1 : 0;
        return int(LUX_OK);
    });
}

int lux_hip_get_render_tuning(lux_render_tuning* t) {
    return guarded([&] {
        if (!t) throw Error(LUX_ERR_INVALID_ARGUMENT, "null argument");
        t->capacity = g_tuning.capacity;
        t->grid = g_tuning.grid;
        t->block = g_tuning.block;
        t->use_graph = g_tuning.use_graph;
        t->preorder_walk = g_tuning.preorder_walk;
        return int(LUX_OK);
    });
}

}  // extern "C"
