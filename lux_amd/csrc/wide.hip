// wide.hip -- builds and refits the 4-wide traversal tree (see wide.hpp) from the
// right-first binary node array of an LBVH.
//
// The collapse runs top-down, one launch per level: a record takes the two
// children of its binary root (right child k+1, left child escape(k+1)) and
// repeatedly replaces the internal child with the largest box surface by its
// two children, in place (so right-first preorder is kept), until it holds 4
// children or only leaves. Internal children get records allocated with one
// atomic per parent and go to the next level's queue with their resume word.
// A refit copies every record's child boxes again from the refit binary
// nodes, in parallel (each record keeps its children's binary positions).
#include <algorithm>

#include "lux_runtime.hpp"
#include "wide.hpp"

namespace lux {
namespace {

constexpr int kBlock = 256;
// The merged Morton + index key has 62 bits, so a Karras tree is at most 63
// levels deep; every record consumes at least one binary level.
constexpr int kWideLevels = 64;

struct Box6 {
    float lx, ly, lz, hx, hy, hz;
};

__device__ __forceinline__ Box6 node_box(const lux_bvh_node* nodes, uint32_t pos) {
    float4 a, b;
    load_node(nodes, int(pos), a, b);
    return Box6{a.x, a.y, a.z, a.w, b.x, b.y};
}

__device__ __forceinline__ bool node_is_leaf(const lux_bvh_node* nodes, uint32_t pos) {
    return nodes[pos].primitive_id >= 0;
}

__device__ __forceinline__ float half_area(const Box6& b) {
    float dx = b.hx - b.lx, dy = b.hy - b.ly, dz = b.hz - b.lz;
    return dx * dy + dy * dz + dz * dx;
}

// Writes the child boxes, child words and slot meta of record `w` from the
// binary nodes at positions src[0..m) (leaf children store their primitive id).
__device__ void fill_record(const lux_bvh_node* nodes, const uint32_t* src, int m, WideNode& w) {
    uint32_t leaf_mask = 0u;
    for (int i = 0; i < kWide; ++i) {
        const bool live = i < m;
        const Box6 b = live ? node_box(nodes, src[i]) : Box6{INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY,
                                                                 -INFINITY};
        w.lo_x[i] = b.lx;
        w.lo_y[i] = b.ly;
        w.lo_z[i] = b.lz;
        w.hi_x[i] = b.hx;
        w.hi_y[i] = b.hy;
        w.hi_z[i] = b.hz;
        if (live && node_is_leaf(nodes, src[i])) {
            leaf_mask |= 1u << i;
            w.child[i] = uint32_t(nodes[src[i]].primitive_id);
        }
    }
    w.meta = uint32_t(m) | (leaf_mask << 8);
}

__device__ __forceinline__ void load_record(const WideNode* wide, uint32_t k, WideNode& w) {
    const float4* p = reinterpret_cast<const float4*>(wide + k);
    float4* q = reinterpret_cast<float4*>(&w);
#pragma unroll
    for (int i = 0; i < 8; ++i) q[i] = p[i];
}

__device__ __forceinline__ void store_record(WideNode* wide, uint32_t k, const WideNode& w) {
    float4* p = reinterpret_cast<float4*>(wide + k);
    const float4* q = reinterpret_cast<const float4*>(&w);
#pragma unroll
    for (int i = 0; i < 8; ++i) p[i] = q[i];
}

// counts: [0..2] level queue sizes (rotating), [3] records allocated.
__global__ void k_wide_init(const lux_bvh_node* __restrict__ nodes, int n, uint4* q0, uint32_t* counts,
                            WideNode* wide, uint32_t* wsrc, DevInfo* info) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    counts[1] = 0;
    counts[2] = 0;
    info->wide_count = 0;
    if (n <= 0) {
        counts[0] = 0;
        counts[3] = 0;
        return;
    }
    counts[3] = 1;  // record 0 = root
    if (n == 1) {   // a single leaf: one record with one leaf child
        WideNode w{};
        uint32_t src = 0;
        fill_record(nodes, &src, 1, w);
        w.esc = kWideEnd;
        store_record(wide, 0, w);
        wsrc[0] = 0;
        counts[0] = 0;
        return;
    }
    q0[0] = make_uint4(0u, 0u, kWideEnd, 0u);  // binary root 0 -> record 0
    counts[0] = 1;
}

__global__ __launch_bounds__(kBlock) void k_wide_level(const lux_bvh_node* __restrict__ nodes, int node_count,
                                                       const uint4* __restrict__ qin, uint4* __restrict__ qout,
                                                       uint32_t* counts, int level, int last_level,
                                                       WideNode* __restrict__ wide, uint32_t* __restrict__ wsrc,
                                                       uint32_t capacity, DevInfo* info) {
    const int ci = level % 3, co = (level + 1) % 3, cz = (level + 2) % 3;
    // a counter may run past the capacity after a reported overflow; never read past it
    const uint32_t n_in = min(counts[ci], capacity);
    if (blockIdx.x == 0 && threadIdx.x == 0) counts[cz] = 0;  // read two levels ago, appended next level
    for (uint32_t it = blockIdx.x * kBlock + threadIdx.x; it < n_in; it += gridDim.x * kBlock) {
        const uint4 item = qin[it];
        const uint32_t k = item.x, rec = item.y, esc = item.z;
        if (k + 1u >= uint32_t(node_count) || rec >= capacity) {
            set_error(info, kErrWideBuild);
            continue;
        }
        // frontier in right-first preorder (right child x + 1, left child escape(x + 1))
        uint32_t f[kWide];
        int m = 2;
        f[0] = k + 1u;
        f[1] = uint32_t(nodes[k + 1u].escape);
        if (f[1] >= uint32_t(node_count)) {
            set_error(info, kErrWideBuild);
            continue;
        }
        bool bad = false;
        for (int round = 0; round < kWide - 2 && !bad; ++round) {
            int pick = -1;
            float area = -1.0f;
            for (int i = 0; i < m; ++i) {
                if (node_is_leaf(nodes, f[i])) continue;
                float a = half_area(node_box(nodes, f[i]));
                if (a > area) {
                    area = a;
                    pick = i;
                }
            }
            if (pick < 0) break;
            const uint32_t x = f[pick];
            if (x + 1u >= uint32_t(node_count)) {
                bad = true;
                break;
            }
            const uint32_t left = uint32_t(nodes[x + 1u].escape);
            if (left >= uint32_t(node_count)) {
                bad = true;
                break;
            }
            for (int i = m; i > pick + 1; --i) f[i] = f[i - 1];
            f[pick] = x + 1u;
            f[pick + 1] = left;
            ++m;
        }
        if (bad) {
            set_error(info, kErrWideBuild);
            continue;
        }
        uint32_t leaf_mask = 0u, n_int = 0u;
        for (int i = 0; i < m; ++i) {
            if (node_is_leaf(nodes, f[i])) leaf_mask |= 1u << i;
            else ++n_int;
        }
        uint32_t base = 0u;
        if (n_int) base = __hip_atomic_fetch_add(&counts[3], n_int, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (base + n_int > capacity) {
            set_error(info, kErrWideBuild);
            continue;
        }
        WideNode w{};
        fill_record(nodes, f, m, w);
        w.esc = esc;
        uint32_t r = 0u;
        for (int i = 0; i < m; ++i) {
            wsrc[size_t(rec) * kWide + i] = f[i];
            if (!((leaf_mask >> i) & 1u)) w.child[i] = base + r++;
        }
        store_record(wide, rec, w);
        if (n_int) {
            if (level == last_level) {
                set_error(info, kErrWideBuild);  // deeper than any Karras tree can be
                continue;
            }
            const uint32_t slot0 =
                __hip_atomic_fetch_add(&counts[co], n_int, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (slot0 + n_int > capacity) {
                set_error(info, kErrWideBuild);
                continue;
            }
            r = 0u;
            for (int i = 0; i < m; ++i) {
                if ((leaf_mask >> i) & 1u) continue;
                const uint32_t resume = (i + 1 < m) ? ((rec << 2) | uint32_t(i + 1)) : esc;
                qout[slot0 + r] = make_uint4(f[i], base + r, resume, 0u);
                ++r;
            }
        }
    }
}

__global__ void k_wide_finish(const uint32_t* counts, uint32_t capacity, DevInfo* info) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint32_t n = counts[3];
        if (n > capacity) info->error |= kErrWideBuild;
        info->wide_count = (info->error & kErrWideBuild) ? 0u : n;
    }
}

// Refit: every record re-copies its child boxes from the refit binary nodes.
__global__ __launch_bounds__(kBlock) void k_wide_refit(const lux_bvh_node* __restrict__ nodes, int node_count,
                                                       const uint32_t* __restrict__ wsrc,
                                                       WideNode* __restrict__ wide, const uint32_t* counts,
                                                       uint32_t capacity, DevInfo* info) {
    const uint32_t n = info->wide_count <= capacity ? info->wide_count : 0u;  // 0 after a failed build
    (void)counts;
    for (uint32_t rec = blockIdx.x * kBlock + threadIdx.x; rec < n; rec += gridDim.x * kBlock) {
        WideNode w;
        load_record(wide, rec, w);
        const int m = int(w.meta & 0xffu);
        if (m < 1 || m > kWide) {
            set_error(info, kErrWideBuild);
            continue;
        }
        uint32_t f[kWide];
        bool bad = false;
        for (int i = 0; i < m; ++i) {
            f[i] = wsrc[size_t(rec) * kWide + i];
            bad |= f[i] >= uint32_t(node_count);
        }
        if (bad) {
            set_error(info, kErrWideBuild);
            continue;
        }
        const uint32_t children[kWide] = {w.child[0], w.child[1], w.child[2], w.child[3]};
        fill_record(nodes, f, m, w);
        for (int i = 0; i < m; ++i)
            if (!((w.meta >> (8 + i)) & 1u)) w.child[i] = children[i];  // record links are unchanged
        store_record(wide, rec, w);
    }
}

}  // namespace

void wide_build(BvhDev& d, hipStream_t s) {
    const int n = d.n;
    const int node_count = n > 0 ? 2 * n - 1 : 0;
    const int need = std::max(1, d.capacity - 1);
    if (d.wide_capacity < need) {  // first use, or the tree grew
        if (d.wide_capacity) hip_check(hipDeviceSynchronize(), "synchronize before wide tree growth");
        void* old[] = {d.wide, d.wsrc, d.wq[0], d.wq[1], d.wcounts};
        for (void* p : old)
            if (p) (void)hipFree(p);
        auto grab = [](size_t bytes) {
            void* p = nullptr;
            hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 16));
            if (e != hipSuccess)
                throw Error(LUX_ERR_OUT_OF_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
            return p;
        };
        d.wide = static_cast<WideNode*>(grab(sizeof(WideNode) * size_t(need)));
        d.wsrc = static_cast<uint32_t*>(grab(4 * size_t(need) * kWide));
        d.wq[0] = static_cast<uint4*>(grab(sizeof(uint4) * size_t(need)));
        d.wq[1] = static_cast<uint4*>(grab(sizeof(uint4) * size_t(need)));
        d.wcounts = static_cast<uint32_t*>(grab(16));
        d.wide_capacity = need;
    }
    LUX_LAUNCH("wide_init", k_wide_init, 1, 64, 0, s, (const lux_bvh_node*)d.nodes, n, d.wq[0], d.wcounts, d.wide,
               d.wsrc, d.info);
    if (n >= 2) {
        const int grid = std::max(1, std::min(1024, div_up(n, 4 * kBlock)));
        for (int level = 0; level < kWideLevels; ++level)
            LUX_LAUNCH("wide_level", k_wide_level, grid, kBlock, 0, s, (const lux_bvh_node*)d.nodes, node_count,
                       (const uint4*)d.wq[level & 1], d.wq[(level + 1) & 1], d.wcounts, level, kWideLevels - 1,
                       d.wide, d.wsrc, uint32_t(d.wide_capacity), d.info);
    }
    LUX_LAUNCH("wide_finish", k_wide_finish, 1, 64, 0, s, (const uint32_t*)d.wcounts, uint32_t(d.wide_capacity),
               d.info);
}

void wide_refit(BvhDev& d, hipStream_t s) {
    if (d.n <= 0) return;
    const int grid = std::max(1, std::min(1024, div_up(d.n, 4 * kBlock)));
    LUX_LAUNCH("wide_refit", k_wide_refit, grid, kBlock, 0, s, (const lux_bvh_node*)d.nodes, 2 * d.n - 1,
               (const uint32_t*)d.wsrc, d.wide, (const uint32_t*)d.wcounts, uint32_t(d.wide_capacity), d.info);
}

}  // namespace lux
