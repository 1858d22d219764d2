// ph_common.h -- internal definitions shared by the HIP sources of libpartia_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/partia_hip.h"

#define PH_LUT_N_ENTRIES 4096  // entries of each line-integrated kernel table

namespace ph {

using i32 = int32_t;
using i64 = int64_t;
using u32 = uint32_t;
using u64 = uint64_t;

constexpr int kWave = 64;        // CDNA wavefront width
constexpr int kBlock = 256;      // 4 waves per workgroup
constexpr int kNumXcd = 8;       // MI355X: 8 accelerator complex dies, 32 CUs each

// Status with message, propagated to ph_last_error.
struct Error {
    int code;
    std::string msg;
};

#define PH_HIP(expr)                                                                                   \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess) {                                                                        \
            if (_e == hipErrorOutOfMemory || _e == hipErrorMemoryAllocation)                           \
                throw ::ph::Error{PH_ERR_OUT_OF_MEMORY, std::string("out of device memory: ") + #expr}; \
            throw ::ph::Error{PH_ERR_HIP, std::string(hipGetErrorString(_e)) + " at " + #expr};          \
        }                                                                                              \
    } while (0)

// Per-context cache of device blocks: a freed block is parked here and handed
// out again to a request of at most its size (and at least half of it), so
// steady-state calls (build, interpolate, destroy, repeated) make no driver
// allocation.  A block released during a call may still be read by work queued
// on the context's stream, and the host-side copies of a call are not ordered
// with that stream, so released blocks wait in `pending` and become available
// only after the stream has drained (settle).  The cache is emptied when the
// context is destroyed, and its parked blocks before retrying a failed
// allocation.
struct BlockCache {
    std::multimap<size_t, void*> parked;  // block size -> pointer, free to hand out
    std::vector<std::pair<size_t, void*>> pending;  // released, stream not yet drained
    size_t parked_bytes = 0;
    BlockCache() = default;
    BlockCache(const BlockCache&) = delete;
    BlockCache& operator=(const BlockCache&) = delete;
    ~BlockCache() { purge(); }
    void* take(size_t b, size_t& got) {
        auto it = parked.lower_bound(b);
        if (it == parked.end() || it->first > 2 * b) return nullptr;
        void* p = it->second;
        got = it->first;
        parked_bytes -= got;
        parked.erase(it);
        return p;
    }
    void put(void* p, size_t b) { pending.emplace_back(b, p); }
    // The stream that may still use the pending blocks has drained.
    void settle() {
        for (auto& e : pending) {
            parked.emplace(e.first, e.second);
            parked_bytes += e.first;
        }
        pending.clear();
    }
    void purge_parked() {
        for (auto& kv : parked) (void)hipFree(kv.second);
        parked.clear();
        parked_bytes = 0;
    }
    void purge() {  // hipFree waits for the device, so pending blocks are idle by then
        purge_parked();
        for (auto& e : pending) (void)hipFree(e.second);
        pending.clear();
    }
};

// Launch configuration of the point / grid kernels (the autotuner's knobs).
// Neither knob changes any sample's arithmetic, so results are identical for
// every configuration:
//   occupancy  fraction of the default workgroups per CU (1, 0.75, 0.5 ...),
//              enforced with unused dynamic LDS per workgroup;
//   xcd_map    block order: 0 hardware order, 1 one contiguous range of tiles
//              per XCD (shared L2), 2..7 runs of 2^(xcd_map+2) blocks dealt
//              round-robin to the XCDs (shared L2 and balanced dense regions);
//              -1: each kernel family's measured best (runs of 16 blocks for
//              point / grid kernels; line kernels: line_xcd);
//   order      tile order: 0 x-fastest rows, 1 bricks of 8^3 (grids) or 8^2
//              (images) tiles; -1: each family's measured best (rows for
//              grids, bricks for images).
struct LaunchCfg {
    float occupancy = 1.f;
    int xcd_map = -1;
    int order = -1;
    int tile = 1;  // FP32C grids: one wave walks its tile's union (0: one walk per lane)
    int leaves = 8;    // grid tile walks: leaves per bucket of the walked tree (4, 8 or 16)
    int block = 256;   // grid tile walks: threads per block (64, 128, 256 or 512)
};
// Configuration of the entry point running on this thread (set by the entry points).
inline thread_local LaunchCfg tl_launch{};
// Block order of a point / grid kernel and of a line kernel under cfg.
inline int point_xcd(const LaunchCfg& c) { return c.xcd_map < 0 ? 2 : c.xcd_map; }
// Line kernels: runs of 512 blocks when the launch holds at least four
// rounds of them (one run per XCD per round), else the hardware order; longer
// runs on smaller images put whole dense regions on a few XCDs.
inline int line_xcd(const LaunchCfg& c, i64 blocks) {
    if (c.xcd_map >= 0) return c.xcd_map;
    return blocks >= 4 * 8 * 512 ? 7 : 0;
}
inline int point_order(const LaunchCfg& c) { return c.order < 0 ? 0 : c.order; }
inline int line_order(const LaunchCfg& c) { return c.order < 0 ? 1 : c.order; }

// Cache of the context whose entry point runs on this thread (set by the entry
// points), or null: then buffers use hipMalloc / hipFree directly.
inline thread_local BlockCache* tl_cache = nullptr;
// Set while an entry point that must not allocate device memory runs: any
// DevBuf allocation then fails with PH_ERR_INTERNAL instead of allocating.
inline thread_local bool tl_no_alloc = false;

// Device buffer owned by the library.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    BlockCache* cache = nullptr;  // where the block returns on release, or null (hipFree)
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            if (cache) cache->put(p, bytes);
            else (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
        cache = nullptr;
    }
    void alloc(size_t b) {
        if (b <= bytes && p) return;
        if (tl_no_alloc) throw Error{PH_ERR_INTERNAL, "device allocation inside an entry point that allocates nothing"};
        release();
        if (b == 0) b = 16;
        b = (b + 255) & ~size_t(255);
        if (tl_cache) {
            size_t got = 0;
            if (void* q = tl_cache->take(b, got)) {
                p = q;
                bytes = got;
                cache = tl_cache;
                return;
            }
        }
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess && tl_cache && tl_cache->parked_bytes) {  // give parked blocks back and retry
            (void)hipGetLastError();
            tl_cache->purge_parked();
            e = hipMalloc(&p, b);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            throw Error{PH_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(b) + " bytes failed"};
        }
        bytes = b;
        cache = tl_cache;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Packed internal-node record for traversal: bounding box, subtree max scale,
// left child and escape (unified IDs).  64 bytes for double, 48 for float.
template <class T> struct alignas(16) NodeRec {
    T lo[3];
    T hi[3];
    T scale;
    i32 left;
    i32 escape;
};
// Leaf record: position and smoothing length of the particle in Morton order.
template <class T> struct alignas(16) LeafRec {
    T x, y, z, h;
};

}  // namespace ph

// Context: one device, one stream.
struct ph_context {
    int device = 0;
    hipStream_t stream = nullptr;
    ph::BlockCache cache;      // parked device blocks of this context's buffers (destroyed last)
    std::string last_error;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // 0-1 call, 2-3 build, 4-5 kernel
    ph_stats stats{};
    int count_neighbours = 0;
    int traversal = PH_TRAVERSAL_AUTO;
    ph::DevBuf lut64, lut32;   // line-integrated tables [6][4096]
    ph::DevBuf counters;       // small device counters (accepted neighbours, rechecks, flags)
    int num_cu = 256;
    std::vector<struct ph_lbvh*> trees;  // live trees; released with the context
    std::vector<struct ph_batch*> batches;  // live batches; their device state is released with the context
    // launch configuration of point / grid calls: explicit, or chosen per
    // workload key by the autotuner (timed on the call's own data)
    int autotune = 0;
    ph::LaunchCfg launch{};
    struct Tuned {
        ph::LaunchCfg cfg;
        double best_ms = 0.0, default_ms = 0.0;
    };
    std::map<uint64_t, Tuned> tuned;
    // device address of this GPU's copy of the library's static workspace
    // (ph_point_samples_device_noalloc), looked up at context creation
    void* noalloc_scratch = nullptr;
};

// Device-resident InterpolationInput after LinearBVH!.
struct ph_lbvh {
    ph_context* ctx = nullptr;
    int64_t n = 0;
    int32_t ncol = 0;
    int32_t kernel = 0;
    int32_t code_bits = 63;
    int32_t precision = PH_PREC_FP64;  // PH_PREC_FP64 or PH_PREC_FP32 columns
    int32_t dims = 3;   // 2: InterpolationInput{2} (2D Morton codes, z column all zeros)
    int32_t root = 0;
    // Morton order columns: x,y,z,m,h,rho then ncol quantity columns, (6+ncol) x n
    ph::DevBuf cols;
    ph::DevBuf codes;   // u64 [n]
    ph::DevBuf order;   // u32 [n], 0-based source index
    ph::DevBuf left, right, escape, parent;  // i32 [2n-1]
    ph::DevBuf range;   // i32 [2*(n-1)]: lo, hi (1-based leaf indices) per internal node
    ph::DevBuf nodes;   // NodeRec<T> [n-1]
    ph::DevBuf leaves;  // LeafRec<T> [n]
    ph::DevBuf leaf_escape;  // i32 [n]
    ph::DevBuf rec;          // packed per-particle records of the last interpolation call
    ph::DevBuf ha;           // per-sample smoothing length of the nearest particle
    ph::DevBuf nn;           // per-sample Morton index of the nearest particle (adjoints)
    // Bucket trees (built with the tree, ensure_buckets): the LBVH in DFS
    // preorder with every subtree of at most 1 (the nearest-h searches),
    // kBucketSmall (4; the per-lane kernels), kBucket (8) or kBucketLarge (16)
    // leaves (the grid tile walks) collapsed into one bucket.
    bool buckets_ready = false;    // all of them (the one-leaf tree if buckets1_ready)
    bool buckets1_ready = false;
    int64_t nrec = 0, nrec1 = 0, nrec8 = 0, nrec16 = 0;
    ph::DevBuf brecs1;    // one-leaf bucket-tree records [nrec1]
    ph::DevBuf brecs8;    // 8-leaf bucket-tree records [nrec8]
    ph::DevBuf brecs16;   // 16-leaf bucket-tree records [nrec16]
    ph::DevBuf brecs;     // 4-leaf bucket-tree records [nrec]
    ph::DevBuf rec32;     // FP32-compensated records of the last call ({m/rho, 1/h}, columns)
    // Records of the per-wave line walks (LineRec<T> [2n-1] in DFS preorder:
    // the boxes and node-only terms of the line pruning test, the leaves; then
    // one of padding and the axis flag), built by the first line call.
    bool lnodes_ready = false;
    ph::DevBuf lnodes;

    // Frees every device buffer (the context is going away; the handle stays
    // valid for ph_lbvh_destroy only).
    void release_device() {
        for (ph::DevBuf* b : {&cols, &codes, &order, &left, &right, &escape, &parent, &range, &nodes, &leaves,
                              &leaf_escape, &rec, &ha, &nn, &brecs, &brecs8, &brecs16, &brecs1, &rec32, &lnodes})
            b->release();
        buckets_ready = false;
        lnodes_ready = false;
        buckets1_ready = false;
        nrec = nrec1 = nrec8 = nrec16 = 0;
    }
};

namespace ph {
struct BatchState;
}
// Batched sampling requests recorded over one tree (ph_batch_create).
struct ph_batch {
    ph_context* ctx = nullptr;  // null once the context was destroyed
    ph_lbvh* B = nullptr;       // null once the tree was destroyed
    ph::BatchState* st = nullptr;
};
