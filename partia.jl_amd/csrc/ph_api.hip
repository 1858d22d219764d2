// ph_api.hip -- extern "C" entry points of libpartia_hip (declared in include/partia_hip.h).
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "ph_build.h"
#include "ph_interp.h"
#include "line_lut_tables.inc"

using namespace ph;

// Static workspace of ph_point_samples_device_noalloc: the per-call buffers of
// the bucket kernels (nearest-particle h per sample, packed particle records)
// for callers that cannot allocate device memory.  Every GPU the library runs
// on holds its own copy until the process ends; in host code `scratch` is only
// a stand-in for the symbol, so the per-device address from
// hipGetSymbolAddress is what kernels receive.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_addr;  // device -> address of its copy of `scratch`

// Address of `scratch` on the current device (which must be `device`), looked
// up once per device.
void* scratch_address(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) return it->second;
    void* ptr = nullptr;
    PH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error{PH_ERR_HIP, "hipGetSymbolAddress returned a null address for the static workspace"};
    g_scratch_addr[device] = ptr;
    return ptr;
}

thread_local std::string g_noctx_error;

int fail(ph_context* ctx, int code, const std::string& msg) {
    if (ctx) ctx->last_error = msg;
    else g_noctx_error = msg;
    return code;
}

// Routes the call's device allocations through the context's block cache.
struct CacheScope {
    BlockCache* saved;
    explicit CacheScope(ph_context* ctx) : saved(tl_cache) { tl_cache = ctx ? &ctx->cache : nullptr; }
    ~CacheScope() { tl_cache = saved; }
};

// Blocks released by the call return to the cache once the stream has drained.
void settle_cache(ph_context* ctx) {
    if (!ctx || ctx->cache.pending.empty()) return;
    if (ctx->stream && hipStreamSynchronize(ctx->stream) != hipSuccess) {
        (void)hipGetLastError();
        return;  // left pending: freed with the context
    }
    ctx->cache.settle();
}

template <class F> int guarded(ph_context* ctx, F&& f) {
    struct Settle {
        ph_context* c;
        ~Settle() { settle_cache(c); }
    } settle{ctx};
    try {
        CacheScope scope(ctx);
        if (ctx) {
            PH_HIP(hipSetDevice(ctx->device));
            ctx->last_error.clear();
        }
        return f();
    } catch (const Error& e) {
        return fail(ctx, e.code, e.msg);
    } catch (const std::bad_alloc&) {
        return fail(ctx, PH_ERR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(ctx, PH_ERR_INTERNAL, e.what());
    }
}

void need(bool ok, int code, const std::string& msg) {
    if (!ok) throw Error{code, msg};
}

float elapsed(ph_context* ctx, int a, int b) {
    float ms = 0.f;
    PH_HIP(hipEventElapsedTime(&ms, ctx->ev[a], ctx->ev[b]));
    return ms;
}

void check_catalog(const ph_lbvh* B, const ph_catalog* cat) {
    // the reference's interpolation kernels are 3D only (README.md:152)
    need(B->dims == 3, PH_ERR_ARGUMENT,
         "interpolation needs a 3D tree: this one was built from an InterpolationInput{2}");
    need(cat != nullptr, PH_ERR_ARGUMENT, "catalog is null");
    need(cat->n_scalar >= 0 && cat->n_grad >= 0 && cat->n_div >= 0 && cat->n_curl >= 0, PH_ERR_ARGUMENT,
         "negative catalog size");
    auto slot_ok = [&](int s, bool allow0) { return (s >= 1 && s <= B->ncol) || (allow0 && s == 0); };
    for (int j = 0; j < cat->n_scalar; ++j)
        need(slot_ok(cat->scalar_slots[j], false), PH_ERR_ARGUMENT, "scalar slot out of range");
    for (int j = 0; j < cat->n_grad; ++j)
        need(slot_ok(cat->grad_slots[j], true), PH_ERR_ARGUMENT, "gradient slot out of range");
    for (int j = 0; j < 3 * cat->n_div; ++j)
        need(slot_ok(cat->div_slots[j], false), PH_ERR_ARGUMENT, "divergence slot out of range");
    for (int j = 0; j < 3 * cat->n_curl; ++j)
        need(slot_ok(cat->curl_slots[j], false), PH_ERR_ARGUMENT, "curl slot out of range");
}

void check_precision(const ph_lbvh* B, int precision) {
    need(precision == PH_PREC_FP64 || precision == PH_PREC_FP32 || precision == PH_PREC_FP32C, PH_ERR_ARGUMENT,
         "unknown precision");
    if (precision == PH_PREC_FP32)
        need(B->precision == PH_PREC_FP32, PH_ERR_ARGUMENT, "PH_PREC_FP32 needs a tree built from Float32 columns");
    else
        need(B->precision == PH_PREC_FP64, PH_ERR_ARGUMENT, "float64 evaluation needs a tree built from Float64 columns");
}

// Sets the thread's launch configuration for the duration of a scope.
struct LaunchScope {
    LaunchCfg saved;
    explicit LaunchScope(const LaunchCfg& c) : saved(tl_launch) { tl_launch = c; }
    ~LaunchScope() { tl_launch = saved; }
};

// Workload key of the autotuner: precision, strategy, kernel, catalog shape,
// sample source and the sizes' binary orders of magnitude.
uint64_t tune_key(const ph_lbvh* B, const CatChunk& c, int strategy, int precision, bool grid, i64 S) {
    const uint64_t shape = (c.ng + c.nd + c.nc == 0 && c.ns <= 2) ? 0
                         : (c.ns == 2 && c.ng == 1 && c.nd == 1 && c.nc == 1) ? 1 : 2;
    auto lg = [](i64 v) {
        uint64_t b = 0;
        while (v > 1) {
            v >>= 1;
            ++b;
        }
        return b;
    };
    return (uint64_t)(precision == PH_PREC_FP64 ? 0 : 1) | (uint64_t)strategy << 2 | (uint64_t)B->kernel << 4 |
           shape << 8 | (uint64_t)(grid ? 1 : 0) << 10 | lg(S) << 12 | lg(B->n) << 20;
}

// Times every candidate configuration on the call itself (outputs are the same
// for all of them) and keeps the fastest per workload key.
LaunchCfg autotuned(ph_context* ctx, ph_lbvh* B, const CatChunk& c, int strategy, int precision, const PointList* pl,
                    const GridAxes* ga, i64 S, void* out) {
    const uint64_t key = tune_key(B, c, strategy, precision, ga != nullptr, S);
    auto it = ctx->tuned.find(key);
    if (it != ctx->tuned.end()) return it->second.cfg;
    // the default first; full occupancy with the block orders that balance
    // dense regions over the XCDs, both tile orders; two lower occupancies
    LaunchCfg first = ctx->launch;
    first.xcd_map = point_xcd(first);
    first.order = point_order(first);
    std::vector<LaunchCfg> cand{first};
    auto add = [&](LaunchCfg a) {
        for (const LaunchCfg& b : cand)
            if (a.occupancy == b.occupancy && a.xcd_map == b.xcd_map && a.order == b.order) return;
        cand.push_back(a);
    };
    for (int order : {0, 1})
        for (int xm : {2, 3, 7, 0}) add(LaunchCfg{1.f, xm, order});
    for (float occ : {0.75f, 0.5f}) add(LaunchCfg{occ, 2, 0});
    std::vector<double> best(cand.size(), 1e300);
    hipEvent_t* ev = &ctx->ev[2];
    auto run = [&](LaunchCfg cf) {
        cf.tile = ctx->traversal != PH_TRAVERSAL_BUCKET;
        LaunchScope scope(cf);
        launch_bucket(B, c, strategy, precision, pl, ga, S, out, nullptr, ctx->stream, ev);
        PH_HIP(hipEventSynchronize(ev[1]));
        float ms = 0.f;
        PH_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
        return (double)ms;
    };
    run(cand[0]);  // warm-up: records, caches
    for (int rep = 0; rep < 2; ++rep)
        for (size_t i = 0; i < cand.size(); ++i) best[i] = std::min(best[i], run(cand[i]));
    size_t w = 0;
    for (size_t i = 1; i < cand.size(); ++i)
        if (best[i] < best[w]) w = i;
    LaunchCfg chosen = cand[w];
    double chosen_ms = best[w];
    // grid tile walks: then the walked tree's bucket size and the block size,
    // one at a time from the best so far
    if (ga && ctx->traversal != PH_TRAVERSAL_BUCKET) {
        auto stage = [&](auto set, std::initializer_list<int> values) {
            LaunchCfg base = chosen;
            for (int v : values) {
                LaunchCfg cf = base;
                set(cf, v);
                if (cf.leaves == base.leaves && cf.block == base.block) continue;
                run(cf);  // warm-up
                const double ms = std::min(run(cf), run(cf));
                if (ms < chosen_ms) {
                    chosen = cf;
                    chosen_ms = ms;
                }
            }
        };
        stage([](LaunchCfg& cf, int v) { cf.leaves = v; }, {4, 8, 16});
        stage([](LaunchCfg& cf, int v) { cf.block = v; }, {128, 256, 512});
    }
    ph_context::Tuned t;
    t.cfg = chosen;
    t.best_ms = chosen_ms;
    t.default_ms = best[0];
    ctx->tuned[key] = t;
    return t.cfg;
}

// Point / grid launches: the bucket traversal for FP32C and (by default) float64,
// the reference-structured per-leaf traversal on request and for Float32 trees.
void launch_point_like(ph_context* ctx, ph_lbvh* B, const CatChunk& c, int strategy, int precision,
                       const PointList* pl, const GridAxes* ga, i64 S, int L, void* out, u64* counter,
                       hipEvent_t* ev) {
    bool bucket = precision == PH_PREC_FP32C;
    if (precision == PH_PREC_FP64) bucket = ctx->traversal != PH_TRAVERSAL_LANE;
    if (bucket) {
        LaunchCfg cfg = ctx->launch;
        if (ctx->autotune && S > 0) cfg = autotuned(ctx, B, c, strategy, precision, pl, ga, S, out);
        cfg.tile = ctx->traversal != PH_TRAVERSAL_BUCKET;  // FP32C grids: per-wave walk unless asked otherwise
        LaunchScope scope(cfg);
        launch_bucket(B, c, strategy, precision, pl, ga, S, out, counter, ctx->stream, ev);
    } else {
        launch_points(B, c, strategy, precision, pl, ga, S, L, out, counter, ctx->stream, ev);
    }
}

size_t out_elem(int precision) { return precision == PH_PREC_FP64 ? sizeof(double) : sizeof(float); }

struct Counters {
    u64* dev;
};

// Runs the interpolation of all catalog chunks; device-side timing into stats:
// interp_ms brackets the whole call, kernel_ms sums the dominant kernel's launches.
template <class Launch> void run_chunks(ph_context* ctx, const ph_catalog* cat, i64 S, Launch&& launch) {
    std::vector<CatChunk> chunks;
    const int L = split_catalog(cat, chunks);
    u64* counter = nullptr;
    if (ctx->count_neighbours) {
        ctx->counters.alloc(64);
        counter = ctx->counters.as<u64>();
        PH_HIP(hipMemsetAsync(counter, 0, 16, ctx->stream));
    }
    PH_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    double kms = 0.0;
    for (const CatChunk& c : chunks) {
        launch(c, L, counter, &ctx->ev[4]);
        if (S > 0) {
            PH_HIP(hipEventSynchronize(ctx->ev[5]));
            kms += elapsed(ctx, 4, 5);
        }
    }
    PH_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    PH_HIP(hipEventSynchronize(ctx->ev[1]));
    ctx->stats.interp_ms = elapsed(ctx, 0, 1);
    ctx->stats.kernel_ms = kms;
    ctx->stats.samples = S;
    ctx->stats.accepted = -1;
    ctx->stats.rechecks = -1;
    if (counter) {
        u64 h[2] = {0, 0};
        PH_HIP(hipMemcpy(h, counter, sizeof(h), hipMemcpyDeviceToHost));
        const u64 nch = chunks.empty() ? 1 : chunks.size();
        ctx->stats.accepted = (int64_t)(h[0] / nch);
        ctx->stats.rechecks = (int64_t)(h[1] / nch);
    }
}

// Copies a device L x S block of T to a host double array.
void download(ph_context* ctx, const void* dev, int precision, int64_t count, double* host) {
    if (precision == PH_PREC_FP64) {
        PH_HIP(hipMemcpy(host, dev, sizeof(double) * count, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(count);
        PH_HIP(hipMemcpy(tmp.data(), dev, sizeof(float) * count, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < count; ++i) host[i] = tmp[i];
    }
}

int catalog_L(const ph_catalog* c) { return c->n_scalar + 3 * c->n_grad + c->n_div + 3 * c->n_curl; }

}  // namespace

extern "C" {

int ph_ctx_create(int device, ph_context** out) {
    if (!out) return fail(nullptr, PH_ERR_ARGUMENT, "out is null");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return fail(nullptr, PH_ERR_NO_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= count) return fail(nullptr, PH_ERR_ARGUMENT, "device index out of range");
    auto ctx = std::make_unique<ph_context>();
    ctx->device = device;
    int rc = guarded(ctx.get(), [&] {
        PH_HIP(hipSetDevice(device));
        PH_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        for (auto& e : ctx->ev) PH_HIP(hipEventCreate(&e));
        hipDeviceProp_t prop;
        PH_HIP(hipGetDeviceProperties(&prop, device));
        ctx->num_cu = prop.multiProcessorCount;
        ctx->noalloc_scratch = scratch_address(device);
        // line-integrated tables, float64 and float32 copies
        const size_t nt = 6 * (size_t)PH_LUT_N;
        ctx->lut64.alloc(sizeof(double) * nt);
        ctx->lut32.alloc(sizeof(float) * nt);
        std::vector<float> t32(nt);
        for (size_t i = 0; i < nt; ++i) t32[i] = (float)(&ph_lut_inorm[0][0])[i];
        PH_HIP(hipMemcpy(ctx->lut64.p, &ph_lut_inorm[0][0], sizeof(double) * nt, hipMemcpyHostToDevice));
        PH_HIP(hipMemcpy(ctx->lut32.p, t32.data(), sizeof(float) * nt, hipMemcpyHostToDevice));
        return PH_OK;
    });
    if (rc != PH_OK) {
        g_noctx_error = ctx->last_error;
        return rc;
    }
    *out = ctx.release();
    return PH_OK;
}

int ph_ctx_destroy(ph_context* ctx) {
    if (!ctx) return PH_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    // batches and trees outliving their context keep a host handle without device state
    for (ph_batch* b : ctx->batches) {
        batch_free(b->st);
        b->st = nullptr;
        b->ctx = nullptr;
        b->B = nullptr;
    }
    ctx->batches.clear();
    for (ph_lbvh* B : ctx->trees) {
        B->release_device();
        B->ctx = nullptr;
    }
    ctx->trees.clear();
    ctx->lut64.release();
    ctx->lut32.release();
    ctx->counters.release();
    ctx->cache.purge();
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PH_OK;
}

const char* ph_last_error(const ph_context* ctx) { return ctx ? ctx->last_error.c_str() : g_noctx_error.c_str(); }

int ph_get_stats(const ph_context* ctx, ph_stats* out) {
    if (!ctx || !out) return PH_ERR_ARGUMENT;
    *out = ctx->stats;
    return PH_OK;
}

int ph_set_count_neighbours(ph_context* ctx, int enable) {
    if (!ctx) return PH_ERR_ARGUMENT;
    ctx->count_neighbours = enable ? 1 : 0;
    return PH_OK;
}

int ph_set_traversal(ph_context* ctx, int mode) {
    if (!ctx) return PH_ERR_ARGUMENT;
    if (mode < PH_TRAVERSAL_AUTO || mode > PH_TRAVERSAL_WAVE)
        return fail(ctx, PH_ERR_ARGUMENT, "unknown traversal mode");
    ctx->traversal = mode;
    return PH_OK;
}

int ph_set_autotune(ph_context* ctx, int enable) {
    if (!ctx) return PH_ERR_ARGUMENT;
    ctx->autotune = enable ? 1 : 0;
    return PH_OK;
}

int ph_set_launch(ph_context* ctx, double occupancy, int32_t xcd_map, int32_t tile_order) {
    if (!ctx) return PH_ERR_ARGUMENT;
    if (!(occupancy > 0.0 && occupancy <= 1.0) || xcd_map < -1 || xcd_map > 7 || tile_order < -1 ||
        tile_order > 1)
        return fail(ctx, PH_ERR_ARGUMENT, "occupancy must be in (0, 1], xcd_map in -1..7 and tile_order in -1..1");
    ctx->launch.occupancy = (float)occupancy;
    ctx->launch.xcd_map = xcd_map;
    ctx->launch.order = tile_order;
    return PH_OK;
}

int ph_set_tile_shape(ph_context* ctx, int32_t bucket_leaves, int32_t block_size) {
    if (!ctx) return PH_ERR_ARGUMENT;
    if (!(bucket_leaves == 0 || bucket_leaves == 4 || bucket_leaves == 8 || bucket_leaves == 16) ||
        !(block_size == 0 || block_size == 64 || block_size == 128 || block_size == 256 || block_size == 512))
        return fail(ctx, PH_ERR_ARGUMENT, "bucket_leaves must be 0, 4, 8 or 16 and block_size 0, 64, 128, 256 or 512");
    ctx->launch.leaves = bucket_leaves ? bucket_leaves : 8;
    ctx->launch.block = block_size ? block_size : kBlock;
    return PH_OK;
}

int ph_autotune_report(ph_context* ctx, int32_t capacity, ph_tune_entry* out, int32_t* count) {
    if (!ctx || capacity < 0 || (capacity > 0 && !out)) return PH_ERR_ARGUMENT;
    int32_t k = 0;
    for (const auto& kv : ctx->tuned) {
        if (k < capacity) {
            ph_tune_entry& e = out[k];
            e.key = kv.first;
            e.occupancy = kv.second.cfg.occupancy;
            e.xcd_map = kv.second.cfg.xcd_map;
            e.tile_order = kv.second.cfg.order;
            e.best_ms = kv.second.best_ms;
            e.default_ms = kv.second.default_ms;
            e.bucket_leaves = kv.second.cfg.leaves;
            e.block_size = kv.second.cfg.block;
        }
        ++k;
    }
    if (count) *count = k;
    return PH_OK;
}

int ph_device_alloc(ph_context* ctx, int64_t bytes, void** dptr) {
    return guarded(ctx, [&] {
        need(ctx && dptr && bytes >= 0, PH_ERR_ARGUMENT, "bad arguments");
        *dptr = nullptr;
        hipError_t e = hipMalloc(dptr, bytes > 0 ? (size_t)bytes : 16);
        if (e != hipSuccess && ctx->cache.parked_bytes) {  // give parked blocks back and retry
            (void)hipGetLastError();
            ctx->cache.purge_parked();
            e = hipMalloc(dptr, bytes > 0 ? (size_t)bytes : 16);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *dptr = nullptr;
            throw Error{PH_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(bytes) + " bytes failed"};
        }
        return PH_OK;
    });
}

int ph_device_free(ph_context* ctx, void* dptr) {
    return guarded(ctx, [&] {
        if (dptr) PH_HIP(hipFree(dptr));
        return PH_OK;
    });
}

int ph_memcpy_htod(ph_context* ctx, void* dst, const void* src, int64_t bytes) {
    return guarded(ctx, [&] {
        need(bytes >= 0 && (bytes == 0 || (dst && src)), PH_ERR_ARGUMENT, "bad copy arguments");
        PH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        PH_HIP(hipStreamSynchronize(ctx->stream));
        return PH_OK;
    });
}

int ph_memcpy_dtoh(ph_context* ctx, void* dst, const void* src, int64_t bytes) {
    return guarded(ctx, [&] {
        need(bytes >= 0 && (bytes == 0 || (dst && src)), PH_ERR_ARGUMENT, "bad copy arguments");
        PH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        PH_HIP(hipStreamSynchronize(ctx->stream));
        return PH_OK;
    });
}

int ph_synchronize(ph_context* ctx) {
    return guarded(ctx, [&] {
        PH_HIP(hipStreamSynchronize(ctx->stream));
        return PH_OK;
    });
}

static int build_common(ph_context* ctx, int64_t n, const double* const* src, bool host_src, int32_t ncol,
                        int32_t kernel, int32_t code_bits, int32_t precision, ph_lbvh** out, int32_t dims = 3) {
    return guarded(ctx, [&] {
        need(ctx != nullptr && out != nullptr, PH_ERR_ARGUMENT, "null context or output");
        *out = nullptr;
        need(n >= 1, PH_ERR_CAPACITY, "BinaryRadixTree: enc.codes must be non-empty (got n=0).");
        need(n <= (int64_t)1073741824, PH_ERR_CAPACITY,
             "BinaryRadixTree: n=" + std::to_string(n) + " exceeds the Int32 node-id capacity (n <= 1073741824).");
        need(ncol >= 0, PH_ERR_ARGUMENT, "negative column count");
        need(kernel >= 0 && kernel <= 5, PH_ERR_ARGUMENT, "unknown kernel");
        need(code_bits == 30 || code_bits == 63, PH_ERR_ARGUMENT, "code_bits must be 30 (UInt32) or 63 (UInt64)");
        need(precision == PH_PREC_FP64 || precision == PH_PREC_FP32, PH_ERR_ARGUMENT,
             "tree precision must be PH_PREC_FP64 or PH_PREC_FP32");
        for (int c = 0; c < 6 + ncol; ++c)
            need(src[c] != nullptr || (c == 2 && dims == 2), PH_ERR_ARGUMENT, "null column pointer");
        auto B = std::make_unique<ph_lbvh>();
        B->ctx = ctx;
        B->dims = dims;
        B->n = n;
        B->ncol = ncol;
        B->kernel = kernel;
        B->code_bits = code_bits;
        B->precision = precision;
        BuildScratch W;
        PH_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
        // the bucket trees of the point / grid kernels, built with the tree so
        // that interpolation calls allocate only per-call buffers; their
        // read-back of the record counts also reads the build's refit check
        const bool buckets = buckets_supported(B.get());
        u32* refit_bad = nullptr;
        build_lbvh(B.get(), src, host_src, ctx->stream, W, buckets ? &refit_bad : nullptr);
        if (buckets) ensure_buckets(B.get(), ctx->stream, refit_bad);
        PH_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
        PH_HIP(hipEventSynchronize(ctx->ev[3]));
        ctx->stats.build_ms = elapsed(ctx, 2, 3);
        ctx->trees.push_back(B.get());
        *out = B.release();
        return PH_OK;
    });
}

int ph_lbvh_build(ph_context* ctx, int64_t n, const double* x, const double* y, const double* z, const double* m,
                  const double* h, const double* rho, int32_t ncol, const double* const* cols, int32_t kernel,
                  int32_t code_bits, int32_t precision, ph_lbvh** out) {
    std::vector<const double*> src = {x, y, z, m, h, rho};
    for (int c = 0; c < ncol; ++c) src.push_back(cols ? cols[c] : nullptr);
    return build_common(ctx, n, src.data(), true, ncol, kernel, code_bits, precision, out);
}

int ph_lbvh_build_device(ph_context* ctx, int64_t n, const double* x, const double* y, const double* z,
                         const double* m, const double* h, const double* rho, int32_t ncol, const double* const* cols,
                         int32_t kernel, int32_t code_bits, int32_t precision, ph_lbvh** out) {
    std::vector<const double*> src = {x, y, z, m, h, rho};
    for (int c = 0; c < ncol; ++c) src.push_back(cols ? cols[c] : nullptr);
    return build_common(ctx, n, src.data(), false, ncol, kernel, code_bits, precision, out);
}

int ph_lbvh_build2d(ph_context* ctx, int64_t n, const double* x, const double* y, const double* m,
                    const double* h, const double* rho, int32_t ncol, const double* const* cols, int32_t kernel,
                    int32_t code_bits, int32_t precision, ph_lbvh** out) {
    std::vector<const double*> src = {x, y, nullptr, m, h, rho};
    for (int c = 0; c < ncol; ++c) src.push_back(cols ? cols[c] : nullptr);
    return build_common(ctx, n, src.data(), true, ncol, kernel, code_bits, precision, out, 2);
}

int ph_lbvh_build2d_device(ph_context* ctx, int64_t n, const double* x, const double* y, const double* m,
                           const double* h, const double* rho, int32_t ncol, const double* const* cols,
                           int32_t kernel, int32_t code_bits, int32_t precision, ph_lbvh** out) {
    std::vector<const double*> src = {x, y, nullptr, m, h, rho};
    for (int c = 0; c < ncol; ++c) src.push_back(cols ? cols[c] : nullptr);
    return build_common(ctx, n, src.data(), false, ncol, kernel, code_bits, precision, out, 2);
}

int ph_lbvh_dims(const ph_lbvh* B, int32_t* dims) {
    if (!B || !dims) return PH_ERR_ARGUMENT;
    *dims = B->dims;
    return PH_OK;
}

int ph_lbvh_destroy(ph_lbvh* B) {
    if (!B) return PH_OK;
    if (ph_context* ctx = B->ctx) {  // null once the context was destroyed (memory already released)
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        auto& v = ctx->trees;
        for (size_t t = 0; t < v.size(); ++t)
            if (v[t] == B) {
                v.erase(v.begin() + (std::ptrdiff_t)t);
                break;
            }
        for (ph_batch* b : ctx->batches)  // batches over this tree can no longer run
            if (b->B == B) b->B = nullptr;
        delete B;  // device blocks return to the context's cache; the stream is idle
        ctx->cache.settle();
        return PH_OK;
    }
    delete B;
    return PH_OK;
}

int ph_lbvh_set_columns(ph_lbvh* B, const double* m, const double* rho, int32_t ncol, const double* const* cols) {
    if (!B || !B->ctx) return fail(nullptr, PH_ERR_ARGUMENT, "null tree or tree of a destroyed context");
    ph_context* ctx = B->ctx;
    return guarded(ctx, [&] {
        need(m && rho && ncol >= 0 && (ncol == 0 || cols), PH_ERR_ARGUMENT, "null column pointer");
        for (int c = 0; c < ncol; ++c) need(cols[c] != nullptr, PH_ERR_ARGUMENT, "null column pointer");
        const i64 n = B->n;
        const bool f64 = B->precision == PH_PREC_FP64;
        const size_t es = f64 ? sizeof(double) : sizeof(float);
        PH_HIP(hipStreamSynchronize(ctx->stream));
        if (ncol != B->ncol) {
            // keep x, y, z, (m), h, (rho); reallocate for the new column count
            DevBuf nc;
            nc.alloc(es * (size_t)(6 + ncol) * n);
                PH_HIP(hipMemcpy(nc.p, B->cols.p, es * 6 * (size_t)n, hipMemcpyDeviceToDevice));
            std::swap(nc.p, B->cols.p);
            std::swap(nc.bytes, B->cols.bytes);
            std::swap(nc.cache, B->cols.cache);
            B->ncol = ncol;
        }
        auto put = [&](int c, const double* src) {
            if (f64) {
                PH_HIP(hipMemcpy(B->cols.as<double>() + (size_t)c * n, src, sizeof(double) * n, hipMemcpyHostToDevice));
            } else {
                std::vector<float> t(src, src + n);
                PH_HIP(hipMemcpy(B->cols.as<float>() + (size_t)c * n, t.data(), sizeof(float) * n,
                                 hipMemcpyHostToDevice));
            }
        };
        put(3, m);
        put(5, rho);
        for (int c = 0; c < ncol; ++c) put(6 + c, cols[c]);
        return PH_OK;
    });
}

int ph_lbvh_info(const ph_lbvh* B, int64_t* n, int32_t* root, int32_t* ncol, int32_t* code_bits, int32_t* precision) {
    if (!B) return PH_ERR_ARGUMENT;
    if (n) *n = B->n;
    if (root) *root = B->root;
    if (ncol) *ncol = B->ncol;
    if (code_bits) *code_bits = B->code_bits;
    if (precision) *precision = B->precision;
    return PH_OK;
}

int ph_lbvh_get(ph_lbvh* B, int32_t which, void* dst) {
    if (!B || !B->ctx) return fail(nullptr, PH_ERR_ARGUMENT, "null tree or tree of a destroyed context");
    ph_context* ctx = B->ctx;
    return guarded(ctx, [&] {
        need(dst != nullptr, PH_ERR_ARGUMENT, "destination is null");
        const i64 n = B->n, ni = n - 1, tot = 2 * n - 1;
        PH_HIP(hipStreamSynchronize(ctx->stream));
        auto tree = [&](const DevBuf& b) { PH_HIP(hipMemcpy(dst, b.p, sizeof(i32) * tot, hipMemcpyDeviceToHost)); };
        const bool f64 = B->precision == PH_PREC_FP64;
        auto col = [&](int c, double* o) {
            if (f64) {
                PH_HIP(hipMemcpy(o, B->cols.as<double>() + (size_t)c * n, sizeof(double) * n, hipMemcpyDeviceToHost));
            } else {
                std::vector<float> t(n);
                PH_HIP(hipMemcpy(t.data(), B->cols.as<float>() + (size_t)c * n, sizeof(float) * n,
                                 hipMemcpyDeviceToHost));
                for (i64 i = 0; i < n; ++i) o[i] = t[i];
            }
        };
        auto nodes = [&](auto tag, int field, double* o) {
            using T = decltype(tag);
            std::vector<NodeRec<T>> v((size_t)(ni > 0 ? ni : 0));
            if (ni > 0)
                PH_HIP(hipMemcpy(v.data(), B->nodes.p, sizeof(NodeRec<T>) * ni, hipMemcpyDeviceToHost));
            for (i64 i = 0; i < ni; ++i) {
                if (field == 0)
                    for (int d = 0; d < 3; ++d) o[d * ni + i] = v[i].lo[d];
                else if (field == 1)
                    for (int d = 0; d < 3; ++d) o[d * ni + i] = v[i].hi[d];
                else
                    o[i] = v[i].scale;
            }
        };
        switch (which) {
        case PH_GET_ORDER: {
            std::vector<u32> t(n);
            PH_HIP(hipMemcpy(t.data(), B->order.p, sizeof(u32) * n, hipMemcpyDeviceToHost));
            u64* o = static_cast<u64*>(dst);
            for (i64 i = 0; i < n; ++i) o[i] = (u64)t[i] + 1;
            break;
        }
        case PH_GET_CODES: PH_HIP(hipMemcpy(dst, B->codes.p, sizeof(u64) * n, hipMemcpyDeviceToHost)); break;
        case PH_GET_LEFT: tree(B->left); break;
        case PH_GET_RIGHT: tree(B->right); break;
        case PH_GET_ESCAPE: tree(B->escape); break;
        case PH_GET_PARENT: tree(B->parent); break;
        case PH_GET_LEAF_COOR:
            for (int d = 0; d < 3; ++d) col(d, static_cast<double*>(dst) + (size_t)d * n);
            break;
        case PH_GET_LEAF_SCALE: col(4, static_cast<double*>(dst)); break;
        case PH_GET_NODE_MIN:
        case PH_GET_NODE_MAX:
        case PH_GET_NODE_SCALE: {
            const int field = which - PH_GET_NODE_MIN;
            if (f64) nodes(double(0), field, static_cast<double*>(dst));
            else nodes(float(0), field, static_cast<double*>(dst));
            break;
        }
        case PH_GET_COLUMNS:
            for (int c = 0; c < 6 + B->ncol; ++c) col(c, static_cast<double*>(dst) + (size_t)c * n);
            break;
        default: throw Error{PH_ERR_ARGUMENT, "unknown selector"};
        }
        return PH_OK;
    });
}

int ph_point_samples_device(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                            int64_t S, const double* px, const double* py, const double* pz, void* out) {
    return guarded(ctx, [&] {
        need(B != nullptr && B->ctx == ctx, PH_ERR_ARGUMENT, "tree does not belong to this context");
        need(strategy >= PH_GATHER && strategy <= PH_SYMMETRIC, PH_ERR_ARGUMENT, "unknown strategy");
        need(S >= 0, PH_ERR_ARGUMENT, "negative sample count");
        check_precision(B, precision);
        check_catalog(B, cat);
        need(S == 0 || (px && py && pz && out), PH_ERR_ARGUMENT, "null sample or output pointer");
        PointList pl{px, py, pz};
        run_chunks(ctx, cat, S, [&](const CatChunk& c, int L, u64* counter, hipEvent_t* ev) {
            launch_point_like(ctx, B, c, strategy, precision, &pl, nullptr, S, L, out, counter, ev);
        });
        return PH_OK;
    });
}

int ph_point_samples_device_noalloc(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy,
                                    int32_t precision, int64_t S, const double* px, const double* py,
                                    const double* pz, void* out) {
    if (!ctx) return fail(nullptr, PH_ERR_ARGUMENT, "context is null");
    return guarded(ctx, [&] {
        need(B != nullptr && B->ctx == ctx, PH_ERR_ARGUMENT, "tree does not belong to this context");
        need(strategy >= PH_GATHER && strategy <= PH_SYMMETRIC, PH_ERR_ARGUMENT, "unknown strategy");
        need(S >= 0, PH_ERR_ARGUMENT, "negative sample count");
        need(precision == PH_PREC_FP64 || precision == PH_PREC_FP32C, PH_ERR_ARGUMENT,
             "ph_point_samples_device_noalloc evaluates in PH_PREC_FP64 or PH_PREC_FP32C");
        check_precision(B, precision);
        check_catalog(B, cat);
        need(S == 0 || (px && py && pz && out), PH_ERR_ARGUMENT, "null sample or output pointer");
        need(B->buckets_ready, PH_ERR_ARGUMENT,
             "ph_point_samples_device_noalloc needs a tree with bucket trees (3D, float64, at most 2^30 particles)");
        need(ctx->noalloc_scratch != nullptr, PH_ERR_INTERNAL, "static workspace address missing");
        if (S == 0) return PH_OK;
        // workspace: ha (gather / symmetric) at offset 0, then the largest
        // chunk's packed records, both 256-byte aligned
        std::vector<CatChunk> chunks;
        split_catalog(cat, chunks);
        const i64 cap = (i64)kNoAllocScratchBytes;
        auto up = [](i64 b) { return (b + 255) & ~i64(255); };
        need(S <= cap / 8, PH_ERR_CAPACITY, "ph_point_samples_device_noalloc: S exceeds the static workspace");
        const i64 ha_bytes = strategy != PH_SCATTER ? up(8 * S) : 0;
        i64 rec_bytes = 0;
        for (const CatChunk& c : chunks) {
            const i64 rw = bucket_record_doubles(c, precision);
            const i64 nrec = bucket_records(B->n, precision);
            need(nrec <= cap / (8 * rw), PH_ERR_CAPACITY,
                 "ph_point_samples_device_noalloc: records exceed the static workspace");
            rec_bytes = std::max(rec_bytes, up(8 * rw * nrec));
        }
        need(ha_bytes + rec_bytes <= cap, PH_ERR_CAPACITY,
             "ph_point_samples_device_noalloc needs " + std::to_string(ha_bytes + rec_bytes) +
                 " bytes of workspace (8 S + 8 rw records); the static workspace holds " + std::to_string(cap) +
                 " bytes: use ph_point_samples_device");
        unsigned char* base = static_cast<unsigned char*>(ctx->noalloc_scratch);
        CallWorkspace ws{reinterpret_cast<double*>(base + ha_bytes), reinterpret_cast<double*>(base)};
        struct NoAlloc {
            bool saved = tl_no_alloc;
            NoAlloc() { tl_no_alloc = true; }
            ~NoAlloc() { tl_no_alloc = saved; }
        } no_alloc;
        PointList pl{px, py, pz};
        LaunchScope scope(ctx->launch);
        for (const CatChunk& c : chunks)
            launch_bucket(B, c, strategy, precision, &pl, nullptr, S, out, nullptr, ctx->stream, nullptr, &ws);
        return PH_OK;
    });
}

int ph_point_samples(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                     int64_t S, const double* px, const double* py, const double* pz, double* out_host) {
    return guarded(ctx, [&] {
        need(S >= 0, PH_ERR_ARGUMENT, "negative sample count");
        need(cat != nullptr, PH_ERR_ARGUMENT, "catalog is null");
        need(S == 0 || (px && py && pz && out_host), PH_ERR_ARGUMENT, "null sample or output pointer");
        const int L = catalog_L(cat);
        DevBuf dp, dout;
        dp.alloc(sizeof(double) * 3 * (size_t)(S > 0 ? S : 1));
        dout.alloc(out_elem(precision) * (size_t)L * (size_t)(S > 0 ? S : 1));
        double* d = dp.as<double>();
        if (S > 0) {
                PH_HIP(hipMemcpy(d, px, sizeof(double) * S, hipMemcpyHostToDevice));
            PH_HIP(hipMemcpy(d + S, py, sizeof(double) * S, hipMemcpyHostToDevice));
            PH_HIP(hipMemcpy(d + 2 * S, pz, sizeof(double) * S, hipMemcpyHostToDevice));
        }
        int rc = ph_point_samples_device(ctx, B, cat, strategy, precision, S, d, d + S, d + 2 * S, dout.p);
        if (rc != PH_OK) return rc;
        if (S > 0) download(ctx, dout.p, precision, (int64_t)L * S, out_host);
        return PH_OK;
    });
}

int ph_structured_grid_device(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                              const double* ax, int64_t nx, const double* ay, int64_t ny, const double* az, int64_t nz,
                              void* out) {
    return ph_structured_grid_coords_device(ctx, B, cat, strategy, precision, PH_COORDS_CARTESIAN, ax, nx, ay, ny, az,
                                            nz, out);
}

int ph_structured_grid_coords_device(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy,
                                     int32_t precision, int32_t coords, const double* ax, int64_t nx,
                                     const double* ay, int64_t ny, const double* az, int64_t nz, void* out) {
    return guarded(ctx, [&] {
        need(coords >= PH_COORDS_CARTESIAN && coords <= PH_COORDS_SPHERICAL, PH_ERR_ARGUMENT,
             "unknown coordinate system");
        need(B != nullptr && B->ctx == ctx, PH_ERR_ARGUMENT, "tree does not belong to this context");
        need(strategy >= PH_GATHER && strategy <= PH_SYMMETRIC, PH_ERR_ARGUMENT, "unknown strategy");
        need(nx >= 0 && ny >= 0 && nz >= 0, PH_ERR_ARGUMENT, "negative grid size");
        check_precision(B, precision);
        check_catalog(B, cat);
        const i64 S = nx * ny * nz;
        need(S == 0 || (ax && ay && az && out), PH_ERR_ARGUMENT, "null axis or output pointer");
        GridAxes ga{ax, ay, az, nx, ny, nz, coords};
        run_chunks(ctx, cat, S, [&](const CatChunk& c, int L, u64* counter, hipEvent_t* ev) {
            launch_point_like(ctx, B, c, strategy, precision, nullptr, &ga, S, L, out, counter, ev);
        });
        return PH_OK;
    });
}

int ph_structured_grid(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                       const double* ax, int64_t nx, const double* ay, int64_t ny, const double* az, int64_t nz,
                       double* out_host) {
    return ph_structured_grid_coords(ctx, B, cat, strategy, precision, PH_COORDS_CARTESIAN, ax, nx, ay, ny, az, nz,
                                     out_host);
}

int ph_structured_grid_coords(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy,
                              int32_t precision, int32_t coords, const double* ax, int64_t nx, const double* ay,
                              int64_t ny, const double* az, int64_t nz, double* out_host) {
    return guarded(ctx, [&] {
        // angle axes hold (sin, cos) pairs
        const int64_t my = coords == PH_COORDS_CARTESIAN ? ny : 2 * ny;
        const int64_t mz = coords == PH_COORDS_SPHERICAL ? 2 * nz : nz;
        need(cat != nullptr, PH_ERR_ARGUMENT, "catalog is null");
        need(nx >= 0 && ny >= 0 && nz >= 0, PH_ERR_ARGUMENT, "negative grid size");
        const i64 S = nx * ny * nz;
        need(S == 0 || (ax && ay && az && out_host), PH_ERR_ARGUMENT, "null axis or output pointer");
        const int L = catalog_L(cat);
        DevBuf da, dout;
        da.alloc(sizeof(double) * (size_t)(nx + my + mz + 1));
        dout.alloc(out_elem(precision) * (size_t)L * (size_t)(S > 0 ? S : 1));
        double* d = da.as<double>();
        if (nx) PH_HIP(hipMemcpy(d, ax, sizeof(double) * nx, hipMemcpyHostToDevice));
        if (my) PH_HIP(hipMemcpy(d + nx, ay, sizeof(double) * my, hipMemcpyHostToDevice));
        if (mz) PH_HIP(hipMemcpy(d + nx + my, az, sizeof(double) * mz, hipMemcpyHostToDevice));
        int rc = ph_structured_grid_coords_device(ctx, B, cat, strategy, precision, coords, d, nx, d + nx, ny,
                                                  d + nx + my, nz, dout.p);
        if (rc != PH_OK) return rc;
        if (S > 0) download(ctx, dout.p, precision, (int64_t)L * S, out_host);
        return PH_OK;
    });
}

static void check_line_catalog(const ph_lbvh* B, const ph_catalog* cat, int32_t strategy) {
    need(strategy == PH_SCATTER, PH_ERR_STRATEGY,
         "LineSamples_interpolation only supports itpScatter. Line-integrated samples do not have a well-defined "
         "query smoothing length ha, so itpGather and itpSymmetric are not supported.");
    check_catalog(B, cat);
    need(cat->n_grad == 0 && cat->n_div == 0 && cat->n_curl == 0, PH_ERR_ARGUMENT,
         "line-integrated interpolation supports scalar quantities only");
}

int ph_line_samples_device(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                           int64_t S, const double* ox, const double* oy, const double* oz, const double* dx,
                           const double* dy, const double* dz, void* out) {
    return guarded(ctx, [&] {
        need(B != nullptr && B->ctx == ctx, PH_ERR_ARGUMENT, "tree does not belong to this context");
        need(S >= 0, PH_ERR_ARGUMENT, "negative sample count");
        check_line_catalog(B, cat, strategy);
        check_precision(B, precision);
        need(S == 0 || (ox && oy && oz && dx && dy && dz && out), PH_ERR_ARGUMENT, "null ray or output pointer");
        LineList ll{ox, oy, oz, dx, dy, dz};
        const void* lut = precision == PH_PREC_FP64 ? ctx->lut64.p : ctx->lut32.p;
        LaunchScope scope(ctx->launch);  // block order
        run_chunks(ctx, cat, S, [&](const CatChunk& c, int L, u64* counter, hipEvent_t* ev) {
            launch_lines(B, c, strategy, precision, &ll, nullptr, S, L, out, counter, lut,
                         ctx->traversal == PH_TRAVERSAL_WAVE, ctx->stream, ev);
        });
        return PH_OK;
    });
}

int ph_line_samples(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision, int64_t S,
                    const double* ox, const double* oy, const double* oz, const double* dx, const double* dy,
                    const double* dz, double* out_host) {
    return guarded(ctx, [&] {
        need(B != nullptr && B->ctx == ctx, PH_ERR_ARGUMENT, "tree does not belong to this context");
        need(S >= 0, PH_ERR_ARGUMENT, "negative sample count");
        check_line_catalog(B, cat, strategy);
        need(S == 0 || (ox && oy && oz && dx && dy && dz && out_host), PH_ERR_ARGUMENT, "null ray or output pointer");
        const int L = catalog_L(cat);
        DevBuf dr, dout;
        dr.alloc(sizeof(double) * 6 * (size_t)(S > 0 ? S : 1));
        dout.alloc(out_elem(precision) * (size_t)L * (size_t)(S > 0 ? S : 1));
        double* d = dr.as<double>();
        const double* srcs[6] = {ox, oy, oz, dx, dy, dz};
        for (int k = 0; k < 6 && S > 0; ++k)
            PH_HIP(hipMemcpy(d + (size_t)k * S, srcs[k], sizeof(double) * S, hipMemcpyHostToDevice));
        int rc = ph_line_samples_device(ctx, B, cat, strategy, precision, S, d, d + S, d + 2 * S, d + 3 * S, d + 4 * S,
                                        d + 5 * S, dout.p);
        if (rc != PH_OK) return rc;
        if (S > 0) download(ctx, dout.p, precision, (int64_t)L * S, out_host);
        return PH_OK;
    });
}

// Image-plane rays: origin + a u + b v along dir, (a, b) = (ax[i], ay[j]) or,
// polar, (ax[i] ay[2j + 1], ax[i] ay[2j]) with ay the (sin, cos) pairs.
static int line_grid(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                     const double* origin3, const double* u3, const double* v3, const double* dir3, const double* ax,
                     int64_t nx, const double* ay, int64_t ny, int polar, void* out) {
    return guarded(ctx, [&] {
        need(B != nullptr && B->ctx == ctx, PH_ERR_ARGUMENT, "tree does not belong to this context");
        need(nx >= 0 && ny >= 0, PH_ERR_ARGUMENT, "negative image size");
        check_line_catalog(B, cat, strategy);
        check_precision(B, precision);
        need(origin3 && u3 && v3 && dir3, PH_ERR_ARGUMENT, "null frame vector");
        const i64 S = nx * ny;
        need(S == 0 || (ax && ay && out), PH_ERR_ARGUMENT, "null axis or output pointer");
        LineGrid lg{};
        for (int d = 0; d < 3; ++d) {
            lg.o[d] = origin3[d];
            lg.u[d] = u3[d];
            lg.v[d] = v3[d];
            lg.d[d] = dir3[d];
        }
        lg.ax = ax;
        lg.ay = ay;
        lg.nx = nx;
        lg.ny = ny;
        lg.polar = polar;
        const void* lut = precision == PH_PREC_FP64 ? ctx->lut64.p : ctx->lut32.p;
        LaunchScope scope(ctx->launch);  // block order
        run_chunks(ctx, cat, S, [&](const CatChunk& c, int L, u64* counter, hipEvent_t* ev) {
            launch_lines(B, c, strategy, precision, nullptr, &lg, S, L, out, counter, lut,
                         ctx->traversal != PH_TRAVERSAL_LANE, ctx->stream, ev);
        });
        return PH_OK;
    });
}

int ph_line_grid_device(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                        const double* origin3, const double* u3, const double* v3, const double* dir3,
                        const double* ax, int64_t nx, const double* ay, int64_t ny, void* out) {
    return line_grid(ctx, B, cat, strategy, precision, origin3, u3, v3, dir3, ax, nx, ay, ny, 0, out);
}

int ph_line_polar_device(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                         const double* origin3, const double* u3, const double* v3, const double* dir3,
                         const double* s_dev, int64_t ns, const double* trig_dev, int64_t nphi, void* out) {
    return line_grid(ctx, B, cat, strategy, precision, origin3, u3, v3, dir3, s_dev, ns, trig_dev, nphi, 1, out);
}

// Adjoint of a one-scalar line projection (shared by the list and image-plane forms).
static void run_line_adjoint(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                             const LineList* ll, const LineGrid* lg, i64 S, const void* cot, double* grad) {
    need(B != nullptr && B->ctx == ctx, PH_ERR_ARGUMENT, "tree does not belong to this context");
    check_line_catalog(B, cat, strategy);
    need(cat->n_scalar == 1, PH_ERR_ARGUMENT, "the line adjoint takes a catalog of exactly one scalar");
    check_precision(B, precision);
    need(precision != PH_PREC_FP32C, PH_ERR_ARGUMENT, "the line adjoint runs in PH_PREC_FP64 or PH_PREC_FP32");
    need(grad != nullptr && (S == 0 || cot != nullptr), PH_ERR_ARGUMENT, "null cotangent or gradient pointer");
    const void* lut = precision == PH_PREC_FP64 ? ctx->lut64.p : ctx->lut32.p;
    const i32 qsrc = 5 + cat->scalar_slots[0];
    const bool shep = cat->scalar_shepard[0] != 0;
    LaunchScope scope(ctx->launch);  // block order
    PH_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    launch_lines_adjoint(B, qsrc, shep, precision, ll, lg, S, cot, lut, grad, ctx->stream, &ctx->ev[4]);
    PH_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    PH_HIP(hipEventSynchronize(ctx->ev[1]));
    ctx->stats.interp_ms = elapsed(ctx, 0, 1);
    ctx->stats.kernel_ms = elapsed(ctx, 4, 5);
    ctx->stats.samples = S;
    ctx->stats.accepted = -1;
    ctx->stats.rechecks = -1;
}

int ph_line_samples_adjoint_device(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy,
                                   int32_t precision, int64_t S, const double* ox, const double* oy, const double* oz,
                                   const double* dx, const double* dy, const double* dz, const void* cot,
                                   double* grad) {
    return guarded(ctx, [&] {
        need(S >= 0, PH_ERR_ARGUMENT, "negative sample count");
        need(S == 0 || (ox && oy && oz && dx && dy && dz), PH_ERR_ARGUMENT, "null line pointer");
        LineList ll{ox, oy, oz, dx, dy, dz};
        run_line_adjoint(ctx, B, cat, strategy, precision, &ll, nullptr, S, cot, grad);
        return PH_OK;
    });
}

static int line_grid_adjoint(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy,
                                int32_t precision, const double* origin3, const double* u3, const double* v3,
                                const double* dir3, const double* ax, int64_t nx, const double* ay, int64_t ny,
                                const void* cot, double* grad, int polar) {
    return guarded(ctx, [&] {
        need(nx >= 0 && ny >= 0, PH_ERR_ARGUMENT, "negative image size");
        need(origin3 && u3 && v3 && dir3, PH_ERR_ARGUMENT, "null frame vector");
        const i64 S = nx * ny;
        need(S == 0 || (ax && ay), PH_ERR_ARGUMENT, "null axis pointer");
        LineGrid lg{};
        for (int d = 0; d < 3; ++d) {
            lg.o[d] = origin3[d];
            lg.u[d] = u3[d];
            lg.v[d] = v3[d];
            lg.d[d] = dir3[d];
        }
        lg.ax = ax;
        lg.ay = ay;
        lg.nx = nx;
        lg.ny = ny;
        lg.polar = polar;
        run_line_adjoint(ctx, B, cat, strategy, precision, nullptr, &lg, S, cot, grad);
        return PH_OK;
    });
}

int ph_line_grid_adjoint_device(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy,
                                int32_t precision, const double* origin3, const double* u3, const double* v3,
                                const double* dir3, const double* ax, int64_t nx, const double* ay, int64_t ny,
                                const void* cot, double* grad) {
    return line_grid_adjoint(ctx, B, cat, strategy, precision, origin3, u3, v3, dir3, ax, nx, ay, ny, cot, grad, 0);
}

int ph_line_polar_adjoint_device(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy,
                                 int32_t precision, const double* origin3, const double* u3, const double* v3,
                                 const double* dir3, const double* s_dev, int64_t ns, const double* trig_dev,
                                 int64_t nphi, const void* cot, double* grad) {
    return line_grid_adjoint(ctx, B, cat, strategy, precision, origin3, u3, v3, dir3, s_dev, ns, trig_dev, nphi, cot,
                             grad, 1);
}

// Adjoint of one output group of point / grid interpolation (shared by both forms).
static void run_points_adjoint(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                               const PointList* pl, const GridAxes* ga, i64 S, const void* cot, double* grad) {
    need(B != nullptr && B->ctx == ctx, PH_ERR_ARGUMENT, "tree does not belong to this context");
    need(strategy >= PH_GATHER && strategy <= PH_SYMMETRIC, PH_ERR_ARGUMENT, "unknown strategy");
    check_catalog(B, cat);
    need(cat->n_scalar + cat->n_grad + cat->n_div + cat->n_curl == 1, PH_ERR_ARGUMENT,
         "the point adjoint takes a catalog of exactly one scalar, gradient, divergence or curl");
    check_precision(B, precision);
    need(precision != PH_PREC_FP32, PH_ERR_ARGUMENT, "the point adjoint runs in PH_PREC_FP64 or PH_PREC_FP32C");
    need(grad != nullptr && (S == 0 || cot != nullptr), PH_ERR_ARGUMENT, "null cotangent or gradient pointer");
    AdjointOut ao;
    if (cat->n_scalar == 1) {
        ao.q[0] = 5 + cat->scalar_slots[0];
        ao.shepard = cat->scalar_shepard[0] != 0;
    } else if (cat->n_grad == 1) {
        need(cat->grad_slots[0] >= 1, PH_ERR_ARGUMENT, "the adjoint of the density gradient (slot 0) is not provided");
        ao.mode = 1;
        ao.q[0] = 5 + cat->grad_slots[0];
    } else {
        const int32_t* sl = cat->n_div == 1 ? cat->div_slots : cat->curl_slots;
        ao.mode = cat->n_div == 1 ? 2 : 3;
        for (int j = 0; j < 3; ++j) ao.q[j] = 5 + sl[j];
    }
    LaunchScope scope(ctx->launch);  // block order
    PH_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    launch_points_adjoint(B, strategy, ao, precision, pl, ga, S, cot, grad, ctx->stream, &ctx->ev[4]);
    PH_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    PH_HIP(hipEventSynchronize(ctx->ev[1]));
    ctx->stats.interp_ms = elapsed(ctx, 0, 1);
    ctx->stats.kernel_ms = S > 0 ? elapsed(ctx, 4, 5) : 0.0;
    ctx->stats.samples = S;
    ctx->stats.accepted = -1;
    ctx->stats.rechecks = -1;
}

int ph_point_samples_adjoint_device(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy,
                                    int32_t precision, int64_t S, const double* px, const double* py,
                                    const double* pz, const void* cot, double* grad) {
    return guarded(ctx, [&] {
        need(S >= 0, PH_ERR_ARGUMENT, "negative sample count");
        need(S == 0 || (px && py && pz), PH_ERR_ARGUMENT, "null sample pointer");
        PointList pl{px, py, pz};
        run_points_adjoint(ctx, B, cat, strategy, precision, &pl, nullptr, S, cot, grad);
        return PH_OK;
    });
}

int ph_structured_grid_adjoint_device(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy,
                                      int32_t precision, const double* ax, int64_t nx, const double* ay, int64_t ny,
                                      const double* az, int64_t nz, const void* cot, double* grad) {
    return guarded(ctx, [&] {
        need(nx >= 0 && ny >= 0 && nz >= 0, PH_ERR_ARGUMENT, "negative grid size");
        const i64 S = nx * ny * nz;
        need(S == 0 || (ax && ay && az), PH_ERR_ARGUMENT, "null axis pointer");
        GridAxes ga{ax, ay, az, nx, ny, nz};
        run_points_adjoint(ctx, B, cat, strategy, precision, nullptr, &ga, S, cot, grad);
        return PH_OK;
    });
}

int ph_nearest_h(ph_context* ctx, ph_lbvh* B, int64_t S, const double* px, const double* py, const double* pz,
                 double* ha) {
    return guarded(ctx, [&] {
        need(B != nullptr && B->ctx == ctx, PH_ERR_ARGUMENT, "tree does not belong to this context");
        // pz may be null for a 2D tree (points in the plane z = 0)
        need(S >= 0 && (S == 0 || (px && py && (pz || B->dims == 2) && ha)), PH_ERR_ARGUMENT, "bad arguments");
        if (S == 0) return PH_OK;
        DevBuf d;
        d.alloc(sizeof(double) * 4 * (size_t)S);
        double* p = d.as<double>();
        PH_HIP(hipMemcpy(p, px, sizeof(double) * S, hipMemcpyHostToDevice));
        PH_HIP(hipMemcpy(p + S, py, sizeof(double) * S, hipMemcpyHostToDevice));
        if (pz) PH_HIP(hipMemcpy(p + 2 * S, pz, sizeof(double) * S, hipMemcpyHostToDevice));
        else PH_HIP(hipMemset(p + 2 * S, 0, sizeof(double) * S));
        PointList pl{p, p + S, p + 2 * S};
        launch_nearest_h(B, pl, S, p + 3 * S, ctx->stream);
        PH_HIP(hipStreamSynchronize(ctx->stream));
        PH_HIP(hipMemcpy(ha, p + 3 * S, sizeof(double) * S, hipMemcpyDeviceToHost));
        return PH_OK;
    });
}

// The query points of the LBVH queries; a 2D tree takes pz = null (zeros).
static PointList query_points(ph_lbvh* B, i64 S, const double* px, const double* py, const double* pz, DevBuf& zero,
                              hipStream_t s) {
    need(B->precision == PH_PREC_FP64, PH_ERR_ARGUMENT, "the LBVH queries need a tree built from Float64 columns");
    need(S >= 0, PH_ERR_ARGUMENT, "negative point count");
    need(S == 0 || (px && py && (pz || B->dims == 2)), PH_ERR_ARGUMENT, "null point pointer");
    if (S > 0 && !pz) {
        zero.alloc(sizeof(double) * (size_t)S);
        PH_HIP(hipMemsetAsync(zero.p, 0, sizeof(double) * (size_t)S, s));
        pz = zero.as<double>();
    }
    return PointList{px, py, pz};
}

int ph_query_device(ph_context* ctx, ph_lbvh* B, int64_t S, const double* px, const double* py, const double* pz,
                    const double* radius, int64_t* count, int64_t* closest, double* closest_d2,
                    const int64_t* offsets, int64_t* pool) {
    return guarded(ctx, [&] {
        need(B != nullptr && B->ctx == ctx, PH_ERR_ARGUMENT, "tree does not belong to this context");
        need(S == 0 || radius, PH_ERR_ARGUMENT, "null radius pointer");
        need(!pool || offsets, PH_ERR_ARGUMENT, "a pool needs offsets");
        DevBuf zero;
        const PointList pl = query_points(B, S, px, py, pz, zero, ctx->stream);
        if (S > 0) launch_query(B, pl, S, radius, count, closest, closest_d2, offsets, pool, ctx->stream);
        PH_HIP(hipStreamSynchronize(ctx->stream));
        return PH_OK;
    });
}

int ph_find_nearest_device(ph_context* ctx, ph_lbvh* B, int64_t S, const double* px, const double* py,
                           const double* pz, int64_t* idx, double* d2) {
    return guarded(ctx, [&] {
        need(B != nullptr && B->ctx == ctx, PH_ERR_ARGUMENT, "tree does not belong to this context");
        need(S == 0 || idx, PH_ERR_ARGUMENT, "null index pointer");
        DevBuf zero;
        const PointList pl = query_points(B, S, px, py, pz, zero, ctx->stream);
        if (S > 0) launch_find_nearest(B, pl, S, idx, d2, ctx->stream);
        PH_HIP(hipStreamSynchronize(ctx->stream));
        return PH_OK;
    });
}

// ---- single-quantity kernels ----

// Launches on the context's stream between the call events; stats as for the
// other interpolation calls (no neighbour count).
extern "C++" template <class F> static void timed_call(ph_context* ctx, i64 S, F&& launch) {
    PH_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    launch();
    PH_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    PH_HIP(hipEventSynchronize(ctx->ev[1]));
    ctx->stats.interp_ms = ctx->stats.kernel_ms = elapsed(ctx, 0, 1);
    ctx->stats.samples = S;
    ctx->stats.accepted = -1;
    ctx->stats.rechecks = -1;
}

static void check_single(ph_context* ctx, const ph_lbvh* B, int32_t strategy, int32_t precision, int64_t S) {
    need(B != nullptr && B->ctx == ctx, PH_ERR_ARGUMENT, "tree does not belong to this context");
    need(B->dims == 3, PH_ERR_ARGUMENT, "interpolation needs a 3D tree: this one was built from an InterpolationInput{2}");
    need(strategy >= PH_GATHER && strategy <= PH_SYMMETRIC, PH_ERR_ARGUMENT, "unknown strategy");
    need(S >= 0, PH_ERR_ARGUMENT, "negative sample count");
    need(precision == PH_PREC_FP64 || precision == PH_PREC_FP32, PH_ERR_ARGUMENT,
         "single-quantity kernels evaluate in the tree's float type: PH_PREC_FP64 or PH_PREC_FP32");
    check_precision(B, precision);
}

int ph_point_single_device(ph_context* ctx, ph_lbvh* B, int32_t kind, int32_t strategy, int32_t precision, int64_t S,
                           const double* px, const double* py, const double* pz, const double* ha, void* out) {
    return guarded(ctx, [&] {
        check_single(ctx, B, strategy, precision, S);
        need(kind >= PH_SINGLE_DENSITY && kind <= PH_SINGLE_GRADIENT_DENSITY, PH_ERR_ARGUMENT,
             "unknown single-quantity kind");
        need(S == 0 || (px && py && pz && out), PH_ERR_ARGUMENT, "null sample or output pointer");
        PointList pl{px, py, pz};
        timed_call(ctx, S, [&] { launch_point_single(B, kind, strategy, pl, ha, S, out, ctx->stream); });
        return PH_OK;
    });
}

int ph_point_single(ph_context* ctx, ph_lbvh* B, int32_t kind, int32_t strategy, int32_t precision, int64_t S,
                    const double* px, const double* py, const double* pz, const double* ha, double* out_host) {
    return guarded(ctx, [&] {
        check_single(ctx, B, strategy, precision, S);
        need(S == 0 || (px && py && pz && out_host), PH_ERR_ARGUMENT, "null sample or output pointer");
        const int L = kind == PH_SINGLE_GRADIENT_DENSITY ? 3 : 1;
        const size_t n = (size_t)(S > 0 ? S : 1);
        DevBuf dp, dout;
        dp.alloc(sizeof(double) * 4 * n);
        dout.alloc(out_elem(precision) * (size_t)L * n);
        double* d = dp.as<double>();
        const double* src[4] = {px, py, pz, ha};
        for (int k = 0; k < 4 && S > 0; ++k)
            if (src[k]) PH_HIP(hipMemcpy(d + (size_t)k * S, src[k], sizeof(double) * S, hipMemcpyHostToDevice));
        int rc = ph_point_single_device(ctx, B, kind, strategy, precision, S, d, d + S, d + 2 * S,
                                        ha ? d + 3 * S : nullptr, dout.p);
        if (rc != PH_OK) return rc;
        if (S > 0) download(ctx, dout.p, precision, (int64_t)L * S, out_host);
        return PH_OK;
    });
}

// Line kernels with the caller's h_a: the catalog's quantities (cat) or, with
// cat null, the line density.
static int line_single_device(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                              int64_t S, const double* ox, const double* oy, const double* oz, const double* dx,
                              const double* dy, const double* dz, const double* ha, void* out) {
    return guarded(ctx, [&] {
        check_single(ctx, B, strategy, precision, S);
        if (cat) {
            check_catalog(B, cat);
            need(cat->n_grad == 0 && cat->n_div == 0 && cat->n_curl == 0, PH_ERR_ARGUMENT,
                 "line-integrated interpolation supports scalar quantities only");
        }
        need(S == 0 || (ox && oy && oz && dx && dy && dz && out), PH_ERR_ARGUMENT, "null line or output pointer");
        need(S == 0 || strategy == PH_SCATTER || ha != nullptr, PH_ERR_ARGUMENT,
             "itpGather and itpSymmetric take the smoothing length ha of every line");
        LineList ll{ox, oy, oz, dx, dy, dz};
        const void* lut = precision == PH_PREC_FP64 ? ctx->lut64.p : ctx->lut32.p;
        std::vector<CatChunk> chunks;
        if (cat) split_catalog(cat, chunks);
        timed_call(ctx, S, [&] {
            if (!cat) launch_line_single(B, nullptr, strategy, ll, ha, S, out, lut, ctx->stream);
            for (const CatChunk& c : chunks) launch_line_single(B, &c, strategy, ll, ha, S, out, lut, ctx->stream);
        });
        return PH_OK;
    });
}

static int line_single_host(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                            int64_t S, const double* ox, const double* oy, const double* oz, const double* dx,
                            const double* dy, const double* dz, const double* ha, double* out_host) {
    return guarded(ctx, [&] {
        need(S >= 0, PH_ERR_ARGUMENT, "negative sample count");
        need(S == 0 || (ox && oy && oz && dx && dy && dz && out_host), PH_ERR_ARGUMENT, "null line or output pointer");
        const int L = cat ? catalog_L(cat) : 1;
        const size_t n = (size_t)(S > 0 ? S : 1);
        DevBuf dr, dout;
        dr.alloc(sizeof(double) * 7 * n);
        dout.alloc(out_elem(precision) * (size_t)(L > 0 ? L : 1) * n);
        double* d = dr.as<double>();
        const double* src[7] = {ox, oy, oz, dx, dy, dz, ha};
        for (int k = 0; k < 7 && S > 0; ++k)
            if (src[k]) PH_HIP(hipMemcpy(d + (size_t)k * S, src[k], sizeof(double) * S, hipMemcpyHostToDevice));
        int rc = line_single_device(ctx, B, cat, strategy, precision, S, d, d + S, d + 2 * S, d + 3 * S, d + 4 * S,
                                    d + 5 * S, ha ? d + 6 * S : nullptr, dout.p);
        if (rc != PH_OK) return rc;
        if (S > 0 && L > 0) download(ctx, dout.p, precision, (int64_t)L * S, out_host);
        return PH_OK;
    });
}

int ph_line_samples_ha_device(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                              int64_t S, const double* ox, const double* oy, const double* oz, const double* dx,
                              const double* dy, const double* dz, const double* ha, void* out) {
    if (!cat) return fail(ctx, PH_ERR_ARGUMENT, "catalog is null");
    return line_single_device(ctx, B, cat, strategy, precision, S, ox, oy, oz, dx, dy, dz, ha, out);
}

int ph_line_samples_ha(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                       int64_t S, const double* ox, const double* oy, const double* oz, const double* dx,
                       const double* dy, const double* dz, const double* ha, double* out_host) {
    if (!cat) return fail(ctx, PH_ERR_ARGUMENT, "catalog is null");
    return line_single_host(ctx, B, cat, strategy, precision, S, ox, oy, oz, dx, dy, dz, ha, out_host);
}

int ph_line_density_device(ph_context* ctx, ph_lbvh* B, int32_t strategy, int32_t precision, int64_t S,
                           const double* ox, const double* oy, const double* oz, const double* dx, const double* dy,
                           const double* dz, const double* ha, void* out) {
    return line_single_device(ctx, B, nullptr, strategy, precision, S, ox, oy, oz, dx, dy, dz, ha, out);
}

int ph_line_density(ph_context* ctx, ph_lbvh* B, int32_t strategy, int32_t precision, int64_t S, const double* ox,
                    const double* oy, const double* oz, const double* dx, const double* dy, const double* dz,
                    const double* ha, double* out_host) {
    return line_single_host(ctx, B, nullptr, strategy, precision, S, ox, oy, oz, dx, dy, dz, ha, out_host);
}

int ph_line_table(int32_t kernel, double* inorm) {
    if (kernel < 0 || kernel > 5 || !inorm) return fail(nullptr, PH_ERR_ARGUMENT, "bad arguments");
    std::memcpy(inorm, ph_lut_inorm[kernel], sizeof(double) * PH_LUT_N);
    return PH_OK;
}

int ph_batch_create(ph_context* ctx, ph_lbvh* B, const ph_catalog* cat, int32_t strategy, int32_t precision,
                    int32_t nreq, const ph_request* reqs, ph_batch** out) {
    if (out) *out = nullptr;
    return guarded(ctx, [&] {
        need(ctx != nullptr && out != nullptr, PH_ERR_ARGUMENT, "null context or output handle");
        need(B != nullptr && B->ctx == ctx, PH_ERR_ARGUMENT, "tree does not belong to this context");
        need(strategy >= PH_GATHER && strategy <= PH_SYMMETRIC, PH_ERR_ARGUMENT, "unknown strategy");
        need(nreq >= 0 && (nreq == 0 || reqs), PH_ERR_ARGUMENT, "bad request list");
        need(precision == PH_PREC_FP64 || precision == PH_PREC_FP32C, PH_ERR_ARGUMENT,
             "batches run in PH_PREC_FP64 or PH_PREC_FP32C");
        check_precision(B, precision);
        check_catalog(B, cat);
        need(B->n <= (int64_t(1) << 30), PH_ERR_CAPACITY, "batches support up to 2^30 particles");
        auto b = std::make_unique<ph_batch>();
        b->ctx = ctx;
        b->B = B;
        b->st = batch_build(ctx, B, cat, strategy, precision, nreq, reqs);
        ctx->batches.push_back(b.get());
        *out = b.release();
        return PH_OK;
    });
}

int ph_batch_run(ph_batch* b) {
    if (!b || !b->ctx) return fail(nullptr, PH_ERR_ARGUMENT, "null batch or batch of a destroyed context");
    ph_context* ctx = b->ctx;
    return guarded(ctx, [&] {
        need(b->B != nullptr, PH_ERR_ARGUMENT, "the batch's tree was destroyed");
        batch_run(b->st, ctx, b->B);
        return PH_OK;
    });
}

int ph_batch_times(ph_batch* b, double* first_us, double* last_us) {
    if (!b || !b->ctx) return fail(nullptr, PH_ERR_ARGUMENT, "null batch or batch of a destroyed context");
    ph_context* ctx = b->ctx;
    return guarded(ctx, [&] {
        batch_times(b->st, ctx, first_us, last_us);
        return PH_OK;
    });
}

int ph_batch_destroy(ph_batch* b) {
    if (!b) return PH_OK;
    if (ph_context* ctx = b->ctx) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        auto& v = ctx->batches;
        for (size_t t = 0; t < v.size(); ++t)
            if (v[t] == b) {
                v.erase(v.begin() + (std::ptrdiff_t)t);
                break;
            }
        batch_free(b->st);  // device blocks return to the context's cache; the stream is idle
        ctx->cache.settle();
    }
    delete b;
    return PH_OK;
}

}  // extern "C"
