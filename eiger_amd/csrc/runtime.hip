// runtime.hip -- errors, type helpers, the device memory manager, kernel
// timing and fill kernels of libeiger_groupby.so.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <chrono>
#include <mutex>
#include <unordered_map>

#include "eg_internal.hpp"

namespace eg {

// ---- errors ---------------------------------------------------------------------
static thread_local std::string g_last_error;

void set_last_error(const std::string& m) { g_last_error = m; }

const char* last_error_cstr() { return g_last_error.c_str(); }

// ---- types ---------------------------------------------------------------------
int type_size(int32_t t)
{
    switch (t) {
    case EG_INT8: case EG_UINT8: case EG_BOOL8: return 1;
    case EG_INT16: case EG_UINT16: return 2;
    case EG_INT32: case EG_UINT32: case EG_FLOAT32: return 4;
    case EG_INT64: case EG_UINT64: case EG_FLOAT64: return 8;
    default: return 0;
    }
}

bool is_signed_int(int32_t t) { return t == EG_INT8 || t == EG_INT16 || t == EG_INT32 || t == EG_INT64; }
bool is_unsigned_int(int32_t t) { return t == EG_UINT8 || t == EG_UINT16 || t == EG_UINT32 || t == EG_UINT64; }

// agg_output_data_type (operators/groupby/aggregate.cu:20-53) with the SUM
// widening of agg_func.cuh:18-36.
int32_t agg_output_type(int32_t agg_kind, int32_t type)
{
    if (agg_kind == EG_AGG_COUNT_ALL) return EG_INT64;
    bool numeric = is_int(type) || is_float(type);
    if (!numeric) return EG_TYPE_EMPTY;
    switch (agg_kind) {
    case EG_AGG_AVG: return EG_FLOAT64;
    case EG_AGG_SUM:
        if (is_signed_int(type)) return EG_INT64;
        if (is_unsigned_int(type)) return EG_UINT64;
        return type;
    case EG_AGG_MAX: case EG_AGG_MIN: case EG_AGG_ANY: return type;
    default: return EG_TYPE_EMPTY;
    }
}

// ---- fill kernels ----------------------------------------------------------------
__global__ void fill_zero_kernel(uint4* p, int64_t n16, uint8_t* tail, int64_t ntail)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) p[i] = make_uint4(0u, 0u, 0u, 0u);
    if (blockIdx.x == 0 && (int64_t)threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

void fill_zero(void* p, size_t bytes, hipStream_t s)
{
    if (!bytes) return;
    int64_t n16 = (int64_t)(bytes / 16);
    int64_t ntail = (int64_t)(bytes - (size_t)n16 * 16);
    int64_t blocks = std::min<int64_t>(std::max<int64_t>((n16 + 255) / 256, 1), 8192);
    KernelTimer timer("fill_zero", s);
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (uint4*)p, n16,
                       (uint8_t*)p + n16 * 16, ntail);
    EG_LAUNCH_CHECK();
}

// ---- device facts ----------------------------------------------------------------
const DeviceInfo& device_info()
{
    static std::mutex mu;
    static std::map<int, DeviceInfo> cache;
    int dev = 0;
    EG_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    DeviceInfo d{};
    d.device = dev;
    EG_HIP(hipDeviceGetAttribute(&d.num_cus, hipDeviceAttributeMultiprocessorCount, dev));
    EG_HIP(hipDeviceGetAttribute(&d.max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    EG_HIP(hipDeviceGetAttribute(&d.warp_size, hipDeviceAttributeWarpSize, dev));
    if (d.warp_size != 64) raise(EG_ERR_UNSUPPORTED, "wave64 device required (gfx950)");
    return cache.emplace(dev, d).first->second;
}

int default_peel_min()
{
    static const int v = [] {
        const char* e = std::getenv("EG_PEEL_MIN");
        const int m = e ? std::atoi(e) : 0;
        return (m >= 2 && m <= 65) ? m : 32;
    }();
    return v;
}

size_t enable_dynamic_lds(const void* kernel)
{
    static std::mutex mu;
    static std::unordered_map<const void*, size_t> done;
    std::lock_guard<std::mutex> lk(mu);
    auto it = done.find(kernel);
    if (it != done.end()) return it->second;
    hipFuncAttributes fa{};
    EG_HIP(hipFuncGetAttributes(&fa, kernel));
    const size_t avail = (size_t)device_info().max_lds - fa.sharedSizeBytes;
    EG_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)avail));
    done.emplace(kernel, avail);
    return avail;
}

// ---- memory manager (mirrors gamax::MemoryManager, memory/mem_manager.hpp:12-158) --
namespace {
struct Alloc {
    size_t bytes;
    hipStream_t stream;
};

struct Pool {
    char* base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_list;  // offset -> length, coalesced
};

std::mutex g_mm_mu;
int g_mode = EG_MM_CACHE;
std::unordered_map<void*, Alloc> g_allocs;
int64_t g_cur = 0, g_peak = 0;
Pool g_pool;
bool g_pool_threshold_set = false;

constexpr size_t kAlign = 256;

// Size classes of the caching allocator: four per power of two (<= 25% slack).
size_t cache_class(size_t bytes)
{
    if (bytes <= 4096) return 4096;
    int lg = 63 - __builtin_clzll((unsigned long long)bytes);
    const size_t step = (size_t)1 << (lg - 2);
    return (bytes + step - 1) / step * step;
}

struct CacheKey {
    hipStream_t stream;
    size_t bytes;
    bool operator<(const CacheKey& o) const
    {
        return stream != o.stream ? std::less<hipStream_t>()(stream, o.stream) : bytes < o.bytes;
    }
};
std::map<CacheKey, std::vector<void*>> g_cache;
int64_t g_cached = 0;

void cache_release_all()
{
    // every cached block may still be referenced by queued work on its stream
    (void)hipDeviceSynchronize();
    for (auto& kv : g_cache)
        for (void* p : kv.second) (void)hipFree(p);
    g_cache.clear();
    g_cached = 0;
}

void* pool_alloc(size_t bytes)
{
    bytes = (bytes + kAlign - 1) / kAlign * kAlign;
    for (auto it = g_pool.free_list.begin(); it != g_pool.free_list.end(); ++it) {
        if (it->second >= bytes) {
            size_t off = it->first, len = it->second;
            g_pool.free_list.erase(it);
            if (len > bytes) g_pool.free_list.emplace(off + bytes, len - bytes);
            return g_pool.base + off;
        }
    }
    return nullptr;
}

void pool_free(void* p, size_t bytes)
{
    bytes = (bytes + kAlign - 1) / kAlign * kAlign;
    size_t off = (size_t)((char*)p - g_pool.base);
    auto it = g_pool.free_list.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != g_pool.free_list.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        g_pool.free_list.erase(nx);
    }
    if (it != g_pool.free_list.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            g_pool.free_list.erase(it);
        }
    }
}
}  // namespace

void* mm_alloc(size_t bytes, hipStream_t s)
{
    if (bytes == 0) bytes = kAlign;
    std::lock_guard<std::mutex> lk(g_mm_mu);
    void* p = nullptr;
    if (g_mode == EG_MM_POOL) {
        p = pool_alloc(bytes);
        if (!p) throw std::bad_alloc();
    } else if (g_mode == EG_MM_CACHE) {
        bytes = cache_class(bytes);
        auto it = g_cache.find(CacheKey{s, bytes});
        if (it != g_cache.end() && !it->second.empty()) {
            p = it->second.back();
            it->second.pop_back();
            g_cached -= (int64_t)bytes;
        } else {
            hipError_t e = hipMalloc(&p, bytes);
            if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
                (void)hipGetLastError();
                cache_release_all();
                e = hipMalloc(&p, bytes);
            }
            if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
                (void)hipGetLastError();
                throw std::bad_alloc();
            }
            if (e != hipSuccess) raise(EG_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
    } else {
        if (!g_pool_threshold_set) {
            int dev = 0;
            hipMemPool_t mp;
            if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&mp, dev) == hipSuccess) {
                uint64_t thr = UINT64_MAX;
                (void)hipMemPoolSetAttribute(mp, hipMemPoolAttrReleaseThreshold, &thr);
            }
            g_pool_threshold_set = true;
        }
        hipError_t e = hipMallocAsync(&p, bytes, s);
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
            (void)hipGetLastError();
            throw std::bad_alloc();
        }
        if (e != hipSuccess) raise(EG_ERR_DEVICE, std::string("hipMallocAsync: ") + hipGetErrorString(e));
    }
    g_allocs[p] = Alloc{bytes, s};
    g_cur += (int64_t)bytes;
    g_peak = std::max(g_peak, g_cur);
    return p;
}

void mm_free(void* p)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_mm_mu);
    auto it = g_allocs.find(p);
    if (it == g_allocs.end()) raise(EG_ERR_INVALID, "eg_free: pointer not allocated by this library");
    Alloc a = it->second;
    g_allocs.erase(it);
    g_cur -= (int64_t)a.bytes;
    if (g_mode == EG_MM_POOL) pool_free(p, a.bytes);
    else if (g_mode == EG_MM_CACHE) {
        g_cache[CacheKey{a.stream, a.bytes}].push_back(p);
        g_cached += (int64_t)a.bytes;
    } else EG_HIP(hipFreeAsync(p, a.stream));
}

void mm_usage(int64_t* cur, int64_t* peak)
{
    std::lock_guard<std::mutex> lk(g_mm_mu);
    if (cur) *cur = g_cur;
    if (peak) *peak = g_peak;
}

void mm_set_mode(int mode, int64_t pool_bytes)
{
    std::lock_guard<std::mutex> lk(g_mm_mu);
    if (!g_allocs.empty()) raise(EG_ERR_INVALID, "memory mode can only change while nothing is allocated");
    if (mode != EG_MM_ASYNC && mode != EG_MM_POOL && mode != EG_MM_CACHE) raise(EG_ERR_INVALID, "unknown memory mode");
    cache_release_all();
    if (g_pool.base) {
        EG_HIP(hipFree(g_pool.base));
        g_pool = Pool{};
    }
    if (mode == EG_MM_POOL) {
        if (pool_bytes <= 0) raise(EG_ERR_INVALID, "pool_bytes must be positive");
        size_t sz = ((size_t)pool_bytes + kAlign - 1) / kAlign * kAlign;
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, sz);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw std::bad_alloc();
        }
        g_pool.base = (char*)p;
        g_pool.size = sz;
        g_pool.free_list.emplace(0, sz);
    }
    g_mode = mode;
    g_peak = g_cur;
}

int64_t* pinned_i64(int count)
{
    static thread_local int64_t* buf = nullptr;
    static thread_local int cap = 0;
    if (count > cap) {
        if (buf) (void)hipHostFree(buf);
        buf = nullptr;
        int c = std::max(count, 64);
        EG_HIP(hipHostMalloc((void**)&buf, sizeof(int64_t) * c, hipHostMallocDefault));
        cap = c;
    }
    return buf;
}

void* pinned_staging(size_t bytes, int slot)
{
    struct Buf {
        void* p = nullptr;
        size_t cap = 0;
    };
    static thread_local Buf bufs[4];
    if (slot < 0 || slot >= 4) raise(EG_ERR_INVALID, "pinned staging slot");
    Buf& b = bufs[slot];
    if (bytes > b.cap) {
        if (b.p) (void)hipHostFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        const size_t c = std::max<size_t>(bytes + bytes / 2, 1 << 16);
        EG_HIP(hipHostMalloc(&b.p, c, hipHostMallocDefault));
        b.cap = c;
    }
    return b.p;
}

// ---- kernel timing -----------------------------------------------------------------
namespace {
struct TimingRec {
    std::string name;
    hipEvent_t a, b;
    double host_ms;  // host phases (a == nullptr)
};
std::mutex g_t_mu;
bool g_timing = false;
std::vector<TimingRec> g_recs;
std::vector<hipEvent_t> g_event_pool;
}  // namespace

bool timing_enabled() { return g_timing; }

// Events come from a pool refilled by timing_reset, so a timed launch costs
// two event records and no event creation.
static hipEvent_t pooled_event()
{
    std::lock_guard<std::mutex> lk(g_t_mu);
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    EG_HIP(hipEventCreate(&e));
    return e;
}

KernelTimer::KernelTimer(const char* name, hipStream_t s) : name_(name), s_(s)
{
    if (!g_timing) return;
    a_ = pooled_event();
    b_ = pooled_event();
    EG_HIP(hipEventRecord(a_, s_));
}

KernelTimer::~KernelTimer()
{
    if (!a_) return;
    if (hipEventRecord(b_, s_) != hipSuccess) return;
    std::lock_guard<std::mutex> lk(g_t_mu);
    g_recs.push_back(TimingRec{name_, a_, b_, 0.0});
}

HostPhase::HostPhase(const char* name, hipStream_t s) : name_(name), s_(s)
{
    static const bool enabled = [] {
        const char* e = std::getenv("EG_TIME_HOST_PHASES");
        return e && std::atoi(e) == 1;
    }();
    if (!g_timing || !enabled) return;
    on_ = true;
    (void)hipStreamSynchronize(s_);
    t0_ = std::chrono::steady_clock::now();
}

HostPhase::~HostPhase()
{
    if (!on_) return;
    (void)hipStreamSynchronize(s_);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
    std::lock_guard<std::mutex> lk(g_t_mu);
    g_recs.push_back(TimingRec{name_, nullptr, nullptr, ms});
}

void timing_count(const char* name)
{
    if (!g_timing) return;
    std::lock_guard<std::mutex> lk(g_t_mu);
    g_recs.push_back(TimingRec{name, nullptr, nullptr, 0.0});
}

int timing_enable(int on)
{
    if (on) {
        // events for the first ~1000 timed launches, created before timing starts
        std::lock_guard<std::mutex> lk(g_t_mu);
        while (g_event_pool.size() < 2048) {
            hipEvent_t e = nullptr;
            EG_HIP(hipEventCreate(&e));
            g_event_pool.push_back(e);
        }
    }
    g_timing = on != 0;
    return EG_OK;
}

int timing_reset()
{
    std::lock_guard<std::mutex> lk(g_t_mu);
    for (auto& r : g_recs) {
        if (!r.a) continue;  // host phase: no events
        // back to the pool once recorded work has passed them
        if (hipEventSynchronize(r.b) == hipSuccess) {
            g_event_pool.push_back(r.a);
            g_event_pool.push_back(r.b);
        } else {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
    }
    g_recs.clear();
    return EG_OK;
}

void timing_query(const char* name, int64_t* launches, double* total_ms)
{
    std::lock_guard<std::mutex> lk(g_t_mu);
    int64_t n = 0;
    double tot = 0.0;
    for (auto& r : g_recs) {
        if (r.name != name) continue;
        if (!r.a) {
            tot += r.host_ms;
            ++n;
            continue;
        }
        EG_HIP(hipEventSynchronize(r.b));
        float ms = 0.f;
        EG_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        tot += ms;
        ++n;
    }
    if (launches) *launches = n;
    if (total_ms) *total_ms = tot;
}

}  // namespace eg

// ---- bandwidth probe ------------------------------------------------------------------
namespace eg {
namespace {
__global__ __launch_bounds__(256) void probe_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                         int64_t n16)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}
}  // namespace
}  // namespace eg

extern "C" int eg_probe_copy(int64_t bytes, int32_t iters, void* stream, double* ms)
{
    // Four 256-thread workgroups per CU, one 16-byte element per thread per
    // iteration: the fastest plain copy among the grid sizes and per-thread
    // widths measured on MI355X (4 GiB: ~5.4 TB/s read + write, against ~4.7
    // with sixteen workgroups per CU).  *ms is the fastest of `iters` launches.
    return eg::guard([&] {
        if (bytes <= 0 || iters <= 0 || !ms) eg::raise(EG_ERR_INVALID, "bad arguments");
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        const int64_t n16 = bytes / 16;
        eg::DevBuf a((size_t)n16 * 16, s), b((size_t)n16 * 16, s);
        eg::fill_zero(a.get(), (size_t)n16 * 16, s);
        const eg::DeviceInfo& di = eg::device_info();
        const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((n16 + 255) / 256, (int64_t)di.num_cus * 4));
        hipEvent_t e0, e1;
        EG_HIP(hipEventCreate(&e0));
        EG_HIP(hipEventCreate(&e1));
        hipLaunchKernelGGL(eg::probe_copy_kernel, dim3((unsigned)grid), dim3(256), 0, s, a.get<const uint4>(),
                           b.get<uint4>(), n16);
        EG_LAUNCH_CHECK();
        float best = 0.f;
        for (int i = 0; i < iters; ++i) {
            eg::KernelTimer t("probe_copy", s);
            EG_HIP(hipEventRecord(e0, s));
            hipLaunchKernelGGL(eg::probe_copy_kernel, dim3((unsigned)grid), dim3(256), 0, s, a.get<const uint4>(),
                               b.get<uint4>(), n16);
            EG_LAUNCH_CHECK();
            EG_HIP(hipEventRecord(e1, s));
            EG_HIP(hipEventSynchronize(e1));
            float t1 = 0.f;
            EG_HIP(hipEventElapsedTime(&t1, e0, e1));
            if (i == 0 || t1 < best) best = t1;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = (double)best;
    });
}
