// eg_internal.hpp -- host-side internals of libeiger_groupby.so: errors,
// the device memory manager, kernel timing, and the lowered group-by plan.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>

#include <cstdint>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "eg_spec.hpp"
#include "eiger_groupby.h"

namespace eg {

// ---- errors ------------------------------------------------------------------
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] inline void raise(int status, const std::string& msg) { throw Error(status, msg); }

// A request list with more aggregation states or value columns than one
// kernel pass holds; the top level splits it (api.hip).
struct WideError : Error {
    explicit WideError(const std::string& m) : Error(EG_ERR_UNSUPPORTED, m) {}
};

#define EG_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t eg_e_ = (call);                                                           \
        if (eg_e_ != hipSuccess)                                                             \
            ::eg::raise(EG_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(eg_e_)); \
    } while (0)

#define EG_LAUNCH_CHECK() EG_HIP(hipGetLastError())

void set_last_error(const std::string& m);

// Runs f, converting exceptions into a status code + last-error message.
template <class F>
int guard(F&& f)
{
    try {
        f();
        return EG_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc& e) {
        set_last_error(std::string("out of memory: ") + e.what());
        return EG_ERR_OOM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return EG_ERR_INVALID;
    }
}

// ---- types ---------------------------------------------------------------------
int type_size(int32_t t);
bool is_signed_int(int32_t t);
bool is_unsigned_int(int32_t t);
inline bool is_int(int32_t t) { return is_signed_int(t) || is_unsigned_int(t); }
inline bool is_float(int32_t t) { return t == EG_FLOAT32 || t == EG_FLOAT64; }
int32_t agg_output_type(int32_t agg_kind, int32_t type);  // EG_TYPE_EMPTY if unsupported

// ---- memory manager -------------------------------------------------------------
void* mm_alloc(size_t bytes, hipStream_t s);
void mm_free(void* p);
void mm_usage(int64_t* cur, int64_t* peak);
void mm_set_mode(int mode, int64_t pool_bytes);

// RAII device buffer from the memory manager.
class DevBuf {
public:
    DevBuf() = default;
    DevBuf(size_t bytes, hipStream_t s) : p_(bytes ? mm_alloc(bytes, s) : nullptr), n_(bytes) {}
    ~DevBuf() { reset(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) { reset(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void reset()
    {
        if (p_) mm_free(p_);
        p_ = nullptr;
        n_ = 0;
    }
    void* release()
    {
        void* p = p_;
        p_ = nullptr;
        n_ = 0;
        return p;
    }
    template <class T = void>
    T* get() const { return static_cast<T*>(p_); }
    size_t size() const { return n_; }

private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// Pinned host scratch for small device->host reads (one per thread).
int64_t* pinned_i64(int count);
// Thread-local pinned staging buffer of at least `bytes` (reused; contents
// valid until the next call on this thread with slot `slot`).
void* pinned_staging(size_t bytes, int slot);

// ---- device facts ------------------------------------------------------------
struct DeviceInfo {
    int device;
    int num_cus;
    int max_lds;      // bytes per workgroup
    int warp_size;
};
const DeviceInfo& device_info();

// Raises the kernel's dynamic LDS limit to what its static LDS leaves of the
// workgroup's LDS; returns that many bytes.  Cached per kernel.
size_t enable_dynamic_lds(const void* kernel);

// ---- kernel timing -----------------------------------------------------------
bool timing_enabled();
// Records hipEvents around a launch on `s` when timing is enabled.
class KernelTimer {
public:
    KernelTimer(const char* name, hipStream_t s);
    ~KernelTimer();

private:
    const char* name_;
    hipStream_t s_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

// Wall time of a host-orchestrated phase, recorded under `name`.  It
// synchronises the stream at both ends, so it is on only while timing is
// enabled and EG_TIME_HOST_PHASES=1 (a diagnostic: with it, kernel timing
// would no longer time the library as it runs).
class HostPhase {
public:
    HostPhase(const char* name, hipStream_t s);
    ~HostPhase();

private:
    const char* name_;
    hipStream_t s_;
    bool on_ = false;
    std::chrono::steady_clock::time_point t0_;
};

// Counts one occurrence of an event (a retry, say) under `name` while timing
// is enabled; queried like a kernel, with 0 ms.
void timing_count(const char* name);

// ---- fills ------------------------------------------------------------------------
// Zero-fill with a kernel (graph-safe, no memset nodes).
void fill_zero(void* p, size_t bytes, hipStream_t s);

// ---- plan --------------------------------------------------------------------------
// One output column of the final table (after the key).
struct OutCol {
    int32_t code;     // OutCode
    int32_t type;     // gamax type id
    int32_t op;
    int32_t op2;
};

struct Plan {
    AggSpec spec{};
    KeyDesc key{};
    int32_t key_type = 0;
    int64_t rows = 0;
    std::vector<OutCol> outs;
};

struct ViewIn {
    const void* data;
    int64_t size;
    int32_t type;
};

// Lowers (key, value columns, requests) to a Plan (key == nullptr: scalar
// aggregate, no key column).  rewrite_avg: the top-level
// groupby() semantics (AVG -> SUM + count, trailing count when agg_count);
// otherwise the strategy-level contract (AVG rejected).
Plan make_plan(const ViewIn* key, const std::vector<ViewIn>& vals,
               const eg_agg_request* reqs, int64_t nreqs, bool agg_count, bool rewrite_avg);

// Result of a strategy: device columns [key, outs...] with num_groups rows.
struct Result {
    int64_t num_groups = 0;
    std::vector<DevBuf> cols;
    std::vector<int32_t> types;
};

// Allocates the output columns for `groups` rows and fills an OutSpec.
void alloc_outputs(const Plan& plan, int64_t groups, hipStream_t s, Result& res, OutSpec& os);
// Fills an OutSpec for output columns the caller owns (out_data: one pointer
// per plan output, in order).
void set_outspec(const Plan& plan, void* key_data, void* const* out_data, OutSpec& os);

// ---- strategies ----------------------------------------------------------------------
// Lanes of a wave on one group before the wave combines their rows (65 =
// never): 32.  Combining costs a DPP reduction per op; below ~32 lanes the
// LDS atomics' same-address serialisation is cheaper (2^28 rows, Zipf keys:
// hash D = 2^10 s = 1.5 2.2 ms at 32 against 2.8 ms at 8; partition
// D = 2^22 s = 1.0 2.1 ms at 32, 4.5 ms never combining).  EG_PEEL_MIN
// overrides it for tuning.
int default_peel_min();

struct HashTuning {
    int lds_slots = 0;        // cap on LDS slots per workgroup, 0 = what fits
    int replicas = 0;         // cap on LDS table replicas, 0 = up to 64
    int lds_max_ratio = 1 << 20;  // use LDS only if group bound <= ratio * slots
    int peel_min = default_peel_min();  // lanes on one group before the wave combines them (65 = off)
    int block = 1024;
    int blocks_per_cu = 0;    // 0 = occupancy
    int64_t table_slots = 0;  // 0 = from the group bound
};

Result run_hash(const Plan& plan, int64_t group_bound, hipStream_t s, const HashTuning& t = HashTuning{});

// The hash strategy into caller-owned outputs (os, `capacity` rows each) with
// its HBM table in the library's static device scratch: no device memory is
// allocated or freed.  Returns the group count; raises EG_ERR_CAPACITY when
// the table would exceed the scratch, when the groups overflow the table the
// scratch holds, or when they exceed `capacity` (outputs then untouched).
// Calls that overlap on one device share the scratch and race.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
int64_t run_hash_noalloc(const Plan& plan, int64_t group_bound, const OutSpec& os, int64_t capacity, hipStream_t s);

// ---- key mapping (keynorm.hip) -------------------------------------------------------
// Single key columns of 1- or 2-byte integers, BOOL8 or floats run on a 4- or
// 8-byte integer column that keeps equality and order (floats: -0.0 = +0.0,
// one NaN); group keys are mapped back.
bool key_needs_mapping(int32_t type);
int32_t mapped_key_type(int32_t type);
DevBuf map_keys(const void* keys, int32_t type, int64_t n, hipStream_t s);
DevBuf unmap_keys(const void* mapped, int64_t n, int32_t type, hipStream_t s);

// ---- generator -------------------------------------------------------------------------
void gen_keys_uniform(void* out, int32_t type, int64_t begin, int64_t count, int64_t n,
                      int64_t distinct, uint64_t seed, hipStream_t s);
void gen_keys_zipf(void* out, int32_t type, int64_t begin, int64_t count, const double* d_cdf,
                   int64_t distinct, uint64_t seed, hipStream_t s);
void gen_values(void* out, int32_t type, int32_t dist, int64_t begin, int64_t count,
                uint64_t seed, uint32_t stream_id, hipStream_t s);

}  // namespace eg
