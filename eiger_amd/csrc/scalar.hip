// scalar.hip -- aggregate without keys (gamax::groupby::aggregate::aggregate,
// operators/groupby/aggregate.cu:131-165): one output row per request.
//
// One pass: each thread folds its rows into register states, a wave combines
// them with cross-lane shuffles, and one lane per wave merges into the global
// states with a device atomic.  The same AggSpec/OutSpec lowering as the
// group-by is used, so COUNT_ALL, AVG and the output typing match it.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "eg_device.hpp"
#include "eg_internal.hpp"

namespace eg {
namespace {

template <int NV>
__global__ __launch_bounds__(256) void scalar_kernel(AggSpec spec, int64_t n, unsigned long long* st)
{
    uint64_t acc[kMaxOps];
#pragma unroll
    for (int a = 0; a < kMaxOps; ++a) acc[a] = 0ull;
    bool has = false;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        uint64_t cv[NV];
#pragma unroll
        for (int c = 0; c < NV; ++c)
            cv[c] = c < spec.nvcols ? dev::canon(dev::load_raw(spec.cols[c].data, spec.cols[c].bytes, r),
                                                 spec.cols[c].conv)
                                    : 0ull;
#pragma unroll
        for (int a = 0; a < kMaxOps; ++a) {
            if (a < spec.nops) {
                const OpDesc od = spec.ops[a];
                uint64_t v = 1ull;
#pragma unroll
                for (int c = 0; c < NV; ++c)
                    if (c == od.col) v = cv[c];
                const uint64_t opnd = dev::operand(od, v);
                acc[a] = has ? dev::combine(dev::class_of(spec, a), acc[a], opnd) : opnd;
            }
        }
        has = true;
    }
    // wave reduction of (has, acc[a])
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const bool oh = __shfl_xor((int)has, off) != 0;
#pragma unroll
        for (int a = 0; a < kMaxOps; ++a) {
            if (a < spec.nops) {
                uint64_t o = __shfl_xor((unsigned long long)acc[a], off);
                if (oh) acc[a] = has ? dev::combine(dev::class_of(spec, a), acc[a], o) : o;
            }
        }
        has = has || oh;
    }
    if (dev::lane_id() == 0 && has) {
#pragma unroll
        for (int a = 0; a < kMaxOps; ++a)
            if (a < spec.nops) dev::apply_atomic(dev::class_of(spec, a), &st[a], acc[a]);
    }
}

__global__ void emit_one_kernel(OutSpec os, const unsigned long long* st)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) dev::emit_group(os, 0, 0ull, (const uint64_t*)st, 1);
}

// Identity values of the reference's functors for an empty input
// (agg_func.cuh:37-207: Sum 0, Max numeric_limits<T>::min(), Min
// numeric_limits<T>::max(), Any 0); AVG = 0/0.
void empty_value(const OutCol& o, unsigned char* out)
{
    std::memset(out, 0, 8);
    auto put = [&](auto v) { std::memcpy(out, &v, sizeof v); };
    switch (o.code) {
    case OUT_MAX_I:
        switch (o.type) {
        case EG_INT8: put((int8_t)INT8_MIN); break;
        case EG_INT16: put((int16_t)INT16_MIN); break;
        case EG_INT32: put((int32_t)INT32_MIN); break;
        default: put((int64_t)INT64_MIN); break;
        }
        break;
    case OUT_MIN_I:
        switch (o.type) {
        case EG_INT8: put((int8_t)INT8_MAX); break;
        case EG_INT16: put((int16_t)INT16_MAX); break;
        case EG_INT32: put((int32_t)INT32_MAX); break;
        default: put((int64_t)INT64_MAX); break;
        }
        break;
    case OUT_MIN_U: std::memset(out, 0xff, 8); break;
    case OUT_MAX_F:
        if (o.type == EG_FLOAT32) put(FLT_MIN); else put(DBL_MIN);
        break;
    case OUT_MIN_F:
        if (o.type == EG_FLOAT32) put(FLT_MAX); else put(DBL_MAX);
        break;
    case OUT_AVG_I: case OUT_AVG_U: case OUT_AVG_F: put(std::nan("")); break;
    default: break;
    }
}

}  // namespace

void scalar_aggregate(const std::vector<ViewIn>& cols, const eg_agg_request* reqs, int64_t nreqs,
                      hipStream_t s, Result& res)
{
    Plan plan = make_plan(nullptr, cols, reqs, nreqs, false, true);
    Result full;
    OutSpec os;
    alloc_outputs(plan, 1, s, full, os);
    if (plan.rows == 0) {
        for (size_t j = 0; j < plan.outs.size(); ++j) {
            unsigned char v[8];
            empty_value(plan.outs[j], v);
            EG_HIP(hipMemcpyAsync(full.cols[j + 1].get(), v, (size_t)type_size(plan.outs[j].type),
                                  hipMemcpyHostToDevice, s));
        }
        EG_HIP(hipStreamSynchronize(s));
    } else {
        DevBuf st(kMaxOps * 8, s);
        fill_zero(st.get(), kMaxOps * 8, s);
        const DeviceInfo& di = device_info();
        const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((plan.rows + 255) / 256, (int64_t)di.num_cus * 8));
        KernelTimer timer("scalar_agg", s);
        const int nv = plan.spec.nvcols;
        auto* stp = st.get<unsigned long long>();
        if (nv <= 1) hipLaunchKernelGGL(scalar_kernel<1>, dim3((unsigned)grid), dim3(256), 0, s, plan.spec, plan.rows, stp);
        else if (nv <= 2) hipLaunchKernelGGL(scalar_kernel<2>, dim3((unsigned)grid), dim3(256), 0, s, plan.spec, plan.rows, stp);
        else if (nv <= 4) hipLaunchKernelGGL(scalar_kernel<4>, dim3((unsigned)grid), dim3(256), 0, s, plan.spec, plan.rows, stp);
        else hipLaunchKernelGGL(scalar_kernel<8>, dim3((unsigned)grid), dim3(256), 0, s, plan.spec, plan.rows, stp);
        EG_LAUNCH_CHECK();
        hipLaunchKernelGGL(emit_one_kernel, dim3(1), dim3(64), 0, s, os, stp);
        EG_LAUNCH_CHECK();
        EG_HIP(hipStreamSynchronize(s));
    }
    res.num_groups = 1;
    res.cols.clear();
    res.types.clear();
    for (size_t j = 1; j < full.cols.size(); ++j) {
        res.cols.push_back(std::move(full.cols[j]));
        res.types.push_back(full.types[j]);
    }
}

}  // namespace eg
