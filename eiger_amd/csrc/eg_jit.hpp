// eg_jit.hpp -- host interface of the run-time kernel specialisation (jit.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "eg_spec.hpp"

namespace eg {
namespace jit {

enum class Kind { HASH_AGG, PART_AGG, PART_AGG_GATHER };

// The kernel specialised for (kind, key bytes, value layout type name, ops of
// `spec`), or nullptr while it is not available: JIT off, `rows` below the
// size threshold, compile pending (async mode) or failed.
// PART_AGG kinds with `os`: the output list is compiled in too (dev::FixOuts).
hipFunction_t lookup(Kind kind, int kb, const std::string& vlayout, const AggSpec& spec, int64_t rows,
                     const OutSpec* os = nullptr);

// Queues the specialisation for `arch` on the background worker without
// touching a device (lookups on a device of that architecture string find
// it); false when it is already known, JIT is off, or the process is exiting.
bool prefetch(Kind kind, int kb, const std::string& vlayout, const AggSpec& spec, const std::string& arch,
              const OutSpec* os = nullptr);

// hipModuleLaunchKernel; false (and the entry disabled) when the launch is
// refused, so the caller launches its generic kernel instead.
bool launch(hipFunction_t f, unsigned grid, unsigned block, size_t lds, hipStream_t s, void** args);

// Compiles the specialisation for `arch` without loading it (no device
// needed); false with the compiler log on failure.
bool compile_check(Kind kind, int kb, const std::string& vlayout, const AggSpec& spec, const std::string& arch,
                   std::string& log, const OutSpec* os = nullptr);

// Resident workgroups per CU of f at this block size and dynamic LDS.
int occupancy(hipFunction_t f, int block, size_t lds);

// Name of a compiled value layout, as the run-time compiler spells it.
template <class VL>
std::string vlayout_name()
{
    return "eg::lds::VLayout<" + std::to_string(VL::NV) + ", " + std::to_string(VL::template width<0>()) + ", " +
           std::to_string(VL::template width<1>()) + ", " + (VL::GEN ? "true" : "false") + ", " +
           std::to_string(VL::template conv<0>()) + ", " + std::to_string(VL::template conv<1>()) + ">";
}

void configure(int mode, int64_t min_rows);
void wait_all();
void stats(int64_t* out6);

}  // namespace jit
}  // namespace eg
