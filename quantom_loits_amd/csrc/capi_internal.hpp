// capi_internal.hpp -- error helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace qloits {

// LDS a single work-group may use on gfx950 (160 KiB per CU), less the
// kernels' static LDS.
constexpr int64_t kMaxLds = 160 * 1024 - 2048;

void set_error(const std::string& msg);
int fail(int status, const std::string& msg);
int check_hip(hipError_t e, const char* what);

}  // namespace qloits
