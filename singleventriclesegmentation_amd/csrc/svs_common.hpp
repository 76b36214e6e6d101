// Host-side helpers shared by the C-ABI entry points: error reporting,
// argument validation and launch bookkeeping.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

namespace svs {

// The last error message of the calling thread (read through svs_last_error).
extern thread_local char g_last_error[1024];

// Records a formatted message and returns the status code the entry points use
// for invalid arguments (1) or device errors (2).
int fail(int code, const char* fmt, ...);

inline int grid_for(int64_t n, int block) { return (int)((n + block - 1) / block); }

inline bool valid_bc(int bc) { return bc >= 0 && bc <= 2; }
inline bool valid_ip(int ip) { return ip >= 0 && ip <= 2; }

}  // namespace svs

#define SVS_STATUS_OK 0
#define SVS_STATUS_INVALID 1
#define SVS_STATUS_DEVICE 2

#define SVS_REQUIRE(cond, ...)                                   \
  do {                                                           \
    if (!(cond)) return ::svs::fail(SVS_STATUS_INVALID, __VA_ARGS__); \
  } while (0)

#define SVS_HIP(expr)                                                                         \
  do {                                                                                        \
    hipError_t svs_e_ = (expr);                                                               \
    if (svs_e_ != hipSuccess)                                                                 \
      return ::svs::fail(SVS_STATUS_DEVICE, "%s failed: %s", #expr, hipGetErrorString(svs_e_)); \
  } while (0)

#define SVS_LAUNCHED() SVS_HIP(hipGetLastError())

// A one-dimensional launch covers at most 2^32 - 256 work-items (the grid's
// work-item count is a 32-bit quantity); kernels that take one work-item per
// element are bounded by it.
#define SVS_MAX_WORKITEMS (((int64_t)1 << 32) - 256)
#define SVS_REQUIRE_WORKITEMS(n, what)                                                                     \
  SVS_REQUIRE((int64_t)(n) <= SVS_MAX_WORKITEMS, "%s: %lld elements exceed one launch's %lld work-items", \
              what, (long long)(n), (long long)SVS_MAX_WORKITEMS)

// Dimensions must be positive and the voxel count must fit the int32 index
// space the reference uses (PackedTensorAccessor32).
#define SVS_REQUIRE_DIMS(NZ, NY, NX)                                                                  \
  SVS_REQUIRE((NZ) > 0 && (NY) > 0 && (NX) > 0, "dimensions must be positive (got %d,%d,%d)", (int)(NZ), \
              (int)(NY), (int)(NX));                                                                   \
  SVS_REQUIRE((int64_t)(NZ) * (NY) * (NX) < ((int64_t)1 << 31), "volume exceeds 2^31 voxels")
