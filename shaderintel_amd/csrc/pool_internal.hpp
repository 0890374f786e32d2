// Library-internal access to a scene pool for the serving loop (not part of the C ABI).
#pragma once
#include <hip/hip_runtime_api.h>

#include "../../include/mpm.h"

namespace mpm {
// The HIP stream every launch of the pool is enqueued on.
hipStream_t pool_stream(mpm_pool* pool);
}  // namespace mpm
