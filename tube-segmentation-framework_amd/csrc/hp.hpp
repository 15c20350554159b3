// Histogram-pyramid compaction pieces shared by the entry points of hp.hip
// and hp_noalloc.hip.
#pragma once
#include "tsf_common.hpp"

namespace tsf {

struct HPPlan {
    int cube_levels;   // levels of the cube key (pyramid size / 16 = 2^cube_levels)
    long long ncubes;  // 8^cube_levels
};

// The pyramid's domain: the smallest 16 * 2^L cube holding the volume.
HPPlan hp_plan(Dims d);

// Argument checks shared by the compaction entry points.
tsf_status hp_check(const uint8_t* mask, tsf_dims dims, int32_t* positions, int64_t capacity, int64_t* count,
                    const char* what);

// Count, scan and scatter on a workspace laid out as
// tsf_hp_compact_workspace_size describes (cube counts, then 256-byte
// aligned offsets), then the total read back to the host (the stream is
// synchronised).  TSF_ERR_CAPACITY when positions is given and the total
// exceeds capacity.
tsf_status hp_run(const uint8_t* mask, Dims d, const HPPlan& p, void* ws, int32_t* positions, int64_t capacity,
                  int64_t* count, hipStream_t stream, const char* what);

}  // namespace tsf
