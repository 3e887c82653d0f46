// vrx_lbvh.h -- node layout and build entry shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vrx_hip.h"

namespace vrx {

// 64-byte BVH2 node holding the boxes of both children (see lbvh.hip).
struct alignas(16) Node {
  float4 a, b, c;
  uint4 d;
};
static_assert(sizeof(Node) == 64, "Node must be 64 bytes");

size_t leaf_record_bytes(int type);
void bvh_sizes(int32_t type, uint32_t n, vrx_bvh_sizes *o);
int bvh_build_impl(const vrx_prims *p, const vrx_bvh *out, void *d_ws, size_t ws_bytes,
    uint32_t *codes_out, float *cbox_out, hipStream_t s);

} // namespace vrx
