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

__device__ __forceinline__ Node gld_node(const Node *nodes, size_t i)
{
  typedef float f4v __attribute__((ext_vector_type(4)));
  typedef unsigned int u4v __attribute__((ext_vector_type(4)));
  const __attribute__((address_space(1))) f4v *q =
      (const __attribute__((address_space(1))) f4v *)(nodes + i);
  f4v a = q[0], b = q[1], c = q[2];
  u4v d = ((const __attribute__((address_space(1))) u4v *)(nodes + i))[3];
  Node n;
  n.a = make_float4(a.x, a.y, a.z, a.w);
  n.b = make_float4(b.x, b.y, b.z, b.w);
  n.c = make_float4(c.x, c.y, c.z, c.w);
  n.d = make_uint4(d.x, d.y, d.z, d.w);
  return n;
}

// 128-byte 4-wide node (one HBM line): the boxes of up to four children in
// SoA form and their links (bit 31 = leaf, as in Node); slots at or beyond
// `count` are unused.
struct alignas(16) Node4 {
  float4 lox, loy, loz, hix, hiy, hiz;
  uint4 link;
  uint4 meta; // x = count (2..4)
};
static_assert(sizeof(Node4) == 128, "Node4 must be 128 bytes");

__device__ __forceinline__ Node4 gld_node4(const Node4 *nodes, size_t i)
{
  typedef float f4v __attribute__((ext_vector_type(4)));
  typedef unsigned int u4v __attribute__((ext_vector_type(4)));
  const __attribute__((address_space(1))) f4v *q =
      (const __attribute__((address_space(1))) f4v *)(nodes + i);
  f4v a = q[0], b = q[1], c = q[2], d = q[3], e = q[4], f = q[5];
  const __attribute__((address_space(1))) u4v *u =
      (const __attribute__((address_space(1))) u4v *)(nodes + i);
  u4v l = u[6], m = u[7];
  Node4 n;
  n.lox = make_float4(a.x, a.y, a.z, a.w);
  n.loy = make_float4(b.x, b.y, b.z, b.w);
  n.loz = make_float4(c.x, c.y, c.z, c.w);
  n.hix = make_float4(d.x, d.y, d.z, d.w);
  n.hiy = make_float4(e.x, e.y, e.z, e.w);
  n.hiz = make_float4(f.x, f.y, f.z, f.w);
  n.link = make_uint4(l.x, l.y, l.z, l.w);
  n.meta = make_uint4(m.x, m.y, m.z, m.w);
  return n;
}

size_t leaf_record_bytes(int type);
void bvh_sizes(int32_t type, uint32_t n, vrx_bvh_sizes *o);
int bvh_build_impl(const vrx_prims *p, const vrx_bvh *out, void *d_ws, size_t ws_bytes,
    uint32_t *codes_out, float *cbox_out, hipStream_t s);

} // namespace vrx
