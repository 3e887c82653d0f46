// vrx_math.h -- fp32 vector arithmetic with a fixed evaluation order.
//
// Every kernel that claims bit-exact agreement with the CPU path (Morton
// codes, closest-hit t and primitive selection, normals, channel values) is
// compiled with -ffp-contract=off and uses only these helpers, so a*b+c stays
// a rounded multiply followed by a rounded add, sqrt and divide are the
// correctly rounded IEEE operations, and dot products associate as
// ((x*x' + y*y') + z*z').
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define VRX_HD __host__ __device__ __forceinline__
#define VRX_GLOBAL __attribute__((address_space(1)))

namespace vrx {

// Loads and stores through global (HBM) pointers.  Pointers fetched from
// scene records in memory are generic to the compiler, which would otherwise
// emit flat_* instructions (slower, and they also tie up the LDS counter the
// traversal stack waits on).
template <typename T>
__device__ __forceinline__ T gld(const T *p)
{
  return *(const VRX_GLOBAL T *)p;
}
template <typename T>
__device__ __forceinline__ T gld(const T *p, size_t i)
{
  return ((const VRX_GLOBAL T *)p)[i];
}
template <typename T>
__device__ __forceinline__ void gst(T *p, size_t i, const T &v)
{
  ((VRX_GLOBAL T *)p)[i] = v;
}
// HIP's vector types are classes whose copy constructors take generic
// references, so 16-byte loads go through native vector types instead.
typedef float vrx_f4v __attribute__((ext_vector_type(4)));
typedef unsigned int vrx_u4v __attribute__((ext_vector_type(4)));
typedef float vrx_f2v_a4 __attribute__((ext_vector_type(2))) __attribute__((aligned(4)));
__device__ __forceinline__ float4 gld(const float4 *p, size_t i)
{
  vrx_f4v v = ((const VRX_GLOBAL vrx_f4v *)p)[i];
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 gld(const uint4 *p, size_t i)
{
  vrx_u4v v = ((const VRX_GLOBAL vrx_u4v *)p)[i];
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst(float4 *p, size_t i, const float4 &v)
{
  vrx_f4v w = {v.x, v.y, v.z, v.w};
  ((VRX_GLOBAL vrx_f4v *)p)[i] = w;
}
// four consecutive floats at 8-byte alignment (one 16-byte gather)
typedef float vrx_f4v_a8 __attribute__((ext_vector_type(4))) __attribute__((aligned(8)));
__device__ __forceinline__ float4 gld_quad8(const VRX_GLOBAL float *p)
{
  vrx_f4v_a8 v = *(const VRX_GLOBAL vrx_f4v_a8 *)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
// two adjacent floats at 4-byte alignment (one 8-byte gather)
__device__ __forceinline__ float2 gld_pair(const VRX_GLOBAL float *p)
{
  vrx_f2v_a4 v = *(const VRX_GLOBAL vrx_f2v_a4 *)p;
  return make_float2(v.x, v.y);
}

struct v3 {
  float x, y, z;
};

VRX_HD v3 mk(float x, float y, float z) { return v3{x, y, z}; }
VRX_HD v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
VRX_HD v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
VRX_HD v3 mulf(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
VRX_HD v3 fmul(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
VRX_HD v3 mulv(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
VRX_HD v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
VRX_HD float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
VRX_HD v3 cross(v3 a, v3 b)
{
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
VRX_HD v3 normalize(v3 v) { return mulf(v, 1.0f / sqrtf(dot(v, v))); }
VRX_HD v3 vmin(v3 a, v3 b) { return mk(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
VRX_HD v3 vmax(v3 a, v3 b) { return mk(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
VRX_HD float comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
VRX_HD float cmax3(v3 v) { return fmaxf(fmaxf(v.x, v.y), v.z); }
VRX_HD v3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
__device__ __forceinline__ v3 gld3(const float *p, size_t i)
{
  const VRX_GLOBAL float *q = (const VRX_GLOBAL float *)p + 3 * i;
  return mk(q[0], q[1], q[2]);
}

// 3x4 column-major transform (4 columns of 3): p' = M[:,0]x + M[:,1]y + M[:,2]z + M[:,3]
VRX_HD v3 xfm_point(const float *m, v3 p)
{
  return mk(((m[0] * p.x + m[3] * p.y) + m[6] * p.z) + m[9],
      ((m[1] * p.x + m[4] * p.y) + m[7] * p.z) + m[10],
      ((m[2] * p.x + m[5] * p.y) + m[8] * p.z) + m[11]);
}
VRX_HD v3 xfm_vector(const float *m, v3 p)
{
  return mk((m[0] * p.x + m[3] * p.y) + m[6] * p.z, (m[1] * p.x + m[4] * p.y) + m[7] * p.z,
      (m[2] * p.x + m[5] * p.y) + m[8] * p.z);
}
// normal transform with the inverse matrix: transpose(inv3x3) * n
VRX_HD v3 xfm_normal(const float *inv, v3 n)
{
  return mk((inv[0] * n.x + inv[1] * n.y) + inv[2] * n.z, (inv[3] * n.x + inv[4] * n.y) + inv[5] * n.z,
      (inv[6] * n.x + inv[7] * n.y) + inv[8] * n.z);
}

// Slack applied to slab-test distances so that fp32 rounding in the slab and
// in the primitive tests can never cull a primitive that would report a hit
// (conservative boxes; the closest hit then equals the brute-force one).
constexpr float kBoxLo = 1.0f - 0x1p-18f;
constexpr float kBoxHi = 1.0f + 0x1p-18f;
constexpr uint32_t kLeafBit = 0x80000000u;

// gpu_util.h:245-250 epsilonFrom (ulpEpsilon = 0x1.fp-21)
VRX_HD float eps_from(v3 P, v3 dir, float t)
{
  float a = fmaxf(fmaxf(fabsf(P.x), fabsf(P.y)), fabsf(P.z));
  float m = fmaxf(fmaxf(fabsf(dir.x), fabsf(dir.y)), fabsf(dir.z)) * t;
  return fmaxf(a, m) * 0x1.fp-21f;
}

// PCG32 (O'Neill 2014): per-pixel stream seeded with (pixel index, frame id).
struct Pcg {
  uint64_t state, inc;
};
VRX_HD uint32_t pcg_next(Pcg &r)
{
  uint64_t old = r.state;
  r.state = old * 6364136223846793005ULL + r.inc;
  uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  uint32_t rot = (uint32_t)(old >> 59u);
  return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
VRX_HD void pcg_seed(Pcg &r, uint64_t initstate, uint64_t initseq)
{
  r.state = 0u;
  r.inc = (initseq << 1u) | 1u;
  pcg_next(r);
  r.state += initstate;
  pcg_next(r);
}
VRX_HD float pcg_float(Pcg &r) { return (float)(pcg_next(r) >> 8) * 0x1p-24f; }

} // namespace vrx
