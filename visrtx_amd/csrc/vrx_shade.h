// vrx_shade.h -- device-side per-pixel programs shared by the per-view and
// the batched multi-view kernels: pinhole rays, the emission-absorption
// volume march, channel packing and the raycast / AO pixel program.
#pragma once

#include <float.h>
#include <math.h>

#include "vrx_trace.h"

namespace vrx {

constexpr int kTile = 16;
constexpr int kBlock = kTile * kTile;
// Hard bounds that keep every loop finite on degenerate input (a zero
// volume step, a surface hit at the ray origin with zero epsilon, volumes
// whose exit distance no longer grows by 1e-3 in fp32).
constexpr int kMaxSurfaceLayers = 4096;
constexpr uint32_t kMaxSamplesPerSegment = 1u << 24;
constexpr int kMaxVolumeSegments = 256;
constexpr int kLookCells = 3; // macro cells looked up together by the skipping march

// --------------------------------------------------------------------------
// camera
// --------------------------------------------------------------------------
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

__device__ __forceinline__ void cam_ray(const vrx_camera &c, float sx, float sy, v3 &o, v3 &d)
{
  sx = mixf(c.region[0], c.region[2], sx);
  sy = mixf(c.region[1], c.region[3], sy);
  o = ld3(c.pos);
  d = normalize(add(add(ld3(c.d00), fmul(sx, ld3(c.du))), fmul(sy, ld3(c.dv))));
}

// --------------------------------------------------------------------------
// volumes (fp32)
// --------------------------------------------------------------------------
__device__ __forceinline__ bool vol_slab(
    const vrx_volume &v, v3 o, v3 d, float tmin, float tmax, float &t0, float &t1)
{
  float lo[3] = {v.origin[0], v.origin[1], v.origin[2]};
  float hi[3];
  float tn = -INFINITY, tf = INFINITY;
  float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    hi[k] = v.origin[k] + ((float)v.dims[k] - 1.0f) * v.spacing[k];
    float inv = 1.0f / dd[k];
    float a = (lo[k] - oo[k]) * inv, b = (hi[k] - oo[k]) * inv;
    tn = fmaxf(tn, fminf(a, b));
    tf = fminf(tf, fmaxf(a, b));
  }
  if (!(tn < tf))
    return false;
  if (tf < tmin || tn > tmax)
    return false;
  t0 = fminf(fmaxf(tn, tmin), tmax);
  t1 = fminf(fmaxf(tf, tmin), tmax);
  return true;
}

// Per-volume constants of the march, computed once per ray segment.
struct VolCtx {
  const float *data;
  int dims[3], imax[3];
  float gmax[3];
  float origin[3], inv_sp[3];
  uint32_t sy, sz;
  float step, expo, lo, hi, inv_range;
  const float *mc; // macro-cell majorants (16^3 voxels per cell) or null
  int mcd[3];
  // the 32-bit gather form (b32): the field has 2 to 2^24 - 1 voxels along x,
  // under 2^30 voxels and under 2^24 x-rows, so a voxel's byte offset fits
  // 32 bits, its row index z * dims1 + y is exact in fp32 and row * dims0 is
  // a 24-bit multiply
  bool b32;
  const float *yp; // the y-pair field (b32 fields under 2^29 voxels) or null
  uint32_t ozp;    // y-pair byte offset of the +1 neighbour in z (0 on a one-voxel axis)
  uint32_t oy, oz;        // byte offsets of the +1 neighbour in y / z (0 on a one-voxel axis)
  float fimax[3], fdims1; // imax and dims[1] as floats
  float tf_scale, tf_bias; // LUT texel coordinate s * tf_scale + tf_bias (clamped to [0, 255])
};

__device__ __forceinline__ float vol_step(const vrx_volume &v)
{
  // StructuredRegularField.cpp:175-178 stepSize() = min(spacing) / 2
  return fminf(fminf(v.spacing[0] / 2.0f, v.spacing[1] / 2.0f), v.spacing[2] / 2.0f);
}

__device__ __forceinline__ VolCtx vol_ctx(const vrx_volume &v, float inv_rate, bool skip = false)
{
  VolCtx c;
  c.data = v.d_data;
  c.mc = skip ? v.d_majorant : nullptr;
#pragma unroll
  for (int k = 0; k < 3; k++)
    c.mcd[k] = ((int)v.dims[k] + 15) >> 4;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    c.dims[k] = (int)v.dims[k];
    c.imax[k] = max((int)v.dims[k] - 2, 0);
    c.gmax[k] = (float)v.dims[k] - 1.0f;
    c.origin[k] = v.origin[k];
    c.inv_sp[k] = 1.0f / v.spacing[k];
  }
  c.sy = v.dims[0];
  c.sz = v.dims[0] * v.dims[1];
  c.step = vol_step(v) * inv_rate;
  c.expo = c.step * (1.0f / v.unit_distance);
  c.lo = v.value_range[0];
  c.hi = v.value_range[1];
  c.inv_range = 1.0f / (c.hi - c.lo);
  const uint64_t rows = (uint64_t)v.dims[1] * v.dims[2];
  const uint64_t nvox = rows * v.dims[0];
  c.b32 = v.dims[0] >= 2 && v.dims[0] < (1u << 24) && nvox < (1ull << 30) && rows < (1ull << 24);
  c.yp = c.b32 && nvox < (1ull << 29) ? v.d_ypairs : nullptr;
  c.ozp = v.dims[2] > 1 ? v.dims[0] * v.dims[1] * 8u : 0u;
  c.oy = v.dims[1] > 1 ? v.dims[0] * 4u : 0u;
  c.oz = v.dims[2] > 1 ? v.dims[0] * v.dims[1] * 4u : 0u;
#pragma unroll
  for (int k = 0; k < 3; k++)
    c.fimax[k] = (float)c.imax[k];
  c.fdims1 = (float)v.dims[1];
  // classify()'s clamp min(max(s, lo), hi) maps every sample of a reversed
  // range (lo > hi) to hi, the last texel (x = 255); an empty range (lo ==
  // hi) reads texel 0 there, and so does this form
  const float sc = 256.0f * c.inv_range;
  const bool fin = sc - sc == 0.0f; // finite
  const bool rev = c.lo > c.hi;
  c.tf_scale = fin && !rev ? sc : 0.0f;
  c.tf_bias = rev ? 255.0f : fin ? -c.lo * sc - 0.5f : 0.0f;
  return c;
}

// Software trilinear at voxel coordinates g = (p - origin) / spacing with edge
// clamp: the structured-regular field's clamp-to-edge linear texture
// (StructuredRegularField.cpp:98-159, sampleSpatialField.h:54-78) with full
// fp32 weights.  I64 selects 64-bit gather offsets for fields of >= 2^31 voxels.
// XP: 1 when the field has at least two voxels along x (one 8-byte gather per
// x-pair), 0 when it has one, -1 to decide per call.  The march decides once
// per segment: a uniform branch between the two forms inside the unrolled
// sample loop makes the compiler wait for each sample's gathers before it
// issues the next sample's.
template <bool I64, int XP = -1>
__device__ __forceinline__ float trilinear(const VolCtx &c, float gx, float gy, float gz)
{
  float g[3] = {gx, gy, gz};
  int i0[3], i1[3];
  float f[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float x = fminf(fmaxf(g[k], 0.0f), c.gmax[k]);
    int i = min((int)x, c.imax[k]);
    i0[k] = i;
    i1[k] = min(i + 1, c.dims[k] - 1);
    f[k] = x - (float)i;
  }
  float c000, c100, c010, c110, c001, c101, c011, c111;
  if (XP == 1 || (XP < 0 && c.dims[0] >= 2)) {
    // x-neighbours are adjacent floats (i0 <= dims-2): one 8-byte gather each
    const VRX_GLOBAL float *D = (const VRX_GLOBAL float *)c.data;
    float2 p00, p10, p01, p11;
    if constexpr (I64) {
      size_t z0 = (size_t)i0[2] * c.sz, z1 = (size_t)i1[2] * c.sz;
      size_t y0 = (size_t)i0[1] * c.sy, y1 = (size_t)i1[1] * c.sy;
      p00 = gld_pair(D + z0 + y0 + i0[0]);
      p10 = gld_pair(D + z0 + y1 + i0[0]);
      p01 = gld_pair(D + z1 + y0 + i0[0]);
      p11 = gld_pair(D + z1 + y1 + i0[0]);
    } else {
      uint32_t z0 = (uint32_t)i0[2] * c.sz, z1 = (uint32_t)i1[2] * c.sz;
      uint32_t y0 = (uint32_t)i0[1] * c.sy, y1 = (uint32_t)i1[1] * c.sy;
      p00 = gld_pair(D + z0 + y0 + i0[0]);
      p10 = gld_pair(D + z0 + y1 + i0[0]);
      p01 = gld_pair(D + z1 + y0 + i0[0]);
      p11 = gld_pair(D + z1 + y1 + i0[0]);
    }
    c000 = p00.x; c100 = p00.y; c010 = p10.x; c110 = p10.y;
    c001 = p01.x; c101 = p01.y; c011 = p11.x; c111 = p11.y;
  } else if constexpr (I64) {
    size_t z0 = (size_t)i0[2] * c.sz, z1 = (size_t)i1[2] * c.sz;
    size_t y0 = (size_t)i0[1] * c.sy, y1 = (size_t)i1[1] * c.sy;
    const VRX_GLOBAL float *D = (const VRX_GLOBAL float *)c.data;
    c000 = c100 = D[z0 + y0 + i0[0]];
    c010 = c110 = D[z0 + y1 + i0[0]];
    c001 = c101 = D[z1 + y0 + i0[0]];
    c011 = c111 = D[z1 + y1 + i0[0]];
  } else {
    uint32_t z0 = (uint32_t)i0[2] * c.sz, z1 = (uint32_t)i1[2] * c.sz;
    uint32_t y0 = (uint32_t)i0[1] * c.sy, y1 = (uint32_t)i1[1] * c.sy;
    const VRX_GLOBAL float *D = (const VRX_GLOBAL float *)c.data;
    c000 = c100 = D[z0 + y0 + i0[0]];
    c010 = c110 = D[z0 + y1 + i0[0]];
    c001 = c101 = D[z1 + y0 + i0[0]];
    c011 = c111 = D[z1 + y1 + i0[0]];
  }
  float c00 = c000 + (c100 - c000) * f[0];
  float c10 = c010 + (c110 - c010) * f[0];
  float c01 = c001 + (c101 - c001) * f[0];
  float c11 = c011 + (c111 - c011) * f[0];
  float c0 = c00 + (c10 - c00) * f[1];
  float c1 = c01 + (c11 - c01) * f[1];
  return c0 + (c1 - c0) * f[2];
}

// classifySample (volumeIntegration.h:46-62): LUT texel coordinate u*256-0.5,
// linear, clamped (TransferFunction1D.cpp:152-185).
template <bool LUT_GLOBAL>
__device__ __forceinline__ float4 classify(const float4 *lut, const VolCtx &c, float s)
{
  float cl = fminf(fmaxf(s, c.lo), c.hi);
  float u = (cl - c.lo) * c.inv_range;
  float x = fminf(fmaxf(u * 256.0f - 0.5f, 0.0f), 255.0f);
  int i = min((int)x, 254);
  float f = x - (float)i;
  float4 a, b;
  if constexpr (LUT_GLOBAL) {
    a = gld(lut, i);
    b = gld(lut, i + 1);
  } else {
    a = lut[i];
    b = lut[i + 1];
  }
  return make_float4(a.x + (b.x - a.x) * f, a.y + (b.y - a.y) * f, a.z + (b.z - a.z) * f,
      a.w + (b.w - a.w) * f);
}

// Forms of trilinear() and classify() for marches compared with the FP64
// oracle within a tolerance (same clamps, same voxel and texel selection):
// lerps as one fused multiply-add each; the LUT coordinate
// (clamp(s) - lo) / range * 256 - 0.5 as s * tf_scale + tf_bias before the
// texel clamp (a sample outside the range lands outside [0, 255] and clamps to
// the same end texel).
__device__ __forceinline__ float fma_lerp(float a, float b, float t)
{
  return __builtin_fmaf(b - a, t, a);
}

// trilinear() on a b32 field: voxel indices kept as floats (x >= 0, so floor
// is the truncation), the row index as one exact FMA, and one 32-bit byte
// offset from the field's base for the four 8-byte x-pair gathers (the y / z
// neighbours are constant offsets: 0 on a one-voxel axis, where the edge clamp
// repeats the voxel), which saves the per-gather 64-bit address arithmetic.
__device__ __forceinline__ float trilinear_b32(const VolCtx &c, float gx, float gy, float gz)
{
  const float g[3] = {gx, gy, gz};
  float i0[3], f[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const float x = fminf(fmaxf(g[k], 0.0f), c.gmax[k]);
    i0[k] = fminf(floorf(x), c.fimax[k]);
    f[k] = x - i0[k];
  }
  const uint32_t row = (uint32_t)__builtin_fmaf(i0[2], c.fdims1, i0[1]);
  const uint32_t bo = ((row & 0xffffffu) * ((uint32_t)c.dims[0] & 0xffffffu) + (uint32_t)i0[0]) * 4u;
  const VRX_GLOBAL char *B = (const VRX_GLOBAL char *)c.data;
  const float2 p00 = gld_pair((const VRX_GLOBAL float *)(B + bo));
  const float2 p10 = gld_pair((const VRX_GLOBAL float *)(B + (bo + c.oy)));
  const float2 p01 = gld_pair((const VRX_GLOBAL float *)(B + (bo + c.oz)));
  const float2 p11 = gld_pair((const VRX_GLOBAL float *)(B + (bo + (c.oy + c.oz))));
  const float c00 = fma_lerp(p00.x, p00.y, f[0]);
  const float c10 = fma_lerp(p10.x, p10.y, f[0]);
  const float c01 = fma_lerp(p01.x, p01.y, f[0]);
  const float c11 = fma_lerp(p11.x, p11.y, f[0]);
  return fma_lerp(fma_lerp(c00, c10, f[1]), fma_lerp(c01, c11, f[1]), f[2]);
}

// trilinear_b32() on the y-pair field: the 16 bytes at voxel (x, y, z)
// hold v(x,y), v(x,y+1), v(x+1,y), v(x+1,y+1) of plane z, so two gathers (z
// and z+1) read the cell; the same values and lerps as trilinear_b32().
__device__ __forceinline__ float trilinear_ypairs(const VolCtx &c, float gx, float gy, float gz)
{
  const float g[3] = {gx, gy, gz};
  float i0[3], f[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    // one v_med3 for the clamp (a NaN coordinate, from a NaN ray only, still
    // gives an index inside the field: fminf below returns fimax)
    const float x = __builtin_amdgcn_fmed3f(g[k], 0.0f, c.gmax[k]);
    i0[k] = fminf(floorf(x), c.fimax[k]);
    f[k] = x - i0[k];
  }
  const uint32_t row = (uint32_t)__builtin_fmaf(i0[2], c.fdims1, i0[1]);
  const uint32_t bo = ((row & 0xffffffu) * ((uint32_t)c.dims[0] & 0xffffffu) + (uint32_t)i0[0]) * 8u;
  const VRX_GLOBAL char *B = (const VRX_GLOBAL char *)c.yp;
  const float4 a = gld_quad8((const VRX_GLOBAL float *)(B + bo));
  const float4 b = gld_quad8((const VRX_GLOBAL float *)(B + (bo + c.ozp)));
  const float c00 = fma_lerp(a.x, a.z, f[0]);
  const float c10 = fma_lerp(a.y, a.w, f[0]);
  const float c01 = fma_lerp(b.x, b.z, f[0]);
  const float c11 = fma_lerp(b.y, b.w, f[0]);
  return fma_lerp(fma_lerp(c00, c10, f[1]), fma_lerp(c01, c11, f[1]), f[2]);
}

template <bool LUT_GLOBAL>
__device__ __forceinline__ float4 classify_fma(const float4 *lut, const VolCtx &c, float s)
{
  const float x = fminf(fmaxf(__builtin_fmaf(s, c.tf_scale, c.tf_bias), 0.0f), 255.0f);
  const int i = min((int)x, 254);
  const float f = x - (float)i;
  float4 a, b;
  if constexpr (LUT_GLOBAL) {
    a = gld(lut, i);
    b = gld(lut, i + 1);
  } else {
    a = lut[i];
    b = lut[i + 1];
  }
  return make_float4(fma_lerp(a.x, b.x, f), fma_lerp(a.y, b.y, f), fma_lerp(a.z, b.z, f),
      fma_lerp(a.w, b.w, f));
}

// classify_fma() on a LUT staged as (entry, next entry - entry) pairs: each
// channel's lerp is one FMA on the stored difference (the same operands as
// fma_lerp, so the same result).  lut2[2i] = lut[i], lut2[2i+1] = lut[i+1] -
// lut[i] (i < 255).
__device__ __forceinline__ float4 classify_delta(const float4 *lut2, const VolCtx &c, float s)
{
  const float x = fminf(fmaxf(__builtin_fmaf(s, c.tf_scale, c.tf_bias), 0.0f), 255.0f);
  const int i = min((int)x, 254);
  const float f = x - (float)i;
  const float4 a = lut2[2 * i], d = lut2[2 * i + 1];
  return make_float4(__builtin_fmaf(d.x, f, a.x), __builtin_fmaf(d.y, f, a.y),
      __builtin_fmaf(d.z, f, a.z), __builtin_fmaf(d.w, f, a.w));
}

// _rayMarchVolume (volumeIntegration.h:64-103), jitter off.  Sample k sits at
// t0 + k*step (not an accumulated sum, which keeps positions within an ulp of
// the exact ones); the per-step transmittance pow(1-a, step/unitDistance) is
// evaluated as exp2(expo * log2(1-a)) with the hardware exp2/log2.  The gathers
// of U consecutive samples are issued before any of them is composited (sample
// positions do not depend on the accumulation), which keeps 8*U loads in
// flight per lane; samples past the exit or the 0.99 cut-off are discarded,
// their clamped gathers always stay inside the field.
template <bool I64, int U, bool LUT_GLOBAL, bool SKIP, int XP>
__device__ __forceinline__ void march_xp(const VolCtx &c, const float4 *lut, v3 o, v3 d, float t0,
    float t1, v3 &color, float &opacity, uint32_t &ns)
{
  // voxel-space ray: g(t) = go + gd * t
  const float gox = (o.x - c.origin[0]) * c.inv_sp[0], gdx = d.x * c.inv_sp[0];
  const float goy = (o.y - c.origin[1]) * c.inv_sp[1], gdy = d.y * c.inv_sp[1];
  const float goz = (o.z - c.origin[2]) * c.inv_sp[2], gdz = d.z * c.inv_sp[2];
  float T = 1.0f;
  // macro-cell walk: per axis the reciprocal direction in voxel units
  [[maybe_unused]] float ig[3] = {0.0f, 0.0f, 0.0f};
  if constexpr (SKIP) {
    ig[0] = gdx != 0.0f ? 1.0f / gdx : 0.0f;
    ig[1] = gdy != 0.0f ? 1.0f / gdy : 0.0f;
    ig[2] = gdz != 0.0f ? 1.0f / gdz : 0.0f;
  }
  for (uint32_t k = 0; k < kMaxSamplesPerSegment; k += U) {
    if (SKIP && c.mc) {
      // empty-space skipping: a macro cell with a 0 majorant adds nothing
      // between the ray's entry and exit (alpha 0 gives a step transmittance
      // of exactly 1), so the march jumps to the first sample past the exit.
      // When the cell of sample k is empty, the next kLookCells - 1 cells
      // along the ray come from a DDA over the macro grid (arithmetic only),
      // so their majorant loads are issued together and a run of empty cells
      // costs one more round trip per kLookCells - 1 cells.  A cell's value
      // range covers one voxel beyond its box, which absorbs the rounding of
      // the exits.
      const float t = t0 + (float)k * c.step;
      if (!(t <= t1))
        break;
      const float g[3] = {gox + gdx * t, goy + gdy * t, goz + gdz * t};
      const float gd[3] = {gdx, gdy, gdz}, go[3] = {gox, goy, goz};
      int cc[3];
#pragma unroll
      for (int a = 0; a < 3; a++)
        cc[a] = min(max((int)floorf(g[a] * 0.0625f), 0), c.mcd[a] - 1);
      if (gld(c.mc + ((size_t)cc[2] * c.mcd[1] + cc[1]) * c.mcd[0] + cc[0]) == 0.0f) {
        float mj[kLookCells], te[kLookCells];
        bool in[kLookCells];
        mj[0] = 0.0f;
        in[0] = true;
#pragma unroll
        for (int q = 0; q < kLookCells; q++) {
          if (q > 0) {
            in[q] = cc[0] >= 0 && cc[0] < c.mcd[0] && cc[1] >= 0 && cc[1] < c.mcd[1] && cc[2] >= 0
                && cc[2] < c.mcd[2];
            mj[q] = in[q] ? gld(c.mc + ((size_t)cc[2] * c.mcd[1] + cc[1]) * c.mcd[0] + cc[0]) : 1.0f;
          }
          // exit of this cell and the axis it leaves through
          float tq = INFINITY;
          int ax = 0;
#pragma unroll
          for (int a = 0; a < 3; a++) {
            if (gd[a] != 0.0f) {
              const float e = (16.0f * (float)(cc[a] + (gd[a] > 0.0f ? 1 : 0)) - go[a]) * ig[a];
              if (e < tq) {
                tq = e;
                ax = a;
              }
            }
          }
          te[q] = tq;
          cc[ax] += gd[ax] > 0.0f ? 1 : -1;
        }
        int nz = 0;
#pragma unroll
        for (int q = 0; q < kLookCells; q++)
          nz += (nz == q && in[q] && mj[q] == 0.0f) ? 1 : 0;
        float tn = te[0];
#pragma unroll
        for (int q = 1; q < kLookCells; q++)
          tn = q < nz ? te[q] : tn;
        const float kn = ceilf((tn - t0) / c.step);
        if (!(kn < (float)kMaxSamplesPerSegment))
          break;
        // k advances by at least one sample; the loop adds U
        k = max((uint32_t)kn, k + 1) - U;
        continue;
      }
    }
    float sv[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const float t = t0 + (float)(k + u) * c.step;
      ok[u] = t <= t1;
      sv[u] = trilinear<I64, XP>(c, gox + gdx * t, goy + gdy * t, goz + gdz * t);
    }
    bool done = false;
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (!done) {
        if (!ok[u] || !(opacity < 0.99f)) {
          done = true;
        } else {
          ns++;
          const float s = sv[u];
          if (s == s) {
            const float4 co = classify<LUT_GLOBAL>(lut, c, s);
            const float st = __builtin_amdgcn_exp2f(c.expo * __builtin_amdgcn_logf(1.0f - co.w));
            const float w = T * (1.0f - st);
            color.x += w * co.x;
            color.y += w * co.y;
            color.z += w * co.z;
            opacity += w;
            T *= st;
          }
        }
      }
    }
    if (done)
      break;
  }
}

template <bool I64, int U, bool LUT_GLOBAL, bool SKIP = true>
__device__ __forceinline__ void march(const VolCtx &c, const float4 *lut, v3 o, v3 d, float t0,
    float t1, v3 &color, float &opacity, uint32_t &ns)
{
  if (!(c.step > 0.0f))
    return;
  if (c.dims[0] >= 2)
    march_xp<I64, U, LUT_GLOBAL, SKIP, 1>(c, lut, o, d, t0, t1, color, opacity, ns);
  else
    march_xp<I64, U, LUT_GLOBAL, SKIP, 0>(c, lut, o, d, t0, t1, color, opacity, ns);
}

__device__ __forceinline__ bool vol_is_large(const vrx_volume &v)
{
  return (uint64_t)v.dims[0] * v.dims[1] * v.dims[2] >= (1ull << 31);
}

// rayMarchAllVolumes (volumeIntegration.h:317-350)
// The first sample of a segment, bt0 plus the two jitters of rayMarchVolume
// (volumeIntegration.h:120, the field's step) and _rayMarchVolume (:83, the
// sampling step) when jit is given: two PCG draws per segment, in march order.
__device__ __forceinline__ float jittered_start(const vrx_volume &V, const VolCtx &c, float t0,
    Pcg *jit)
{
  if (!jit)
    return t0;
  t0 += vol_step(V) * pcg_float(*jit);
  return t0 + c.step * pcg_float(*jit);
}

__device__ float march_all(const vrx_scene &sc, v3 o, v3 d, float tlower, float tfar,
    float inv_rate, v3 &color, float &opacity, uint32_t &obj, uint32_t &inst, uint32_t &ns,
    bool skip, Pcg *jit = nullptr)
{
  float depth = tfar;
  bool first = true;
  int last = -1;
  float tl = tlower;
  for (int seg = 0; seg < kMaxVolumeSegments; seg++) {
    int best = -1;
    float bt0 = 0.0f, bt1 = 0.0f;
    for (uint32_t i = 0; i < sc.num_volumes; i++) {
      if ((int)i == last)
        continue;
      float t0, t1;
      if (vol_slab(sc.d_volumes[i], o, d, tl, tfar, t0, t1)) {
        if (best < 0 || t0 < bt0) {
          best = (int)i;
          bt0 = t0;
          bt1 = t1;
        }
      }
    }
    if (best < 0)
      break;
    const vrx_volume &V = sc.d_volumes[best];
    if (first) {
      obj = V.id;
      inst = V.inst_id;
      first = false;
    }
    depth = fminf(depth, bt0);
    const VolCtx c = vol_ctx(V, inv_rate, skip);
    const float ts = jittered_start(V, c, bt0, jit);
    if (vol_is_large(V))
      march<true, 2, true>(c, (const float4 *)V.d_lut, o, d, ts, bt1, color, opacity, ns);
    else
      march<false, 2, true>(c, (const float4 *)V.d_lut, o, d, ts, bt1, color, opacity, ns);
    tl = bt1 + 1e-3f;
    last = best;
    if (!(opacity < 0.99f))
      break;
  }
  return depth;
}

// --------------------------------------------------------------------------
// channel packing (gpu_util.h:323-389, glm packUnorm4x8 / convertLinearToSRGB)
// --------------------------------------------------------------------------
__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

__device__ __forceinline__ uint32_t pack_unorm4x8(float x, float y, float z, float w)
{
  uint32_t r = (uint32_t)(unsigned char)roundf(clamp01(x) * 255.0f);
  r |= (uint32_t)(unsigned char)roundf(clamp01(y) * 255.0f) << 8;
  r |= (uint32_t)(unsigned char)roundf(clamp01(z) * 255.0f) << 16;
  r |= (uint32_t)(unsigned char)roundf(clamp01(w) * 255.0f) << 24;
  return r;
}

__device__ __forceinline__ float to_srgb(float v)
{
  float c = clamp01(v);
  return c < 0.0031308f ? c * 12.92f : powf(c, 0.41666f) * 1.055f - 0.055f;
}

__device__ __forceinline__ void write_color(const vrx_frame &f, size_t idx, float4 c)
{
  if (f.d_color_f32)
    ((float4 *)f.d_color_f32)[idx] = c;
  if (f.d_color_u32) {
    if (f.color_format == VRX_COLOR_SRGB8)
      f.d_color_u32[idx] = pack_unorm4x8(to_srgb(c.x), to_srgb(c.y), to_srgb(c.z), c.w);
    else
      f.d_color_u32[idx] = pack_unorm4x8(c.x, c.y, c.z, c.w);
  }
}

// accumResults / writeOutputColor (gpu_util.h:356-443): the accumulated
// tonemapped colour, its inverse tonemap, depth and (when depth improved) ids.
__device__ __forceinline__ void write_pixel(const vrx_frame &f, size_t idx, float4 accum, int spp,
    float dcur, bool ids_changed, uint32_t pid, uint32_t oid, uint32_t iid)
{
  if (f.d_accum)
    ((float4 *)f.d_accum)[idx] = accum;
  float div = (float)(f.frame_id + (uint32_t)spp);
  float4 avg = make_float4(accum.x / div, accum.y / div, accum.z / div, accum.w / div);
  float s2 = fmaxf(1e-12f, 1.0f - cmax3(mk(avg.x, avg.y, avg.z)));
  float4 outv = make_float4(avg.x / s2, avg.y / s2, avg.z / s2, avg.w);
  write_color(f, idx, outv);
  if (f.d_depth)
    f.d_depth[idx] = dcur;
  if (ids_changed) {
    if (f.d_prim_id)
      f.d_prim_id[idx] = pid;
    if (f.d_obj_id)
      f.d_obj_id[idx] = oid;
    if (f.d_inst_id)
      f.d_inst_id[idx] = iid;
  }
}

// Adds a wave's sum of v to frame counter c (0 rays, 1 volume samples): one
// atomic per wave, into the counter slot of the workgroup (slots spread the
// atomics of a launch over VRX_COUNTER_SLOTS cache lines).
__device__ __forceinline__ void wave_count_add(uint64_t *ctrs, uint32_t c, uint32_t v)
{
  uint32_t s = v;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0 && s) {
    const uint32_t slot = (blockIdx.x + blockIdx.y * gridDim.x) % VRX_COUNTER_SLOTS;
    atomicAdd((unsigned long long *)(ctrs + slot * VRX_COUNTER_STRIDE + c), (unsigned long long)s);
  }
}

// --------------------------------------------------------------------------
// shadow rays, ambient occlusion, lights
// --------------------------------------------------------------------------
// surfaceAttenuation (intersectRay.h:105-111) through the any-hit in trace():
// the opacity accumulated by the surfaces on [0, tmax].  A ray whose opacity
// reached 0.99 counts as fully blocked: which hits the reference visits before
// it stops depends on its traversal order, so its value then lies anywhere in
// [0.99, 1]; below 0.99 every hit is visited and the value is the same.
template <int PT, bool TWO, bool WIDE>
__device__ __forceinline__ float surface_atten(
    const vrx_scene &sc, v3 o, v3 d, float tmax, Stack &st, uint32_t &rays)
{
  Hit hh;
  rays++;
  const bool term = TWO ? trace<PT, true, true, WIDE>(sc, o, d, 0.0f, tmax, false, st, hh)
                        : trace<PT, false, true, WIDE>(sc, o, d, 0.0f, tmax, false, st, hh);
  return term ? 1.0f : hh.a0;
}

// computeAO (gpu/computeAO.h:38-60): n cosine-weighted hemisphere rays from
// the geometric-normal offset point, four PCG draws each.
template <int PT, bool TWO, bool WIDE>
__device__ float ao_factor(const vrx_scene &sc, const SurfHit &s, float dist, int n, Pcg &rs,
    Stack &st, uint32_t &rays)
{
  float weights = 0.0f, hits = 0.0f;
  for (int k = 0; k < n; k++) {
    v3 ao_o = add(s.hp, fmul(s.eps, s.ng));
    float q0 = pcg_float(rs), q1 = pcg_float(rs), q2 = pcg_float(rs);
    (void)pcg_float(rs);
    v3 dir = normalize(mk(2.0f * q0 - 1.0f, 2.0f * q1 - 1.0f, 2.0f * q2 - 1.0f));
    if (!(dot(dir, s.ns) > 0.0f))
      dir = neg(dir);
    float wgt = fmaxf(0.0f, dot(dir, s.ns));
    weights += wgt;
    if (wgt != 0.0f)
      hits += wgt * surface_atten<PT, TWO, WIDE>(sc, ao_o, dir, dist, st, rays);
  }
  return weights > 0.0f ? 1.0f - hits / weights : 0.0f;
}

// volumeAttenuation (DirectLight_ptx.cu:62-69 with the shadow any-hit at
// :270-279): every volume the shadow ray crosses on [0, tmax] is marched into
// one opacity (rayMarchVolume without colour, volumeIntegration.h:64-103),
// volumes in index order, stopping at 0.99.
__device__ float volume_atten(const vrx_scene &sc, v3 o, v3 d, float tmax, float inv_rate,
    bool skip, uint32_t &ns, Pcg *jit = nullptr)
{
  float opacity = 0.0f;
  for (uint32_t i = 0; i < sc.num_volumes && opacity < 0.99f; i++) {
    const vrx_volume &V = sc.d_volumes[i];
    float t0, t1;
    if (!vol_slab(V, o, d, 0.0f, tmax, t0, t1))
      continue;
    v3 unused = mk(0.0f, 0.0f, 0.0f);
    const VolCtx c = vol_ctx(V, inv_rate, skip);
    const float ts = jittered_start(V, c, t0, jit);
    if (vol_is_large(V))
      march<true, 2, true>(c, (const float4 *)V.d_lut, o, d, ts, t1, unused, opacity, ns);
    else
      march<false, 2, true>(c, (const float4 *)V.d_lut, o, d, ts, t1, unused, opacity, ns);
  }
  return opacity;
}

// sampleLight (gpu/sampleLight.h:58-100) for a light already in world space.
struct LightSample {
  v3 radiance, dir;
  float dist;
};

__device__ __forceinline__ LightSample sample_light(const vrx_light *L, v3 hp)
{
  LightSample ls;
  const int type = gld(&L->type);
  const v3 rad = mk(gld(&L->radiance[0]), gld(&L->radiance[1]), gld(&L->radiance[2]));
  if (type == VRX_LIGHT_DIRECTIONAL) {
    ls.dir = mk(gld(&L->dir[0]), gld(&L->dir[1]), gld(&L->dir[2]));
    ls.dist = INFINITY;
    ls.radiance = rad;
  } else {
    const v3 pos = mk(gld(&L->pos[0]), gld(&L->pos[1]), gld(&L->pos[2]));
    if (type == VRX_LIGHT_POINT) {
      const v3 v = sub(pos, hp);
      ls.dist = sqrtf(dot(v, v));
      ls.dir = normalize(v);
      ls.radiance = rad;
    } else {
      // spot: the reference measures the distance after normalising the
      // direction, so its shadow rays end about one unit from the surface
      ls.dir = normalize(sub(pos, hp));
      ls.dist = sqrtf(dot(ls.dir, ls.dir));
      const v3 sd = normalize(mk(gld(&L->dir[0]), gld(&L->dir[1]), gld(&L->dir[2])));
      const float co = gld(&L->cos_outer), ci = gld(&L->cos_inner);
      float spot = dot(sd, neg(ls.dir));
      if (spot < co)
        spot = 0.0f;
      else if (spot > ci)
        spot = 1.0f;
      else {
        spot = (spot - co) / (ci - co);
        spot = spot * spot * (3.0f - 2.0f * spot);
      }
      ls.radiance = mulf(rad, spot);
    }
  }
  return ls;
}

// shadeSurface of the directLight renderer (DirectLight_ptx.cu:71-225) for
// matte surfaces: AO factor, ambient term times the tint, and per light the
// Lambertian term (MatteShader_ptx.cu:71-79: base/pi * max(0, Ns.L) *
// radiance / pdf, pdf = 1) times the surface and volume transmission of the
// shadow ray from the shading point.  Matte surfaces spawn no bounce ray
// (their nextRay is zero), so the bounce loop ends at once.  Returns the
// colour and the material opacity.
template <int PT, bool TWO, bool WIDE>
__device__ float4 shade_direct(const vrx_scene &sc, const vrx_renderer &r, const SurfHit &s,
    int aos, Pcg &rs, Stack &st, uint32_t &rays, uint32_t &vsamples)
{
  const v3 sp = add(s.hp, fmul(s.eps, s.ns));
  const float aof =
      aos > 0 ? ao_factor<PT, TWO, WIDE>(sc, s, r.ao_distance, aos, rs, st, rays) : 1.0f;
  v3 contrib = mk(0.0f, 0.0f, 0.0f);
  if (r.ambient_radiance > 0.0f)
    contrib = mulv(mulf(ld3(r.ambient_color), r.ambient_radiance), s.color);
  const v3 kd = mulf(s.color, 0.318309886183790672f);
  for (uint32_t l = 0; l < sc.num_lights; l++) {
    const LightSample ls = sample_light(sc.d_lights + l, s.hp);
    const float ndl = fmaxf(0.0f, dot(s.ns, ls.dir));
    const v3 c = mulv(mulf(kd, ndl), ls.radiance);
    if (c.x != c.x || c.y != c.y || c.z != c.z)
      continue;
    // a light at grazing or back-facing incidence adds exactly zero whatever
    // its shadow ray finds, so that ray is not traced
    if (c.x == 0.0f && c.y == 0.0f && c.z == 0.0f)
      continue;
    const float so = 1.0f - surface_atten<PT, TWO, WIDE>(sc, sp, ls.dir, ls.dist, st, rays);
    float vo = 1.0f;
    if (PT < 0 && sc.num_volumes)
      vo = 1.0f - volume_atten(sc, sp, ls.dir, ls.dist, r.inv_volume_sampling_rate,
                          r.skip_empty != 0, vsamples, r.volume_jitter ? &rs : nullptr);
    const float att = so * vo;
    contrib = add(contrib, mulf(c, att));
  }
  contrib = mulf(contrib, aof);
  return make_float4(contrib.x, contrib.y, contrib.z, s.opacity);
}

// --------------------------------------------------------------------------
// long primary rays (raycast renderer, one-level scenes of one primitive type)
// --------------------------------------------------------------------------
// A launch waits for its slowest wave, and on scenes seen at grazing angles
// (a heightfield's horizon) a few rays walk far more nodes than the rest,
// one dependent fetch at a time.  With a step budget, the first pass gives
// up on a ray after `budget` loop iterations and queues its pixel; a second
// launch gives each queued ray 16 lanes that walk its BVH together
// (trace_group) and takes the closest hit over the lanes with take()'s rule
// (smaller t, then lower entry and primitive).  Every intersection test is
// the same, so the pixel comes out as the single walk would give it.
struct LongRays {
  uint32_t budget;  // loop iterations before a primary walk is deferred
  uint32_t *count;  // queued pixels
  uint32_t *queue;  // one slot per pixel
  uint32_t *gs;     // second pass: this lane group's shared stack (LDS)
  uint32_t gcap;    // its capacity in entries
};

// queue pixel idx (one atomic per wave: the lanes deferring together)
__device__ __forceinline__ void defer_pixel(const LongRays &lr, uint32_t idx)
{
  const uint64_t m = __ballot(1);
  const uint32_t n = (uint32_t)__popcll(m);
  const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
      __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
  const int leader = __ffsll((long long)m) - 1;
  uint32_t base = 0;
  if (rank == 0)
    base = atomicAdd(lr.count, n);
  base = __shfl(base, leader, 64);
  gst(lr.queue, base + rank, idx);
}

// Closest hit of one ray walked by the 16 lanes of an aligned lane group
// that share one LIFO stack in LDS (gs, gcap entries): each step, every lane
// without a node takes one from its private stack or the shared one, visits
// it, keeps its nearest hit child and pushes the other hit children to the
// shared stack; boxes are culled against the group's closest hit so far.
// Each lane keeps its own closest hit by take() (so ties resolve as in one
// walk) and the group takes the closest over its lanes.  A push that would
// overflow the shared stack goes to the lanes' private stacks instead.
template <int PT, bool WIDE>
__device__ __forceinline__ bool trace_group(const vrx_scene &sc, v3 wo, v3 wd, float tmin,
    float tmax, bool cull, Stack &st, Hit &h, uint32_t *gs, uint32_t gcap)
{
  h.found = false;
  h.t = tmax;
  h.entry = 0;
  h.prim = 0;
  h.a0 = h.a1 = h.a2 = 0.0f;
  if (sc.num_entries == 0)
    return false;
  st.cnt = 0;
  BlasCtx b;
  v3 o, d;
  uint32_t root, n;
  load_blas(sc, 0u, wo, wd, cull, false, b, o, d, root, n);
  if (n == 0)
    return false;
  const v3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  WT wt;
  if (blas_type<PT>(b) == VRX_TRIANGLE)
    wt = wt_setup(d);
  const uint32_t lane = threadIdx.x & 63u, s = lane & 15u;
  const uint64_t gmask = 0xFFFFull << (lane & ~15u);
  const uint64_t below = (1ull << lane) - 1ull;
  bool has = s == 0;
  uint32_t node = root;
  uint32_t top = 0;  // shared stack entries (the same in all 16 lanes)
  float tb = tmax;   // the group's closest hit so far
  for (;;) {
    if (!has && st.cnt > 0) {
      node = st.pop();
      has = true;
    }
    const uint64_t need = __ballot(!has) & gmask;
    const uint32_t take = min((uint32_t)__popcll(need), top);
    if (!has) {
      const uint32_t rank = (uint32_t)__popcll(need & below);
      if (rank < take) {
        node = gs[top - 1u - rank];
        has = true;
      }
    }
    top -= take;
    if ((__ballot(has) & gmask) == 0ull)
      break;
    // visit: up to three children to hand out (the nearest is kept)
    uint32_t out[3];
    uint32_t nout = 0;
    if (has) {
      if (WIDE && !(node & kLeafBit)) {
        const Node4 nd = gld_node4(b.wnodes, node);
        const float lx[4] = {nd.lox.x, nd.lox.y, nd.lox.z, nd.lox.w};
        const float ly[4] = {nd.loy.x, nd.loy.y, nd.loy.z, nd.loy.w};
        const float lz[4] = {nd.loz.x, nd.loz.y, nd.loz.z, nd.loz.w};
        const float hx[4] = {nd.hix.x, nd.hix.y, nd.hix.z, nd.hix.w};
        const float hy[4] = {nd.hiy.x, nd.hiy.y, nd.hiy.z, nd.hiy.w};
        const float hz[4] = {nd.hiz.x, nd.hiz.y, nd.hiz.z, nd.hiz.w};
        uint32_t l[4] = {nd.link.x, nd.link.y, nd.link.z, nd.link.w};
        float t[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          float tn;
          const bool hit = (uint32_t)c < nd.meta.x
              && slab(lx[c], ly[c], lz[c], hx[c], hy[c], hz[c], o, inv, tmin, tb, tn);
          t[c] = hit ? tn : INFINITY;
        }
        cswap(t[0], l[0], t[1], l[1]);
        cswap(t[2], l[2], t[3], l[3]);
        cswap(t[0], l[0], t[2], l[2]);
        cswap(t[1], l[1], t[3], l[3]);
        cswap(t[1], l[1], t[2], l[2]);
        has = t[0] < INFINITY;
        node = l[0];
        if (t[3] < INFINITY)
          out[nout++] = l[3];
        if (t[2] < INFINITY)
          out[nout++] = l[2];
        if (t[1] < INFINITY)
          out[nout++] = l[1];
      } else if (!(node & kLeafBit)) {
        const Node nd = gld_node(b.nodes, node);
        float tn0, tn1;
        const bool h0 = slab(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, o, inv, tmin, tb, tn0);
        const bool h1 = slab(nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, o, inv, tmin, tb, tn1);
        has = h0 || h1;
        if (h0 && h1) {
          const bool swap = tn1 < tn0;
          node = swap ? nd.d.y : nd.d.x;
          out[nout++] = swap ? nd.d.x : nd.d.y;
        } else {
          node = h0 ? nd.d.x : nd.d.y;
        }
      } else {
        float tcur = h.t;
        (void)test_leaf<PT, false>(b, node & ~kLeafBit, o, d, wt, tmin, tcur, h);
        has = false;
      }
    }
    // hand the other children to the group: exclusive prefix sum over the 16 lanes
    uint32_t incl = nout;
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) {
      const uint32_t v = (uint32_t)__shfl_up((int)incl, (unsigned)m, 16);
      if (s >= (uint32_t)m)
        incl += v;
    }
    const uint32_t total = (uint32_t)__shfl((int)incl, 15, 16);
    if (top + total <= gcap) {
      const uint32_t at = top + incl - nout;
      for (uint32_t j = 0; j < nout; j++)
        gs[at + j] = out[j];
      top += total;
    } else {
      for (uint32_t j = 0; j < nout; j++)
        st.push(out[j]);
    }
    // the group's culling bound: the closest hit of any lane
    float t = h.found ? h.t : tmax;
#pragma unroll
    for (int m = 1; m < 16; m <<= 1)
      t = fminf(t, __shfl_xor(t, m, 64));
    tb = t;
  }
  bool found = h.found;
#pragma unroll
  for (int m = 1; m < 16; m <<= 1) {
    const bool of = __shfl_xor((int)found, m, 64) != 0;
    const float ot = __shfl_xor(h.t, m, 64);
    const uint32_t oe = (uint32_t)__shfl_xor((int)h.entry, m, 64);
    const uint32_t op = (uint32_t)__shfl_xor((int)h.prim, m, 64);
    const float oa0 = __shfl_xor(h.a0, m, 64), oa1 = __shfl_xor(h.a1, m, 64),
                oa2 = __shfl_xor(h.a2, m, 64);
    const bool better = of
        && (!found || ot < h.t || (ot == h.t && (oe < h.entry || (oe == h.entry && op < h.prim))));
    if (better) {
      found = true;
      h.t = ot;
      h.entry = oe;
      h.prim = op;
      h.a0 = oa0;
      h.a1 = oa1;
      h.a2 = oa2;
    }
  }
  h.found = found;
  return found;
}

// --------------------------------------------------------------------------
// the per-pixel program
// --------------------------------------------------------------------------
// One pixel of the raycast / AO / directLight program (Raycast_ptx.cu:60-179,
// AmbientOcclusion_ptx.cu:86-220, DirectLight_ptx.cu:311-418): all pixel
// samples, the surface/volume layer loop, accumulation and channel writes at
// output index idx.
// LONG (raycast, one-level scenes of one primitive type only): 1 = walks
// over lr.budget iterations queue their pixel and write nothing; 2 = the
// pixel's walks are shared by its 16-lane group (trace_group), and the
// group's first lane writes it.
template <int PT, bool TWO, int KIND, bool WIDE = false, int LONG = 0>
__device__ __forceinline__ void render_pixel(const vrx_scene &sc, const vrx_camera &cam,
    const vrx_renderer &r, const vrx_frame &f, uint32_t px, uint32_t py, size_t idx, Stack &st,
    uint32_t &rays, uint32_t &vsamples, const LongRays *lr = nullptr)
{
  static_assert(LONG == 0 || (KIND == VRX_RENDER_RAYCAST && !TWO && PT >= 0),
      "long-ray passes are for the one-level primitive raycast");
  const float iw = 1.0f / (float)f.width, ih = 1.0f / (float)f.height;
  constexpr bool kSampled = KIND == VRX_RENDER_AO || KIND == VRX_RENDER_DIRECT;
  const int spp = kSampled ? max(r.pixel_samples, 1) : 1;
  const int aos = min(max(r.ambient_samples, 0), 256);
  const bool cull = r.cull_backfaces != 0;
  Pcg rs;
  // the sequence is keyed by the pixel within its view, so a view rendered
  // in a batch draws the same numbers as the same view rendered alone
  if (kSampled || r.volume_jitter)
    pcg_seed(rs, (uint64_t)px + (uint64_t)py * f.width, (uint64_t)f.frame_id);
  Pcg *jit = r.volume_jitter ? &rs : nullptr;

  float4 accum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float dcur = FLT_MAX;
  uint32_t pid = 0, oid = 0, iid = 0;
  bool ids_changed = f.frame_id == 0;
  if (f.frame_id != 0) {
    accum = ((const float4 *)f.d_accum)[idx];
    if (f.d_depth)
      dcur = f.d_depth[idx];
  }
  for (int i = 0; i < spp; i++) {
    v3 o, d;
    if (kSampled) {
      float r0 = pcg_float(rs), r1 = pcg_float(rs);
      (void)pcg_float(rs);
      (void)pcg_float(rs);
      cam_ray(cam, ((float)px + r0) * iw, ((float)py + r1) * ih, o, d);
    } else {
      cam_ray(cam, (float)px * iw, (float)py * ih, o, d);
    }
    float tlower = 0.0f;
    const float tmax = FLT_MAX;
    v3 outc = mk(0.0f, 0.0f, 0.0f);
    float outo = 0.0f, depth = 1e30f;
    uint32_t prim = ~0u, obj = ~0u, inst = ~0u;
    bool first = true;
    for (int layer = 0; layer < kMaxSurfaceLayers && outo < 0.99f; layer++) {
      Hit h;
      rays++;
      bool found;
      if constexpr (LONG == 1) {
        bool deferred = false;
        found = trace<PT, false, false, WIDE, true>(sc, o, d, tlower, tmax, cull, st, h, lr->budget,
            &deferred);
        if (deferred) {
          // the second pass renders the whole pixel and counts its rays
          defer_pixel(*lr, (uint32_t)idx);
          rays -= (uint32_t)(layer + 1);
          return;
        }
      } else if constexpr (LONG == 2) {
        found = trace_group<PT, WIDE>(sc, o, d, tlower, tmax, cull, st, h, lr->gs, lr->gcap);
      } else {
        found = TWO ? trace<PT, true, false, WIDE>(sc, o, d, tlower, tmax, cull, st, h)
                    : trace<PT, false, false, WIDE>(sc, o, d, tlower, tmax, cull, st, h);
      }
      v3 color = mk(0.0f, 0.0f, 0.0f);
      float opacity = 0.0f;
      if (found) {
        SurfHit s;
        populate(sc, h, o, d, s);
        uint32_t vo = ~0u, vi = ~0u;
        float vdepth = s.t;
        if (PT < 0 && sc.num_volumes) // scenes with volumes use the generic kernel
          vdepth = march_all(sc, o, d, tlower, s.t, r.inv_volume_sampling_rate, color, opacity,
              vo, vi, vsamples, r.skip_empty != 0, jit);
        if (first) {
          if (vdepth < s.t) {
            depth = vdepth;
            prim = 0;
            obj = vo;
            inst = vi;
          } else {
            depth = s.t;
            prim = s.prim_out;
            obj = s.obj;
            inst = s.inst;
          }
          first = false;
        }
        if (KIND == VRX_RENDER_DIRECT) {
          const float4 sr = shade_direct<PT, TWO, WIDE>(sc, r, s, aos, rs, st, rays, vsamples);
          // DirectLight_ptx.cu:372-381: the shading result replaces the
          // colour and opacity marched through volumes in front of the
          // surface, and is then accumulated onto itself
          if (sr.x != sr.x || sr.y != sr.y || sr.z != sr.z) {
            color = mk(0.0f, 0.0f, 0.0f);
            opacity = 0.0f;
          } else {
            color = mk(sr.x, sr.y, sr.z);
            opacity = sr.w;
          }
          float w = 1.0f - opacity;
          color = mk(color.x + sr.x * w, color.y + sr.y * w, color.z + sr.z * w);
          opacity = opacity + sr.w * (1.0f - opacity);
        } else {
          v3 lighting;
          if (KIND == VRX_RENDER_AO) {
            const float ao =
                aos > 0 ? ao_factor<PT, TWO, WIDE>(sc, s, r.ao_distance, aos, rs, st, rays) : 1.0f;
            float scl = ao * r.ambient_radiance;
            lighting = fmul(scl, ld3(r.ambient_color));
          } else {
            float scl = fabsf(dot(d, s.ns));
            lighting = fmul(scl, ld3(r.ambient_color));
          }
          v3 bl = mulv(s.color, lighting);
          float w = 1.0f - opacity;
          color = mk(color.x + bl.x * w, color.y + bl.y * w, color.z + bl.z * w);
          opacity = opacity + s.opacity * (1.0f - opacity);
        }
        color = mulf(color, opacity);
        float wo = 1.0f - outo;
        outc = mk(outc.x + color.x * wo, outc.y + color.y * wo, outc.z + color.z * wo);
        outo = outo + opacity * (1.0f - outo);
        tlower = s.t + s.eps;
      } else {
        uint32_t vo = ~0u, vi = ~0u;
        float vdepth = tmax;
        if (PT < 0 && sc.num_volumes) // scenes with volumes use the generic kernel
          vdepth = march_all(sc, o, d, tlower, tmax, r.inv_volume_sampling_rate, color, opacity,
              vo, vi, vsamples, r.skip_empty != 0, jit);
        if (first) {
          depth = fminf(depth, vdepth);
          prim = 0;
          obj = vo;
          inst = vi;
        }
        color = mulf(color, opacity);
        float w = 1.0f - opacity;
        color = mk(color.x + r.background[0] * w, color.y + r.background[1] * w,
            color.z + r.background[2] * w);
        opacity = opacity + r.background[3] * (1.0f - opacity);
        float wo = 1.0f - outo;
        outc = mk(outc.x + color.x * wo, outc.y + color.y * wo, outc.z + color.z * wo);
        outo = outo + opacity * (1.0f - outo);
        break;
      }
    }
    // accumResults (gpu_util.h:393-443)
    float sc_ = 1.0f + fmaxf(0.0f, cmax3(outc));
    accum.x = accum.x + outc.x / sc_;
    accum.y = accum.y + outc.y / sc_;
    accum.z = accum.z + outc.z / sc_;
    accum.w = accum.w + outo;
    if (depth < dcur) {
      dcur = depth;
      pid = prim;
      oid = obj;
      iid = inst;
      ids_changed = true;
    }
  }
  if (LONG != 2 || (threadIdx.x & 15u) == 0)
    write_pixel(f, idx, accum, spp, dcur, ids_changed, pid, oid, iid);
}

} // namespace vrx
