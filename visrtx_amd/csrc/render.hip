// render.hip -- raycast, ambient-occlusion and directLight renderers, volume marching,
// frame clears, camera setup and the instance-level BVH build.
//
// Per-pixel programs restated from the reference raygen programs:
//   renderer/Raycast_ptx.cu:60-179          (raycast: first-hit ids, facing-ratio shading)
//   renderer/AmbientOcclusion_ptx.cu:86-220 (spp loop, AO any-hit rays, gpu/computeAO.h:38-60)
//   renderer/DirectLight_ptx.cu:62-420      (lights, shadow rays, gpu/sampleLight.h:58-100)
//   gpu/volumeIntegration.h:46-165,317-350  (emission-absorption march, all volumes)
//   gpu/gpu_util.h:323-443                  (tonemap accumulation, channel writes)
//   gpu/cameraCreateRay.h:38-81             (pinhole rays; raycast uses the pixel corner)
// Launch shape: one 256-lane workgroup per 16x16 pixel tile, each wave an
// 8x8 tile (coherent rays share node fetches); the traversal short stack
// lives in dynamic LDS, laid out [slot][lane] so a wave's pushes hit 64
// distinct banks.
#include <float.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "vrx_internal.h"
#include "vrx_shade.h"

namespace vrx {

template <int PT, bool TWO, int KIND, bool WIDE, int LONG = 0>
__device__ __forceinline__ void render_tile(const vrx_scene &sc, const vrx_camera &cam,
    const vrx_renderer &r, const vrx_frame &f, uint32_t sd, uint32_t ovf_cap, uint32_t *scratch,
    const LongRays *lr = nullptr)
{
  extern __shared__ uint32_t s_stack[];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const uint32_t px = blockIdx.x * kTile + (wave & 1) * 8 + (lane & 7);
  const uint32_t py = blockIdx.y * kTile + (wave >> 1) * 8 + (lane >> 3);
  const bool inside = px < f.width && py < f.height;

  Stack st;
  st.lds = s_stack + tid;
  st.stride = kBlock;
  st.sd = sd;
  st.sd_mask = sd - 1;
  const uint32_t gthreads = gridDim.x * gridDim.y * kBlock;
  const uint32_t gtid = (blockIdx.y * gridDim.x + blockIdx.x) * kBlock + tid;
  st.ovf = scratch + 64 + gtid;
  st.ovf_stride = gthreads;
  st.ovf_cap = ovf_cap;
  st.err = scratch;
  st.cnt = 0;

  uint32_t rays = 0, vsamples = 0;
  if (inside)
    render_pixel<PT, TWO, KIND, WIDE, LONG>(sc, cam, r, f, px, py, (size_t)px + (size_t)py * f.width,
        st, rays, vsamples, lr);
  if (f.d_counters) {
    wave_count_add(f.d_counters, 0, rays);
    wave_count_add(f.d_counters, 1, vsamples);
  }
}

template <int PT, bool TWO, int KIND, bool WIDE>
__global__ __launch_bounds__(kBlock) void k_render(vrx_scene sc, vrx_camera cam, vrx_renderer r,
    vrx_frame f, uint32_t sd, uint32_t ovf_cap, uint32_t *scratch)
{
  render_tile<PT, TWO, KIND, WIDE>(sc, cam, r, f, sd, ovf_cap, scratch);
}

// The same program held to 80 VGPRs (6 waves per SIMD), for the
// primitive-specialised AO and directLight programs and two-level raycast:
// their secondary rays and instance walks are latency-bound, and the extra
// waves outweigh the registers spilled to scratch (10M spheres: AO 16 spp
// 13.1 -> 11.9 ms, directLight 4 spp 6.4 -> 5.7 ms; 1M instances +7 %).
// Single-level raycast (66-88 VGPRs) and the generic kernel (130-150 VGPRs)
// keep k_render.
template <int PT, bool TWO, int KIND, bool WIDE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_render_occ6(
    vrx_scene sc, vrx_camera cam, vrx_renderer r, vrx_frame f, uint32_t sd, uint32_t ovf_cap,
    uint32_t *scratch)
{
  render_tile<PT, TWO, KIND, WIDE>(sc, cam, r, f, sd, ovf_cap, scratch);
}

// The one-level primitive raycast with a step budget: walks longer than
// lr.budget loop iterations queue their pixel for k_render_long.
template <int PT, bool WIDE>
__global__ __launch_bounds__(kBlock) void k_render_defer(vrx_scene sc, vrx_camera cam,
    vrx_renderer r, vrx_frame f, uint32_t sd, uint32_t ovf_cap, uint32_t *scratch, LongRays lr)
{
  render_tile<PT, false, VRX_RENDER_RAYCAST, WIDE, 1>(sc, cam, r, f, sd, ovf_cap, scratch, &lr);
}

// The queued pixels, 16 lanes each (four per wave), over a resident grid.
// Dynamic LDS: the lanes' private short stacks (sd entries each), then one
// shared stack of lr.gcap entries per lane group.
template <int PT, bool WIDE>
__global__ __launch_bounds__(kBlock) void k_render_long(vrx_scene sc, vrx_camera cam,
    vrx_renderer r, vrx_frame f, uint32_t sd, uint32_t ovf_cap, uint32_t *scratch, LongRays lr)
{
  extern __shared__ uint32_t s_stack[];
  const int tid = threadIdx.x;
  lr.gs = s_stack + sd * kBlock + (uint32_t)(tid >> 4) * lr.gcap;
  Stack st;
  st.lds = s_stack + tid;
  st.stride = kBlock;
  st.sd = sd;
  st.sd_mask = sd - 1;
  const uint32_t gthreads = gridDim.x * kBlock;
  st.ovf = scratch + 64 + blockIdx.x * kBlock + tid;
  st.ovf_stride = gthreads;
  st.ovf_cap = ovf_cap;
  st.err = scratch;
  st.cnt = 0;
  // the queue holds at most one entry per pixel; entries are pixel indices
  const uint32_t npix = f.width * f.height;
  const uint32_t n = min(gld(lr.count, 0), npix);
  const uint32_t groups = gridDim.x * (kBlock / 16);
  uint32_t rays = 0, vsamples = 0;
  for (uint32_t i = blockIdx.x * (kBlock / 16) + ((uint32_t)tid >> 4); i < n; i += groups) {
    const uint32_t idx = gld(lr.queue, i);
    if (idx >= npix)
      continue;
    const uint32_t py = idx / f.width, px = idx - py * f.width;
    uint32_t rr = 0;
    render_pixel<PT, false, VRX_RENDER_RAYCAST, WIDE, 2>(sc, cam, r, f, px, py, idx, st, rr,
        vsamples, &lr);
    if ((tid & 15) == 0)
      rays += rr;
  }
  if (f.d_counters)
    wave_count_add(f.d_counters, 0, rays);
}

// Raycast renderer on a world holding exactly one volume and no surfaces:
// the same per-pixel program as k_render's miss branch (the march of
// rayMarchAllVolumes, the renderer's colour-times-opacity, background), with
// the transfer-function LUT staged in LDS and no traversal state, so the
// kernel needs about half the registers of the general one.  One lane per ray
// (march()): the kernel for volumes with a macro-cell grid, whose
// empty-space skipping walks cells one dependent lookup at a time, so the
// launch wants as many rays in flight as it can hold.
template <bool I64, int U, bool SKIP>
__global__ __launch_bounds__(kBlock) void k_render_volume_seq(vrx_scene sc, vrx_camera cam,
    vrx_renderer r, vrx_frame f)
{
  __shared__ float4 s_lut[256];
  const vrx_volume &V = sc.d_volumes[0];
  s_lut[threadIdx.x] = gld((const float4 *)V.d_lut, threadIdx.x);
  __syncthreads();
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const uint32_t px = blockIdx.x * kTile + (wave & 1) * 8 + (lane & 7);
  const uint32_t py = blockIdx.y * kTile + (wave >> 1) * 8 + (lane >> 3);
  uint32_t vsamples = 0;
  if (px < f.width && py < f.height) {
    const size_t idx = (size_t)px + (size_t)py * f.width;
    const float iw = 1.0f / (float)f.width, ih = 1.0f / (float)f.height;
    v3 o, d;
    cam_ray(cam, (float)px * iw, (float)py * ih, o, d);
    v3 color = mk(0.0f, 0.0f, 0.0f);
    float opacity = 0.0f;
    uint32_t vo = ~0u, vi = ~0u;
    float vdepth = FLT_MAX;
    float t0, t1;
    if (vol_slab(V, o, d, 0.0f, FLT_MAX, t0, t1)) {
      vo = V.id;
      vi = V.inst_id;
      vdepth = t0;
      march<I64, U, false, SKIP>(vol_ctx(V, r.inv_volume_sampling_rate, SKIP), s_lut, o, d, t0, t1, color, opacity,
          vsamples);
    }
    const float depth = fminf(1e30f, vdepth);
    color = mulf(color, opacity);
    float w = 1.0f - opacity;
    color = mk(color.x + r.background[0] * w, color.y + r.background[1] * w,
        color.z + r.background[2] * w);
    opacity = opacity + r.background[3] * (1.0f - opacity);
    v3 outc = mk(0.0f + color.x * 1.0f, 0.0f + color.y * 1.0f, 0.0f + color.z * 1.0f);
    float outo = 0.0f + opacity * 1.0f;
    float4 accum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float dcur = FLT_MAX;
    if (f.frame_id != 0) {
      accum = ((const float4 *)f.d_accum)[idx];
      if (f.d_depth)
        dcur = f.d_depth[idx];
    }
    float sc_ = 1.0f + fmaxf(0.0f, cmax3(outc));
    accum.x = accum.x + outc.x / sc_;
    accum.y = accum.y + outc.y / sc_;
    accum.z = accum.z + outc.z / sc_;
    accum.w = accum.w + outo;
    bool ids_changed = f.frame_id == 0;
    uint32_t pid = 0, oid = 0, iid = 0;
    if (depth < dcur) {
      dcur = depth;
      pid = 0;
      oid = vo;
      iid = vi;
      ids_changed = true;
    }
    write_pixel(f, idx, accum, 1, dcur, ids_changed, pid, oid, iid);
  }
  if (f.d_counters) {
    wave_count_add(f.d_counters, 0, px < f.width && py < f.height ? 1u : 0u);
    wave_count_add(f.d_counters, 1, vsamples);
  }
}

// Quad-lane helpers: lanes 4q..4q+3 of a wave hold the four consecutive
// samples of one ray; quad_perm DPP moves read within the quad.
#define VRX_QUAD_MOV(v, ctrl)                                                                    \
  __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, (v)), (ctrl), 0xF, \
                                0xF, false))
constexpr int kQuadShr1 = 0x90;  // lane j reads lane max(j-1, 0)
constexpr int kQuadShr2 = 0x44;  // lane j reads lane j-2 (lanes 0, 1 read themselves)
constexpr int kQuadLast = 0xFF;  // every lane reads lane 3
constexpr int kQuadXor1 = 0xB1;  // lane j reads lane j^1
constexpr int kQuadXor2 = 0x4E;  // lane j reads lane j^2
constexpr uint32_t kVolTile = 8; // pixels per side of a workgroup's tile (four 4x4 waves)

// _rayMarchVolume (volumeIntegration.h:64-103), jitter off, with four lanes
// per ray: lane j of a ray's quad takes samples 4m + j (m = 0, 1, ...), so a
// wave's 64 lanes gather for 16 neighbouring rays at 4 consecutive positions
// each, about half the distinct cache lines per gather instruction of one lane
// per ray (the march is bound by the L1 tag rate of its gathers).  Sample
// positions, the trilinear gather and the classification are those of
// march(); each block of four consecutive samples is composited at once:
//   P_j = st_0 ... st_j (the block's step transmittances, a quad scan);
//   the opacity before sample j is opacity + T (1 - P_{j-1}), and sample j is
//   taken while every sample up to it lies inside the segment and found that
//   opacity below 0.99 (the sequential march's test);
//   a taken sample adds w_j = T P_{j-1} (1 - st_j) to its lane's colour and
//   opacity partials (summed over the quad once, at the end);
//   a fully taken block moves T to T P_3 and opacity to opacity + T (1 - P_3).
// This regroups the sequential products and sums of march() (rounding-level
// differences, bounded by the FP64 oracle like march() itself).  Sixteen rays
// per wave suit dense fields, where a ray's cost is its samples; volumes with
// a macro-cell grid take the one-lane-per-ray kernel.
template <bool I64, int U, int XP>
__device__ __forceinline__ void march_quad_xp(const VolCtx &c, const float4 *lut, v3 o, v3 d,
    float t0, float t1, uint32_t j, uint32_t lane, v3 &color, float &opacity, uint32_t &ns)
{
  const float gox = (o.x - c.origin[0]) * c.inv_sp[0], gdx = d.x * c.inv_sp[0];
  const float goy = (o.y - c.origin[1]) * c.inv_sp[1], gdy = d.y * c.inv_sp[1];
  const float goz = (o.z - c.origin[2]) * c.inv_sp[2], gdz = d.z * c.inv_sp[2];
  float T = 1.0f, op = 0.0f; // the ray's transmittance and opacity before the block
  for (uint32_t k = 0; k < kMaxSamplesPerSegment; k += 4 * U) {
    float sv[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const float t = __builtin_fmaf((float)(k + 4 * u + j), c.step, t0);
      ok[u] = t <= t1;
      const float gx = __builtin_fmaf(gdx, t, gox), gy = __builtin_fmaf(gdy, t, goy),
                  gz = __builtin_fmaf(gdz, t, goz);
      if constexpr (XP == 3)
        sv[u] = trilinear_ypairs(c, gx, gy, gz);
      else if constexpr (XP == 2)
        sv[u] = trilinear_b32(c, gx, gy, gz);
      else
        sv[u] = trilinear<I64, XP>(c, gx, gy, gz);
    }
    bool stop = false;
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (!stop) {
        const float s = sv[u];
        float st = 1.0f;
        float4 co = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (ok[u] && s == s) {
          co = classify_delta(lut, c, s);
          st = __builtin_amdgcn_exp2f(c.expo * __builtin_amdgcn_logf(1.0f - co.w));
        }
        // inclusive and exclusive products of st over the quad
        float p = st;
        {
          const float y = VRX_QUAD_MOV(p, kQuadShr1);
          p = j >= 1 ? p * y : p;
        }
        {
          const float y = VRX_QUAD_MOV(p, kQuadShr2);
          p = j >= 2 ? p * y : p;
        }
        const float y = VRX_QUAD_MOV(p, kQuadShr1);
        const float pe = j >= 1 ? y : 1.0f;
        const float tb = T * pe; // transmittance before this sample
        const bool cont = ok[u] && op + (T - tb) < 0.99f;
        // a sample is taken when it and every earlier one of its block continue
        const uint64_t m = __ballot(cont);
        uint64_t a = m & ((m << 1) | 0x1111111111111111ull);
        a &= (a << 2) | 0x3333333333333333ull;
        if ((a >> lane) & 1ull) {
          ns++;
          const float w = tb * (1.0f - st);
          color.x = __builtin_fmaf(w, co.x, color.x);
          color.y = __builtin_fmaf(w, co.y, color.y);
          color.z = __builtin_fmaf(w, co.z, color.z);
          opacity += w;
        }
        if (((a >> (lane & ~3u)) & 0xFull) != 0xFull) {
          stop = true;
        } else {
          const float pl = VRX_QUAD_MOV(p, kQuadLast);
          op = op + (T - T * pl);
          T = T * pl;
        }
      }
    }
    if (stop)
      break;
  }
}

template <bool I64, int U>
__device__ __forceinline__ void march_quad(const VolCtx &c, const float4 *lut, v3 o, v3 d, float t0,
    float t1, uint32_t j, uint32_t lane, v3 &color, float &opacity, uint32_t &ns)
{
  if (!(c.step > 0.0f))
    return;
  // XP: 3 the y-pair gathers, 2 the b32 gathers, 1 x-pair gathers with
  // 64-bit offsets, 0 a field one voxel wide in x (one branch per segment,
  // outside the sample loop)
  if (!I64 && c.yp)
    march_quad_xp<I64, U, 3>(c, lut, o, d, t0, t1, j, lane, color, opacity, ns);
  else if (!I64 && c.b32)
    march_quad_xp<I64, U, 2>(c, lut, o, d, t0, t1, j, lane, color, opacity, ns);
  else if (c.dims[0] >= 2)
    march_quad_xp<I64, U, 1>(c, lut, o, d, t0, t1, j, lane, color, opacity, ns);
  else
    march_quad_xp<I64, U, 0>(c, lut, o, d, t0, t1, j, lane, color, opacity, ns);
}

__device__ __forceinline__ float quad_sum(float v)
{
  v = v + VRX_QUAD_MOV(v, kQuadXor1);
  return v + VRX_QUAD_MOV(v, kQuadXor2);
}

// The same program for a volume without a macro-cell grid (dense fields): a
// 256-lane workgroup covers an 8x8 pixel tile, each wave a 4x4 tile with four
// lanes per ray (march_quad); lane 0 of each quad writes the pixel.
template <bool I64, int U>
__global__ __launch_bounds__(kBlock) void k_render_volume_quad(vrx_scene sc, vrx_camera cam,
    vrx_renderer r, vrx_frame f)
{
  // the LUT as (entry, next - entry) pairs (classify_delta)
  __shared__ float4 s_lut[512];
  const vrx_volume &V = sc.d_volumes[0];
  {
    const float4 a = gld((const float4 *)V.d_lut, threadIdx.x);
    const float4 b = gld((const float4 *)V.d_lut, min(threadIdx.x + 1u, 255u));
    s_lut[2 * threadIdx.x] = a;
    s_lut[2 * threadIdx.x + 1] = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, b.w - a.w);
  }
  __syncthreads();
  const uint32_t tid = threadIdx.x;
  const uint32_t wave = tid >> 6, lane = tid & 63, q = lane >> 2, j = lane & 3;
  const uint32_t px = blockIdx.x * kVolTile + (wave & 1) * 4 + (q & 3);
  const uint32_t py = blockIdx.y * kVolTile + (wave >> 1) * 4 + (q >> 2);
  const bool inside = px < f.width && py < f.height;
  uint32_t vsamples = 0;
  if (inside) {
    const size_t idx = (size_t)px + (size_t)py * f.width;
    const float iw = 1.0f / (float)f.width, ih = 1.0f / (float)f.height;
    v3 o, d;
    cam_ray(cam, (float)px * iw, (float)py * ih, o, d);
    v3 color = mk(0.0f, 0.0f, 0.0f);
    float opacity = 0.0f;
    uint32_t vo = ~0u, vi = ~0u;
    float vdepth = FLT_MAX;
    float t0, t1;
    if (vol_slab(V, o, d, 0.0f, FLT_MAX, t0, t1)) {
      vo = V.id;
      vi = V.inst_id;
      vdepth = t0;
      march_quad<I64, U>(vol_ctx(V, r.inv_volume_sampling_rate), s_lut, o, d, t0, t1, j, lane,
          color, opacity, vsamples);
    }
    // the quad's partial sums (every lane of a quad reaches this point)
    color = mk(quad_sum(color.x), quad_sum(color.y), quad_sum(color.z));
    opacity = quad_sum(opacity);
    if (j == 0) {
      const float depth = fminf(1e30f, vdepth);
      color = mulf(color, opacity);
      float w = 1.0f - opacity;
      color = mk(color.x + r.background[0] * w, color.y + r.background[1] * w,
          color.z + r.background[2] * w);
      opacity = opacity + r.background[3] * (1.0f - opacity);
      v3 outc = mk(0.0f + color.x * 1.0f, 0.0f + color.y * 1.0f, 0.0f + color.z * 1.0f);
      float outo = 0.0f + opacity * 1.0f;
      float4 accum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      float dcur = FLT_MAX;
      if (f.frame_id != 0) {
        accum = ((const float4 *)f.d_accum)[idx];
        if (f.d_depth)
          dcur = f.d_depth[idx];
      }
      float sc_ = 1.0f + fmaxf(0.0f, cmax3(outc));
      accum.x = accum.x + outc.x / sc_;
      accum.y = accum.y + outc.y / sc_;
      accum.z = accum.z + outc.z / sc_;
      accum.w = accum.w + outo;
      bool ids_changed = f.frame_id == 0;
      uint32_t pid = 0, oid = 0, iid = 0;
      if (depth < dcur) {
        dcur = depth;
        pid = 0;
        oid = vo;
        iid = vi;
        ids_changed = true;
      }
      write_pixel(f, idx, accum, 1, dcur, ids_changed, pid, oid, iid);
    }
  }
  if (f.d_counters) {
    wave_count_add(f.d_counters, 0, inside && j == 0 ? 1u : 0u);
    wave_count_add(f.d_counters, 1, vsamples);
  }
}

__global__ void k_clear_frame(vrx_frame f)
{
  size_t n = (size_t)f.width * f.height;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    ((float4 *)f.d_accum)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (f.d_depth)
      f.d_depth[i] = FLT_MAX;
    if (f.d_prim_id)
      f.d_prim_id[i] = 0u;
    if (f.d_obj_id)
      f.d_obj_id[i] = 0u;
    if (f.d_inst_id)
      f.d_inst_id[i] = 0u;
  }
}

// --------------------------------------------------------------------------
// instance level: world boxes of entries
// --------------------------------------------------------------------------
// Macro-cell majorants: one 256-lane workgroup per 16^3-voxel cell reduces
// the min / max over the cell's voxels extended by one voxel on every side
// (every voxel a trilinear sample within a voxel of the cell can read), then
// takes the largest LUT alpha over the entries that range classifies to
// (the same clamp / scale / floor as classify()).  NaN voxels are ignored
// (the march skips NaN samples); a cell of only NaN gets majorant 0.
__global__ __launch_bounds__(256) void k_majorants(vrx_volume v, float *out)
{
  __shared__ float s_lo[4], s_hi[4];
  const int mx = ((int)v.dims[0] + 15) >> 4, my = ((int)v.dims[1] + 15) >> 4;
  const int cell = blockIdx.x;
  const int cx = cell % mx, cy = (cell / mx) % my, cz = cell / (mx * my);
  int b0[3], n[3];
  const int cc[3] = {cx, cy, cz};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    b0[a] = max(16 * cc[a] - 1, 0);
    const int b1 = min(16 * cc[a] + 17, (int)v.dims[a] - 1);
    n[a] = b1 - b0[a] + 1;
  }
  float lo = INFINITY, hi = -INFINITY;
  const int total = n[0] * n[1] * n[2];
  for (int i = threadIdx.x; i < total; i += 256) {
    const int x = i % n[0], y = (i / n[0]) % n[1], z = i / (n[0] * n[1]);
    const float val = gld(v.d_data, (size_t)(b0[0] + x) + (size_t)v.dims[0] *
        ((size_t)(b0[1] + y) + (size_t)v.dims[1] * (size_t)(b0[2] + z)));
    lo = fminf(lo, val); // fminf / fmaxf return the other operand for NaN
    hi = fmaxf(hi, val);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo = fminf(lo, __shfl_xor(lo, off, 64));
    hi = fmaxf(hi, __shfl_xor(hi, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    s_lo[threadIdx.x >> 6] = lo;
    s_hi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x != 0)
    return;
  lo = fminf(fminf(s_lo[0], s_lo[1]), fminf(s_lo[2], s_lo[3]));
  hi = fmaxf(fmaxf(s_hi[0], s_hi[1]), fmaxf(s_hi[2], s_hi[3]));
  float m = 0.0f;
  if (lo <= hi) {
    const float vlo = v.value_range[0], vhi = v.value_range[1];
    const float inv = 1.0f / (vhi - vlo);
    int e[2];
    const float ends[2] = {lo, hi};
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const float cl = fminf(fmaxf(ends[q], vlo), vhi);
      const float x = fminf(fmaxf((cl - vlo) * inv * 256.0f - 0.5f, 0.0f), 255.0f);
      e[q] = min((int)x, 254);
    }
    const int i0 = min(e[0], e[1]), i1 = max(e[0], e[1]) + 1;
    const float4 *lut = (const float4 *)v.d_lut;
    for (int i = i0; i <= i1; i++)
      m = fmaxf(m, gld(lut, i).w);
  }
  out[cell] = m;
}

// Woodcock (delta) tracking of one ray per lane through a volume with a
// majorant grid (volumeIntegration.h:167-228 with dda.h:43-122): the ray
// walks the 16^3-voxel macro cells it crosses; in a cell of majorant m it
// takes tentative collisions at exponential distances of mean step / m
// (step = the field's step, min(spacing)/2), and a tentative collision is a
// real one when the LUT alpha there is at least u * m.  PCG32 per ray,
// keyed (ray index, seed).  t_out is the collision distance (or the end of
// the ray's interval), tr 0 on a collision else 1, albedo the LUT colour at
// the last sample taken.  Loops are bounded by the cell count and by a cap
// on tentative collisions per cell.
__global__ __launch_bounds__(256) void k_sample_distance(vrx_volume v, uint32_t n,
    const float *org, const float *dir, const float *tmm, uint64_t seed, float *out_t,
    float *out_albedo, float *out_tr)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const v3 o = mk(gld(org, 3 * (size_t)i), gld(org, 3 * (size_t)i + 1), gld(org, 3 * (size_t)i + 2));
  const v3 d = mk(gld(dir, 3 * (size_t)i), gld(dir, 3 * (size_t)i + 1), gld(dir, 3 * (size_t)i + 2));
  const float tmin = gld(tmm, 2 * (size_t)i), tmax = gld(tmm, 2 * (size_t)i + 1);
  float t_out = tmax, tr = 1.0f;
  float3 alb = make_float3(0.0f, 0.0f, 0.0f);
  float t0, t1;
  if (vol_slab(v, o, d, tmin, tmax, t0, t1)) {
    const VolCtx c = vol_ctx(v, 1.0f, false);
    const float step = c.step;
    const float *mc = v.d_majorant;
    const float g0[3] = {(o.x - c.origin[0]) * c.inv_sp[0], (o.y - c.origin[1]) * c.inv_sp[1],
        (o.z - c.origin[2]) * c.inv_sp[2]};
    const float gd[3] = {d.x * c.inv_sp[0], d.y * c.inv_sp[1], d.z * c.inv_sp[2]};
    Pcg rs;
    pcg_seed(rs, (uint64_t)i, seed);
    float t = t0;
    const int max_cells = 3 * (c.mcd[0] + c.mcd[1] + c.mcd[2]) + 8;
    bool hit = false;
    for (int cell = 0; cell < max_cells && t < t1 && !hit; cell++) {
      // the cell of the segment starting at t: the point a little past t
      const float tm = fminf(t + 1e-3f * step, 0.5f * (t + t1));
      int ci[3];
#pragma unroll
      for (int a = 0; a < 3; a++)
        ci[a] = min(max((int)floorf((g0[a] + gd[a] * tm) * 0.0625f), 0), c.mcd[a] - 1);
      float te = t1;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        if (gd[a] > 0.0f)
          te = fminf(te, (16.0f * (float)(ci[a] + 1) - g0[a]) / gd[a]);
        else if (gd[a] < 0.0f)
          te = fminf(te, (16.0f * (float)ci[a] - g0[a]) / gd[a]);
      }
      if (!(te > t))
        te = tm; // a cell whose exit rounds to t: step past it
      const float m = gld(mc + ((size_t)ci[2] * c.mcd[1] + ci[1]) * c.mcd[0] + ci[0]);
      if (m > 0.0f) {
        for (int k = 0; k < 4096; k++) {
          t -= logf(1.0f - pcg_float(rs)) / m * step;
          if (t >= te)
            break;
          const float s = trilinear<false>(c, g0[0] + gd[0] * t, g0[1] + gd[1] * t,
              g0[2] + gd[2] * t);
          if (s == s) {
            const float4 co = classify<true>((const float4 *)v.d_lut, c, s);
            alb = make_float3(co.x, co.y, co.z);
            const float u = pcg_float(rs);
            if (co.w >= u * m) {
              tr = 0.0f;
              t_out = t;
              hit = true;
              break;
            }
          }
        }
      }
      t = te;
    }
  }
  out_t[i] = t_out;
  out_tr[i] = tr;
  out_albedo[3 * (size_t)i] = alb.x;
  out_albedo[3 * (size_t)i + 1] = alb.y;
  out_albedo[3 * (size_t)i + 2] = alb.z;
}

__global__ void k_entry_boxes(vrx_scene sc, float *boxes)
{
  uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= sc.num_entries)
    return;
  const vrx_entry &E = sc.d_entries[e];
  const vrx_geom &G = sc.d_geoms[E.geom];
  const vrx_bvh_header *H = (const vrx_bvh_header *)G.bvh.d_header;
  v3 lo = mk(H->lo[0], H->lo[1], H->lo[2]), hi = mk(H->hi[0], H->hi[1], H->hi[2]);
  v3 wlo = lo, whi = hi;
  if (!E.identity && G.n > 0) {
    wlo = mk(INFINITY, INFINITY, INFINITY);
    whi = mk(-INFINITY, -INFINITY, -INFINITY);
#pragma unroll
    for (int c = 0; c < 8; c++) {
      v3 p = mk((c & 1) ? hi.x : lo.x, (c & 2) ? hi.y : lo.y, (c & 4) ? hi.z : lo.z);
      v3 q = xfm_point(E.xfm, p);
      wlo = vmin(wlo, q);
      whi = vmax(whi, q);
    }
  }
  float *b = boxes + 6 * (size_t)e;
  b[0] = wlo.x; b[1] = wlo.y; b[2] = wlo.z;
  b[3] = whi.x; b[4] = whi.y; b[5] = whi.z;
}

// --------------------------------------------------------------------------
// launch helpers
// --------------------------------------------------------------------------
static uint32_t pick_sd(const vrx_renderer *r)
{
  int s = r && r->stack_lds > 0 ? r->stack_lds : 32;
  uint32_t p = 8;
  while (p < (uint32_t)s && p < 64)
    p <<= 1;
  return p;
}

static bool wide_of(const vrx_renderer *r)
{
  return r && r->node_width == 4;
}

static uint32_t max_stack(const vrx_scene *sc, bool wide)
{
  // BVH2: one push per level (+2); 4-wide: up to three pushes per 4-wide
  // level, and a 4-wide level spans two BVH2 levels
  uint32_t m = wide ? 3 * ((sc->max_blas_height + 1) / 2) + 2 : sc->max_blas_height + 2;
  if (sc->num_entries > 1)
    m += wide ? 3 * ((sc->tlas_height + 1) / 2) + 1 : sc->tlas_height + 1;
  return m;
}

static uint32_t ovf_cap_of(const vrx_scene *sc, uint32_t sd, bool wide)
{
  uint32_t m = max_stack(sc, wide);
  return m > sd ? m - sd : 0u;
}

// byte offset of the long-ray queue in the frame scratch: after the header
// and the stack-spill area of one lane per pixel of the 16x16-tile grid
static size_t long_queue_offset(uint32_t width, uint32_t height, uint32_t cap)
{
  const size_t tiles = (size_t)((width + kTile - 1) / kTile) * ((height + kTile - 1) / kTile);
  return 256 + tiles * kBlock * (size_t)cap * 4;
}

template <int PT, bool TWO, int KIND, bool WIDE>
static void launch_one(const vrx_scene &sc, const vrx_camera &cam, const vrx_renderer &r,
    const vrx_frame &f, uint32_t sd, uint32_t cap, hipStream_t s)
{
  dim3 grid((f.width + kTile - 1) / kTile, (f.height + kTile - 1) / kTile);
  size_t lds = (size_t)sd * kBlock * 4;
  if constexpr (PT >= 0 && !TWO && KIND == VRX_RENDER_RAYCAST) {
    if (r.defer_steps > 0) {
      LongRays lr;
      lr.budget = (uint32_t)r.defer_steps;
      // shared-stack capacity of a lane group: each step adds at most three
      // entries per lane (one with BVH2 nodes) above the entries of the
      // levels still open; pushes past it go to the lanes' private stacks
      const uint32_t levels = WIDE ? (sc.max_blas_height + 1) / 2 : sc.max_blas_height;
      const uint32_t room = (uint32_t)((size_t(128) << 10) - lds) / (4u * (kBlock / 16)); // 128 KiB of LDS
      lr.gcap = std::min(std::max(((WIDE ? 48u : 16u) * (levels + 2) + 63u) & ~63u, 64u), room & ~63u);
      lr.gs = nullptr;
      lr.count = (uint32_t *)f.d_scratch + 1; // header word 1
      lr.queue = (uint32_t *)((char *)f.d_scratch + long_queue_offset(f.width, f.height, cap));
      // the queue count is reset by a kernel on the stream (a failed launch
      // is reported by the caller's launch check)
      (void)vrx_fill_u32(lr.count, 0u, 1, (vrx_stream)s);
      hipLaunchKernelGGL((k_render_defer<PT, WIDE>), grid, dim3(kBlock), lds, s, sc, cam, r, f, sd,
          cap, (uint32_t *)f.d_scratch, lr);
      // a resident grid for the queued rays, no larger than the first pass
      // (whose lanes the spill area is sized for)
      const size_t lds2 = lds + (size_t)(kBlock / 16) * lr.gcap * 4;
      int per_cu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)&k_render_long<PT, WIDE>,
              kBlock, lds2) != hipSuccess || per_cu < 1)
        per_cu = 1;
      const uint32_t blocks = std::min((uint32_t)(per_cu * device_cus()), grid.x * grid.y);
      hipLaunchKernelGGL((k_render_long<PT, WIDE>), dim3(blocks), dim3(kBlock), lds2, s, sc, cam,
          r, f, sd, cap, (uint32_t *)f.d_scratch, lr);
      return;
    }
  }
  if constexpr (PT >= 0 && (KIND != VRX_RENDER_RAYCAST || TWO))
    hipLaunchKernelGGL((k_render_occ6<PT, TWO, KIND, WIDE>), grid, dim3(kBlock), lds, s, sc, cam, r,
        f, sd, cap, (uint32_t *)f.d_scratch);
  else
    hipLaunchKernelGGL((k_render<PT, TWO, KIND, WIDE>), grid, dim3(kBlock), lds, s, sc, cam, r, f,
        sd, cap, (uint32_t *)f.d_scratch);
}

template <int PT, bool TWO, bool WIDE>
static void launch_kw(const vrx_scene &sc, const vrx_camera &cam, const vrx_renderer &r,
    const vrx_frame &f, uint32_t sd, uint32_t cap, hipStream_t s)
{
  if (r.kind == VRX_RENDER_DIRECT)
    launch_one<PT, TWO, VRX_RENDER_DIRECT, WIDE>(sc, cam, r, f, sd, cap, s);
  else if (r.kind == VRX_RENDER_AO)
    launch_one<PT, TWO, VRX_RENDER_AO, WIDE>(sc, cam, r, f, sd, cap, s);
  else
    launch_one<PT, TWO, VRX_RENDER_RAYCAST, WIDE>(sc, cam, r, f, sd, cap, s);
}

template <int PT, bool TWO>
static int launch_kind(const vrx_scene &sc, const vrx_camera &cam, const vrx_renderer &r,
    const vrx_frame &f, uint32_t sd, uint32_t cap, hipStream_t s)
{
  if (wide_of(&r))
    launch_kw<PT, TWO, true>(sc, cam, r, f, sd, cap, s);
  else
    launch_kw<PT, TWO, false>(sc, cam, r, f, sd, cap, s);
  VRX_LAUNCH_CHECK();
  return VRX_OK;
}

template <int PT>
static int launch_pt(const vrx_scene &sc, const vrx_camera &cam, const vrx_renderer &r,
    const vrx_frame &f, uint32_t sd, uint32_t cap, hipStream_t s)
{
  if (sc.num_entries > 1)
    return launch_kind<PT, true>(sc, cam, r, f, sd, cap, s);
  return launch_kind<PT, false>(sc, cam, r, f, sd, cap, s);
}

} // namespace vrx

using namespace vrx;

extern "C" {

size_t vrx_render_scratch_bytes(const vrx_scene *scene, const vrx_renderer *r, uint32_t width,
    uint32_t height)
{
  if (!scene)
    return 0;
  uint32_t sd = pick_sd(r);
  uint32_t cap = ovf_cap_of(scene, sd, wide_of(r));
  const size_t queue = r && r->defer_steps > 0 ? (size_t)width * height * 4 : 0;
  return long_queue_offset(width, height, cap) + queue;
}

int vrx_frame_clear(const vrx_frame *f, vrx_stream s)
{
  VRX_ARG(f && f->d_accum, "frame/accum");
  size_t n = (size_t)f->width * f->height;
  unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
  if (blocks == 0)
    return VRX_OK;
  hipLaunchKernelGGL(k_clear_frame, dim3(blocks), dim3(256), 0, (hipStream_t)s, *f);
  VRX_LAUNCH_CHECK();
  return VRX_OK;
}

int vrx_render(const vrx_scene *scene, const vrx_camera *cam, const vrx_renderer *r,
    const vrx_frame *f, vrx_stream s)
{
  VRX_ARG(scene && cam && r && f, "null argument");
  VRX_ARG(f->width > 0 && f->height > 0, "frame size");
  VRX_ARG(f->d_accum, "frame accumulation buffer");
  VRX_ARG(scene->num_entries == 0 || (scene->d_entries && scene->d_geoms), "scene arrays");
  VRX_ARG(scene->num_entries <= 1 || scene->tlas.d_leaf_ids, "instance BVH not built");
  VRX_ARG(scene->num_volumes == 0 || scene->d_volumes, "volume array");
  VRX_ARG(r->kind >= VRX_RENDER_RAYCAST && r->kind <= VRX_RENDER_DIRECT, "renderer kind");
  VRX_ARG(scene->num_lights == 0 || scene->d_lights, "light array");
  VRX_ARG(r->node_width == 0 || r->node_width == 2 || r->node_width == 4, "node width 2 or 4");
  VRX_ARG(r->defer_steps >= 0, "defer_steps");
  VRX_ARG(!wide_of(r) || scene->num_entries == 0 || scene->wide,
      "node width 4 needs 4-wide nodes for every BVH (vrx_bvh_widen, scene->wide)");
  uint32_t sd = pick_sd(r);
  uint32_t cap = ovf_cap_of(scene, sd, wide_of(r));
  size_t need = vrx_render_scratch_bytes(scene, r, f->width, f->height);
  if (!f->d_scratch || f->scratch_bytes < need) {
    set_error("render scratch too small: %zu < %zu", f->scratch_bytes, need);
    return VRX_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)s;
  // the volume-only kernel marches from the segment entry; jittered starts
  // take the generic kernel
  if (scene->num_entries == 0 && scene->num_volumes == 1 && r->kind == VRX_RENDER_RAYCAST
      && !r->volume_jitter) {
    const bool big = scene->max_volume_voxels >= (1ull << 31);
    if (scene->num_majorant_grids > 0) {
      // a volume with a macro-cell grid: one lane per ray, with empty-space
      // skipping when the renderer asks for it (the frame is the same)
      dim3 grid((f->width + kTile - 1) / kTile, (f->height + kTile - 1) / kTile);
      const int u = r->vol_unroll > 0 ? r->vol_unroll : 4;
      const bool skip = r->skip_empty != 0;
#define VRX_VOL_LAUNCH(B, U)                                                                    \
  do {                                                                                          \
    if (skip)                                                                                   \
      hipLaunchKernelGGL((k_render_volume_seq<B, U, true>), grid, dim3(kBlock), 0, st, *scene,   \
          *cam, *r, *f);                                                                        \
    else                                                                                        \
      hipLaunchKernelGGL((k_render_volume_seq<B, U, false>), grid, dim3(kBlock), 0, st, *scene,  \
          *cam, *r, *f);                                                                        \
  } while (0)
      if (big)
        VRX_VOL_LAUNCH(true, 2);
      else if (u <= 1)
        VRX_VOL_LAUNCH(false, 1);
      else if (u == 2)
        VRX_VOL_LAUNCH(false, 2);
      else if (u <= 4)
        VRX_VOL_LAUNCH(false, 4);
      else
        VRX_VOL_LAUNCH(false, 8);
#undef VRX_VOL_LAUNCH
    } else {
      // dense field: four lanes per ray, volumeUnroll blocks of four samples
      // per iteration (default 2)
      dim3 grid((f->width + kVolTile - 1) / kVolTile, (f->height + kVolTile - 1) / kVolTile);
      const int u = r->vol_unroll > 0 ? r->vol_unroll : 2;
#define VRX_VOL_LAUNCH(B, U)                                                                    \
  hipLaunchKernelGGL((k_render_volume_quad<B, U>), grid, dim3(kBlock), 0, st, *scene, *cam, *r, \
      *f)
      if (big)
        VRX_VOL_LAUNCH(true, 2);
      else if (u <= 1)
        VRX_VOL_LAUNCH(false, 1);
      else if (u <= 2)
        VRX_VOL_LAUNCH(false, 2);
      else
        VRX_VOL_LAUNCH(false, 4);
#undef VRX_VOL_LAUNCH
    }
    VRX_LAUNCH_CHECK();
    return VRX_OK;
  }
  // the primitive-specialised kernels carry no volume code: scenes with
  // volumes take the generic one
  switch (scene->num_volumes ? -1 : scene->single_type) {
  case VRX_TRIANGLE:
    return launch_pt<VRX_TRIANGLE>(*scene, *cam, *r, *f, sd, cap, st);
  case VRX_SPHERE:
    return launch_pt<VRX_SPHERE>(*scene, *cam, *r, *f, sd, cap, st);
  case VRX_CYLINDER:
    return launch_pt<VRX_CYLINDER>(*scene, *cam, *r, *f, sd, cap, st);
  default:
    return launch_pt<-1>(*scene, *cam, *r, *f, sd, cap, st);
  }
}

size_t vrx_volume_grid_cells(const vrx_volume *v)
{
  if (!v)
    return 0;
  return (size_t)((v->dims[0] + 15) / 16) * ((v->dims[1] + 15) / 16) * ((v->dims[2] + 15) / 16);
}

// One wave: lane i sums slot i of the two frame counters, a wave reduction
// gives the totals, and the slots are cleared for the next frame.
__global__ __launch_bounds__(64) void k_frame_counters(uint64_t *ctrs, const uint32_t *flag,
    uint64_t *out)
{
  const uint32_t lane = threadIdx.x;
  uint64_t r = 0, v = 0;
  if (ctrs && lane < VRX_COUNTER_SLOTS) {
    r = ctrs[lane * VRX_COUNTER_STRIDE + 0];
    v = ctrs[lane * VRX_COUNTER_STRIDE + 1];
    ctrs[lane * VRX_COUNTER_STRIDE + 0] = 0;
    ctrs[lane * VRX_COUNTER_STRIDE + 1] = 0;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    r += __shfl_xor(r, off, 64);
    v += __shfl_xor(v, off, 64);
  }
  if (lane == 0) {
    out[0] = r;
    out[1] = v;
    out[2] = flag ? (uint64_t)*flag : 0ull;
  }
}

int vrx_frame_collect_counters(const vrx_frame *f, uint64_t *out, vrx_stream s)
{
  VRX_ARG(f && out, "frame and output");
  static_assert(VRX_COUNTER_SLOTS <= 64, "one lane per counter slot");
  hipLaunchKernelGGL(k_frame_counters, dim3(1), dim3(64), 0, (hipStream_t)s, f->d_counters,
      (const uint32_t *)f->d_scratch, out);
  VRX_LAUNCH_CHECK();
  return VRX_OK;
}

// y-pairs: out[2i] = v[i], out[2i+1] = v[i + nx] (the voxel one row up, or
// itself on the last row), as one 8-byte store per voxel.
__global__ __launch_bounds__(256) void k_ypairs(const float *__restrict__ v,
    float2 *__restrict__ out, uint32_t nx, uint32_t ny, uint64_t n)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t y = (uint32_t)((i / nx) % ny);
    const float a = gld(v, i);
    const float b = y + 1 < ny ? gld(v, i + nx) : a;
    out[i] = make_float2(a, b);
  }
}

int vrx_volume_build_ypairs(const vrx_volume *v, float *d_ypairs, vrx_stream s)
{
  VRX_ARG(v && v->d_data && d_ypairs, "volume with data and output");
  const uint64_t n = (uint64_t)v->dims[0] * v->dims[1] * v->dims[2];
  VRX_ARG(n > 0 && n < (1ull << 29), "y-pairs need 0 < voxels < 2^29");
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_ypairs, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)s, v->d_data,
      (float2 *)d_ypairs, v->dims[0], v->dims[1], n);
  VRX_LAUNCH_CHECK();
  return VRX_OK;
}

int vrx_volume_majorant_grid(const vrx_volume *v, float *d_majorant, vrx_stream s)
{
  VRX_ARG(v && d_majorant, "null argument");
  VRX_ARG(v->d_data && v->d_lut, "volume data / LUT");
  VRX_ARG(v->dims[0] > 0 && v->dims[1] > 0 && v->dims[2] > 0, "volume dims");
  const size_t cells = vrx_volume_grid_cells(v);
  VRX_ARG(cells < (1u << 31), "too many macro cells");
  hipLaunchKernelGGL(k_majorants, dim3((unsigned)cells), dim3(256), 0, (hipStream_t)s, *v,
      d_majorant);
  VRX_LAUNCH_CHECK();
  return VRX_OK;
}

int vrx_volume_sample_distance(const vrx_volume *v, uint32_t n, const float *d_org,
    const float *d_dir, const float *d_tminmax, uint64_t seed, float *d_t, float *d_albedo,
    float *d_tr, vrx_stream s)
{
  VRX_ARG(v && v->d_data && v->d_lut && v->d_majorant, "volume with data, LUT and majorant grid");
  VRX_ARG(v->dims[0] >= 2 && v->dims[1] >= 1 && v->dims[2] >= 1, "volume dims");
  VRX_ARG((uint64_t)v->dims[0] * v->dims[1] * v->dims[2] < (1ull << 31), "volume too large");
  VRX_ARG(n == 0 || (d_org && d_dir && d_tminmax && d_t && d_albedo && d_tr), "ray arrays");
  if (n == 0)
    return VRX_OK;
  hipLaunchKernelGGL(k_sample_distance, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)s, *v,
      n, d_org, d_dir, d_tminmax, seed, d_t, d_albedo, d_tr);
  VRX_LAUNCH_CHECK();
  return VRX_OK;
}

int vrx_camera_perspective(const float pos[3], const float dir[3], const float up[3], float fovy,
    float aspect, const float region[4], vrx_camera *out)
{
  VRX_ARG(pos && dir && up && out, "null argument");
  v3 d = normalize(ld3(dir));
  v3 u = normalize(ld3(up));
  float h = 2.0f * tanf(0.5f * fovy);
  float w = h * aspect;
  v3 du = mulf(normalize(cross(d, u)), w);
  v3 dv = mulf(normalize(cross(du, d)), h);
  v3 d00 = sub(sub(d, fmul(0.5f, du)), fmul(0.5f, dv));
  memcpy(out->pos, pos, 12);
  out->d00[0] = d00.x; out->d00[1] = d00.y; out->d00[2] = d00.z;
  out->du[0] = du.x; out->du[1] = du.y; out->du[2] = du.z;
  out->dv[0] = dv.x; out->dv[1] = dv.y; out->dv[2] = dv.z;
  if (region)
    memcpy(out->region, region, 16);
  else {
    out->region[0] = 0.0f; out->region[1] = 0.0f; out->region[2] = 1.0f; out->region[3] = 1.0f;
  }
  return VRX_OK;
}

int vrx_tlas_query_sizes(uint32_t num_entries, vrx_bvh_sizes *out)
{
  VRX_ARG(out, "out");
  bvh_sizes(VRX_BOX, num_entries, out);
  out->workspace += ((size_t)num_entries * 24 + 255) & ~size_t(255);
  return VRX_OK;
}

int vrx_tlas_build(const vrx_scene *scene, const vrx_bvh *out, void *d_workspace,
    size_t workspace_bytes, vrx_stream s)
{
  VRX_ARG(scene && out, "null argument");
  vrx_bvh_sizes sz;
  vrx_tlas_query_sizes(scene->num_entries, &sz);
  if (!d_workspace || workspace_bytes < sz.workspace) {
    set_error("tlas workspace too small");
    return VRX_ERR_ARG;
  }
  size_t box_bytes = ((size_t)scene->num_entries * 24 + 255) & ~size_t(255);
  float *boxes = (float *)d_workspace;
  hipStream_t st = (hipStream_t)s;
  if (scene->num_entries) {
    hipLaunchKernelGGL(k_entry_boxes, dim3((scene->num_entries + 255) / 256), dim3(256), 0, st,
        *scene, boxes);
    VRX_LAUNCH_CHECK();
  }
  vrx_prims p;
  memset(&p, 0, sizeof(p));
  p.type = VRX_BOX;
  p.n = scene->num_entries;
  p.d_position = boxes;
  return bvh_build_impl(&p, out, (char *)d_workspace + box_bytes, workspace_bytes - box_bytes,
      nullptr, nullptr, st);
}

} // extern "C"
