// diffvol.hip -- differentiable emission-absorption volume integral.
//
// Forward: the plain integral of gpu/volumeIntegration.h:64-103 (jitter off)
// over a structured-regular density grid classified by a 256-entry RGBA LUT:
//   C = sum_k T_k (1 - tau_k) c_k,  O = sum_k T_k (1 - tau_k),
//   T_{k+1} = T_k tau_k,  tau_k = (1 - a_k)^(step / unitDistance).
// The reference has only this forward pass; the adjoint is new.  The forward
// stores, per ray, the sample count, the final transmittance, the
// transmittance every `checkpoint` (64) steps and the FP64 sums of w_k c_k
// and w_k -- never every sample.
//
// Adjoint: with v_k = G.c_k + gO (G = dL/dC, gO = dL/dO) and the suffix
//   R_k = sum_{j>k} (w_j / T_{k+1}) v_j,   R_{k-1} = (1 - tau_k) v_k + tau_k R_k,
// the gradients are dL/dtau_k = T_k (R_k - v_k), dL/dc_k = w_k G.  Two
// schedules compute them (vrx_diffvol.adjoint):
//  * replay: the march is replayed backward from the last sample; T_k is
//    re-anchored at every checkpoint and recovered inside a segment as
//    T_{k+1} / tau_k (a forward replay from the segment's checkpoint is used
//    when tau_k is too small for that division);
//  * suffix (default): the march is replayed forward, in the forward's own
//    order, and T_{k+1} R_k = sum_{j>k} w_j v_j is the forward's FP64 total
//    minus the FP64 prefix, so no division by tau and no checkpoint reads.
// Both add each sample's eight trilinear corner gradients to a per-wave LDS
// window of the grid; LUT gradients are pre-reduced in LDS per workgroup.
#include <float.h>
#include <limits.h>
#include <string.h>

#include <algorithm>

#include "vrx_internal.h"
#include "vrx_math.h"

namespace vrx {

constexpr int kDvBlock = 256;
constexpr float kMinTauForDivision = 1e-3f;

struct DvCtx {
  const float *density;
  int dims[3], imax[3];
  float gmax[3];
  float origin[3], inv_sp[3], lo_b[3], hi_b[3];
  uint32_t sy, sz;
  float step, expo, lo, hi, inv_range;
  uint32_t nck;        // checkpoint slots per ray (2 + segments); the FP64
                       // totals follow in slots nck .. nck + 7
  int ckpt;            // steps per checkpoint
  int early;
  // ray layout of the adjoint's workgroups: 16x16 pixel tiles of images of
  // img_w x img_h when img_w > 0, else 256 consecutive rays
  uint32_t img_w, img_h, tiles_x, tiles_per_img;
};

static DvCtx make_ctx(const vrx_diffvol &v, uint32_t tile_w = 16)
{
  DvCtx c;
  c.density = v.d_density;
  for (int k = 0; k < 3; k++) {
    c.dims[k] = (int)v.dims[k];
    c.imax[k] = std::max((int)v.dims[k] - 2, 0);
    c.gmax[k] = (float)v.dims[k] - 1.0f;
    c.origin[k] = v.origin[k];
    c.inv_sp[k] = 1.0f / v.spacing[k];
    c.lo_b[k] = v.origin[k];
    c.hi_b[k] = v.origin[k] + ((float)v.dims[k] - 1.0f) * v.spacing[k];
  }
  c.sy = v.dims[0];
  c.sz = v.dims[0] * v.dims[1];
  c.step = v.step;
  c.expo = v.step * (1.0f / v.unit_distance);
  c.lo = v.value_range[0];
  c.hi = v.value_range[1];
  c.inv_range = 1.0f / (c.hi - c.lo);
  c.ckpt = v.checkpoint > 0 ? v.checkpoint : 64;
  c.nck = 2 + (v.max_steps + c.ckpt - 1) / c.ckpt;
  c.early = v.early_stop;
  c.img_w = v.image_w;
  c.img_h = v.image_h;
  c.tiles_x = (v.image_w + tile_w - 1) / tile_w;
  c.tiles_per_img = c.tiles_x * ((v.image_h + 15) / 16);
  return c;
}

__device__ __forceinline__ bool dv_slab(const DvCtx &c, v3 o, v3 d, float &t0, float &t1)
{
  float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  float tn = -INFINITY, tf = INFINITY;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float inv = 1.0f / dd[k];
    float a = (c.lo_b[k] - oo[k]) * inv, b = (c.hi_b[k] - oo[k]) * inv;
    tn = fmaxf(tn, fminf(a, b));
    tf = fminf(tf, fmaxf(a, b));
  }
  if (!(tn < tf) || tf < 0.0f)
    return false;
  t0 = fmaxf(tn, 0.0f);
  t1 = tf;
  return true;
}

struct Sample {
  float s;          // density
  float fx, fy, fz; // trilinear fractions
  uint32_t base;    // index of the (i0, j0, k0) corner
  int ci, cj, ck;   // the cell (i0, j0, k0)
};

__device__ __forceinline__ Sample dv_sample(const DvCtx &c, float gx, float gy, float gz)
{
  float g[3] = {gx, gy, gz};
  int i0[3];
  float f[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float x = fminf(fmaxf(g[k], 0.0f), c.gmax[k]);
    int i = min((int)x, c.imax[k]);
    i0[k] = i;
    f[k] = x - (float)i;
  }
  const uint32_t base = (uint32_t)i0[0] + (uint32_t)i0[1] * c.sy + (uint32_t)i0[2] * c.sz;
  const VRX_GLOBAL float *D = (const VRX_GLOBAL float *)c.density + base;
  // grids here have every dimension >= 2, so the +1 neighbours exist
  float2 p00 = gld_pair(D), p10 = gld_pair(D + c.sy), p01 = gld_pair(D + c.sz),
         p11 = gld_pair(D + c.sy + c.sz);
  float c00 = p00.x + (p00.y - p00.x) * f[0];
  float c10 = p10.x + (p10.y - p10.x) * f[0];
  float c01 = p01.x + (p01.y - p01.x) * f[0];
  float c11 = p11.x + (p11.y - p11.x) * f[0];
  float c0 = c00 + (c10 - c00) * f[1];
  float c1 = c01 + (c11 - c01) * f[1];
  Sample sm;
  sm.s = c0 + (c1 - c0) * f[2];
  sm.fx = f[0];
  sm.fy = f[1];
  sm.fz = f[2];
  sm.base = base;
  sm.ci = i0[0];
  sm.cj = i0[1];
  sm.ck = i0[2];
  return sm;
}

struct Cls {
  float4 co;
  float4 slope; // lut[i + 1] - lut[i]
  int i;
  float f;
  float du_ds;  // d(texel coordinate)/ds, 0 where clamped
};

__device__ __forceinline__ Cls dv_classify(const float4 *lut, const DvCtx &c, float s)
{
  Cls r;
  float cl = fminf(fmaxf(s, c.lo), c.hi);
  float u = (cl - c.lo) * c.inv_range;
  float xr = u * 256.0f - 0.5f;
  float x = fminf(fmaxf(xr, 0.0f), 255.0f);
  int i = min((int)x, 254);
  float f = x - (float)i;
  float4 a = lut[i], b = lut[i + 1];
  r.co = make_float4(a.x + (b.x - a.x) * f, a.y + (b.y - a.y) * f, a.z + (b.z - a.z) * f,
      a.w + (b.w - a.w) * f);
  r.i = i;
  r.f = f;
  const bool inside = s > c.lo && s < c.hi && xr > 0.0f && xr < 255.0f;
  r.du_ds = inside ? 256.0f * c.inv_range : 0.0f;
  r.slope = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, b.w - a.w);
  return r;
}

__device__ __forceinline__ float dv_tau(const DvCtx &c, float a)
{
  return __builtin_amdgcn_exp2f(c.expo * __builtin_amdgcn_logf(1.0f - a));
}

__global__ __launch_bounds__(kDvBlock) void k_diffvol_fwd(DvCtx c, const float *lut_g, uint32_t n,
    const float *org, const float *dir, float *rgba, float *ckpt)
{
  __shared__ float4 s_lut[256];
  s_lut[threadIdx.x] = gld((const float4 *)lut_g, threadIdx.x);
  __syncthreads();
  const uint32_t r = blockIdx.x * kDvBlock + threadIdx.x;
  if (r >= n)
    return;
  const v3 o = gld3(org, r), d = gld3(dir, r);
  float C0 = 0.0f, C1 = 0.0f, C2 = 0.0f, O = 0.0f, T = 1.0f;
  double D0 = 0.0, D1 = 0.0, D2 = 0.0, DO = 0.0; // FP64 sums of w c and w
  uint32_t N = 0;
  float t0, t1;
  if (dv_slab(c, o, d, t0, t1) && c.step > 0.0f) {
    const float gox = (o.x - c.origin[0]) * c.inv_sp[0], gdx = d.x * c.inv_sp[0];
    const float goy = (o.y - c.origin[1]) * c.inv_sp[1], gdy = d.y * c.inv_sp[1];
    const float goz = (o.z - c.origin[2]) * c.inv_sp[2], gdz = d.z * c.inv_sp[2];
    const uint32_t kmax = (c.nck - 2) * (uint32_t)c.ckpt;
    for (uint32_t k = 0; k < kmax; k++) {
      const float t = t0 + (float)k * c.step;
      if (!(t <= t1) || (c.early && !(O < 0.99f)))
        break;
      if (k % (uint32_t)c.ckpt == 0)
        gst(ckpt, (size_t)(2 + k / c.ckpt) * n + r, T);
      const Sample sm = dv_sample(c, gox + gdx * t, goy + gdy * t, goz + gdz * t);
      const Cls cl = dv_classify(s_lut, c, sm.s);
      const float tau = dv_tau(c, cl.co.w);
      const float w = T * (1.0f - tau);
      C0 += w * cl.co.x;
      C1 += w * cl.co.y;
      C2 += w * cl.co.z;
      O += w;
      const double wd = (double)w;
      D0 = fma(wd, (double)cl.co.x, D0);
      D1 = fma(wd, (double)cl.co.y, D1);
      D2 = fma(wd, (double)cl.co.z, D2);
      DO += wd;
      T *= tau;
      N = k + 1;
    }
  }
  gst(ckpt, (size_t)0 * n + r, __uint_as_float(N));
  gst(ckpt, (size_t)1 * n + r, T);
  const double tot[4] = {D0, D1, D2, DO};
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(tot[q]);
    gst(ckpt, (size_t)(c.nck + 2 * q) * n + r, __uint_as_float((uint32_t)b));
    gst(ckpt, (size_t)(c.nck + 2 * q + 1) * n + r, __uint_as_float((uint32_t)(b >> 32)));
  }
  gst((float4 *)rgba, r, make_float4(C0, C1, C2, O));
}

// ---------------------------------------------------------------------------
// Backward.  The LUT gradient of a ray is carried in registers for the two
// entries (i, i+1) of the current LUT interval and moves with the interval
// when it changes by one; other changes flush the carried entries.
// ---------------------------------------------------------------------------
struct LutAcc {
  float lo[4], hi[4]; // gradients of entries i and i+1
  int i;
};

// move the carried interval to i; Flush(e, g[4]) adds g to LUT entry e
template <class Flush>
__device__ __forceinline__ void lut_enter(Flush flush, LutAcc &L, int i)
{
  if (i == L.i)
    return;
  const bool up = L.i >= 0 && i == L.i + 1;   // entry L.i+1 is kept as the new lo
  const bool down = L.i >= 0 && i == L.i - 1; // entry L.i is kept as the new hi
  // the entry that moves with the interval is not flushed (its share would
  // add zero); only lanes with something to add issue the LDS atomics
  if (L.i >= 0 && !down)
    flush(L.i, L.lo);
  if (L.i >= 0 && !up)
    flush(L.i + 1, L.hi);
#pragma unroll
  for (int ch = 0; ch < 4; ch++) {
    const float nlo = up ? L.hi[ch] : 0.0f, nhi = down ? L.lo[ch] : 0.0f;
    L.lo[ch] = nlo;
    L.hi[ch] = nhi;
  }
  L.i = i;
}

// ---------------------------------------------------------------------------
// The adjoint kernel, both schedules.  A workgroup is kSfxWaves waves; each
// wave traces one 8x8 pixel tile and steps its rays in lockstep: the j-th
// sample a ray processes (k = j forward, k = N-1-j for the replay) runs at
// step off + j.  off = 0 (rays start together at their first / last sample)
// unless those samples' cells spread over more than kAlignBox voxels, as on
// a face seen at a grazing angle; then the rays are aligned by distance from
// the camera.  Every chunk of up to kSfxChunk steps, the wave takes the box of
// voxels its samples touch; if the box fits its LDS window (kWinCap
// entries), each sample's eight corner gradients are added there, as 64-bit
// fixed-point integers -- LDS integer atomics are cheap, LDS float atomics
// are not -- and the window is added to the grid at the end of the chunk
// with row-contiguous float atomics.  A box that does not fit is halved in
// steps, down to one step; beyond that the corners go to the grid.  The
// fixed-point scale of each workgroup comes from a bound on any single
// addition built from its rays' largest output gradient and the LUT range:
// an addition is below 2^50 units, so the sums that can meet in one word
// stay far from overflow, and the resolution is 2^-50 of the bound.
// ---------------------------------------------------------------------------
constexpr int kSfxWaves = 8;
constexpr int kSfxThreads = 64 * kSfxWaves;
constexpr int kWinCap = 1024;
constexpr int kSfxChunk = 16;
constexpr int kAlignBox = 256;

// x scale (a power of two) rounded to an integer, as a 64-bit two's
// complement word.  For |x scale| < 2^51, adding 1.5 2^52 leaves the integer
// in the low bits of the double's mantissa, so the word is the double's bits
// minus those of 1.5 2^52.
__device__ __forceinline__ unsigned long long to_fix(float x, double scale)
{
  const double d = fma((double)x, scale, 6755399441055744.0);
  return (unsigned long long)__double_as_longlong(d) - 0x4338000000000000ull;
}

__device__ __forceinline__ float from_fix(unsigned long long q, double inv_scale)
{
  return (float)((double)(long long)q * inv_scale);
}

struct WinSink {
  unsigned long long *win; // this wave's window
  int ox, oy, oz;          // box origin (voxel coordinates)
  int sx, sxy;             // box row and slice pitch
  bool lds;                // the box fits the window
};

__device__ __forceinline__ int win_index(const WinSink &S, int i, int j, int k)
{
  return (i - S.ox) + (j - S.oy) * S.sx + (k - S.oz) * S.sxy;
}

// the trilinear cell of a sample position, exactly as dv_sample finds it
__device__ __forceinline__ void cell_of(const DvCtx &c, float gx, float gy, float gz, int cell[3])
{
  const float g[3] = {gx, gy, gz};
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const float x = fminf(fmaxf(g[q], 0.0f), c.gmax[q]);
    cell[q] = min((int)x, c.imax[q]);
  }
}

// Wave-wide min / max with every lane active: within each 16-lane row by DPP
// (quad_perm xor 1 and xor 2, then row_ror 4 and 8), then across the four rows
// by readlane, so the reductions stay off the LDS pipe that carries the
// scatter's atomics.  A source lane DPP cannot read leaves the lane's own
// value, which min / max absorb.
template <class Op>
__device__ __forceinline__ int wave_reduce_i(int x, Op op)
{
  x = op(x, __builtin_amdgcn_update_dpp(x, x, 0xB1, 0xF, 0xF, false));  // quad_perm [1,0,3,2]
  x = op(x, __builtin_amdgcn_update_dpp(x, x, 0x4E, 0xF, 0xF, false));  // quad_perm [2,3,0,1]
  x = op(x, __builtin_amdgcn_update_dpp(x, x, 0x124, 0xF, 0xF, false)); // row_ror:4
  x = op(x, __builtin_amdgcn_update_dpp(x, x, 0x128, 0xF, 0xF, false)); // row_ror:8
  return op(op(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 16)),
      op(__builtin_amdgcn_readlane(x, 32), __builtin_amdgcn_readlane(x, 48)));
}

__device__ __forceinline__ int wave_min(int x)
{
  return wave_reduce_i(x, [](int a, int b) { return min(a, b); });
}

__device__ __forceinline__ int wave_max(int x)
{
  return wave_reduce_i(x, [](int a, int b) { return max(a, b); });
}

__device__ __forceinline__ float wave_fmax(float x)
{
  return __int_as_float(wave_reduce_i(__float_as_int(x),
      [](int a, int b) { return __float_as_int(fmaxf(__int_as_float(a), __int_as_float(b))); }));
}

__device__ __forceinline__ void wave_box(int bl[3], int bh[3])
{
#pragma unroll
  for (int q = 0; q < 3; q++) {
    bl[q] = wave_min(bl[q]);
    bh[q] = wave_max(bh[q]);
  }
}

// LDS operations of this wave complete before the next ones are issued
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// exponent e with bound * 2^e < 2^50 (bound 0: e = 0)
__device__ __forceinline__ int fix_exponent(float bound)
{
  if (!(bound > 0.0f))
    return 0;
  if (!(bound < 3.0e38f))
    return -120;
  int e;
  frexpf(bound, &e); // bound < 2^e
  return max(-120, min(120, 50 - e));
}

// REPLAY = false: suffix schedule; true: replay from the T checkpoints
template <bool REPLAY>
__global__ __launch_bounds__(kSfxThreads) void k_diffvol_bwd(DvCtx c, const float *lut_g,
    uint32_t n, const float *org, const float *dir, const float *grad_rgba, const float *ckpt,
    float *grad_density, float *lut_partials)
{
  __shared__ float4 s_lut[256];
  __shared__ unsigned long long s_glut[1024]; // channel-major: [ch * 256 + entry]
  __shared__ unsigned long long s_win[kSfxWaves * kWinCap];
  __shared__ uint32_t s_stat[4]; // max |G|, max |LUT|, max alpha, max(-alpha, 0) as float bits
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x < 4)
    s_stat[threadIdx.x] = 0;
  for (int i = threadIdx.x; i < 1024; i += kSfxThreads)
    s_glut[i] = 0;
  for (int i = threadIdx.x; i < kSfxWaves * kWinCap; i += kSfxThreads)
    s_win[i] = 0;
  unsigned long long *win = s_win + wave * kWinCap;
  __syncthreads();
  if (threadIdx.x < 256) {
    const float4 e = gld((const float4 *)lut_g, threadIdx.x);
    s_lut[threadIdx.x] = e;
    const float lm = wave_fmax(fmaxf(fmaxf(fabsf(e.x), fabsf(e.y)), fmaxf(fabsf(e.z), fabsf(e.w))));
    const float am = wave_fmax(e.w), an = wave_fmax(-e.w);
    if (lane == 0) {
      atomicMax(&s_stat[1], __float_as_uint(lm));
      atomicMax(&s_stat[2], __float_as_uint(fmaxf(am, 0.0f)));
      atomicMax(&s_stat[3], __float_as_uint(fmaxf(an, 0.0f)));
    }
  }

  uint32_t r;
  bool valid;
  if (c.img_w > 0) {
    // workgroup tile 32x16 pixels: 4x2 wave tiles of 8x8
    const uint32_t img = blockIdx.x / c.tiles_per_img, t = blockIdx.x - img * c.tiles_per_img;
    const uint32_t ty = t / c.tiles_x, tx = t - ty * c.tiles_x;
    const uint32_t x = tx * 32 + (wave & 3) * 8 + (lane & 7), y = ty * 16 + (wave >> 2) * 8 + (lane >> 3);
    valid = x < c.img_w && y < c.img_h;
    r = (uint32_t)((size_t)img * c.img_w * c.img_h + (size_t)y * c.img_w + x);
    valid = valid && r < n;
  } else {
    r = blockIdx.x * kSfxThreads + threadIdx.x;
    valid = r < n;
  }

  uint32_t N = 0;
  float t0 = 0.0f, t1 = 0.0f;
  float gx = 0.0f, gy = 0.0f, gz = 0.0f, gO = 0.0f;
  float gox = 0.0f, gdx = 0.0f, goy = 0.0f, gdy = 0.0f, goz = 0.0f, gdz = 0.0f;
  float T = 1.0f;     // suffix: T before the next sample
  double Ssuf = 0.0;  // suffix: sum over the samples not yet replayed of w_j v_j
  float Tn = 0.0f;    // replay: T after the sample replayed last
  float R = 0.0f;     // replay: R of the sample replayed last
  if (valid) {
    N = __float_as_uint(gld(ckpt, (size_t)0 * n + r));
    const v3 o = gld3(org, r), d = gld3(dir, r);
    if (N > 0 && dv_slab(c, o, d, t0, t1)) {
      const float4 G4 = gld((const float4 *)grad_rgba, r);
      gx = G4.x;
      gy = G4.y;
      gz = G4.z;
      gO = G4.w;
      gox = (o.x - c.origin[0]) * c.inv_sp[0];
      gdx = d.x * c.inv_sp[0];
      goy = (o.y - c.origin[1]) * c.inv_sp[1];
      gdy = d.y * c.inv_sp[1];
      goz = (o.z - c.origin[2]) * c.inv_sp[2];
      gdz = d.z * c.inv_sp[2];
      if (REPLAY) {
        Tn = gld(ckpt, (size_t)1 * n + r); // T after the last sample
      } else {
        double tot[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const uint32_t lo = __float_as_uint(gld(ckpt, (size_t)(c.nck + 2 * q) * n + r));
          const uint32_t hi = __float_as_uint(gld(ckpt, (size_t)(c.nck + 2 * q + 1) * n + r));
          tot[q] = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
        }
        Ssuf = ((double)gx * tot[0] + (double)gy * tot[1]) + ((double)gz * tot[2] + (double)gO * tot[3]);
      }
    } else {
      N = 0;
    }
  }
  {
    const float gm = wave_fmax(fmaxf(fmaxf(fabsf(gx), fabsf(gy)), fmaxf(fabsf(gz), fabsf(gO))));
    if (lane == 0)
      atomicMax(&s_stat[0], __float_as_uint(gm));
  }
  __syncthreads();

  // Fixed-point scales.  With Gm = max |G|, Lm = max |LUT|, alpha in
  // [-an, am]: |w| <= Wm = Tm (1 + (1 + an)^expo), Tm = (1 + an)^(expo kmax)
  // bounds the transmittance, |v| <= Vm = Gm (3 Lm + 1), |suffix|, |T R|
  // and |T v| <= Wm Vm, |da| <= Dm = 2 expo Wm Vm / max(1 - am, 1e-8); a sample
  // adds at most Pm = max(Gm Wm, Dm) to a LUT scalar and at most
  // 256/range 2 Lm (3 Gm Wm + Dm) to ds, so to any corner.
  // A carried LUT scalar holds at most the ray's kmax samples.  A window word
  // gets at most kSfxChunk additions per ray of the wave in a chunk, a LUT
  // word at most kmax Pm per ray of the workgroup: with additions below 2^50
  // units, no word exceeds 2^60.
  double dscale, lscale;
  double dinv, linv;
  {
    const float Gm = __uint_as_float(s_stat[0]), Lm = __uint_as_float(s_stat[1]);
    const float am = __uint_as_float(s_stat[2]), an = __uint_as_float(s_stat[3]);
    const float kmax = (float)((c.nck - 2) * (uint32_t)c.ckpt);
    const float Tm = an > 0.0f ? powf(1.0f + an, c.expo * kmax) : 1.0f;
    const float Wm = Tm * (1.0f + powf(1.0f + an, c.expo));
    const float Vm = Gm * (3.0f * Lm + 1.0f);
    const float Dm = 2.0f * c.expo * Wm * Vm / fmaxf(1.0f - am, 1e-8f);
    const float Pm = fmaxf(Gm * Wm, Dm);
    const float dsm = 256.0f * fabsf(c.inv_range) * 2.0f * Lm * (3.0f * Gm * Wm + Dm);
    const int ed = fix_exponent(dsm);
    const int el = fix_exponent(Pm * kmax);
    dscale = ldexp(1.0, ed);
    lscale = ldexp(1.0, el);
    dinv = ldexp(1.0, -ed);
    linv = ldexp(1.0, -el);
  }

  // lockstep offset of this wave's rays
  int off = 0;
  {
    int bl[3] = {INT_MAX, INT_MAX, INT_MAX}, bh[3] = {INT_MIN, INT_MIN, INT_MIN};
    const float tf = t0 + (float)(REPLAY ? (int)N - 1 : 0) * c.step; // first sample replayed
    if (N > 0) {
      int ce[3];
      cell_of(c, gox + gdx * tf, goy + gdy * tf, goz + gdz * tf, ce);
#pragma unroll
      for (int q = 0; q < 3; q++) {
        bl[q] = ce[q];
        bh[q] = ce[q] + 1;
      }
    }
    wave_box(bl, bh);
    const int ev = bh[0] >= bl[0] ? (bh[0] - bl[0] + 1) * (bh[1] - bl[1] + 1) * (bh[2] - bl[2] + 1) : 0;
    if (ev > kAlignBox) {
      // distance of the first replayed sample, in steps
      const int e = N > 0 ? (int)(t0 / c.step) + (REPLAY ? (int)N - 1 : 0) : 0;
      if (REPLAY) {
        const int emax = wave_max(N > 0 ? e : INT_MIN); // the farthest starts first
        off = N > 0 ? emax - e : 0;
      } else {
        off = e;
      }
    }
  }
  const int gbeg = wave_min(N > 0 ? off : INT_MAX), gend = wave_max(N > 0 ? off + (int)N : INT_MIN);

  LutAcc L;
  L.i = -1;
#pragma unroll
  for (int ch = 0; ch < 4; ch++)
    L.lo[ch] = L.hi[ch] = 0.0f;
  const auto lut_flush = [&](int e, const float g[4]) {
#pragma unroll
    for (int ch = 0; ch < 4; ch++)
      __hip_atomic_fetch_add(s_glut + ch * 256 + e, to_fix(g[ch], lscale), __ATOMIC_RELAXED,
          __HIP_MEMORY_SCOPE_WORKGROUP);
  };

  int CL = kSfxChunk;
  for (int g0 = gbeg; g0 < gend; g0 += CL) {
    WinSink S;
    S.win = win;
    int vol;
    for (CL = kSfxChunk;; CL >>= 1) {
      const int g1 = min(gend, g0 + CL);
      // cells this ray's samples of the chunk touch (replay indices j)
      const int ja = max(g0 - off, 0), jb = min((int)N - 1, g1 - 1 - off);
      int bl[3] = {INT_MAX, INT_MAX, INT_MAX}, bh[3] = {INT_MIN, INT_MIN, INT_MIN};
      if (N > 0 && jb >= ja) {
        const int ka = REPLAY ? (int)N - 1 - jb : ja, kb = REPLAY ? (int)N - 1 - ja : jb;
        const float ta = t0 + (float)ka * c.step, tb = t0 + (float)kb * c.step;
        int ca[3], cb[3];
        cell_of(c, gox + gdx * ta, goy + gdy * ta, goz + gdz * ta, ca);
        cell_of(c, gox + gdx * tb, goy + gdy * tb, goz + gdz * tb, cb);
#pragma unroll
        for (int q = 0; q < 3; q++) {
          bl[q] = min(ca[q], cb[q]);
          bh[q] = max(ca[q], cb[q]) + 1;
        }
      }
      wave_box(bl, bh);
      const bool any = bh[0] >= bl[0];
      S.ox = bl[0];
      S.oy = bl[1];
      S.oz = bl[2];
      // odd row and slice pitches: the corners of neighbouring lanes' cells
      // then fall in different LDS banks more often (16% fewer conflict
      // cycles than the tight box at the converged 128^3 state)
      S.sx = any ? (bh[0] - bl[0] + 1) | 1 : 0;
      S.sxy = any ? S.sx * ((bh[1] - bl[1] + 1) | 1) : 0;
      vol = any ? S.sxy * (bh[2] - bl[2] + 1) : 0;
      S.lds = any && vol <= kWinCap;
      if (S.lds || !any || CL == 1)
        break;
    }
    const int g1 = min(gend, g0 + CL);
    for (int g = max(g0, off); g < g1 && g < off + (int)N; g++) {
      const int j = g - off;
      const int k = REPLAY ? (int)N - 1 - j : j;
      const float t = t0 + (float)k * c.step;
      const Sample sm = dv_sample(c, gox + gdx * t, goy + gdy * t, goz + gdz * t);
      const Cls cl = dv_classify(s_lut, c, sm.s);
      const float tau = dv_tau(c, cl.co.w);
      const float v = (gx * cl.co.x + gy * cl.co.y + gz * cl.co.z) + gO;
      const float oma = fmaxf(1.0f - cl.co.w, 1e-8f);
      float w, da;
      if (REPLAY) {
        // T_k: the checkpoint, T_{k+1} / tau_k, or a forward replay from the
        // segment's checkpoint where tau_k is too small for the division
        const uint32_t seg = (uint32_t)k / (uint32_t)c.ckpt;
        if ((uint32_t)k % (uint32_t)c.ckpt == 0) {
          T = gld(ckpt, (size_t)(2 + seg) * n + r);
        } else if (tau >= kMinTauForDivision && Tn > 0.0f) {
          T = Tn * __builtin_amdgcn_rcpf(tau);
        } else {
          T = gld(ckpt, (size_t)(2 + seg) * n + r);
          for (uint32_t q = seg * (uint32_t)c.ckpt; q < (uint32_t)k; q++) {
            const float tq = t0 + (float)q * c.step;
            const Sample sq = dv_sample(c, gox + gdx * tq, goy + gdy * tq, goz + gdz * tq);
            T *= dv_tau(c, dv_classify(s_lut, c, sq.s).co.w);
          }
        }
        w = T * (1.0f - tau);
        const float dtau = T * (R - v);
        da = dtau * (-c.expo) * tau * __builtin_amdgcn_rcpf(oma);
        R = (1.0f - tau) * v + tau * R;
        Tn = T;
      } else {
        w = T * (1.0f - tau);
        // sum over the later samples of w_j v_j = T_{k+1} R_k, from the FP64
        // total minus the FP64 prefix
        const double vd = fma((double)gx, (double)cl.co.x,
            fma((double)gy, (double)cl.co.y, fma((double)gz, (double)cl.co.z, (double)gO)));
        Ssuf = fma(-(double)w, vd, Ssuf);
        const float Tk1 = T * tau;
        da = ((float)Ssuf - Tk1 * v) * (-c.expo) * __builtin_amdgcn_rcpf(oma);
        T = Tk1;
      }
      const float dr = w * gx, dg = w * gy, db = w * gz;
      lut_enter(lut_flush, L, cl.i);
      const float wf0 = 1.0f - cl.f, wf1 = cl.f;
      L.lo[0] = fmaf(wf0, dr, L.lo[0]);
      L.lo[1] = fmaf(wf0, dg, L.lo[1]);
      L.lo[2] = fmaf(wf0, db, L.lo[2]);
      L.lo[3] = fmaf(wf0, da, L.lo[3]);
      L.hi[0] = fmaf(wf1, dr, L.hi[0]);
      L.hi[1] = fmaf(wf1, dg, L.hi[1]);
      L.hi[2] = fmaf(wf1, db, L.hi[2]);
      L.hi[3] = fmaf(wf1, da, L.hi[3]);
      // density: d(channel)/ds through the LUT slope, then trilinear weights
      const float4 sl = cl.slope;
      const float ds = fmaf(dr, sl.x, fmaf(dg, sl.y, fmaf(db, sl.z, da * sl.w))) * cl.du_ds;
      // the eight trilinear corners
      const float ux = 1.0f - sm.fx, uy = 1.0f - sm.fy, uz = 1.0f - sm.fz;
      const float d0 = ds * ux, d1 = ds * sm.fx;
      const float w00 = uy * uz, w10 = sm.fy * uz, w01 = uy * sm.fz, w11 = sm.fy * sm.fz;
      const float wq[8] = {d0 * w00, d1 * w00, d0 * w10, d1 * w10, d0 * w01, d1 * w01, d0 * w11, d1 * w11};
      if (S.lds) {
        const int lb = win_index(S, sm.ci, sm.cj, sm.ck);
#pragma unroll
        for (int q = 0; q < 8; q++)
          __hip_atomic_fetch_add(S.win + lb + (q & 1) + ((q >> 1) & 1) * S.sx + ((q >> 2) & 1) * S.sxy,
              to_fix(wq[q], dscale), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      } else if (ds != 0.0f) {
#pragma unroll
        for (int q = 0; q < 8; q++)
          unsafeAtomicAdd(grad_density + sm.base + (uint32_t)(q & 1) + (uint32_t)((q >> 1) & 1) * c.sy
              + (uint32_t)((q >> 2) & 1) * c.sz, wq[q]);
      }
    }
    if (S.lds) {
      // add the window to the grid, row by row (x fastest)
      wave_lds_sync();
      const float inv_sxy = 1.0f / (float)S.sxy, inv_sx = 1.0f / (float)S.sx;
      for (int i = lane; i < vol; i += 64) {
        const unsigned long long q = win[i];
        if (q != 0) {
          // i, sxy < 2^12: the float quotients are off by at most one
          int lz = (int)((float)i * inv_sxy);
          lz += (lz + 1) * S.sxy <= i ? 1 : 0;
          lz -= lz * S.sxy > i ? 1 : 0;
          const int rem = i - lz * S.sxy;
          int ly = (int)((float)rem * inv_sx);
          ly += (ly + 1) * S.sx <= rem ? 1 : 0;
          ly -= ly * S.sx > rem ? 1 : 0;
          const int lx = rem - ly * S.sx;
          unsafeAtomicAdd(grad_density + (uint32_t)(S.ox + lx) + (uint32_t)(S.oy + ly) * c.sy
              + (uint32_t)(S.oz + lz) * c.sz, from_fix(q, dinv));
          win[i] = 0;
        }
      }
      wave_lds_sync();
    }
  }
  if (L.i >= 0) {
    lut_flush(L.i, L.lo);
    lut_flush(L.i + 1, L.hi);
  }
  __syncthreads();
  float *dst = lut_partials + (size_t)blockIdx.x * 1024;
  for (int i = threadIdx.x; i < 1024; i += kSfxThreads)
    gst(dst, i, from_fix(s_glut[(i & 3) * 256 + (i >> 2)], linv));
}

// LUT gradient: the per-workgroup partials are summed in two fixed-order
// stages (slices of workgroups, then the slices), so the result is
// deterministic and the reduction runs on kLutSlices x 4 workgroups.
constexpr uint32_t kLutSlices = 128;

__global__ __launch_bounds__(256) void k_reduce_lut_slices(const float *partials, uint32_t nblocks,
    float *slices)
{
  const uint32_t j = blockIdx.x * 256 + threadIdx.x; // LUT scalar 0..1023
  const uint32_t sl = blockIdx.y;
  const uint32_t per = (nblocks + kLutSlices - 1) / kLutSlices;
  const uint32_t b0 = sl * per, b1 = min(nblocks, b0 + per);
  float s = 0.0f;
  for (uint32_t b = b0; b < b1; b++)
    s += gld(partials, (size_t)b * 1024 + j);
  gst(slices, (size_t)sl * 1024 + j, s);
}

__global__ __launch_bounds__(256) void k_reduce_lut(const float *slices, float *grad_lut)
{
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  float s = 0.0f;
  for (uint32_t sl = 0; sl < kLutSlices; sl++)
    s += gld(slices, (size_t)sl * 1024 + j);
  grad_lut[j] += s;
}

// up to 32 cameras travel by value in the kernel arguments (2 KiB), so each
// launch owns its cameras and concurrent calls on any streams cannot race
constexpr uint32_t kCamsPerLaunch = 32;
struct CamBlock {
  vrx_camera c[kCamsPerLaunch];
};

// pinhole rays through the pixel corners (gpu/cameraCreateRay.h:74-81,
// raycast variant), row 0 = bottom; ray index = view*w*h + y*w + x, for the
// views first_view .. first_view + nviews - 1 (cams.c[view - first_view])
__global__ void k_make_rays(CamBlock cams, uint32_t first_view, uint32_t nviews, uint32_t w,
    uint32_t h, float *org, float *dir)
{
  const size_t per = (size_t)w * h;
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= per * nviews)
    return;
  const uint32_t lv = (uint32_t)(j / per);
  const uint32_t rem = (uint32_t)(j - (size_t)lv * per);
  const uint32_t y = rem / w, x = rem - y * w;
  const size_t i = (size_t)(first_view + lv) * per + rem;
  const vrx_camera &c = cams.c[lv];
  float sx = (float)x * (1.0f / (float)w), sy = (float)y * (1.0f / (float)h);
  sx = c.region[0] * (1.0f - sx) + c.region[2] * sx;
  sy = c.region[1] * (1.0f - sy) + c.region[3] * sy;
  v3 dd = normalize(add(add(ld3(c.d00), fmul(sx, ld3(c.du))), fmul(sy, ld3(c.dv))));
  gst(org, 3 * i, c.pos[0]);
  gst(org, 3 * i + 1, c.pos[1]);
  gst(org, 3 * i + 2, c.pos[2]);
  gst(dir, 3 * i, dd.x);
  gst(dir, 3 * i + 1, dd.y);
  gst(dir, 3 * i + 2, dd.z);
}

static int check_dv(const vrx_diffvol *v)
{
  VRX_ARG(v && v->d_density && v->d_lut, "diffvol arguments");
  VRX_ARG(v->dims[0] >= 2 && v->dims[1] >= 2 && v->dims[2] >= 2, "grid dims must be >= 2");
  VRX_ARG((uint64_t)v->dims[0] * v->dims[1] * v->dims[2] < (1ull << 31), "grid too large");
  VRX_ARG(v->step > 0.0f, "step must be > 0");
  VRX_ARG(v->max_steps > 0, "max_steps must be > 0");
  return VRX_OK;
}

} // namespace vrx

using namespace vrx;

extern "C" {

size_t vrx_diffvol_ckpt_floats(const vrx_diffvol *v, uint32_t n)
{
  if (!v)
    return 0;
  int ck = v->checkpoint > 0 ? v->checkpoint : 64;
  // count, final T, T every ck steps, then the FP64 totals (8 words)
  return (size_t)(2 + (v->max_steps + ck - 1) / ck + 8) * n;
}

int vrx_diffvol_forward(const vrx_diffvol *v, uint32_t n, const float *d_org, const float *d_dir,
    float *d_rgba_out, float *d_ckpt, vrx_stream s)
{
  int st = check_dv(v);
  if (st)
    return st;
  if (n == 0)
    return VRX_OK;
  VRX_ARG(d_org && d_dir && d_rgba_out && d_ckpt, "ray/output buffers");
  DvCtx c = make_ctx(*v);
  hipLaunchKernelGGL(k_diffvol_fwd, dim3((n + kDvBlock - 1) / kDvBlock), dim3(kDvBlock), 0,
      (hipStream_t)s, c, v->d_lut, n, d_org, d_dir, d_rgba_out, d_ckpt);
  VRX_LAUNCH_CHECK();
  return VRX_OK;
}

int vrx_diffvol_backward(const vrx_diffvol *v, uint32_t n, const float *d_org, const float *d_dir,
    const float *d_rgba, const float *d_grad_rgba, const float *d_ckpt, float *d_grad_density,
    float *d_grad_lut, vrx_stream s)
{
  (void)d_rgba;
  int st = check_dv(v);
  if (st)
    return st;
  if (n == 0)
    return VRX_OK;
  VRX_ARG(d_org && d_dir && d_grad_rgba && d_ckpt && d_grad_density && d_grad_lut, "buffers");
  VRX_ARG(v->adjoint == VRX_ADJOINT_SUFFIX || v->adjoint == VRX_ADJOINT_REPLAY, "unknown adjoint schedule");
  const bool suffix = v->adjoint == VRX_ADJOINT_SUFFIX;
  DvCtx c = make_ctx(*v, 32); // workgroup tiles of 32x16 pixels
  if (c.img_w > 0) {
    const size_t per = (size_t)c.img_w * c.img_h;
    VRX_ARG(c.img_h > 0 && n % per == 0, "image_w/image_h do not divide the ray count");
  }
  const uint32_t threads = (uint32_t)kSfxThreads;
  const uint32_t nb = c.img_w > 0 ? (uint32_t)(n / ((size_t)c.img_w * c.img_h)) * c.tiles_per_img
                                  : (n + threads - 1) / threads;
  // per-workgroup LUT partials: stream-ordered memory, so calls on other
  // streams or devices never share it
  const size_t need = ((size_t)nb + kLutSlices) * 1024 * 4;
  void *partials = nullptr;
  VRX_HIP(hipMallocAsync(&partials, need, (hipStream_t)s));
  if (suffix)
    hipLaunchKernelGGL(k_diffvol_bwd<false>, dim3(nb), dim3(kSfxThreads), 0, (hipStream_t)s, c,
        v->d_lut, n, d_org, d_dir, d_grad_rgba, d_ckpt, d_grad_density, (float *)partials);
  else
    hipLaunchKernelGGL(k_diffvol_bwd<true>, dim3(nb), dim3(kSfxThreads), 0, (hipStream_t)s, c,
        v->d_lut, n, d_org, d_dir, d_grad_rgba, d_ckpt, d_grad_density, (float *)partials);
  float *slices = (float *)partials + (size_t)nb * 1024;
  hipLaunchKernelGGL(k_reduce_lut_slices, dim3(4, kLutSlices), dim3(256), 0, (hipStream_t)s,
      (const float *)partials, nb, slices);
  hipLaunchKernelGGL(k_reduce_lut, dim3(4), dim3(256), 0, (hipStream_t)s, (const float *)slices,
      d_grad_lut);
  const hipError_t launched = hipGetLastError();
  VRX_HIP(hipFreeAsync(partials, (hipStream_t)s));
  VRX_HIP(launched);
  return VRX_OK;
}

int vrx_make_rays(const vrx_camera *cams, uint32_t nviews, uint32_t w, uint32_t h, float *d_org,
    float *d_dir, vrx_stream s)
{
  VRX_ARG(cams && d_org && d_dir, "null argument");
  VRX_ARG(nviews >= 1 && nviews <= 256, "1..256 views");
  const size_t per = (size_t)w * h;
  if (per == 0)
    return VRX_OK;
  for (uint32_t v0 = 0; v0 < nviews; v0 += kCamsPerLaunch) {
    const uint32_t nv = nviews - v0 < kCamsPerLaunch ? nviews - v0 : kCamsPerLaunch;
    CamBlock cb;
    memset(&cb, 0, sizeof(cb));
    memcpy(cb.c, cams + v0, sizeof(vrx_camera) * nv);
    const size_t total = per * nv;
    hipLaunchKernelGGL(k_make_rays, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
        (hipStream_t)s, cb, v0, nv, (uint32_t)w, (uint32_t)h, d_org, d_dir);
    VRX_LAUNCH_CHECK();
  }
  return VRX_OK;
}

} // extern "C"
