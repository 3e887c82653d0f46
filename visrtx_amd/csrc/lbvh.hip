// lbvh.hip -- software BVH build for gfx950.
//
// Replaces the closed OptiX acceleration build of the reference
// (optix_visrtx.cpp:37-121, called from scene/Group.cpp:189-266 and
// scene/World.cpp:202-254).  Pipeline, all integer work bit-exact:
//   1. primitive bounds -> centroid bounds (exact min/max reduction)
//   2. 30-bit Morton codes (10 bits per axis)
//   3. stable LSD radix sort of (code, index), 4 passes of 8 bits; each pass
//      is histogram -> per-digit row scan (256 workgroups) -> LDS-staged
//      scatter with wave-ballot stable ranking (tile = 256 lanes x 16 keys)
//   4. Karras (HPG 2012) hierarchy, ties broken with index-augmented keys
//   5. refit without atomics, in three launches.  A node's common-prefix
//      length (delta) is always shorter than its children's, so refitting in
//      decreasing delta visits children first.  Nodes whose leaf range lies
//      in one chunk of kRefitChunk sorted leaves are refit by that chunk's
//      workgroup, delta by delta, with boxes in LDS; nodes crossing a chunk
//      border but inside one span of kRefitSpan leaves by that span's
//      workgroup; the few crossing a span border by one last workgroup.
//      Every node is written once, with plain stores; workgroup barriers
//      order the deltas inside a launch, kernel boundaries the launches.
// Bounds inputs follow scene/surface/geometry/Sphere.cu:82-110 (c +- r),
// Cylinder.cpp:53-102 (min/max(v0,v1) +- |r|) and Triangle.cpp:112-138.
//
// Node layout (64 B, one 128-byte line holds two nodes): the boxes of BOTH
// children plus the child links, so one node fetch tests two children:
//   a = (lo0.x lo0.y lo0.z hi0.x) b = (hi0.y hi0.z lo1.x lo1.y)
//   c = (lo1.z hi1.x hi1.y hi1.z) d = (child0, child1, height0, height1)
#include "vrx_internal.h"
#include "vrx_lbvh.h"
#include "vrx_math.h"

namespace vrx {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kScanThreads = 1024;

static inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

static uint32_t sort_blocks(uint32_t n) { return (n + kSortTile - 1) / kSortTile; }

// workgroups of the centroid-box reduction (one partial box each)
constexpr uint32_t kCboxBlocks = 1024;

// refit granularity: leaves per chunk (pass 1, LDS) and per span (pass 2)
constexpr uint32_t kRefitChunk = 512;
constexpr uint32_t kRefitSpan = 1u << 16;
constexpr int kRefitThreads = 512;
constexpr int kSpanThreads = 1024;
// a Karras tree is at most 64 internal nodes deep (delta grows by at least
// one per level and is below 64), so at most 64 nodes contain any two
// neighbouring leaves: a span holds at most 64 (kRefitSpan / kRefitChunk - 1)
// nodes that cross a chunk border, and the top pass at most 64 per span border
constexpr uint32_t kSpanListCap = 64u * (kRefitSpan / kRefitChunk);

// level byte of an internal node: delta in bits 0..5, refit pass in bits 6..7
// (1: inside a chunk, 2: inside a span, 3: crossing a span border)
constexpr uint32_t kPassShift = 6;

struct Workspace {
  uint32_t *keys_a, *vals_a, *keys_b, *vals_b;
  uint32_t *hist;      // 256 x nblocks, digit-major
  uint32_t *row_total; // 256
  uint32_t *cbox;      // 6 ordered-uint floats
  uint32_t *cbox_part; // 6 x kCboxBlocks per-workgroup partial bounds
  uint2 *child;        // n-1 child pairs from the hierarchy pass
  uint8_t *level;      // n-1 level bytes (delta | pass << 6)
  uint32_t *top;       // nodes of the top pass
  uint32_t *top_count; // 1
  void *rec;           // n leaf records in input order (written coalesced by
                       // the Morton pass, gathered into sorted order later)
};

static uint32_t top_capacity(uint32_t n)
{
  return 64u * ((n + kRefitSpan - 1) / kRefitSpan + 1);
}

static size_t workspace_layout(int type, uint32_t n, void *base, Workspace *w)
{
  size_t off = 0;
  char *b = (char *)base;
  auto take = [&](size_t bytes) -> void * {
    void *p = b ? b + off : nullptr;
    off += align256(bytes);
    return p;
  };
  size_t nn = n ? n : 1;
  Workspace ws;
  ws.keys_a = (uint32_t *)take(4 * nn);
  ws.vals_a = (uint32_t *)take(4 * nn);
  ws.keys_b = (uint32_t *)take(4 * nn);
  ws.vals_b = (uint32_t *)take(4 * nn);
  ws.hist = (uint32_t *)take(4 * (size_t)256 * sort_blocks((uint32_t)nn));
  ws.row_total = (uint32_t *)take(4 * 256);
  ws.cbox = (uint32_t *)take(4 * 8);
  ws.cbox_part = (uint32_t *)take(4 * 6 * (size_t)kCboxBlocks);
  ws.child = (uint2 *)take(8 * nn);
  ws.level = (uint8_t *)take(nn);
  ws.top = (uint32_t *)take(4 * (size_t)top_capacity((uint32_t)nn));
  ws.top_count = (uint32_t *)take(4 * 64);
  ws.rec = take(leaf_record_bytes(type) * nn);
  if (w)
    *w = ws;
  return off;
}

size_t leaf_record_bytes(int type)
{
  switch (type) {
  case VRX_TRIANGLE:
    return 48;
  case VRX_SPHERE:
    return 16;
  case VRX_CYLINDER:
  case VRX_CONE:
  case VRX_CURVE:
    return 32;
  default:
    return 0;
  }
}

// --------------------------------------------------------------------------
// ordered-uint encoding of floats for atomic min/max
// --------------------------------------------------------------------------
__device__ __forceinline__ uint32_t f2o(float f)
{
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t u)
{
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

template <int T>
__device__ __forceinline__ void prim_box(const vrx_prims &p, uint32_t i, v3 &lo, v3 &hi)
{
  if constexpr (T == VRX_SPHERE) {
    uint32_t vi = p.d_index ? p.d_index[i] : i;
    v3 c = ld3(p.d_position + 3 * (size_t)vi);
    float r = p.d_radius ? p.d_radius[vi] : p.global_radius;
    lo = mk(c.x - r, c.y - r, c.z - r);
    hi = mk(c.x + r, c.y + r, c.z + r);
  } else if constexpr (T == VRX_CYLINDER) {
    uint32_t a = p.d_index ? p.d_index[2 * (size_t)i] : 2 * i;
    uint32_t b = p.d_index ? p.d_index[2 * (size_t)i + 1] : 2 * i + 1;
    v3 p0 = ld3(p.d_position + 3 * (size_t)a), p1 = ld3(p.d_position + 3 * (size_t)b);
    float r = fabsf(p.d_radius ? p.d_radius[i] : p.global_radius);
    v3 mn = vmin(p0, p1), mx = vmax(p0, p1);
    lo = mk(mn.x - r, mn.y - r, mn.z - r);
    hi = mk(mx.x + r, mx.y + r, mx.z + r);
  } else if constexpr (T == VRX_CONE) {
    // Cone.cpp:98-105: the end points' box grown by the larger end radius
    uint32_t a = p.d_index ? p.d_index[2 * (size_t)i] : 2 * i;
    uint32_t b = p.d_index ? p.d_index[2 * (size_t)i + 1] : 2 * i + 1;
    v3 p0 = ld3(p.d_position + 3 * (size_t)a), p1 = ld3(p.d_position + 3 * (size_t)b);
    float r = fmaxf(p.d_radius[a], p.d_radius[b]);
    v3 mn = vmin(p0, p1), mx = vmax(p0, p1);
    lo = mk(mn.x - r, mn.y - r, mn.z - r);
    hi = mk(mx.x + r, mx.y + r, mx.z + r);
  } else if constexpr (T == VRX_CURVE) {
    // a sphere swept from (p0, r0) to (p1, r1): its box is the two end
    // spheres' box, since p -/+ r is linear along the segment
    uint32_t a = p.d_index[i];
    v3 p0 = ld3(p.d_position + 3 * (size_t)a), p1 = ld3(p.d_position + 3 * (size_t)a + 3);
    float r0 = p.d_radius ? p.d_radius[a] : p.global_radius;
    float r1 = p.d_radius ? p.d_radius[a + 1] : p.global_radius;
    lo = vmin(mk(p0.x - r0, p0.y - r0, p0.z - r0), mk(p1.x - r1, p1.y - r1, p1.z - r1));
    hi = vmax(mk(p0.x + r0, p0.y + r0, p0.z + r0), mk(p1.x + r1, p1.y + r1, p1.z + r1));
  } else if constexpr (T == VRX_TRIANGLE) {
    uint32_t a = p.d_index ? p.d_index[3 * (size_t)i] : 3 * i;
    uint32_t b = p.d_index ? p.d_index[3 * (size_t)i + 1] : 3 * i + 1;
    uint32_t c = p.d_index ? p.d_index[3 * (size_t)i + 2] : 3 * i + 2;
    v3 v0 = ld3(p.d_position + 3 * (size_t)a), v1 = ld3(p.d_position + 3 * (size_t)b),
       v2 = ld3(p.d_position + 3 * (size_t)c);
    lo = vmin(vmin(v0, v1), v2);
    hi = vmax(vmax(v0, v1), v2);
  } else {
    const float *bx = p.d_position + 6 * (size_t)i;
    lo = mk(bx[0], bx[1], bx[2]);
    hi = mk(bx[3], bx[4], bx[5]);
  }
}

__device__ __forceinline__ v3 centroid(v3 lo, v3 hi)
{
  return mk((lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f);
}


__device__ __forceinline__ void cbox_add(v3 lo, v3 hi, uint32_t mn[3], uint32_t mx[3])
{
  const v3 c = centroid(lo, hi);
  const float cc[3] = {c.x, c.y, c.z};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (cc[k] == cc[k]) { // NaN centroids are ignored, as fminf/fmaxf do
      const uint32_t o = f2o(cc[k]);
      mn[k] = min(mn[k], o);
      mx[k] = max(mx[k], o);
    }
  }
}

// Centroid box: grid-stride, four primitives in flight per lane, then a wave
// reduction and one LDS step across the waves: one partial box per workgroup.
template <int T>
__global__ __launch_bounds__(256) void k_centroid_bounds(vrx_prims p, uint32_t *part)
{
  __shared__ uint32_t red[6][4];
  uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
  const uint32_t stride = gridDim.x * blockDim.x;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  for (; (uint64_t)i + 3ull * stride < p.n; i += 4 * stride) {
    v3 lo[4], hi[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
      prim_box<T>(p, i + (uint32_t)u * stride, lo[u], hi[u]);
#pragma unroll
    for (int u = 0; u < 4; u++)
      cbox_add(lo[u], hi[u], mn, mx);
  }
  for (; i < p.n; i += stride) {
    v3 lo, hi;
    prim_box<T>(p, i, lo, hi);
    cbox_add(lo, hi, mn, mx);
  }
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      mn[k] = min(mn[k], (uint32_t)__shfl_xor((int)mn[k], m, 64));
      mx[k] = max(mx[k], (uint32_t)__shfl_xor((int)mx[k], m, 64));
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      red[k][wave] = mn[k];
      red[3 + k][wave] = mx[k];
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    uint32_t v = red[k][0];
#pragma unroll
    for (int w = 1; w < 4; w++)
      v = k < 3 ? min(v, red[k][w]) : max(v, red[k][w]);
    part[(size_t)k * gridDim.x + blockIdx.x] = v;
  }
}

// the partial boxes of the nblocks workgroups -> the centroid box (a second
// launch instead of same-address atomics from every workgroup)
__global__ __launch_bounds__(256) void k_cbox_final(const uint32_t *part, uint32_t nblocks,
    uint32_t *cbox)
{
  __shared__ uint32_t red[6][4];
  uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
  for (uint32_t b = threadIdx.x; b < nblocks; b += 256) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      mn[k] = min(mn[k], part[(size_t)k * nblocks + b]);
      mx[k] = max(mx[k], part[(size_t)(3 + k) * nblocks + b]);
    }
  }
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      mn[k] = min(mn[k], (uint32_t)__shfl_xor((int)mn[k], m, 64));
      mx[k] = max(mx[k], (uint32_t)__shfl_xor((int)mx[k], m, 64));
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      red[k][wave] = mn[k];
      red[3 + k][wave] = mx[k];
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    uint32_t v = red[k][0];
#pragma unroll
    for (int w = 1; w < 4; w++)
      v = k < 3 ? min(v, red[k][w]) : max(v, red[k][w]);
    cbox[k] = v;
  }
}

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v)
{
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__device__ __forceinline__ uint32_t quant10(float c, float lo, float scale)
{
  float x = (c - lo) * scale;
  if (x >= 1023.0f)
    return 1023u;
  if (x > 0.0f)
    return (uint32_t)x;
  return 0u;
}

template <int T>
__device__ __forceinline__ void write_leaf(
    const vrx_prims &p, uint32_t orig, uint32_t k, void *leaves);

template <int T>
__global__ __launch_bounds__(256) void k_morton(
    vrx_prims p, const uint32_t *cbox, uint32_t *keys, uint32_t *vals, float *cbox_out, void *rec)
{
  float lo[3], hi[3], scale[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lo[k] = o2f(cbox[k]);
    hi[k] = o2f(cbox[3 + k]);
    float ext = hi[k] - lo[k];
    scale[k] = ext > 0.0f ? 1024.0f / ext : 0.0f;
  }
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (cbox_out && i == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      cbox_out[k] = lo[k];
      cbox_out[3 + k] = hi[k];
    }
  }
  if (i >= p.n)
    return;
  v3 blo, bhi;
  prim_box<T>(p, i, blo, bhi);
  v3 c = centroid(blo, bhi);
  uint32_t qx = quant10(c.x, lo[0], scale[0]);
  uint32_t qy = quant10(c.y, lo[1], scale[1]);
  uint32_t qz = quant10(c.z, lo[2], scale[2]);
  keys[i] = expand_bits10(qx) * 4u + expand_bits10(qy) * 2u + expand_bits10(qz);
  vals[i] = i;
  if constexpr (T != VRX_BOX)
    write_leaf<T>(p, i, i, rec); // input order: a coalesced write
}

// --------------------------------------------------------------------------
// radix sort
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(
    const uint32_t *keys, uint32_t n, uint32_t shift, uint32_t nblocks, uint32_t *hist)
{
  __shared__ uint32_t h[kSortThreads / 64][256];
  const int wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < (kSortThreads / 64) * 256; i += kSortThreads)
    (&h[0][0])[i] = 0u;
  __syncthreads();
  const uint32_t base = blockIdx.x * (uint32_t)kSortTile;
#pragma unroll
  for (int k = 0; k < kSortItems; k++) {
    uint32_t idx = base + (uint32_t)k * kSortThreads + threadIdx.x;
    if (idx < n)
      atomicAdd(&h[wave][(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  uint32_t d = threadIdx.x; // kSortThreads == 256 digits
  uint32_t s = 0;
#pragma unroll
  for (int w = 0; w < kSortThreads / 64; w++)
    s += h[w][d];
  hist[(size_t)d * nblocks + blockIdx.x] = s;
}

// Row scan of the digit-major histogram: workgroup d turns row d (one count
// per sort block) into exclusive block offsets within the digit and writes
// the digit total.  The digit bases are scanned inside the scatter kernel.
__global__ __launch_bounds__(256) void k_scan_rows(uint32_t *hist, uint32_t nblocks, uint32_t *row_total)
{
  __shared__ uint32_t sums[256];
  uint32_t *row = hist + (size_t)blockIdx.x * nblocks;
  const uint32_t chunk = (nblocks + 255) / 256;
  const uint32_t beg = min(nblocks, threadIdx.x * chunk);
  const uint32_t end = min(nblocks, beg + chunk);
  uint32_t s = 0;
  for (uint32_t i = beg; i < end; i++)
    s += row[i];
  sums[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    uint32_t v = (int)threadIdx.x >= off ? sums[threadIdx.x - off] : 0u;
    __syncthreads();
    sums[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x ? sums[threadIdx.x - 1] : 0u;
  for (uint32_t i = beg; i < end; i++) {
    uint32_t v = row[i];
    row[i] = run;
    run += v;
  }
  if (threadIdx.x == 255)
    row_total[blockIdx.x] = sums[255];
}

// Stable LDS-staged scatter.  Element order inside a tile is (item k, wave,
// lane), which is the original index order, so ranks are stable.
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t *keys_in,
    const uint32_t *vals_in, uint32_t *keys_out, uint32_t *vals_out, uint32_t n, uint32_t shift,
    uint32_t nblocks, const uint32_t *offsets, const uint32_t *row_total)
{
  constexpr int W = kSortThreads / 64;
  __shared__ uint32_t digit_base[256];
  __shared__ uint32_t digit_total[256];
  __shared__ uint32_t round_cnt[W][256];
  __shared__ uint32_t digit_start[256];
  __shared__ uint32_t stage_k[kSortTile];
  __shared__ uint32_t stage_v[kSortTile];

  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t base = blockIdx.x * (uint32_t)kSortTile;

  digit_total[tid] = 0u;
#pragma unroll
  for (int w = 0; w < W; w++)
    round_cnt[w][tid] = 0u;
  // exclusive scan of the digit totals: global start of each digit
  if (tid < 64) {
    uint32_t c0 = row_total[4 * tid], c1 = row_total[4 * tid + 1];
    uint32_t c2 = row_total[4 * tid + 2], c3 = row_total[4 * tid + 3];
    uint32_t sm = c0 + c1 + c2 + c3;
    uint32_t incl = sm;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      uint32_t v = __shfl_up(incl, off, 64);
      if (lane >= off)
        incl += v;
    }
    uint32_t ex = incl - sm;
    digit_base[4 * tid] = ex;
    digit_base[4 * tid + 1] = ex + c0;
    digit_base[4 * tid + 2] = ex + c0 + c1;
    digit_base[4 * tid + 3] = ex + c0 + c1 + c2;
  }

  uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
  for (int k = 0; k < kSortItems; k++) {
    uint32_t idx = base + (uint32_t)k * kSortThreads + tid;
    bool ok = idx < n;
    key[k] = ok ? keys_in[idx] : 0xFFFFFFFFu;
    val[k] = ok ? vals_in[idx] : 0u;
  }
  __syncthreads();

#pragma unroll
  for (int k = 0; k < kSortItems; k++) {
    uint32_t idx = base + (uint32_t)k * kSortThreads + tid;
    bool ok = idx < n;
    uint32_t d = (key[k] >> shift) & 255u;
    // lanes of this wave holding the same digit (valid ones only)
    uint64_t m = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      uint64_t bb = __ballot((d >> b) & 1u);
      m &= ((d >> b) & 1u) ? bb : ~bb;
    }
    uint32_t wrank = (uint32_t)__popcll(m & lt_mask);
    uint32_t wcount = (uint32_t)__popcll(m);
    if (ok && wrank == 0)
      round_cnt[wave][d] = wcount;
    __syncthreads();
    if (ok) {
      uint32_t pre = digit_total[d];
      for (int w = 0; w < wave; w++)
        pre += round_cnt[w][d];
      rank[k] = pre + wrank;
    }
    __syncthreads();
    {
      uint32_t s = 0;
#pragma unroll
      for (int w = 0; w < W; w++) {
        s += round_cnt[w][tid];
        round_cnt[w][tid] = 0u;
      }
      digit_total[tid] += s;
    }
    __syncthreads();
  }

  // exclusive scan of the block's digit counts -> digit_start
  if (tid < 64) {
    uint32_t c0 = digit_total[4 * tid], c1 = digit_total[4 * tid + 1];
    uint32_t c2 = digit_total[4 * tid + 2], c3 = digit_total[4 * tid + 3];
    uint32_t s = c0 + c1 + c2 + c3;
    uint32_t incl = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      uint32_t v = __shfl_up(incl, off, 64);
      if (lane >= off)
        incl += v;
    }
    uint32_t ex = incl - s;
    digit_start[4 * tid] = ex;
    digit_start[4 * tid + 1] = ex + c0;
    digit_start[4 * tid + 2] = ex + c0 + c1;
    digit_start[4 * tid + 3] = ex + c0 + c1 + c2;
  }
  __syncthreads();

#pragma unroll
  for (int k = 0; k < kSortItems; k++) {
    uint32_t idx = base + (uint32_t)k * kSortThreads + tid;
    if (idx < n) {
      uint32_t d = (key[k] >> shift) & 255u;
      uint32_t slot = digit_start[d] + rank[k];
      stage_k[slot] = key[k];
      stage_v[slot] = val[k];
    }
  }
  __syncthreads();

  const uint32_t valid = min((uint32_t)kSortTile, n - base);
  for (uint32_t i = tid; i < valid; i += kSortThreads) {
    uint32_t kk = stage_k[i];
    uint32_t d = (kk >> shift) & 255u;
    uint32_t pos = digit_base[d] + offsets[(size_t)d * nblocks + blockIdx.x] + (i - digit_start[d]);
    keys_out[pos] = kk;
    vals_out[pos] = stage_v[i];
  }
}

// --------------------------------------------------------------------------
// Karras hierarchy
// --------------------------------------------------------------------------

// One internal node per lane.  The sorted codes around the
// workgroup's nodes (kKarrasHalo on either side) are staged in LDS: a node's
// searches probe keys at most twice its range length away, so all but the
// largest nodes search in LDS; other probes read memory.
constexpr int kKarrasThreads = 1024;
constexpr int64_t kKarrasHalo = 1024;

// delta(i, j) of the Karras paper with index-augmented keys; j is given as
// the distance x from i in direction d (> 0: i + x, else i - x), so all index
// arithmetic is 32-bit (n < 2^31): -1 when j is outside [0, n).
struct KarrasProbe {
  const uint32_t *sc;
  const uint32_t *tile;
  uint32_t t0, tn, n, i, ci;
  int d;
  __device__ __forceinline__ int operator()(uint32_t x) const
  {
    if (d > 0 ? x >= n - i : x > i)
      return -1;
    const uint32_t j = d > 0 ? i + x : i - x;
    // the tile is always read (clamped index), memory only outside it: a
    // select between the two addresses would become one slow flat load
    const uint32_t tj = j - t0;
    uint32_t b = tile[tj < tn ? tj : (j < t0 ? 0u : tn - 1u)];
    if (tj >= tn)
      b = gld(sc, j);
    if (ci != b)
      return __clz(ci ^ b);
    return 32 + __clz(i ^ j);
  }
};

__global__ __launch_bounds__(kKarrasThreads) void k_karras(
    const uint32_t *sc, uint32_t n, uint2 *child, uint8_t *level)
{
  __shared__ uint32_t tile[kKarrasThreads + 2 * kKarrasHalo];
  const uint32_t b0 = blockIdx.x * (uint32_t)kKarrasThreads;
  const uint32_t t0 = b0 > (uint32_t)kKarrasHalo ? b0 - (uint32_t)kKarrasHalo : 0u;
  const uint32_t t1 = (uint32_t)min((uint64_t)b0 + kKarrasThreads + kKarrasHalo, (uint64_t)n);
  for (uint32_t q = t0 + threadIdx.x; q < t1; q += kKarrasThreads)
    tile[q - t0] = sc[q];
  __syncthreads();
  const uint32_t i = b0 + threadIdx.x;
  if (i >= n - 1)
    return;
  KarrasProbe kd{sc, tile, t0, t1 - t0, n, i, tile[i - t0], 1};
  // direction: towards the neighbour sharing the longer prefix
  const int up = kd(1u);
  kd.d = -1;
  const int down = kd(1u);
  const int d = up - down > 0 ? 1 : -1;
  kd.d = -d;
  const int dmin = kd(1u);
  kd.d = d;
  uint32_t lmax = 2;
  while (kd(lmax) > dmin)
    lmax <<= 1;
  uint32_t l = 0;
  for (uint32_t t = lmax >> 1; t >= 1; t >>= 1)
    if (kd(l + t) > dmin)
      l += t;
  const uint32_t j = d > 0 ? i + l : i - l;
  const int dnode = kd(l);
  uint32_t s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (kd(s + t) > dnode)
      s += t;
  } while (t > 1);
  const uint32_t gamma = d > 0 ? i + s : i - s - 1;
  const uint32_t lo = min(i, j), hi = max(i, j);
  const uint32_t left = (lo == gamma) ? (kLeafBit | gamma) : gamma;
  const uint32_t right = (hi == gamma + 1) ? (kLeafBit | (gamma + 1)) : (gamma + 1);
  child[i] = make_uint2(left, right);
  // common-prefix length of the node's key range (0..63: distinct keys
  // differ within 32 bits, equal keys within 31 index bits); every child
  // range has a strictly longer one.  The refit pass follows from where
  // the leaf range [lo, hi] lies.
  const uint32_t lv = (uint32_t)min(max(dnode, 0), 63);
  const uint32_t pass = (lo / kRefitChunk == hi / kRefitChunk) ? 1u
      : (lo / kRefitSpan == hi / kRefitSpan) ? 2u : 3u;
  level[i] = (uint8_t)(lv | (pass << kPassShift));
}

template <int T>
__device__ __forceinline__ void write_leaf(
    const vrx_prims &p, uint32_t orig, uint32_t k, void *leaves)
{
  if constexpr (T == VRX_SPHERE) {
    uint32_t vi = p.d_index ? p.d_index[orig] : orig;
    v3 c = ld3(p.d_position + 3 * (size_t)vi);
    float r = p.d_radius ? p.d_radius[vi] : p.global_radius;
    ((float4 *)leaves)[k] = make_float4(c.x, c.y, c.z, r);
  } else if constexpr (T == VRX_CYLINDER) {
    uint32_t a = p.d_index ? p.d_index[2 * (size_t)orig] : 2 * orig;
    uint32_t b = p.d_index ? p.d_index[2 * (size_t)orig + 1] : 2 * orig + 1;
    v3 p0 = ld3(p.d_position + 3 * (size_t)a), p1 = ld3(p.d_position + 3 * (size_t)b);
    float r = fabsf(p.d_radius ? p.d_radius[orig] : p.global_radius);
    float4 *L = (float4 *)leaves + 2 * (size_t)k;
    L[0] = make_float4(p0.x, p0.y, p0.z, r);
    L[1] = make_float4(p1.x, p1.y, p1.z, 0.0f);
  } else if constexpr (T == VRX_CONE) {
    uint32_t a = p.d_index ? p.d_index[2 * (size_t)orig] : 2 * orig;
    uint32_t b = p.d_index ? p.d_index[2 * (size_t)orig + 1] : 2 * orig + 1;
    v3 p0 = ld3(p.d_position + 3 * (size_t)a), p1 = ld3(p.d_position + 3 * (size_t)b);
    float4 *L = (float4 *)leaves + 2 * (size_t)k;
    L[0] = make_float4(p0.x, p0.y, p0.z, p.d_radius[a]);
    L[1] = make_float4(p1.x, p1.y, p1.z, p.d_radius[b]);
  } else if constexpr (T == VRX_CURVE) {
    uint32_t a = p.d_index[orig];
    v3 p0 = ld3(p.d_position + 3 * (size_t)a), p1 = ld3(p.d_position + 3 * (size_t)a + 3);
    float4 *L = (float4 *)leaves + 2 * (size_t)k;
    L[0] = make_float4(p0.x, p0.y, p0.z, p.d_radius ? p.d_radius[a] : p.global_radius);
    L[1] = make_float4(p1.x, p1.y, p1.z, p.d_radius ? p.d_radius[a + 1] : p.global_radius);
  } else if constexpr (T == VRX_TRIANGLE) {
    uint32_t a = p.d_index ? p.d_index[3 * (size_t)orig] : 3 * orig;
    uint32_t b = p.d_index ? p.d_index[3 * (size_t)orig + 1] : 3 * orig + 1;
    uint32_t c = p.d_index ? p.d_index[3 * (size_t)orig + 2] : 3 * orig + 2;
    v3 v0 = ld3(p.d_position + 3 * (size_t)a), v1 = ld3(p.d_position + 3 * (size_t)b),
       v2 = ld3(p.d_position + 3 * (size_t)c);
    float4 *L = (float4 *)leaves + 3 * (size_t)k;
    L[0] = make_float4(v0.x, v0.y, v0.z, 0.0f);
    L[1] = make_float4(v1.x, v1.y, v1.z, 0.0f);
    L[2] = make_float4(v2.x, v2.y, v2.z, 0.0f);
  }
}

// float4s per leaf record
template <int T>
constexpr int record_quads()
{
  return T == VRX_SPHERE ? 1
      : (T == VRX_CYLINDER || T == VRX_CONE || T == VRX_CURVE) ? 2 : (T == VRX_TRIANGLE ? 3 : 0);
}

// box of a leaf record (same bounds rule as prim_box)
template <int T>
__device__ __forceinline__ void record_box(const float4 *L, v3 &lo, v3 &hi)
{
  if constexpr (T == VRX_SPHERE) {
    const float4 a = L[0];
    lo = mk(a.x - a.w, a.y - a.w, a.z - a.w);
    hi = mk(a.x + a.w, a.y + a.w, a.z + a.w);
  } else if constexpr (T == VRX_CYLINDER) {
    const float4 a = L[0], b = L[1];
    v3 mn = vmin(mk(a.x, a.y, a.z), mk(b.x, b.y, b.z)), mx = vmax(mk(a.x, a.y, a.z), mk(b.x, b.y, b.z));
    lo = mk(mn.x - a.w, mn.y - a.w, mn.z - a.w);
    hi = mk(mx.x + a.w, mx.y + a.w, mx.z + a.w);
  } else if constexpr (T == VRX_CONE) {
    const float4 a = L[0], b = L[1];
    v3 mn = vmin(mk(a.x, a.y, a.z), mk(b.x, b.y, b.z)), mx = vmax(mk(a.x, a.y, a.z), mk(b.x, b.y, b.z));
    float r = fmaxf(a.w, b.w);
    lo = mk(mn.x - r, mn.y - r, mn.z - r);
    hi = mk(mx.x + r, mx.y + r, mx.z + r);
  } else if constexpr (T == VRX_CURVE) {
    const float4 a = L[0], b = L[1];
    lo = vmin(mk(a.x - a.w, a.y - a.w, a.z - a.w), mk(b.x - b.w, b.y - b.w, b.z - b.w));
    hi = vmax(mk(a.x + a.w, a.y + a.w, a.z + a.w), mk(b.x + b.w, b.y + b.w, b.z + b.w));
  } else if constexpr (T == VRX_TRIANGLE) {
    const float4 a = L[0], b = L[1], c = L[2];
    v3 v0 = mk(a.x, a.y, a.z), v1 = mk(b.x, b.y, b.z), v2 = mk(c.x, c.y, c.z);
    lo = vmin(vmin(v0, v1), v2);
    hi = vmax(vmax(v0, v1), v2);
  }
}

// box of sorted leaf k, from its leaf record
template <int T>
__device__ __forceinline__ void leaf_box(
    const vrx_prims &p, const void *leaves, const uint32_t *ids, uint32_t k, v3 &lo, v3 &hi)
{
  if constexpr (record_quads<T>() > 0) {
    const float4 *L = (const float4 *)leaves + (size_t)record_quads<T>() * k;
    float4 r[record_quads<T>()];
#pragma unroll
    for (int q = 0; q < record_quads<T>(); q++)
      r[q] = L[q];
    record_box<T>(r, lo, hi);
  } else {
    prim_box<T>(p, ids[k], lo, hi);
  }
}

template <int T>
__global__ __launch_bounds__(256) void k_leaves(
    vrx_prims p, const uint32_t *order, const void *rec, void *leaves, uint32_t *leaf_ids)
{
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= p.n)
    return;
  const uint32_t orig = order[k];
  leaf_ids[k] = orig;
  // gather the record the Morton pass wrote for primitive `orig`: one
  // contiguous 16/32/48-byte read instead of separate index, vertex and
  // radius gathers
  constexpr int nq = record_quads<T>();
  const float4 *src = (const float4 *)rec + (size_t)nq * orig;
  float4 *dst = (float4 *)leaves + (size_t)nq * k;
#pragma unroll
  for (int q = 0; q < nq; q++)
    gst(dst, q, gld(src, q));
}

// the box and height of child link c of a node, from the leaf records or
// from a finished internal node in memory
template <int T>
__device__ __forceinline__ void child_box_global(const vrx_prims &p, const void *leaves,
    const uint32_t *ids, const Node *nodes, uint32_t c, v3 &lo, v3 &hi, uint32_t &h)
{
  if (c & kLeafBit) {
    leaf_box<T>(p, leaves, ids, c & ~kLeafBit, lo, hi);
    h = 1u;
  } else {
    const Node C = nodes[c];
    lo = vmin(mk(C.a.x, C.a.y, C.a.z), mk(C.b.z, C.b.w, C.c.x));
    hi = vmax(mk(C.a.w, C.b.x, C.b.y), mk(C.c.y, C.c.z, C.c.w));
    h = 1u + max(C.d.z, C.d.w);
  }
}

__device__ __forceinline__ Node make_node(uint2 ch, const v3 lo[2], const v3 hi[2], const uint32_t h[2])
{
  Node N;
  N.a = make_float4(lo[0].x, lo[0].y, lo[0].z, hi[0].x);
  N.b = make_float4(hi[0].y, hi[0].z, lo[1].x, lo[1].y);
  N.c = make_float4(lo[1].z, hi[1].x, hi[1].y, hi[1].z);
  N.d = make_uint4(ch.x, ch.y, h[0], h[1]);
  return N;
}

// boxes and heights of an inside node's children, from the chunk's LDS
__device__ __forceinline__ void chunk_children(uint2 ch, uint32_t base, const float (*s_leaf)[6],
    const float (*s_box)[6], const uint8_t *s_h, v3 lo[2], v3 hi[2], uint32_t h[2])
{
#pragma unroll
  for (int sd = 0; sd < 2; sd++) {
    const uint32_t c = sd ? ch.y : ch.x;
    const uint32_t cl = (c & ~kLeafBit) - base;
    const float *bx = (c & kLeafBit) ? s_leaf[cl] : s_box[cl];
    lo[sd] = mk(bx[0], bx[1], bx[2]);
    hi[sd] = mk(bx[3], bx[4], bx[5]);
    h[sd] = (c & kLeafBit) ? 1u : (uint32_t)s_h[cl];
  }
}

// Pass 1: one workgroup per chunk of kRefitChunk sorted leaves.  The chunk's
// leaf boxes, child pairs and level bytes are loaded into LDS at once; each
// inside node records itself as its children's parent, and every leaf then
// walks up: at each inside parent the first of its two children to arrive
// stops (LDS arrival counter), the second refits the parent from the two
// boxes in LDS and goes on.  The walk stops below a parent that is not an
// inside node.  No barrier between tree levels.
template <int T>
__global__ __launch_bounds__(kRefitThreads) void k_refit_chunk(vrx_prims p, const void *rec,
    void *leaves, const uint32_t *ids, const uint2 *child, const uint8_t *level, Node *nodes)
{
  __shared__ float s_leaf[kRefitChunk][6];
  __shared__ float s_box[kRefitChunk][6];
  __shared__ uint2 s_child[kRefitChunk];
  __shared__ uint32_t s_arrive[kRefitChunk];
  __shared__ uint16_t s_lpar[kRefitChunk], s_npar[kRefitChunk];
  __shared__ uint8_t s_h[kRefitChunk], s_pass[kRefitChunk];
  constexpr int kPer = (int)kRefitChunk / kRefitThreads;
  constexpr uint16_t kNone = 0xFFFFu;
  const uint32_t n = p.n;
  const uint32_t base = blockIdx.x * kRefitChunk;
  const uint32_t nleaf = min(kRefitChunk, n - base);
  // candidates base .. base + nnode - 1 (none in a last chunk of one leaf)
  const uint32_t nnode = base + 1 < n ? min(kRefitChunk, n - 1 - base) : 0u;
  const int tid = threadIdx.x;
#pragma unroll
  for (int m = 0; m < kPer; m++) {
    const uint32_t q = (uint32_t)tid + (uint32_t)m * kRefitThreads;
    s_lpar[q] = kNone;
    s_npar[q] = kNone;
    s_arrive[q] = 0u;
    s_pass[q] = 0u;
    if (q < nleaf) {
      // the leaf record of sorted leaf base + q: gathered from the record
      // the Morton pass wrote for the primitive, stored in sorted order
      const uint32_t orig = ids[base + q];
      v3 lo, hi;
      if constexpr (record_quads<T>() > 0) {
        constexpr int nq = record_quads<T>();
        float4 r[nq];
#pragma unroll
        for (int k = 0; k < nq; k++) {
          r[k] = gld((const float4 *)rec + (size_t)nq * orig, k);
          gst((float4 *)leaves + (size_t)nq * (base + q), k, r[k]);
        }
        record_box<T>(r, lo, hi);
      } else {
        prim_box<T>(p, orig, lo, hi);
      }
      s_leaf[q][0] = lo.x; s_leaf[q][1] = lo.y; s_leaf[q][2] = lo.z;
      s_leaf[q][3] = hi.x; s_leaf[q][4] = hi.y; s_leaf[q][5] = hi.z;
    }
    if (q < nnode) {
      s_child[q] = child[base + q];
      s_pass[q] = (uint8_t)(level[base + q] >> kPassShift);
    }
  }
  __syncthreads();
  // an inside node's children are leaves or inside nodes of this chunk
#pragma unroll
  for (int m = 0; m < kPer; m++) {
    const uint32_t q = (uint32_t)tid + (uint32_t)m * kRefitThreads;
    if (q < nnode && s_pass[q] == 1u) {
      const uint2 ch = s_child[q];
#pragma unroll
      for (int sd = 0; sd < 2; sd++) {
        const uint32_t c = sd ? ch.y : ch.x;
        const uint32_t cl = (c & ~kLeafBit) - base;
        if (c & kLeafBit)
          s_lpar[cl] = (uint16_t)q;
        else
          s_npar[cl] = (uint16_t)q;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < kPer; m++) {
    const uint32_t q = (uint32_t)tid + (uint32_t)m * kRefitThreads;
    uint32_t par = q < nleaf ? s_lpar[q] : kNone;
    while (par != kNone) {
      // the box of the arriving child was stored before this add; the
      // second arrival sees both
      if (__hip_atomic_fetch_add(&s_arrive[par], 1u, __ATOMIC_ACQ_REL,
              __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
        break;
      v3 lo[2], hi[2];
      uint32_t h[2];
      chunk_children(s_child[par], base, s_leaf, s_box, s_h, lo, hi, h);
      const v3 ulo = vmin(lo[0], lo[1]), uhi = vmax(hi[0], hi[1]);
      s_box[par][0] = ulo.x; s_box[par][1] = ulo.y; s_box[par][2] = ulo.z;
      s_box[par][3] = uhi.x; s_box[par][4] = uhi.y; s_box[par][5] = uhi.z;
      s_h[par] = (uint8_t)(1u + max(h[0], h[1]));
      par = s_npar[par];
    }
  }
  __syncthreads();
  // the chunk's inside nodes from LDS, written lane-contiguous (whole lines)
#pragma unroll
  for (int m = 0; m < kPer; m++) {
    const uint32_t q = (uint32_t)tid + (uint32_t)m * kRefitThreads;
    if (q < nnode && s_pass[q] == 1u) {
      v3 lo[2], hi[2];
      uint32_t h[2];
      const uint2 ch = s_child[q];
      chunk_children(ch, base, s_leaf, s_box, s_h, lo, hi, h);
      nodes[base + q] = make_node(ch, lo, hi, h);
    }
  }
}

// refit one node from its children in memory (passes 2 and 3)
template <int T>
__device__ __forceinline__ void refit_node_global(const vrx_prims &p, const void *leaves,
    const uint32_t *ids, uint2 ch, Node *nodes, uint32_t node)
{
  v3 lo[2], hi[2];
  uint32_t h[2];
  child_box_global<T>(p, leaves, ids, nodes, ch.x, lo[0], hi[0], h[0]);
  child_box_global<T>(p, leaves, ids, nodes, ch.y, lo[1], hi[1], h[1]);
  nodes[node] = make_node(ch, lo, hi, h);
}

// Passes 2 and 3 share this: `cnt` entries (key << 6 | delta) in s_in, node
// node_of(key), are sorted
// by delta into s_node (LDS counting sort), their child pairs loaded into
// s_ch, and the nodes are refit delta by delta, longest first; a node's
// children are finished nodes in memory (earlier launches, or this
// workgroup's own nodes of a longer delta, visible after the barrier that
// ends their delta) or leaves.
template <int T, class NodeOf>
__device__ __forceinline__ void refit_sorted(const vrx_prims &p, const void *leaves,
    const uint32_t *ids, const uint2 *child, Node *nodes, const uint32_t *s_in, uint32_t cnt,
    NodeOf node_of, uint32_t *s_node, uint2 *s_ch, uint32_t *s_cnt, uint32_t *s_off)
{
  const uint32_t nt = blockDim.x, tid = threadIdx.x;
  if (tid < 64)
    s_cnt[tid] = 0u;
  __syncthreads();
  for (uint32_t q = tid; q < cnt; q += nt)
    atomicAdd(&s_cnt[s_in[q] & 63u], 1u);
  __syncthreads();
  if (tid < 64) {
    const uint32_t c = s_cnt[tid];
    uint32_t incl = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t v = __shfl_up(incl, off, 64);
      if ((int)tid >= off)
        incl += v;
    }
    s_off[tid] = incl - c;
    s_off[64 + tid] = incl - c; // fill cursor
  }
  __syncthreads();
  for (uint32_t q = tid; q < cnt; q += nt) {
    const uint32_t slot = atomicAdd(&s_off[64 + (s_in[q] & 63u)], 1u);
    const uint32_t node = node_of(s_in[q] >> 6);
    s_node[slot] = node;
    s_ch[slot] = child[node];
  }
  __syncthreads();
  for (int lv = 63; lv >= 0; lv--) {
    const uint32_t c = s_cnt[lv];
    if (!c)
      continue;
    const uint32_t off = s_off[lv];
    for (uint32_t q = tid; q < c; q += nt)
      refit_node_global<T>(p, leaves, ids, s_ch[off + q], nodes, s_node[off + q]);
    __syncthreads();
  }
}

// Pass 2: one workgroup per span of kRefitSpan leaves refits the span's
// nodes that cross a chunk border and lists those crossing a span border for
// the top pass.  The span's nodes are listed in index order (ordered
// compaction of the level bytes); each finds its parent's list slot by a
// binary search, and the nodes are refit by bottom-up walks with LDS
// arrival counters, as in pass 1, with the boxes in memory: a child node is
// a pass-1 node of the previous launch, or a pass-2 node this workgroup
// wrote before the arrival add that lets its parent go.
template <int T>
__global__ __launch_bounds__(kSpanThreads) void k_refit_span(vrx_prims p, const void *leaves,
    const uint32_t *ids, const uint2 *child, const uint8_t *level, Node *nodes, uint32_t *top,
    uint32_t *top_count, uint32_t top_cap)
{
  constexpr uint32_t kBytes = kRefitSpan / kSpanThreads; // level bytes per lane (128)
  constexpr uint32_t kSlots = kSpanListCap / kSpanThreads; // list slots per lane (8)
  constexpr uint16_t kNone = 0xFFFFu;
  __shared__ uint32_t s_list[kSpanListCap];
  __shared__ uint32_t s_need[kSpanListCap];
  __shared__ uint16_t s_par[kSpanListCap];
  __shared__ uint32_t s_wsum[kSpanThreads / 64];
  __shared__ uint32_t s_total;
  const uint32_t n = p.n;
  const uint32_t base = blockIdx.x * kRefitSpan;
  const uint32_t nnode = min(kRefitSpan, n - 1 - base);
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // 1. this lane's kBytes level bytes (16-byte loads; the span start and the
  //    level array are 16-byte aligned)
  const uint32_t b0 = tid * kBytes;
  // 16 level bytes of this lane's range (zero past the span's last node)
  auto bytes16 = [&](uint32_t v) {
    const uint32_t q0 = b0 + v * 16;
    if (q0 + 16 <= nnode)
      return ((const uint4 *)(level + base))[(b0 >> 4) + v];
    uint32_t t[4] = {0u, 0u, 0u, 0u};
    for (uint32_t k = 0; k < 16 && q0 + k < nnode; k++)
      t[k >> 2] |= (uint32_t)level[base + q0 + k] << (8 * (k & 3));
    return make_uint4(t[0], t[1], t[2], t[3]);
  };
  auto pass_in = [](const uint4 &w, uint32_t k) {
    const uint32_t x = k < 4 ? w.x : k < 8 ? w.y : k < 12 ? w.z : w.w;
    return (x >> (8 * (k & 3) + kPassShift)) & 3u;
  };
  uint32_t mine = 0;
  for (uint32_t v = 0; v < kBytes / 16; v++) {
    const uint4 w = bytes16(v);
#pragma unroll
    for (uint32_t k = 0; k < 16; k++) {
      const uint32_t pass = pass_in(w, k);
      mine += pass == 2u ? 1u : 0u;
      if (pass == 3u) {
        const uint32_t t = atomicAdd(top_count, 1u);
        if (t < top_cap)
          top[t] = base + b0 + v * 16 + k;
      }
    }
  }
  // exclusive scan of the per-lane counts (wave scan, then wave totals)
  uint32_t incl = mine;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t v = __shfl_up(incl, off, 64);
    if ((int)lane >= off)
      incl += v;
  }
  if (lane == 63)
    s_wsum[wave] = incl;
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (uint32_t w = 0; w < kSpanThreads / 64; w++) {
      const uint32_t c = s_wsum[w];
      s_wsum[w] = run;
      run += c;
    }
    s_total = run;
  }
  __syncthreads();
  uint32_t slot = s_wsum[wave] + incl - mine;
  for (uint32_t v = 0; v < kBytes / 16; v++) {
    const uint4 w = bytes16(v);
#pragma unroll
    for (uint32_t k = 0; k < 16; k++)
      if (pass_in(w, k) == 2u) {
        if (slot < kSpanListCap)
          s_list[slot] = b0 + v * 16 + k;
        slot++;
      }
  }
  const uint32_t cnt = min(s_total, kSpanListCap);
  for (uint32_t i = tid; i < kSpanListCap; i += kSpanThreads)
    s_par[i] = kNone;
  __syncthreads();
  // 2. parent links among the listed nodes and the number of listed
  //    children each waits for
  uint32_t start = 0; // bit m: slot tid + m * kSpanThreads has no listed child
#pragma unroll
  for (uint32_t m = 0; m < kSlots; m++) {
    const uint32_t i = tid + m * kSpanThreads;
    if (i >= cnt)
      continue;
    const uint2 ch = child[base + s_list[i]];
    uint32_t need = 0;
#pragma unroll
    for (int sd = 0; sd < 2; sd++) {
      const uint32_t c = sd ? ch.y : ch.x;
      if (c & kLeafBit)
        continue;
      const uint32_t cl = c - base; // in the span: the child's range is in its parent's
      uint32_t lo = 0, hi = cnt;      // first slot with s_list >= cl
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_list[mid] < cl)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo < cnt && s_list[lo] == cl) {
        s_par[lo] = (uint16_t)i;
        need++;
      }
    }
    s_need[i] = need;
    start |= need == 0 ? (1u << m) : 0u;
  }
  __syncthreads();
  // 3. walks from the nodes with no listed child
#pragma unroll
  for (uint32_t m = 0; m < kSlots; m++) {
    if (!((start >> m) & 1u))
      continue;
    uint32_t i = tid + m * kSpanThreads;
    for (;;) {
      const uint32_t node = base + s_list[i];
      refit_node_global<T>(p, leaves, ids, child[node], nodes, node);
      const uint32_t par = s_par[i];
      if (par == kNone)
        break;
      // the node store above happens before this add (release); the parent
      // reads its children after the add that brings its count to zero
      if (__hip_atomic_fetch_sub(&s_need[par], 1u, __ATOMIC_ACQ_REL,
              __HIP_MEMORY_SCOPE_WORKGROUP) != 1u)
        break;
      i = par;
    }
  }
}

// Pass 3: one workgroup refits the listed nodes that cross a span border,
// kSpanListCap at a time (a batch holds whole deltas only if the list fits;
// beyond that the list is refit delta by delta from memory).
template <int T>
__global__ __launch_bounds__(kSpanThreads) void k_refit_top(vrx_prims p, const void *leaves,
    const uint32_t *ids, const uint2 *child, const uint8_t *level, Node *nodes,
    const uint32_t *top, const uint32_t *top_count, uint32_t top_cap)
{
  __shared__ uint32_t s_in[kSpanListCap];
  __shared__ uint32_t s_node[kSpanListCap];
  __shared__ uint2 s_ch[kSpanListCap];
  __shared__ uint32_t s_cnt[64], s_off[128];
  const uint32_t cnt = min(*top_count, top_cap);
  if (cnt <= kSpanListCap) {
    for (uint32_t q = threadIdx.x; q < cnt; q += kSpanThreads)
      s_in[q] = (q << 6) | (level[top[q]] & 63u);
    __syncthreads();
    refit_sorted<T>(p, leaves, ids, child, nodes, s_in, cnt, [top](uint32_t k) { return top[k]; },
        s_node, s_ch, s_cnt, s_off);
    return;
  }
  for (int lv = 63; lv >= 0; lv--) {
    for (uint32_t q = threadIdx.x; q < cnt; q += kSpanThreads) {
      const uint32_t node = top[q];
      if ((level[node] & 63u) == (uint32_t)lv)
        refit_node_global<T>(p, leaves, ids, child[node], nodes, node);
    }
    __syncthreads();
  }
}

__device__ __forceinline__ void write_header(
    vrx_bvh_header *h, v3 lo, v3 hi, uint32_t height, uint32_t root, uint32_t n, int type)
{
  h->lo[0] = lo.x; h->lo[1] = lo.y; h->lo[2] = lo.z;
  h->hi[0] = hi.x; h->hi[1] = hi.y; h->hi[2] = hi.z;
  h->height = height;
  h->root = root;
  h->n = n;
  h->type = type;
}

template <int T>
__global__ void k_root_header(vrx_prims p, const void *leaves, const uint32_t *ids,
    const Node *nodes, vrx_bvh_header *header)
{
  if (threadIdx.x != 0)
    return;
  if (p.n == 1) {
    v3 lo, hi;
    leaf_box<T>(p, leaves, ids, 0u, lo, hi);
    write_header(header, lo, hi, 1u, kLeafBit, 1u, T);
    return;
  }
  const Node C = nodes[0];
  v3 lo = vmin(mk(C.a.x, C.a.y, C.a.z), mk(C.b.z, C.b.w, C.c.x));
  v3 hi = vmax(mk(C.a.w, C.b.x, C.b.y), mk(C.c.y, C.c.z, C.c.w));
  write_header(header, lo, hi, 1u + max(C.d.z, C.d.w), 0u, p.n, T);
}

// 4-wide record i from BVH2 node i: each child that is an internal node is
// replaced by its two children (boxes taken from that node), a leaf child
// keeps its slot.  Every BVH2 node gets a record; traversal from the root
// only reaches those two BVH2 levels apart.
__device__ __forceinline__ void node_child(const Node &nd, int j, float b[6], uint32_t &link)
{
  if (j == 0) {
    b[0] = nd.a.x; b[1] = nd.a.y; b[2] = nd.a.z; b[3] = nd.a.w; b[4] = nd.b.x; b[5] = nd.b.y;
    link = nd.d.x;
  } else {
    b[0] = nd.b.z; b[1] = nd.b.w; b[2] = nd.c.x; b[3] = nd.c.y; b[4] = nd.c.z; b[5] = nd.c.w;
    link = nd.d.y;
  }
}

__global__ __launch_bounds__(256) void k_widen(const Node *nodes, uint32_t ninner, Node4 *wide)
{
  // records are staged in LDS and written out lane-contiguous, so each store
  // instruction covers whole 128-byte lines
  __shared__ float4 s_rec[256 * 8];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t b0 = blockIdx.x * blockDim.x;
  const uint32_t nrec = min(256u, ninner - b0);
  if (i < ninner) {
    const Node nd = gld_node(nodes, i);
    // the one or two slots each child contributes (fixed register indices:
    // a dynamically indexed slot array would spill to scratch)
    float cb[2][2][6];
    uint32_t cl[2][2], cc[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
      uint32_t link;
      node_child(nd, j, cb[j][0], link);
      if (link & kLeafBit) {
        cl[j][0] = link;
        cl[j][1] = kLeafBit;
#pragma unroll
        for (int q = 0; q < 6; q++)
          cb[j][1][q] = 0.0f;
        cc[j] = 1;
      } else {
        const Node cn = gld_node(nodes, link);
        node_child(cn, 0, cb[j][0], cl[j][0]);
        node_child(cn, 1, cb[j][1], cl[j][1]);
        cc[j] = 2;
      }
    }
    const bool two = cc[0] == 2;
    const uint32_t cnt = cc[0] + cc[1];
    float bx[4][6];
    uint32_t lk[4];
#pragma unroll
    for (int q = 0; q < 6; q++) {
      bx[0][q] = cb[0][0][q];
      bx[1][q] = two ? cb[0][1][q] : cb[1][0][q];
      bx[2][q] = two ? cb[1][0][q] : cb[1][1][q];
      bx[3][q] = two ? cb[1][1][q] : 0.0f;
    }
    lk[0] = cl[0][0];
    lk[1] = two ? cl[0][1] : cl[1][0];
    lk[2] = two ? cl[1][0] : cl[1][1];
    lk[3] = two ? cl[1][1] : kLeafBit;
    float4 *w = s_rec + threadIdx.x * 8;
    w[0] = make_float4(bx[0][0], bx[1][0], bx[2][0], bx[3][0]);
    w[1] = make_float4(bx[0][1], bx[1][1], bx[2][1], bx[3][1]);
    w[2] = make_float4(bx[0][2], bx[1][2], bx[2][2], bx[3][2]);
    w[3] = make_float4(bx[0][3], bx[1][3], bx[2][3], bx[3][3]);
    w[4] = make_float4(bx[0][4], bx[1][4], bx[2][4], bx[3][4]);
    w[5] = make_float4(bx[0][5], bx[1][5], bx[2][5], bx[3][5]);
    w[6] = make_float4(__uint_as_float(lk[0]), __uint_as_float(lk[1]), __uint_as_float(lk[2]),
        __uint_as_float(lk[3]));
    w[7] = make_float4(__uint_as_float(cnt), 0.0f, 0.0f, 0.0f);
  }
  __syncthreads();
  float4 *out = (float4 *)(wide + b0);
  for (uint32_t q = threadIdx.x; q < nrec * 8; q += blockDim.x)
    gst(out, q, s_rec[q]);
}

__global__ void k_empty_header(vrx_bvh_header *h, int type)
{
  write_header(h, mk(INFINITY, INFINITY, INFINITY), mk(-INFINITY, -INFINITY, -INFINITY), 0u,
      kLeafBit, 0u, type);
}

__global__ void k_fill_zero64(uint32_t *p)
{
  if (threadIdx.x < 64)
    p[threadIdx.x] = 0u;
}

__global__ void k_copy_u32(const uint32_t *src, uint32_t *dst, uint32_t n)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    dst[i] = src[i];
}

template <int T>
static int build_typed(const vrx_prims &p, const vrx_bvh &out, const Workspace &ws,
    uint32_t *codes_out, float *cbox_out, hipStream_t s)
{
  const uint32_t n = p.n;
  const unsigned g256 = (n + 255) / 256;
  const unsigned gb = min(g256, kCboxBlocks);
  hipLaunchKernelGGL(k_centroid_bounds<T>, dim3(gb), dim3(256), 0, s, p, ws.cbox_part);
  hipLaunchKernelGGL(k_cbox_final, dim3(1), dim3(256), 0, s, ws.cbox_part, gb, ws.cbox);
  hipLaunchKernelGGL(k_morton<T>, dim3(g256), dim3(256), 0, s, p, ws.cbox, ws.keys_a, ws.vals_a,
      cbox_out, ws.rec);
  VRX_LAUNCH_CHECK();
  const uint32_t nb = sort_blocks(n);
  uint32_t *ki = ws.keys_a, *vi = ws.vals_a, *ko = ws.keys_b, *vo = ws.vals_b;
  for (int pass = 0; pass < 4; pass++) {
    uint32_t shift = 8u * (uint32_t)pass;
    if (pass == 3)
      vo = out.d_leaf_ids; // the sorted order is the leaf -> primitive map
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kSortThreads), 0, s, ki, n, shift, nb, ws.hist);
    hipLaunchKernelGGL(k_scan_rows, dim3(256), dim3(256), 0, s, ws.hist, nb, ws.row_total);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kSortThreads), 0, s, ki, vi, ko, vo, n,
        shift, nb, ws.hist, ws.row_total);
    uint32_t *t;
    t = ki; ki = ko; ko = t;
    t = vi; vi = vo; vo = t;
  }
  VRX_LAUNCH_CHECK();
  // after 4 passes the sorted keys are back in keys_a, the order in leaf_ids
  if (codes_out)
    hipLaunchKernelGGL(k_copy_u32, dim3(g256), dim3(256), 0, s, ki, codes_out, n);
  if (n < 2) // otherwise the chunk refit gathers the leaf records
    hipLaunchKernelGGL((k_leaves<T>), dim3(g256), dim3(256), 0, s, p, vi, ws.rec, out.d_leaves,
        out.d_leaf_ids);
  Node *nodes = (Node *)out.d_nodes;
  if (n >= 2) {
    const uint32_t ni = n - 1;
    const unsigned gi = (ni + 255) / 256;
    hipLaunchKernelGGL(k_fill_zero64, dim3(1), dim3(64), 0, s, ws.top_count);
    hipLaunchKernelGGL(k_karras, dim3((ni + kKarrasThreads - 1) / kKarrasThreads),
        dim3(kKarrasThreads), 0, s, ki, n, ws.child, ws.level);
    hipLaunchKernelGGL((k_refit_chunk<T>), dim3((n + kRefitChunk - 1) / kRefitChunk),
        dim3(kRefitThreads), 0, s, p, ws.rec, out.d_leaves, out.d_leaf_ids, ws.child, ws.level,
        nodes);
    hipLaunchKernelGGL((k_refit_span<T>), dim3((ni + kRefitSpan - 1) / kRefitSpan),
        dim3(kSpanThreads), 0, s, p, out.d_leaves, out.d_leaf_ids, ws.child, ws.level, nodes,
        ws.top, ws.top_count, top_capacity(n));
    hipLaunchKernelGGL((k_refit_top<T>), dim3(1), dim3(kSpanThreads), 0, s, p, out.d_leaves,
        out.d_leaf_ids, ws.child, ws.level, nodes, ws.top, ws.top_count, top_capacity(n));
  }
  hipLaunchKernelGGL((k_root_header<T>), dim3(1), dim3(64), 0, s, p, out.d_leaves, out.d_leaf_ids,
      nodes, (vrx_bvh_header *)out.d_header);
  VRX_LAUNCH_CHECK();
  return VRX_OK;
}

int bvh_build_impl(const vrx_prims *p, const vrx_bvh *out, void *d_ws, size_t ws_bytes,
    uint32_t *codes_out, float *cbox_out, hipStream_t s)
{
  VRX_ARG(p && out, "prims/out");
  VRX_ARG(p->n < 0x7FFFFFFFu, "too many primitives (max 2^31-1)");
  VRX_ARG(out->d_header, "out->d_header");
  if (p->n == 0) {
    hipLaunchKernelGGL(k_empty_header, dim3(1), dim3(1), 0, s, (vrx_bvh_header *)out->d_header,
        p->type);
    VRX_LAUNCH_CHECK();
    return VRX_OK;
  }
  VRX_ARG(p->d_position, "d_position");
  VRX_ARG(out->d_leaf_ids, "out->d_leaf_ids");
  VRX_ARG(p->n < 2 || out->d_nodes, "out->d_nodes");
  VRX_ARG(leaf_record_bytes(p->type) == 0 || out->d_leaves, "out->d_leaves");
  size_t need = workspace_layout(p->type, p->n, nullptr, nullptr);
  if (!d_ws || ws_bytes < need) {
    set_error("workspace too small: %zu < %zu", ws_bytes, need);
    return VRX_ERR_ARG;
  }
  Workspace ws;
  workspace_layout(p->type, p->n, d_ws, &ws);
  switch (p->type) {
  case VRX_TRIANGLE:
    return build_typed<VRX_TRIANGLE>(*p, *out, ws, codes_out, cbox_out, s);
  case VRX_SPHERE:
    return build_typed<VRX_SPHERE>(*p, *out, ws, codes_out, cbox_out, s);
  case VRX_CYLINDER:
    return build_typed<VRX_CYLINDER>(*p, *out, ws, codes_out, cbox_out, s);
  case VRX_BOX:
    return build_typed<VRX_BOX>(*p, *out, ws, codes_out, cbox_out, s);
  case VRX_CONE:
    VRX_ARG(p->d_radius, "cones need per-vertex radii (d_radius)");
    return build_typed<VRX_CONE>(*p, *out, ws, codes_out, cbox_out, s);
  case VRX_CURVE:
    VRX_ARG(p->d_index, "curves need segment indices (d_index)");
    return build_typed<VRX_CURVE>(*p, *out, ws, codes_out, cbox_out, s);
  default:
    set_error("unknown primitive type %d", p->type);
    return VRX_ERR_ARG;
  }
}

void bvh_sizes(int32_t type, uint32_t n, vrx_bvh_sizes *o)
{
  o->nodes = (size_t)(n > 1 ? n - 1 : 1) * sizeof(Node);
  o->leaves = (size_t)(n ? n : 1) * (leaf_record_bytes(type) ? leaf_record_bytes(type) : 16);
  o->leaf_ids = (size_t)(n ? n : 1) * 4;
  o->header = sizeof(vrx_bvh_header);
  o->workspace = workspace_layout(type, n, nullptr, nullptr);
}

} // namespace vrx

using namespace vrx;

extern "C" {

int vrx_bvh_query_sizes(int32_t type, uint32_t n, vrx_bvh_sizes *out)
{
  VRX_ARG(out, "out");
  VRX_ARG(type >= 0 && type <= VRX_CURVE, "type");
  bvh_sizes(type, n, out);
  return VRX_OK;
}

int vrx_bvh_build(const vrx_prims *prims, const vrx_bvh *out, void *d_workspace,
    size_t workspace_bytes, vrx_stream s)
{
  return bvh_build_impl(prims, out, d_workspace, workspace_bytes, nullptr, nullptr, (hipStream_t)s);
}

int vrx_bvh_build_debug(const vrx_prims *prims, const vrx_bvh *out, void *d_workspace,
    size_t workspace_bytes, uint32_t *d_codes_out, float *d_cbox_out, vrx_stream s)
{
  return bvh_build_impl(
      prims, out, d_workspace, workspace_bytes, d_codes_out, d_cbox_out, (hipStream_t)s);
}

size_t vrx_bvh_wide_bytes(uint32_t n)
{
  return n > 1 ? (size_t)(n - 1) * sizeof(Node4) : 0;
}

int vrx_bvh_widen(const vrx_bvh *bvh, uint32_t n, vrx_stream s)
{
  VRX_ARG(bvh, "bvh");
  if (n < 2)
    return VRX_OK; // a single leaf (or nothing): no inner nodes
  VRX_ARG(bvh->d_nodes && bvh->d_wide, "bvh nodes / wide buffer");
  const uint32_t ninner = n - 1;
  hipLaunchKernelGGL(k_widen, dim3((ninner + 255) / 256), dim3(256), 0, (hipStream_t)s,
      (const Node *)bvh->d_nodes, ninner, (Node4 *)bvh->d_wide);
  VRX_LAUNCH_CHECK();
  return VRX_OK;
}

} // extern "C"
