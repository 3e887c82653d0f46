// batch.hip -- many views of one scene in one persistent launch.
//
// The reference renders one view per optixLaunch (frame/Frame.cu:282-289);
// it has no batched form.  Here every wave of a persistent grid pulls 64-pixel
// tiles of any view from one atomic work queue and runs the same per-pixel
// raycast program as the per-view kernel (vrx_shade.h), so one launch renders
// all views with no per-view tail, and the frame can be captured once in a
// HIP graph (the queue counter is reset by a fill kernel inside the graph,
// never by a memset node).  vrx_batch_autotune times tile shape, workgroup
// size and LDS stack depth on the actual scene and keeps the fastest.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "vrx_internal.h"
#include "vrx_shade.h"

namespace vrx {

constexpr int kBatchMaxBlock = 256;
constexpr uint32_t kMaxThreadsPerCU = 2048;

__device__ __forceinline__ vrx_camera load_camera(const vrx_camera *cams, uint32_t view)
{
  vrx_camera c;
  const float *src = (const float *)(cams + view);
  float *dst = (float *)&c;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(vrx_camera) / 4); i++)
    dst[i] = gld(src, i);
  return c;
}

template <int PT, bool TWO, bool WIDE, int KIND>
__device__ __forceinline__ void render_batch(const vrx_scene &sc, const vrx_renderer &r,
    const vrx_batch &b, uint32_t sd, uint32_t cap, uint32_t *spill, uint32_t tiles_x,
    uint32_t tiles_y)
{
  extern __shared__ uint32_t s_stack[];
  const int tid = threadIdx.x, lane = tid & 63;
  Stack st;
  st.lds = s_stack + tid;
  st.stride = blockDim.x;
  st.sd = sd;
  st.sd_mask = sd - 1;
  const uint32_t gthreads = gridDim.x * blockDim.x;
  const uint32_t gtid = blockIdx.x * blockDim.x + tid;
  st.ovf = spill + 64 + gtid;
  st.ovf_stride = gthreads;
  st.ovf_cap = cap;
  st.err = spill;
  st.cnt = 0;

  vrx_frame f = {};
  f.width = b.width;
  f.height = b.height;
  f.color_format = b.d_color_f32 ? VRX_COLOR_F32 : VRX_COLOR_SRGB8;
  f.d_color_f32 = b.d_color_f32;
  f.d_color_u32 = b.d_color_u32;
  f.d_depth = b.d_depth;
  f.d_prim_id = b.d_prim_id;
  f.d_obj_id = b.d_obj_id;
  f.d_inst_id = b.d_inst_id;

  const uint32_t per_view = tiles_x * tiles_y;
  const uint32_t total = per_view * b.num_views;
  const uint32_t tw = (uint32_t)b.tile_w;
  uint32_t rays = 0, vsamples = 0;
  for (;;) {
    uint32_t t = 0;
    if (lane == 0)
      t = atomicAdd(b.d_work_counter, 1u);
    t = __shfl(t, 0, 64);
    if (t >= total)
      break;
    const uint32_t view = t / per_view;
    const uint32_t rem = t - view * per_view;
    const uint32_t ty = rem / tiles_x, tx = rem - ty * tiles_x;
    const uint32_t px = tx * tw + (uint32_t)lane % tw;
    const uint32_t py = ty * (uint32_t)b.tile_h + (uint32_t)lane / tw;
    if (px < b.width && py < b.height) {
      const vrx_camera cam = load_camera(b.d_cams, view);
      const size_t idx = (size_t)view * b.width * b.height + (size_t)py * b.width + px;
      render_pixel<PT, TWO, KIND, WIDE>(sc, cam, r, f, px, py, idx, st, rays, vsamples);
    }
  }
}

template <int PT, bool TWO, bool WIDE, int KIND>
__global__ __launch_bounds__(kBatchMaxBlock) void k_render_batch(vrx_scene sc, vrx_renderer r,
    vrx_batch b, uint32_t sd, uint32_t cap, uint32_t *spill, uint32_t tiles_x, uint32_t tiles_y)
{
  render_batch<PT, TWO, WIDE, KIND>(sc, r, b, sd, cap, spill, tiles_x, tiles_y);
}

// The raycast program on one primitive type, held to 96 VGPRs (5 waves per
// SIMD): the traversal is latency-bound on L2 misses (every view reads most
// of the scene), and on the 64-view scene the fifth wave outweighs the few
// registers it spills (14.8 -> 13.6 ms).  AO and directLight would spill
// 160-320 B per lane and keep the unconstrained kernel.
template <int PT, bool TWO, bool WIDE>
__global__ __launch_bounds__(kBatchMaxBlock) __attribute__((amdgpu_waves_per_eu(5, 8))) void
k_render_batch_raycast(vrx_scene sc, vrx_renderer r, vrx_batch b, uint32_t sd, uint32_t cap,
    uint32_t *spill, uint32_t tiles_x, uint32_t tiles_y)
{
  render_batch<PT, TWO, WIDE, VRX_RENDER_RAYCAST>(sc, r, b, sd, cap, spill, tiles_x, tiles_y);
}

static int num_cus()
{
  return device_cus();
}

static uint32_t pick_sd(int want)
{
  uint32_t p = 8;
  while (p < (uint32_t)(want > 0 ? want : 32) && p < 64)
    p <<= 1;
  return p;
}

static uint32_t cap_for(const vrx_scene &sc, uint32_t sd, bool wide)
{
  // as the per-view renderer: one push per BVH2 level, up to three per
  // 4-wide level (which spans two BVH2 levels)
  uint32_t m = wide ? 3 * ((sc.max_blas_height + 1) / 2) + 2 : sc.max_blas_height + 2;
  if (sc.num_entries > 1)
    m += wide ? 3 * ((sc.tlas_height + 1) / 2) + 1 : sc.tlas_height + 1;
  return m > sd ? m - sd : 0u;
}

template <int PT, bool WIDE, int KIND>
static const void *kernel_ptr_k(bool two)
{
  if constexpr (KIND == VRX_RENDER_RAYCAST && PT >= 0)
    return two ? (const void *)&k_render_batch_raycast<PT, true, WIDE>
               : (const void *)&k_render_batch_raycast<PT, false, WIDE>;
  else
    return two ? (const void *)&k_render_batch<PT, true, WIDE, KIND>
               : (const void *)&k_render_batch<PT, false, WIDE, KIND>;
}

template <int PT, bool WIDE>
static const void *kernel_ptr(bool two, int kind)
{
  if (kind == VRX_RENDER_AO)
    return kernel_ptr_k<PT, WIDE, VRX_RENDER_AO>(two);
  if (kind == VRX_RENDER_DIRECT)
    return kernel_ptr_k<PT, WIDE, VRX_RENDER_DIRECT>(two);
  return kernel_ptr_k<PT, WIDE, VRX_RENDER_RAYCAST>(two);
}

template <bool WIDE>
static const void *pick_kernel_w(const vrx_scene &sc, int kind)
{
  const bool two = sc.num_entries > 1;
  switch (sc.num_volumes ? -1 : sc.single_type) { // volumes: generic kernel
  case VRX_TRIANGLE:
    return kernel_ptr<VRX_TRIANGLE, WIDE>(two, kind);
  case VRX_SPHERE:
    return kernel_ptr<VRX_SPHERE, WIDE>(two, kind);
  case VRX_CYLINDER:
    return kernel_ptr<VRX_CYLINDER, WIDE>(two, kind);
  default:
    return kernel_ptr<-1, WIDE>(two, kind);
  }
}

static const void *pick_kernel(const vrx_scene &sc, bool wide, int kind)
{
  return wide ? pick_kernel_w<true>(sc, kind) : pick_kernel_w<false>(sc, kind);
}

__global__ void k_reset_counter(uint32_t *c)
{
  if (threadIdx.x == 0)
    *c = 0u;
}

// Queue reset + persistent launch.  `spill` is the stack-spill area (error
// word, then `cap` entries per resident lane from word 64 on) and `counter`
// the tile queue; with spill_limit > 0 the grid is shrunk so that the spill
// area of every resident lane fits in spill_limit bytes.
static int launch_batch(const vrx_scene &sc, const vrx_renderer &r, const vrx_batch &b, hipStream_t s,
    uint32_t *spill, uint32_t *counter, size_t spill_limit = 0)
{
  const int block = std::min(std::max(b.block_threads > 0 ? b.block_threads : 256, 64), 256) & ~63;
  const bool wide = r.node_width == 4;
  const uint32_t sd = pick_sd(r.stack_lds);
  const uint32_t cap = cap_for(sc, sd, wide);
  const size_t lds = (size_t)sd * block * 4;
  const void *kern = pick_kernel(sc, wide, r.kind);
  int per_cu = 0;
  VRX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, lds));
  if (per_cu < 1)
    per_cu = 1;
  uint32_t blocks = b.persistent_blocks > 0 ? (uint32_t)b.persistent_blocks
                                            : (uint32_t)(num_cus() * per_cu);
  const uint32_t max_blocks = (uint32_t)num_cus() * kMaxThreadsPerCU / (uint32_t)block;
  blocks = std::min(blocks, max_blocks);
  if (spill_limit) {
    const size_t per_block = (size_t)cap * 4 * block;
    const size_t fit = per_block ? (spill_limit - 256) / per_block : (size_t)blocks;
    if (fit < 1) {
      set_error("stack spill of one workgroup (%zu B) exceeds the static scratch (%zu B)",
          per_block + 256, spill_limit);
      return VRX_ERR_ARG;
    }
    blocks = (uint32_t)std::min<size_t>(blocks, fit);
  }
  const uint32_t tw = (uint32_t)b.tile_w, th = (uint32_t)b.tile_h;
  const uint32_t tiles_x = (b.width + tw - 1) / tw, tiles_y = (b.height + th - 1) / th;
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, s, counter);
  vrx_batch bq = b;
  bq.d_work_counter = counter;
  void *args[] = {(void *)&sc, (void *)&r, (void *)&bq, (void *)&sd, (void *)&cap,
      (void *)&spill, (void *)&tiles_x, (void *)&tiles_y};
  VRX_HIP(hipLaunchKernel(kern, dim3(blocks), dim3(block), args, lds, s));
  return VRX_OK;
}

static int launch_batch(const vrx_scene &sc, const vrx_renderer &r, const vrx_batch &b, hipStream_t s)
{
  return launch_batch(sc, r, b, s, (uint32_t *)b.d_scratch, b.d_work_counter);
}

static int validate_common(const vrx_scene *scene, const vrx_renderer *r, const vrx_batch *b)
{
  VRX_ARG(scene && r && b, "null argument");
  VRX_ARG(b->num_views > 0 && b->width > 0 && b->height > 0, "batch dimensions");
  VRX_ARG(b->d_cams, "cameras");
  VRX_ARG(b->tile_w > 0 && b->tile_h > 0 && b->tile_w * b->tile_h == 64, "tile must hold 64 pixels");
  VRX_ARG(scene->num_entries == 0 || (scene->d_entries && scene->d_geoms), "scene arrays");
  VRX_ARG(scene->num_entries <= 1 || scene->tlas.d_leaf_ids, "instance BVH not built");
  VRX_ARG(scene->num_volumes == 0 || scene->d_volumes, "volume array");
  VRX_ARG(r->kind >= VRX_RENDER_RAYCAST && r->kind <= VRX_RENDER_DIRECT, "renderer kind");
  VRX_ARG(scene->num_lights == 0 || scene->d_lights, "light array");
  VRX_ARG(r->node_width == 0 || r->node_width == 2 || r->node_width == 4, "node width 2 or 4");
  VRX_ARG(r->node_width != 4 || scene->num_entries == 0 || scene->wide,
      "node width 4 needs 4-wide nodes for every BVH (scene->wide)");
  return VRX_OK;
}

static int validate(const vrx_scene *scene, const vrx_renderer *r, const vrx_batch *b)
{
  int st = validate_common(scene, r, b);
  if (st)
    return st;
  VRX_ARG(b->d_work_counter, "work counter");
  size_t need = vrx_batch_scratch_bytes(scene, r, b);
  if (!b->d_scratch || b->scratch_bytes < need) {
    set_error("batch scratch too small: %zu < %zu", b->scratch_bytes, need);
    return VRX_ERR_ARG;
  }
  return VRX_OK;
}

// ---- the allocation-free variant -------------------------------------------
// One array per GPU, in device memory the code object reserves at load time:
// 256 B for the tile-queue counter, then the stack-spill area.  In host code
// `scratch` names a host-side stand-in, never a device address: kernels get
// the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) uint8_t scratch[VRX_BATCH_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static void *g_scratch_addr[kMaxDevices];
static std::mutex g_scratch_mu;

static int static_scratch(uint8_t **out)
{
  int dev = 0;
  VRX_HIP(hipGetDevice(&dev));
  VRX_ARG(dev >= 0 && dev < kMaxDevices, "device index");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void *p = nullptr;
    VRX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    VRX_ARG(p, "static scratch address");
    g_scratch_addr[dev] = p;
  }
  *out = (uint8_t *)g_scratch_addr[dev];
  return VRX_OK;
}

} // namespace vrx

using namespace vrx;

extern "C" {

size_t vrx_batch_scratch_bytes(const vrx_scene *scene, const vrx_renderer *r, const vrx_batch *b)
{
  (void)b;
  if (!scene)
    return 0;
  // sized for the smallest LDS stack the autotuner tries (16), the deeper
  // stack of 4-wide traversal when the scene has 4-wide nodes, and a grid of
  // at most 2048 resident lanes per CU
  const uint32_t sd = std::min(pick_sd(r ? r->stack_lds : 0), 16u);
  const bool wide = scene->wide || (r && r->node_width == 4);
  const uint32_t cap = std::max(cap_for(*scene, sd, false), wide ? cap_for(*scene, sd, true) : 0u);
  return 256 + (size_t)num_cus() * kMaxThreadsPerCU * cap * 4;
}

int vrx_render_batch(const vrx_scene *scene, const vrx_renderer *r, const vrx_batch *b, vrx_stream s)
{
  int st = validate(scene, r, b);
  if (st)
    return st;
  st = launch_batch(*scene, *r, *b, (hipStream_t)s);
  if (st)
    return st;
  VRX_LAUNCH_CHECK();
  return VRX_OK;
}

int vrx_render_batch_static(const vrx_scene *scene, const vrx_renderer *r, const vrx_batch *b,
    vrx_stream s)
{
  int st = validate_common(scene, r, b);
  if (st)
    return st;
  uint8_t *base = nullptr;
  if ((st = static_scratch(&base)))
    return st;
  // the queue counter is reset by a kernel on s; the spill area needs no
  // clearing (entries are written before they are read, the error word is
  // only ever set)
  st = launch_batch(*scene, *r, *b, (hipStream_t)s, (uint32_t *)(base + 256), (uint32_t *)base,
      VRX_BATCH_STATIC_SCRATCH_BYTES - 256);
  if (st)
    return st;
  VRX_LAUNCH_CHECK();
  return VRX_OK;
}

int vrx_batch_graph_create(const vrx_scene *scene, const vrx_renderer *r, const vrx_batch *b,
    vrx_stream s, vrx_graph *out)
{
  VRX_ARG(out && s, "a non-default stream and an output handle are required");
  int st = validate(scene, r, b);
  if (st)
    return st;
  hipStream_t hs = (hipStream_t)s;
  hipGraph_t g = nullptr;
  VRX_HIP(hipStreamBeginCapture(hs, hipStreamCaptureModeThreadLocal));
  int lst = launch_batch(*scene, *r, *b, hs);
  hipError_t e = hipStreamEndCapture(hs, &g);
  if (lst)
    return lst;
  if (e != hipSuccess)
    return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
  hipGraphExec_t ge = nullptr;
  e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess)
    return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
  *out = (vrx_graph)ge;
  return VRX_OK;
}

int vrx_graph_launch(vrx_graph g, vrx_stream s)
{
  VRX_ARG(g, "graph");
  VRX_HIP(hipGraphLaunch((hipGraphExec_t)g, (hipStream_t)s));
  return VRX_OK;
}

int vrx_graph_destroy(vrx_graph g)
{
  if (g)
    VRX_HIP(hipGraphExecDestroy((hipGraphExec_t)g));
  return VRX_OK;
}

int vrx_batch_autotune(const vrx_scene *scene, vrx_renderer *r, vrx_batch *b, vrx_stream s,
    float *best_ms_out)
{
  int st = validate(scene, r, b);
  if (st)
    return st;
  hipStream_t hs = (hipStream_t)s;
  struct Cand {
    int tw, th, block, sd, width;
  };
  std::vector<Cand> cands;
  const int tiles[][2] = {{8, 8}, {16, 4}, {32, 2}, {64, 1}};
  const bool can_wide = scene->wide || scene->num_entries == 0;
  for (int width : {2, 4}) {
    if (width == 4 && !can_wide)
      continue;
    for (auto &t : tiles)
      for (int block : {64, 128, 256})
        for (int sd : {16, 32})
          cands.push_back({t[0], t[1], block, sd, width});
  }
  hipEvent_t e0, e1;
  VRX_HIP(hipEventCreate(&e0));
  VRX_HIP(hipEventCreate(&e1));
  float best = 3.4e38f;
  Cand bestc = cands[0];
  vrx_renderer rr = *r;
  vrx_batch bb = *b;
  bb.persistent_blocks = 0;
  for (const Cand &c : cands) {
    rr.stack_lds = c.sd;
    rr.node_width = c.width;
    bb.tile_w = c.tw;
    bb.tile_h = c.th;
    bb.block_threads = c.block;
    if ((st = launch_batch(*scene, rr, bb, hs)))
      return st; // warm-up
    VRX_HIP(hipEventRecord(e0, hs));
    for (int rep = 0; rep < 3; rep++)
      if ((st = launch_batch(*scene, rr, bb, hs)))
        return st;
    VRX_HIP(hipEventRecord(e1, hs));
    VRX_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    VRX_HIP(hipEventElapsedTime(&ms, e0, e1));
    ms /= 3.0f;
    if (ms < best) {
      best = ms;
      bestc = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  r->stack_lds = bestc.sd;
  r->node_width = bestc.width;
  b->tile_w = bestc.tw;
  b->tile_h = bestc.th;
  b->block_threads = bestc.block;
  b->persistent_blocks = 0;
  if (best_ms_out)
    *best_ms_out = best;
  return VRX_OK;
}

} // extern "C"
