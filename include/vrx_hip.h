/*
 * vrx_hip.h -- thin C ABI over the gfx950 HIP kernels of the visrtx_amd device.
 *
 * Plain pointers, sizes and opaque stream handles only; every call returns an
 * int status (0 = success, <0 = error, message via vrx_last_error()).  All
 * pointers named d_* are device (HBM) pointers owned by the caller.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * VisRTX source tree):
 *   vrx_bvh_build        optix_visrtx.cpp:37-121 buildOptixBVH (OptiX accel build;
 *                        called from scene/Group.cpp:189-266, scene/World.cpp:202-254)
 *   vrx_tlas_*           scene/World.cpp:256-341 populateOptixInstances + TLAS build
 *   vrx_render           frame/Frame.cu:282-289 optixLaunch of the raygen programs
 *                        renderer/Raycast_ptx.cu:60-179 and
 *                        renderer/AmbientOcclusion_ptx.cu:86-220
 *   vrx_frame_clear      frame/Frame.cu:590-660 Frame::newFrame thrust::fill_n clears
 *   vrx_stream_* / vrx_event_* / vrx_launch_host_func
 *                        the cudaStream / cudaEvent / cudaLaunchHostFunc use in
 *                        frame/Frame.cu:49-53,255,295-307 and optix_visrtx.h:165
 *   vrx_diffvol_*        no reference counterpart (the reference has only the
 *                        forward integral, gpu/volumeIntegration.h:64-165); the
 *                        adjoint is new, for torch.autograd
 *   vrx_batch_*          no reference counterpart (one optixLaunch per view);
 *                        multi-view persistent launch
 */
#ifndef VRX_HIP_H
#define VRX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *vrx_stream; /* hipStream_t; NULL = default stream */
typedef void *vrx_event;  /* hipEvent_t */
typedef void *vrx_graph;  /* hipGraphExec_t */

enum vrx_status {
  VRX_OK = 0,
  VRX_ERR_HIP = -1,
  VRX_ERR_ARG = -2,
  VRX_ERR_NOMEM = -3,
  VRX_ERR_LIMIT = -4,
};

enum vrx_prim_type {
  VRX_TRIANGLE = 0,
  VRX_SPHERE = 1,
  VRX_CYLINDER = 2,
  VRX_BOX = 3, /* plain AABB input (used for the instance level) */
  VRX_CONE = 4, /* capped cones between vertex pairs, per-vertex radii */
  VRX_CURVE = 5 /* round linear curve segments: vertex index[i] to index[i]+1 */
};

enum vrx_memcpy_kind { VRX_H2D = 0, VRX_D2H = 1, VRX_D2D = 2 };

/* ---- runtime plumbing ---------------------------------------------------- */
const char *vrx_last_error(void);
int vrx_version(void);
int vrx_device_count(int *count);
int vrx_set_device(int device);
int vrx_get_device(int *device);
int vrx_device_info(int device, char *name, size_t name_len, int *num_cus, size_t *hbm_bytes);
int vrx_malloc(void **d_ptr, size_t bytes);
int vrx_free(void *d_ptr);
int vrx_host_malloc(void **h_ptr, size_t bytes);
int vrx_host_free(void *h_ptr);
/* The device-side address of pinned host memory from vrx_host_malloc. */
int vrx_host_device_pointer(void *h_ptr, void **d_ptr);
/* Where an application pointer lives (array/Array.cpp:38-65
 * getPointerAddressSpace): host (pageable or pinned), device memory of the
 * current device, managed memory, or device memory of another device. */
enum vrx_pointer_kind { VRX_PTR_HOST = 0, VRX_PTR_DEVICE = 1, VRX_PTR_MANAGED = 2,
  VRX_PTR_FOREIGN = 3 };
int vrx_pointer_kind(const void *p, int *kind);
int vrx_memcpy(void *dst, const void *src, size_t bytes, int kind, vrx_stream s);
int vrx_memcpy_sync(void *dst, const void *src, size_t bytes, int kind);
int vrx_fill_u32(void *d_ptr, uint32_t value, size_t count, vrx_stream s);
/* plain streaming copy kernel (bytes % 16 == 0): the HBM bandwidth probe and
   the calibration kernel for the HBM byte counters */
int vrx_copy_kernel(void *d_dst, const void *d_src, size_t bytes, vrx_stream s);
int vrx_stream_create(vrx_stream *s);
int vrx_stream_destroy(vrx_stream s);
int vrx_stream_sync(vrx_stream s);
int vrx_device_sync(void);
int vrx_event_create(vrx_event *e);
int vrx_event_destroy(vrx_event e);
int vrx_event_record(vrx_event e, vrx_stream s);
int vrx_event_sync(vrx_event e);
int vrx_event_query(vrx_event e); /* 1 = complete, 0 = pending, <0 error */
int vrx_event_elapsed_ms(vrx_event start, vrx_event end, float *ms);
int vrx_launch_host_func(vrx_stream s, void (*fn)(void *), void *user);

/* ---- LBVH build ---------------------------------------------------------- */
typedef struct vrx_prims {
  int32_t type;           /* vrx_prim_type */
  uint32_t n;             /* number of primitives */
  const float *d_position;/* float3 vertices / centers (VRX_BOX: float[6] boxes) */
  const uint32_t *d_index;/* uint3 / uint2 / uint primitive.index or NULL (curve: required,
                             one first-vertex index per segment) */
  const float *d_radius;  /* sphere, cone, curve: per-vertex radius; cylinder: per-primitive;
                             or NULL */
  float global_radius;
} vrx_prims;

/* Output buffers of one BVH.  Sizes (bytes) from vrx_bvh_sizes. */
typedef struct vrx_bvh {
  void *d_nodes;          /* (n-1) * 64 B: both child boxes + children + heights */
  void *d_leaves;         /* n * leaf record (48/16/32/0/32/32 B: tri/sphere/cyl/box/cone/curve) */
  uint32_t *d_leaf_ids;   /* n: sorted leaf -> original primitive index */
  void *d_header;         /* 64 B: root box (6 f32), height, root, n, type */
  void *d_wide;           /* optional (n-1) * 128 B 4-wide nodes (vrx_bvh_widen), or NULL */
} vrx_bvh;

typedef struct vrx_bvh_sizes {
  size_t nodes, leaves, leaf_ids, header, workspace;
} vrx_bvh_sizes;

int vrx_bvh_query_sizes(int32_t type, uint32_t n, vrx_bvh_sizes *out);
/* Morton -> LDS-staged stable radix sort -> Karras -> refit by common-prefix
 * level (no atomics). */
int vrx_bvh_build(const vrx_prims *prims, const vrx_bvh *out, void *d_workspace,
    size_t workspace_bytes, vrx_stream s);
/* 4-wide nodes from a built BVH2 (no reference counterpart: OptiX picks its
 * own node format).  Record i holds the up-to-4 grandchildren of BVH2 node i
 * (a leaf child stays one slot), so the 4-wide tree shares the BVH2's leaves
 * and root.  Writes bvh->d_wide, which must hold vrx_bvh_wide_bytes(n). */
size_t vrx_bvh_wide_bytes(uint32_t n);
int vrx_bvh_widen(const vrx_bvh *bvh, uint32_t n, vrx_stream s);
/* Debug/parity hooks: read back the sorted Morton codes (n) and the order. */
int vrx_bvh_build_debug(const vrx_prims *prims, const vrx_bvh *out, void *d_workspace,
    size_t workspace_bytes, uint32_t *d_codes_out, float *d_cbox_out, vrx_stream s);

/* header layout (host view, 64 B) */
typedef struct vrx_bvh_header {
  float lo[3], hi[3];
  uint32_t height;        /* max leaf depth + 1 (a single leaf has height 1) */
  uint32_t root;          /* 0 or 0x80000000 (single leaf) */
  uint32_t n;
  int32_t type;
  uint32_t pad[6];
} vrx_bvh_header;

/* ---- scene description for rendering -------------------------------------- */
typedef struct vrx_geom {
  int32_t type;
  uint32_t n;
  vrx_bvh bvh;
  /* original attribute arrays (device) for hit population */
  const float *d_position;
  const uint32_t *d_index;
  const float *d_radius;
  float global_radius;
  const uint32_t *d_primitive_id;
  const float *d_vertex_normal;
  int32_t caps;
  int32_t cull_backfaces;
  /* matte material (VALUE path) */
  float color[3];
  float opacity;
  uint32_t surface_id;
  uint32_t pad;
} vrx_geom;

typedef struct vrx_entry {
  float xfm[12];          /* object -> world, column-major 3x4 */
  float inv[12];          /* world -> object */
  uint32_t identity;
  uint32_t geom;
  uint32_t inst_id;
  uint32_t pad;
} vrx_entry;

typedef struct vrx_volume {
  const float *d_data;    /* x fastest */
  uint32_t dims[3];
  float origin[3];
  float spacing[3];
  const float *d_lut;     /* 256 x RGBA (alpha already scaled by opacity) */
  float value_range[2];
  float unit_distance;
  uint32_t id;
  uint32_t inst_id;
  uint32_t pad;
  /* NULL, or the macro-cell majorant grid from vrx_volume_majorant_grid: the
   * march skips the samples inside cells whose majorant is 0 (they add
   * nothing), when the renderer's skip_empty is set */
  const float *d_majorant;
  /* NULL, or the field as y-pairs from vrx_volume_build_ypairs: per voxel the
   * float2 (v(x,y,z), v(x,min(y+1,ny-1),z)), x fastest, 8 bytes per voxel.
   * The dense volume-only march then reads a cell's four corners of one z
   * plane with one 16-byte gather (fields of 2 to 2^24-1 voxels along x,
   * under 2^29 voxels and under 2^24 rows); same values, same frame */
  const float *d_ypairs;
} vrx_volume;

/* One light in world space, as the directLight renderer samples it
 * (gpu/sampleLight.h:58-100; scene/light/{Directional,Point,Spot}.cpp).
 * The host applies the owning instance's transform:
 *   directional: dir = xfm (linear part) * -direction (direction normalised
 *                at commit), the direction towards the light, not
 *                renormalised; radiance = color * irradiance
 *   point:       pos = xfm * position; radiance = color * intensity
 *   spot:        pos = xfm * position; dir = direction as given (the
 *                reference does not transform it); radiance = color *
 *                intensity; cos_outer = cos(openingAngle), cos_inner =
 *                cos(openingAngle - 2 falloffAngle) */
enum vrx_light_type { VRX_LIGHT_DIRECTIONAL = 0, VRX_LIGHT_POINT = 1, VRX_LIGHT_SPOT = 2 };

typedef struct vrx_light {
  int32_t type;
  float radiance[3];
  float pos[3];
  float dir[3];
  float cos_outer, cos_inner;
} vrx_light;

typedef struct vrx_scene {
  const vrx_geom *d_geoms;      /* device array */
  uint32_t num_geoms;
  const vrx_entry *d_entries;   /* device array */
  uint32_t num_entries;
  vrx_bvh tlas;                 /* over entries; unused when num_entries == 1 */
  uint32_t tlas_height;
  uint32_t max_blas_height;
  const vrx_volume *d_volumes;  /* device array */
  uint32_t num_volumes;
  int32_t single_type;          /* prim type if every geom shares it, else -1 */
  uint64_t max_volume_voxels;   /* largest field's voxel count (>= 2^31 selects 64-bit gathers) */
  uint32_t wide;                /* 1 = every BLAS (and the TLAS if any) has d_wide */
  uint32_t num_lights;
  const vrx_light *d_lights;    /* device array (directLight renderer) */
  uint32_t num_majorant_grids;  /* volumes that carry d_majorant */
  uint32_t pad;
} vrx_scene;

typedef struct vrx_camera {
  float pos[3];
  float d00[3];
  float du[3];
  float dv[3];
  float region[4];
} vrx_camera;

enum vrx_renderer_kind { VRX_RENDER_RAYCAST = 0, VRX_RENDER_AO = 1, VRX_RENDER_DIRECT = 2 };
enum vrx_color_format { VRX_COLOR_F32 = 0, VRX_COLOR_SRGB8 = 1, VRX_COLOR_UNORM8 = 2 };

typedef struct vrx_renderer {
  int32_t kind;
  float background[4];
  float ambient_color[3];
  float ambient_radiance;
  int32_t pixel_samples;
  int32_t ambient_samples;
  float ao_distance;
  int32_t cull_backfaces;
  float inv_volume_sampling_rate;
  /* launch tuning (autotuned; 0 = defaults) */
  int32_t stack_lds;      /* LDS short-stack entries per lane: 8/16/32/64 */
  int32_t vol_unroll;     /* volume samples gathered per step: 1/2/4/8 */
  int32_t node_width;     /* BVH node width traversed: 2 (default) or 4 (needs scene->wide) */
  int32_t skip_empty;     /* 1: skip empty macro cells of volumes that carry d_majorant */
  int32_t volume_jitter;  /* 1: each volume segment starts at a random offset, the two
                             jitters of volumeIntegration.h:83,120 (PCG32 per pixel and
                             frame, as the AO directions); 0: the first sample on the
                             segment entry (deterministic) */
  int32_t defer_steps;    /* raycast of a one-level scene of one primitive type: a primary
                             walk still running after this many steps is finished by a
                             second launch with 16 lanes per ray over t-segments (same
                             pixels); 0 = off.  The frame scratch then also holds one u32
                             per pixel (vrx_render_scratch_bytes) */
} vrx_renderer;

typedef struct vrx_frame {
  uint32_t width, height;
  int32_t color_format;
  uint32_t frame_id;      /* 0 = first frame after a reset */
  float *d_accum;         /* width*height*4 */
  float *d_color_f32;     /* or NULL */
  uint32_t *d_color_u32;  /* or NULL */
  float *d_depth;         /* or NULL */
  uint32_t *d_prim_id;
  uint32_t *d_obj_id;
  uint32_t *d_inst_id;
  void *d_scratch;        /* traversal stack overflow; size from vrx_render_scratch_bytes */
  size_t scratch_bytes;
  uint64_t *d_counters;   /* optional: VRX_COUNTER_SLOTS slots of VRX_COUNTER_STRIDE u64;
                             slot s holds [s*VRX_COUNTER_STRIDE + 0] rays traced and
                             [... + 1] volume samples, added by atomics (the caller zeroes
                             them); the frame's totals are the sums over the slots */
} vrx_frame;

/* Frame counters are spread over slots (a workgroup adds into slot
   workgroup-index mod VRX_COUNTER_SLOTS, one 128-byte line per slot) so that
   the per-wave atomics of a launch do not all meet at one address. */
#define VRX_COUNTER_SLOTS 64
#define VRX_COUNTER_STRIDE 16
#define VRX_COUNTER_BYTES (VRX_COUNTER_SLOTS * VRX_COUNTER_STRIDE * 8)

/* After a frame: writes out[0] = rays traced and out[1] = volume samples
   (the sums over f->d_counters' slots) and out[2] = the first 32-bit word of
   f->d_scratch (the traversal stack-overflow flag, 0 when clean), then zeroes
   the counter slots for the next frame, in one small kernel on stream s.
   out must be device-accessible: device memory or pinned host memory from
   vrx_host_malloc (read it after the stream has passed this call).  No
   reference counterpart (Frame.cu reads its counters with copies); it
   replaces a clear and two read-back copies per frame. */
int vrx_frame_collect_counters(const vrx_frame *f, uint64_t *out, vrx_stream s);

size_t vrx_render_scratch_bytes(const vrx_scene *scene, const vrx_renderer *r,
    uint32_t width, uint32_t height);
/* newFrame clears (accumulation, depth = FLT_MAX, ids = 0). */
int vrx_frame_clear(const vrx_frame *f, vrx_stream s);
/* One frame (all pixel samples) of the raycast, AO or directLight renderer
 * (Raycast_ptx.cu:60-179, AmbientOcclusion_ptx.cu:86-220,
 * DirectLight_ptx.cu:228-420).  directLight reads scene->d_lights. */
int vrx_render(const vrx_scene *scene, const vrx_camera *cam, const vrx_renderer *r,
    const vrx_frame *f, vrx_stream s);
/* Macro-cell majorant grid of a volume (scene/volume/space_skipping/
 * UniformGrid.cu:44-258: cells of 16^3 voxels, ceil(dims/16) per axis, the
 * largest transfer-function opacity over each cell's value range).  This
 * grid is conservative, where the reference's is sampled: cell c's value
 * range covers every voxel a trilinear sample inside the cell (or within one
 * voxel of it) can read, and its majorant is the largest LUT alpha of the
 * entries those values interpolate between, so a 0 majorant means no sample
 * in the cell adds colour or opacity.  d_majorant holds
 * vrx_volume_grid_cells(v) floats, x fastest; v->d_data and v->d_lut are read. */
size_t vrx_volume_grid_cells(const vrx_volume *v);
int vrx_volume_majorant_grid(const vrx_volume *v, float *d_majorant, vrx_stream s);
/* Writes v's field as y-pairs (see vrx_volume.d_ypairs) to d_ypairs, which
 * holds dims[0]*dims[1]*dims[2] float2.  No reference counterpart (the
 * reference samples a CUDA 3D texture, whose layout the driver chooses); the
 * layout halves the dense march's gather instructions.  Fields of 2^29
 * voxels or more are refused (their offsets would not fit 32 bits). */
int vrx_volume_build_ypairs(const vrx_volume *v, float *d_ypairs, vrx_stream s);
/* Woodcock (delta) tracking of n rays through v over the majorant grid
 * v->d_majorant (volumeIntegration.h:167-228 _sampleDistance, dda.h:43-122):
 * per ray (origin, direction, [tmin, tmax]) the collision distance d_t (tmax
 * if none), d_tr = 0 on a collision else 1, d_albedo (3 per ray) the LUT
 * colour at the last sample taken.  Random numbers: PCG32 keyed (ray index,
 * seed).  Step = min(spacing)/2, extinction = LUT alpha per step.  Fields
 * below 2^31 voxels with dims[0] >= 2. */
int vrx_volume_sample_distance(const vrx_volume *v, uint32_t n, const float *d_org,
    const float *d_dir, const float *d_tminmax, uint64_t seed, float *d_t, float *d_albedo,
    float *d_tr, vrx_stream s);
/* Host-side camera setup (camera/Perspective.cpp:40-72). */
int vrx_camera_perspective(const float pos[3], const float dir[3], const float up[3],
    float fovy, float aspect, const float region[4], vrx_camera *out);
/* Instance-level build: world boxes of entries + LBVH over them. */
int vrx_tlas_query_sizes(uint32_t num_entries, vrx_bvh_sizes *out);
int vrx_tlas_build(const vrx_scene *scene, const vrx_bvh *out, void *d_workspace,
    size_t workspace_bytes, vrx_stream s);

/* ---- differentiable volume (no reference counterpart) --------------------- */
typedef struct vrx_diffvol {
  const float *d_density;  /* dims product */
  uint32_t dims[3];
  float origin[3];
  float spacing[3];
  const float *d_lut;      /* 256 x 4 */
  float value_range[2];
  float unit_distance;
  float step;              /* world-space step */
  int32_t early_stop;      /* 1 = stop at opacity 0.99 (as the reference) */
  int32_t checkpoint;      /* steps between transmittance checkpoints (64) */
  uint32_t max_steps;      /* per ray upper bound (for checkpoint storage) */
  /* optional layout hint: rays are images of image_w x image_h, row-major,
   * one after another (as vrx_make_rays writes them); the adjoint then gives
   * each workgroup a 16x16 pixel tile so neighbouring rays share gradient
   * scatter.  0 = rays in no particular order. */
  uint32_t image_w, image_h;
  int32_t adjoint;         /* vrx_adjoint: backward schedule */
} vrx_diffvol;

/* Backward schedules of vrx_diffvol_backward (same gradients):
 *  SUFFIX  the march is replayed forward; the suffix sum that the adjoint
 *          needs is the forward's FP64 total minus an FP64 prefix (default);
 *  REPLAY  the march is replayed backward from the transmittance stored
 *          every `checkpoint` steps. */
enum vrx_adjoint { VRX_ADJOINT_SUFFIX = 0, VRX_ADJOINT_REPLAY = 1 };

/* rays: n origins/directions (float3) */
int vrx_diffvol_forward(const vrx_diffvol *v, uint32_t n, const float *d_org, const float *d_dir,
    float *d_rgba_out, float *d_ckpt, vrx_stream s);
int vrx_diffvol_backward(const vrx_diffvol *v, uint32_t n, const float *d_org, const float *d_dir,
    const float *d_rgba, const float *d_grad_rgba, const float *d_ckpt,
    float *d_grad_density, float *d_grad_lut, vrx_stream s);
size_t vrx_diffvol_ckpt_floats(const vrx_diffvol *v, uint32_t n);
/* pinhole rays for a set of views: cams[nviews], w x h each (row 0 = bottom) */
int vrx_make_rays(const vrx_camera *cams, uint32_t nviews, uint32_t w, uint32_t h,
    float *d_org, float *d_dir, vrx_stream s);

/* ---- batched multi-view (no reference counterpart) ------------------------ */
typedef struct vrx_batch {
  uint32_t num_views;
  uint32_t width, height;
  const vrx_camera *d_cams;  /* device array of num_views cameras */
  float *d_depth;            /* views*w*h or NULL */
  uint32_t *d_prim_id;       /* views*w*h or NULL */
  uint32_t *d_obj_id;
  uint32_t *d_inst_id;
  float *d_color_f32;        /* views*w*h*4 or NULL */
  uint32_t *d_color_u32;     /* views*w*h or NULL (SRGB8) */
  void *d_scratch;
  size_t scratch_bytes;
  uint32_t *d_work_counter;  /* 1 u32, persistent-launch tile queue */
  int32_t tile_w, tile_h;    /* pixels per tile, tile_w * tile_h == 64 (one wave) */
  int32_t block_threads;     /* 64, 128 or 256 */
  int32_t persistent_blocks; /* 0 = fill the device at the kernel's occupancy */
} vrx_batch;

/* Renders num_views views of one scene (the raycast, AO or directLight
 * program: color, depth and id channels; each view's random numbers are those
 * of the same view rendered alone) in ONE persistent launch: every wave pulls 64-pixel tiles of
 * any view from the d_work_counter queue.  The counter is reset by a kernel
 * on stream s, so the call is graph-capturable.  Output index of pixel
 * (x, y) of view v is v*w*h + y*w + x, row 0 at the bottom. */
size_t vrx_batch_scratch_bytes(const vrx_scene *scene, const vrx_renderer *r, const vrx_batch *b);
int vrx_render_batch(const vrx_scene *scene, const vrx_renderer *r, const vrx_batch *b,
    vrx_stream s);
/* No counterpart in the reference: vrx_render_batch for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  b->d_scratch, b->scratch_bytes and
 * b->d_work_counter are ignored: the tile-queue counter and the stack-spill
 * area live in a file-scope device array of VRX_BATCH_STATIC_SCRATCH_BYTES
 * that each GPU the library runs on holds until the process ends.  The
 * persistent grid is shrunk until the spill area of every resident lane fits;
 * a scene whose spill for one workgroup does not fit is refused (VRX_ERR_ARG).
 * Nothing is allocated.  Every call on a GPU shares that GPU's array, so calls
 * that overlap on one GPU (on any streams) race: serialise them. */
#define VRX_BATCH_STATIC_SCRATCH_BYTES ((size_t)64 << 20)
int vrx_render_batch_static(const vrx_scene *scene, const vrx_renderer *r, const vrx_batch *b,
    vrx_stream s);
/* Captures vrx_render_batch on stream s (not the default stream) into an
 * instantiated HIP graph; the graph holds the structs' values at capture. */
int vrx_batch_graph_create(const vrx_scene *scene, const vrx_renderer *r, const vrx_batch *b,
    vrx_stream s, vrx_graph *out);
int vrx_graph_launch(vrx_graph g, vrx_stream s);
int vrx_graph_destroy(vrx_graph g);
/* Times tile shape, workgroup size, LDS stack depth and BVH node width (4 only
 * when scene->wide) on this scene and writes the fastest into r->stack_lds,
 * r->node_width and b->tile_w/tile_h/block_threads. */
int vrx_batch_autotune(const vrx_scene *scene, vrx_renderer *r, vrx_batch *b, vrx_stream s,
    float *best_ms_out);

/* ---- synthetic scene generators (host, OpenMP) ---------------------------- */
uint32_t vrx_gen_pcg32(uint64_t *state, uint64_t inc);
void vrx_gen_spheres(uint64_t seed, uint32_t n, float rmin, float rmax, float *centers, float *radii);
void vrx_gen_heightfield(uint64_t seed, uint32_t nx, uint32_t ny, float amplitude, float *verts,
    uint32_t *indices);
void vrx_gen_noise_volume(uint64_t seed, uint32_t nx, uint32_t ny, uint32_t nz, int octaves,
    float base_freq, float *out);
/* noise-displaced UV sphere with ~ntri triangles; returns actual counts */
void vrx_gen_sphere_mesh_counts(uint32_t target_tris, uint32_t *nverts, uint32_t *ntris);
void vrx_gen_sphere_mesh(uint64_t seed, uint32_t target_tris, float displacement, float *verts,
    uint32_t *indices);
float vrx_gen_value_noise3(uint64_t seed, float x, float y, float z);

#ifdef __cplusplus
}
#endif
#endif
