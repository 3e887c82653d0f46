/*
 * oracle.h -- CPU restatement of the VisRTX hot path, used ONLY as a checker.
 *
 * This is test infrastructure.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it, and only to check or to time a CPU
 * baseline; the product (visrtx_amd/) never links, loads or calls it.
 *
 * What it restates (reference paths relative to the VisRTX source tree):
 *   - primitive bounds        devices/rtx/scene/surface/geometry/Sphere.cu:82-110,
 *                             Cylinder.cpp:53-102, Triangle.cpp:112-138
 *   - LBVH build              replaces optix_visrtx.cpp:37-121 (OptiX build is
 *                             closed); Morton-30 + stable sort + Karras (HPG 2012)
 *                             + exact min/max refit
 *   - camera                  camera/Perspective.cpp:40-72, camera/Camera.cpp:68-76,
 *                             gpu/cameraCreateRay.h:38-81
 *   - intersectors            scene/Intersectors_ptx.cu:74-190 (sphere, cylinder),
 *                             watertight triangle (Woop/Benthin/Wald, JCGT 2013)
 *                             in place of the closed OptiX builtin
 *   - hit population          gpu/populateHit.h:150-368, gpu/gpu_util.h:245-315
 *   - raycast renderer        renderer/Raycast_ptx.cu:60-179
 *   - AO renderer             renderer/AmbientOcclusion_ptx.cu:86-220, gpu/computeAO.h:38-60
 *   - volume march (FP64)     gpu/volumeIntegration.h:46-165,317-350,
 *                             gpu/sampleSpatialField.h:54-78
 *   - TF discretisation       scene/volume/TransferFunction1D.cpp:101-150,
 *                             utility/colorMapHelpers.h:43-73
 *   - channel packing         gpu/gpu_util.h:323-443, glm func_packing.inl:67-83,
 *                             glm color_space.inl:7-55
 *
 * Parity pins: each function is pinned by tests/test_oracle_*.py against
 * closed forms, brute force, FP64 cross-checks and invariants.  Unpinned
 * parts are listed in the test files' module docstrings.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  All fp32
 * arithmetic is written with an explicit operation order and no contraction,
 * so that it matches the device code bit for bit where bit-exactness is
 * claimed (Morton codes, topology, t / primitiveId selection).
 */
#ifndef VRX_ORACLE_H
#define VRX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_TRIANGLE = 0, ORC_SPHERE = 1, ORC_CYLINDER = 2, ORC_BOXES = 3, ORC_CONE = 4,
       ORC_CURVE = 5 }; /* ORC_BOXES: pos holds 6 floats (lo, hi) per primitive */

/* ---- PCG32 (O'Neill 2014) ------------------------------------------------ */
typedef struct { uint64_t state, inc; } orc_pcg;
void orc_pcg_seed(orc_pcg *r, uint64_t initstate, uint64_t initseq);
uint32_t orc_pcg_next(orc_pcg *r);
float orc_pcg_float(orc_pcg *r);

/* ---- LBVH ----------------------------------------------------------------- */
/* boxes: n * 6 floats (lo.xyz, hi.xyz) */
void orc_bounds(int type, uint32_t n, const float *pos, const uint32_t *index,
    const float *radius, float global_radius, float *boxes);
/* codes[n]; cbox[6] = centroid bounds */
void orc_morton(const float *boxes, uint32_t n, uint32_t *codes, float *cbox);
uint32_t orc_expand_bits10(uint32_t v);
/* stable sort of (code, index): sorted_codes[n], order[n] (sorted -> original) */
void orc_sort(const uint32_t *codes, uint32_t n, uint32_t *sorted_codes, uint32_t *order);
/* Karras hierarchy. child[2*(n-1)]; leaf children are 0x80000000|k (k sorted leaf).
   parent[(n-1) + n]: parents of internal nodes then of leaves (root parent = ~0u). */
void orc_karras(const uint32_t *sorted_codes, uint32_t n, uint32_t *child, uint32_t *parent);
/* node_boxes[(n-1)*12]: per internal node, (lo0, hi0, lo1, hi1) of its children. */
void orc_refit(const float *boxes, const uint32_t *order, const uint32_t *child,
    uint32_t n, float *node_boxes, float *root_box);

typedef struct {
  uint32_t n;
  uint32_t *order;      /* n */
  uint32_t *child;      /* 2*(n-1) */
  float *node_boxes;    /* 12*(n-1) */
  float root_box[6];
  float *leaf_boxes;    /* 6*n, original order */
} orc_bvh;

int orc_bvh_build(int type, uint32_t n, const float *pos, const uint32_t *index,
    const float *radius, float global_radius, orc_bvh *out);
void orc_bvh_free(orc_bvh *b);
/* Max depth (root = 0) of any leaf. */
uint32_t orc_bvh_depth(const orc_bvh *b);

/* ---- scene ---------------------------------------------------------------- */
typedef struct {
  int32_t type;
  uint32_t num_prims;
  const float *position;        /* vertex.position float3 */
  const uint32_t *index;        /* primitive.index (uint3 / uint2 / uint) or NULL; curves:
                                   first vertex of each segment (required) */
  const float *radius;          /* sphere, cone vertex.radius / cylinder primitive.radius or NULL */
  float global_radius;
  int32_t caps;
  const uint32_t *primitive_id; /* or NULL */
  const float *vertex_normal;   /* triangles, or NULL */
  int32_t cull_backfaces;
  float color[3];
  float opacity;                /* already adjusted by alpha mode */
  uint32_t surface_id;
} orc_geom;

typedef struct {
  float xfm[12];                /* object->world, column-major 3x4 (4 columns of 3) */
  uint32_t id;                  /* instanceId */
  uint32_t geom;                /* index into geoms */
} orc_entry;

typedef struct {
  uint32_t dims[3];
  float origin[3];
  float spacing[3];
  const float *data;            /* x fastest */
  const float *lut;             /* 256 * 4 RGBA, alpha already times opacity */
  float value_range[2];
  float unit_distance;
  uint32_t id;                  /* volume id -> objectId */
  uint32_t inst_id;             /* instance id (world id for world-level) */
} orc_volume;

/* A light in world space (gpu/sampleLight.h:58-100): the host has applied
 * the instance transform.  directional: dir = direction towards the light,
 * radiance = color*irradiance; point: pos, radiance = color*intensity;
 * spot: pos, dir = cone axis (untransformed), radiance = color*intensity,
 * cos_outer/cos_inner of the opening angle and of the inner angle. */
enum { ORC_LIGHT_DIRECTIONAL = 0, ORC_LIGHT_POINT = 1, ORC_LIGHT_SPOT = 2 };

typedef struct {
  int32_t type;
  float radiance[3];
  float pos[3];
  float dir[3];
  float cos_outer, cos_inner;
} orc_light;

typedef struct {
  uint32_t num_geoms;
  const orc_geom *geoms;
  orc_bvh *bvhs;                /* filled by orc_scene_build, one per geom */
  uint32_t num_entries;
  const orc_entry *entries;
  uint32_t num_volumes;
  const orc_volume *volumes;
  uint32_t num_lights;
  const orc_light *lights;
  void *tlas;                   /* filled by orc_scene_build when num_entries > 1 */
} orc_scene;

int orc_scene_build(orc_scene *s);
void orc_scene_free(orc_scene *s);

typedef struct {
  float pos[3], dir[3], up[3];
  float fovy, aspect;
  float region[4];
} orc_camera;

enum { ORC_RAYCAST = 0, ORC_AO = 1, ORC_DIRECT = 2 };
enum { ORC_COLOR_F32 = 0, ORC_COLOR_SRGB8 = 1, ORC_COLOR_UNORM8 = 2 };

typedef struct {
  int32_t kind;
  float background[4];
  float ambient_color[3];
  float ambient_radiance;
  int32_t pixel_samples;
  int32_t ambient_samples;
  float ao_distance;
  int32_t cull_backfaces;
  float volume_sampling_rate;
  int32_t volume_jitter; /* 1: jittered volume segment starts (two PCG draws each) */
} orc_renderer;

typedef struct {
  uint32_t width, height;
  int32_t color_format;
  float *color_f32;             /* w*h*4 or NULL */
  uint32_t *color_u32;          /* w*h or NULL */
  float *depth;                 /* or NULL */
  uint32_t *prim_id, *obj_id, *inst_id;
  uint32_t frame_id;            /* accumulation restarts at 0 each call */
  uint64_t *stats;              /* optional [0]=internal nodes visited,[1]=prims tested,
                                   [2]=rays traced, [3]=volume samples */
  int32_t threads;              /* OpenMP threads (0 = default) */
  uint8_t *ambiguous;           /* optional: 1 where a volume sample count sits on an fp32
                                   decision boundary (see oracle.c AMB_*) */
} orc_frame;

int orc_render(const orc_scene *s, const orc_camera *cam, const orc_renderer *r, orc_frame *f);

/* Macro-cell majorant grid (ceil(dims/16) cells, x fastest) and Woodcock
   distance sampling over it (see oracle.c). */
void orc_majorant_grid(const orc_volume *v, float *out);
void orc_sample_distance(const orc_volume *v, const float *majorant, uint32_t n,
    const float *org, const float *dir, const float *tmm, uint64_t seed, float *t_out,
    float *albedo, float *tr_out);

/* Closest-hit query on a built scene (world space).  Returns 1 on hit. */
typedef struct {
  float t;
  uint32_t entry, prim;
  float ng[3], ns[3];
  float u, v;
} orc_hit;
int orc_trace_closest(const orc_scene *s, const float org[3], const float dir[3],
    float tmin, float tmax, int cull, orc_hit *hit, uint64_t *stats);
int orc_trace_brute(const orc_scene *s, const float org[3], const float dir[3],
    float tmin, float tmax, int cull, orc_hit *hit);
int orc_trace_any(const orc_scene *s, const float org[3], const float dir[3],
    float tmin, float tmax, uint64_t *stats);
/* surfaceAttenuation (intersectRay.h:105-111): opacity accumulated over every
 * surface hit on [tmin, tmax], o += a (1 - o), 1 once it reaches 0.99. */
float orc_surface_attenuation(const orc_scene *s, const float org[3], const float dir[3],
    float tmin, float tmax);

/* Batched closest-hit for arrays of rays (OpenMP). Out arrays n each. */
void orc_trace_rays(const orc_scene *s, uint32_t n, const float *org, const float *dir,
    float tmin, float tmax, float *t_out, uint32_t *entry_out, uint32_t *prim_out,
    uint64_t *stats);

/* ---- primitives (exposed for tests) --------------------------------------- */
int orc_intersect_sphere(const float o[3], const float d[3], const float c[3], float r, float *t);
int orc_intersect_triangle(const float o[3], const float d[3], const float v0[3],
    const float v1[3], const float v2[3], float *t, float *u, float *v);
double orc_intersect_triangle_f64(const double o[3], const double d[3], const double v0[3],
    const double v1[3], const double v2[3]);
int orc_intersect_cylinder(const float o[3], const float d[3], const float p0[3],
    const float p1[3], float radius, int caps, float *ts /*3*/, float *ns /*9*/, int *valid /*3*/);
int orc_intersect_cone(const float o[3], const float d[3], const float p0[3], const float p1[3],
    float ra, float rb, float *ts /*2*/, float *ns /*6*/, int *valid /*2*/);
int orc_intersect_round_cone(const float o[3], const float d[3], const float pa[3],
    const float pb[3], float ra, float rb, float *t, float n[3]);

void orc_camera_ray(const orc_camera *c, uint32_t w, uint32_t h, float px, float py,
    float org[3], float dir[3]);

/* ---- volume (FP64) -------------------------------------------------------- */
double orc_sample_field(const orc_volume *v, const double p[3]);
void orc_classify(const orc_volume *v, double s, double rgba[4]);
/* Marches one segment; returns entry depth (or tmax if missed).  color/opacity
   are accumulated in place.  nsamples incremented. */
double orc_volume_march(const orc_volume *v, const double org[3], const double dir[3],
    double tmin, double tmax, double inv_sampling_rate, int early_stop,
    double color[3], double *opacity, uint64_t *nsamples);
/* Whole image of the plain volume integral (C rgb, O) per pixel, FP64. */
void orc_volume_image(const orc_volume *v, const orc_camera *cam, uint32_t w, uint32_t h,
    double inv_sampling_rate, int early_stop, double *rgba_out, uint64_t *nsamples,
    int32_t threads, uint8_t *ambiguous);

/* TF discretisation to 256 RGBA entries. color: ncolor*ccomp floats (3 or 4 comp),
   opacity: nopacity floats (or NULL for uniform). */
void orc_tf_discretize(const float *color, uint32_t ncolor, uint32_t ccomp,
    const float *opacity, uint32_t nopacity, const float uniform_color[4],
    float uniform_opacity, const float value_range[2], float *lut_out);

/* ---- channel packing ------------------------------------------------------ */
uint32_t orc_pack_unorm4x8(const float v[4]);
void orc_linear_to_srgb(const float in[4], float out[4]);

/* ---- differentiable volume (FP64) ----------------------------------------- */
/* Plain EA integral (no renderer post-multiply), jitter off.  density: dims^3.
   Rays given explicitly.  rgba_out: n*4 (C.rgb, O). */
void orc_diffvol_forward(const uint32_t dims[3], const float origin[3], const float spacing[3],
    const double *density, const double *lut, const double value_range[2], double unit_distance,
    double step, uint32_t n, const float *org, const float *dir, double *rgba_out,
    int early_stop, int32_t threads, uint8_t *ambiguous);
/* Loss = sum over rays and channels of w[i*4+c] * rgba[i*4+c]; returns loss. */
double orc_diffvol_loss(const uint32_t dims[3], const float origin[3], const float spacing[3],
    const double *density, const double *lut, const double value_range[2], double unit_distance,
    double step, uint32_t n, const float *org, const float *dir, const double *w,
    int early_stop, int32_t threads);

#ifdef __cplusplus
}
#endif
#endif
