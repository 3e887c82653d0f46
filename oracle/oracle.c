/*
 * oracle.c -- CPU restatement of the VisRTX hot path (checker only).
 *
 * Test infrastructure: see oracle.h for what loads it and which reference
 * files each part follows.  fp32 code here uses an explicit operation order
 * (compiled with -ffp-contract=off) so that the bit-exact claims
 * (Morton codes, BVH topology, closest-hit t and primitive ids) can be
 * checked against the device; the volume and adjoint paths are FP64.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LEAF_BIT 0x80000000u
/* loop bounds shared with the device (degenerate-input guards) */
#define MAX_SURFACE_LAYERS 4096
#define MAX_SAMPLES_PER_SEGMENT (1u << 24)
#define MAX_VOLUME_SEGMENTS 256
#define BOX_LO (1.0f - 0x1p-18f)
#define BOX_HI (1.0f + 0x1p-18f)

/* ------------------------------------------------------------------------- */
/* fp32 vector helpers with a fixed evaluation order                         */
/* ------------------------------------------------------------------------- */
typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mulf(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 fmul(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline v3 mulv(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline v3 cross(v3 a, v3 b)
{
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 normalize(v3 v) { return mulf(v, 1.0f / sqrtf(dot(v, v))); }
static inline v3 vmin(v3 a, v3 b) { return mk(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
static inline v3 vmax(v3 a, v3 b) { return mk(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
static inline float comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

/* ------------------------------------------------------------------------- */
/* PCG32                                                                      */
/* ------------------------------------------------------------------------- */
uint32_t orc_pcg_next(orc_pcg *r)
{
  uint64_t old = r->state;
  r->state = old * 6364136223846793005ULL + r->inc;
  uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  uint32_t rot = (uint32_t)(old >> 59u);
  return (xs >> rot) | (xs << ((-rot) & 31u));
}

void orc_pcg_seed(orc_pcg *r, uint64_t initstate, uint64_t initseq)
{
  r->state = 0u;
  r->inc = (initseq << 1u) | 1u;
  orc_pcg_next(r);
  r->state += initstate;
  orc_pcg_next(r);
}

float orc_pcg_float(orc_pcg *r) { return (float)(orc_pcg_next(r) >> 8) * 0x1p-24f; }

/* ------------------------------------------------------------------------- */
/* Primitive bounds (Sphere.cu:82-110, Cylinder.cpp:53-102, Cone.cpp:98-105, */
/* Triangle.cpp)                                                             */
/* ------------------------------------------------------------------------- */
static void prim_box(int type, uint32_t i, const float *pos, const uint32_t *index,
    const float *radius, float global_radius, v3 *lo, v3 *hi)
{
  if (type == ORC_SPHERE) {
    uint32_t vi = index ? index[i] : i;
    v3 c = ld3(pos + 3 * (size_t)vi);
    float r = radius ? radius[vi] : global_radius;
    *lo = mk(c.x - r, c.y - r, c.z - r);
    *hi = mk(c.x + r, c.y + r, c.z + r);
  } else if (type == ORC_CYLINDER) {
    uint32_t a = index ? index[2 * (size_t)i] : 2 * i;
    uint32_t b = index ? index[2 * (size_t)i + 1] : 2 * i + 1;
    v3 p0 = ld3(pos + 3 * (size_t)a), p1 = ld3(pos + 3 * (size_t)b);
    float r = fabsf(radius ? radius[i] : global_radius);
    v3 mn = vmin(p0, p1), mx = vmax(p0, p1);
    *lo = mk(mn.x - r, mn.y - r, mn.z - r);
    *hi = mk(mx.x + r, mx.y + r, mx.z + r);
  } else if (type == ORC_CONE) {
    uint32_t a = index ? index[2 * (size_t)i] : 2 * i;
    uint32_t b = index ? index[2 * (size_t)i + 1] : 2 * i + 1;
    v3 p0 = ld3(pos + 3 * (size_t)a), p1 = ld3(pos + 3 * (size_t)b);
    float r = fmaxf(radius[a], radius[b]);
    v3 mn = vmin(p0, p1), mx = vmax(p0, p1);
    *lo = mk(mn.x - r, mn.y - r, mn.z - r);
    *hi = mk(mx.x + r, mx.y + r, mx.z + r);
  } else if (type == ORC_CURVE) {
    /* Curve.cpp:83 (round linear): a sphere swept from (p0, r0) to (p1, r1) */
    uint32_t a = index[i];
    v3 p0 = ld3(pos + 3 * (size_t)a), p1 = ld3(pos + 3 * (size_t)a + 3);
    float r0 = radius ? radius[a] : global_radius, r1 = radius ? radius[a + 1] : global_radius;
    *lo = vmin(mk(p0.x - r0, p0.y - r0, p0.z - r0), mk(p1.x - r1, p1.y - r1, p1.z - r1));
    *hi = vmax(mk(p0.x + r0, p0.y + r0, p0.z + r0), mk(p1.x + r1, p1.y + r1, p1.z + r1));
  } else {
    uint32_t a = index ? index[3 * (size_t)i] : 3 * i;
    uint32_t b = index ? index[3 * (size_t)i + 1] : 3 * i + 1;
    uint32_t c = index ? index[3 * (size_t)i + 2] : 3 * i + 2;
    v3 v0 = ld3(pos + 3 * (size_t)a), v1 = ld3(pos + 3 * (size_t)b), v2 = ld3(pos + 3 * (size_t)c);
    *lo = vmin(vmin(v0, v1), v2);
    *hi = vmax(vmax(v0, v1), v2);
  }
}

void orc_bounds(int type, uint32_t n, const float *pos, const uint32_t *index,
    const float *radius, float global_radius, float *boxes)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    v3 lo, hi;
    prim_box(type, (uint32_t)i, pos, index, radius, global_radius, &lo, &hi);
    float *b = boxes + 6 * i;
    b[0] = lo.x; b[1] = lo.y; b[2] = lo.z; b[3] = hi.x; b[4] = hi.y; b[5] = hi.z;
  }
}

/* ------------------------------------------------------------------------- */
/* Morton-30                                                                 */
/* ------------------------------------------------------------------------- */
uint32_t orc_expand_bits10(uint32_t v)
{
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

static inline uint32_t quant10(float c, float lo, float scale)
{
  float x = (c - lo) * scale;
  if (x >= 1023.0f)
    return 1023u;
  if (x > 0.0f)
    return (uint32_t)x;
  return 0u;
}

void orc_morton(const float *boxes, uint32_t n, uint32_t *codes, float *cbox)
{
  v3 lo = mk(INFINITY, INFINITY, INFINITY), hi = mk(-INFINITY, -INFINITY, -INFINITY);
  for (uint32_t i = 0; i < n; i++) {
    const float *b = boxes + 6 * (size_t)i;
    v3 c = mk((b[0] + b[3]) * 0.5f, (b[1] + b[4]) * 0.5f, (b[2] + b[5]) * 0.5f);
    lo = vmin(lo, c);
    hi = vmax(hi, c);
  }
  v3 ext = sub(hi, lo);
  v3 scale = mk(ext.x > 0.0f ? 1024.0f / ext.x : 0.0f, ext.y > 0.0f ? 1024.0f / ext.y : 0.0f,
      ext.z > 0.0f ? 1024.0f / ext.z : 0.0f);
  if (cbox) {
    cbox[0] = lo.x; cbox[1] = lo.y; cbox[2] = lo.z;
    cbox[3] = hi.x; cbox[4] = hi.y; cbox[5] = hi.z;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const float *b = boxes + 6 * i;
    v3 c = mk((b[0] + b[3]) * 0.5f, (b[1] + b[4]) * 0.5f, (b[2] + b[5]) * 0.5f);
    uint32_t qx = quant10(c.x, lo.x, scale.x);
    uint32_t qy = quant10(c.y, lo.y, scale.y);
    uint32_t qz = quant10(c.z, lo.z, scale.z);
    codes[i] = orc_expand_bits10(qx) * 4u + orc_expand_bits10(qy) * 2u + orc_expand_bits10(qz);
  }
}

/* Stable LSD radix sort of (code, index); 4 passes of 8 bits. */
void orc_sort(const uint32_t *codes, uint32_t n, uint32_t *sorted_codes, uint32_t *order)
{
  uint32_t *ka = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *va = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *kb = sorted_codes, *vb = order;
  for (uint32_t i = 0; i < n; i++) {
    ka[i] = codes[i];
    va[i] = i;
  }
  /* 4 passes: ka->kb, kb->ka, ka->kb, kb->ka ; final result in ka -> copy */
  for (int pass = 0; pass < 4; pass++) {
    uint32_t shift = 8u * (uint32_t)pass;
    uint32_t *sk = (pass & 1) ? kb : ka, *sv = (pass & 1) ? vb : va;
    uint32_t *dk = (pass & 1) ? ka : kb, *dv = (pass & 1) ? va : vb;
    size_t count[257];
    memset(count, 0, sizeof(count));
    for (uint32_t i = 0; i < n; i++)
      count[((sk[i] >> shift) & 255u) + 1]++;
    for (int d = 0; d < 256; d++)
      count[d + 1] += count[d];
    for (uint32_t i = 0; i < n; i++) {
      size_t p = count[(sk[i] >> shift) & 255u]++;
      dk[p] = sk[i];
      dv[p] = sv[i];
    }
  }
  memcpy(sorted_codes, ka, sizeof(uint32_t) * n);
  memcpy(order, va, sizeof(uint32_t) * n);
  free(ka);
  free(va);
}

/* ------------------------------------------------------------------------- */
/* Karras 2012 hierarchy with index-augmented keys                           */
/* ------------------------------------------------------------------------- */
static inline int kdelta(const uint32_t *c, int64_t n, int64_t i, int64_t j)
{
  if (j < 0 || j >= n)
    return -1;
  uint32_t a = c[i], b = c[j];
  if (a != b)
    return __builtin_clz(a ^ b);
  return 32 + __builtin_clz((uint32_t)i ^ (uint32_t)j);
}

void orc_karras(const uint32_t *sc, uint32_t n_, uint32_t *child, uint32_t *parent)
{
  int64_t n = n_;
  if (n < 2) {
    if (parent && n == 1)
      parent[0] = ~0u;
    return;
  }
  if (parent)
    parent[0] = ~0u;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n - 1; i++) {
    int d = (kdelta(sc, n, i, i + 1) - kdelta(sc, n, i, i - 1)) > 0 ? 1 : -1;
    int dmin = kdelta(sc, n, i, i - d);
    int64_t lmax = 2;
    while (kdelta(sc, n, i, i + lmax * d) > dmin)
      lmax <<= 1;
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
      if (kdelta(sc, n, i, i + (l + t) * d) > dmin)
        l += t;
    int64_t j = i + l * d;
    int dnode = kdelta(sc, n, i, j);
    int64_t s = 0, t = l;
    do {
      t = (t + 1) >> 1;
      if (kdelta(sc, n, i, i + (s + t) * d) > dnode)
        s += t;
    } while (t > 1);
    int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    uint32_t left = (lo == gamma) ? (LEAF_BIT | (uint32_t)gamma) : (uint32_t)gamma;
    uint32_t right = (hi == gamma + 1) ? (LEAF_BIT | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
    child[2 * i] = left;
    child[2 * i + 1] = right;
    if (parent) {
      parent[(left & LEAF_BIT) ? (uint32_t)(n - 1) + (left & ~LEAF_BIT) : left] = (uint32_t)i;
      parent[(right & LEAF_BIT) ? (uint32_t)(n - 1) + (right & ~LEAF_BIT) : right] = (uint32_t)i;
    }
  }
}

typedef struct { float lo[3], hi[3]; } box6;

static box6 refit_rec(uint32_t node, const float *boxes, const uint32_t *order,
    const uint32_t *child, float *nb)
{
  box6 r;
  if (node & LEAF_BIT) {
    const float *b = boxes + 6 * (size_t)order[node & ~LEAF_BIT];
    for (int k = 0; k < 3; k++) {
      r.lo[k] = b[k];
      r.hi[k] = b[3 + k];
    }
    return r;
  }
  box6 a = refit_rec(child[2 * (size_t)node], boxes, order, child, nb);
  box6 b = refit_rec(child[2 * (size_t)node + 1], boxes, order, child, nb);
  float *o = nb + 12 * (size_t)node;
  for (int k = 0; k < 3; k++) {
    o[k] = a.lo[k];
    o[3 + k] = a.hi[k];
    o[6 + k] = b.lo[k];
    o[9 + k] = b.hi[k];
    r.lo[k] = fminf(a.lo[k], b.lo[k]);
    r.hi[k] = fmaxf(a.hi[k], b.hi[k]);
  }
  return r;
}

void orc_refit(const float *boxes, const uint32_t *order, const uint32_t *child,
    uint32_t n, float *node_boxes, float *root_box)
{
  if (n == 0)
    return;
  box6 r = refit_rec(n == 1 ? LEAF_BIT : 0u, boxes, order, child, node_boxes);
  if (root_box)
    for (int k = 0; k < 3; k++) {
      root_box[k] = r.lo[k];
      root_box[3 + k] = r.hi[k];
    }
}

int orc_bvh_build(int type, uint32_t n, const float *pos, const uint32_t *index,
    const float *radius, float global_radius, orc_bvh *out)
{
  memset(out, 0, sizeof(*out));
  out->n = n;
  if (n == 0)
    return 0;
  size_t ni = n > 1 ? n - 1 : 1;
  out->leaf_boxes = (float *)malloc(sizeof(float) * 6 * (size_t)n);
  out->order = (uint32_t *)malloc(sizeof(uint32_t) * n);
  out->child = (uint32_t *)malloc(sizeof(uint32_t) * 2 * ni);
  out->node_boxes = (float *)malloc(sizeof(float) * 12 * ni);
  uint32_t *codes = (uint32_t *)malloc(sizeof(uint32_t) * n);
  uint32_t *sc = (uint32_t *)malloc(sizeof(uint32_t) * n);
  if (!out->leaf_boxes || !out->order || !out->child || !out->node_boxes || !codes || !sc)
    return -1;
  if (type == ORC_BOXES)
    memcpy(out->leaf_boxes, pos, sizeof(float) * 6 * (size_t)n);
  else
    orc_bounds(type, n, pos, index, radius, global_radius, out->leaf_boxes);
  orc_morton(out->leaf_boxes, n, codes, NULL);
  orc_sort(codes, n, sc, out->order);
  orc_karras(sc, n, out->child, NULL);
  orc_refit(out->leaf_boxes, out->order, out->child, n, out->node_boxes, out->root_box);
  free(codes);
  free(sc);
  return 0;
}

void orc_bvh_free(orc_bvh *b)
{
  free(b->order);
  free(b->child);
  free(b->node_boxes);
  free(b->leaf_boxes);
  memset(b, 0, sizeof(*b));
}

static uint32_t depth_rec(const orc_bvh *b, uint32_t node, uint32_t d)
{
  if (node & LEAF_BIT)
    return d;
  uint32_t a = depth_rec(b, b->child[2 * (size_t)node], d + 1);
  uint32_t c = depth_rec(b, b->child[2 * (size_t)node + 1], d + 1);
  return a > c ? a : c;
}

uint32_t orc_bvh_depth(const orc_bvh *b)
{
  if (b->n < 2)
    return 0;
  return depth_rec(b, 0, 0);
}

/* ------------------------------------------------------------------------- */
/* Intersectors                                                              */
/* ------------------------------------------------------------------------- */

/* Intersectors_ptx.cu:74-99: only the near root is reported. */
static inline int isect_sphere(v3 o, v3 d, v3 c, float r, float *t)
{
  float rd2 = 1.0f / dot(d, d);
  v3 CO = sub(c, o);
  float projCO = dot(CO, d) * rd2;
  v3 perp = sub(CO, fmul(projCO, d));
  float l2 = dot(perp, perp);
  float r2 = r * r;
  if (l2 > r2)
    return 0;
  float td = sqrtf((r2 - l2) * rd2);
  *t = projCO - td;
  return 1;
}

int orc_intersect_sphere(const float o[3], const float d[3], const float c[3], float r, float *t)
{
  return isect_sphere(ld3(o), ld3(d), ld3(c), r, t);
}

/* Intersectors_ptx.cu:101-190.  Up to three reports in order (tin, tout, cap). */
static inline int isect_cylinder(v3 ro, v3 rd, v3 p0, v3 p1, float radius, int caps,
    float ts[3], v3 ns[3], int valid[3])
{
  valid[0] = valid[1] = valid[2] = 0;
  v3 s = sub(p1, p0);
  v3 sxd = cross(s, rd);
  float a = dot(sxd, sxd);
  if (a == 0.0f)
    return 0;
  v3 f = sub(p0, ro);
  v3 sxf = cross(s, f);
  float ra = 1.0f / a;
  float tsv = dot(sxd, sxf) * ra;
  v3 fp = sub(f, fmul(tsv, rd));
  float s2 = dot(s, s);
  v3 perp = cross(s, fp);
  float c = (radius * radius) * s2 - dot(perp, perp);
  if (c < 0.0f)
    return 0;
  float td = sqrtf(c * ra);
  float tin = tsv - td;
  float tout = tsv + td;
  float sf = dot(s, f);
  float sd = dot(s, rd);
  float inv_s2 = 1.0f / s2;
  float u_in = (tin * sd - sf) * inv_s2;
  v3 N_in = sub(sub(fmul(-td, rd), fp), fmul(u_in, s));
  float u_out = (tout * sd - sf) * inv_s2;
  v3 N_out = sub(sub(fmul(td, rd), fp), fmul(u_out, s));
  if (sd != 0.0f) {
    float rsd = 1.0f / sd;
    float tA = sf * rsd;
    float tB = tA + s2 * rsd;
    float cap_tin = fminf(tA, tB);
    float cap_tout = fmaxf(tA, tB);
    if (tin > cap_tin && tin < cap_tout) {
      ts[0] = tin; ns[0] = N_in; valid[0] = 1;
    }
    if (tout > cap_tin && tout < cap_tout) {
      ts[1] = tout; ns[1] = N_out; valid[1] = 1;
    }
  } else if (sf <= 0.0f && sf >= -s2) {
    ts[0] = tin; ns[0] = N_in; valid[0] = 1;
    ts[1] = tout; ns[1] = N_out; valid[1] = 1;
  }
  if (caps) {
    float ca = s2 - sd * sd;
    float cb = s2 * dot(neg(f), rd) + (sf * sd);
    float cc = s2 * dot(neg(f), neg(f)) - (sf * sf) - (radius * radius * s2);
    float h = cb * cb - ca * cc;
    if (h >= 0.0f) {
      h = sqrtf(h);
      float y = ((-cb - h) / ca) * sd - sf;
      float dd = ((y < 0.0f ? 0.0f : s2) + sf) / sd;
      if (fabsf(cb + ca * dd) < h) {
        float sg = y > 0.0f ? 1.0f : (y < 0.0f ? -1.0f : 0.0f);
        ts[2] = dd;
        ns[2] = mk((s.x * sg) / s2, (s.y * sg) / s2, (s.z * sg) / s2);
        valid[2] = 1;
      }
    }
  }
  return valid[0] | valid[1] | valid[2];
}

/* Intersectors_ptx.cu:192-246: the cap disc at the end beyond which the ray
 * origin projects (m1 < 0: the p0 end, m2 > 0: the p1 end), then the lateral
 * surface of the capped cone (I. Quilez's closed form; the quadratic
 * k2 t^2 + 2 k1 t + k0 = 0 at its smaller root, accepted for 0 < y < m0).
 * Up to two reports in order (cap, side); normals unnormalised. */
static inline int isect_cone(v3 ro, v3 rd, v3 p0, v3 p1, float ra, float rb, float ts[2],
    v3 ns[2], int valid[2])
{
  valid[0] = valid[1] = 0;
  const v3 ba = sub(p1, p0), oa = sub(ro, p0), ob = sub(ro, p1);
  const float m0 = dot(ba, ba), m1 = dot(oa, ba), m2 = dot(ob, ba), m3 = dot(rd, ba);
  const float rinv = 1.0f / sqrtf(m0);
  if (m1 < 0.0f) {
    const v3 q = sub(fmul(m3, oa), fmul(m1, rd));
    if (dot(q, q) < (ra * ra) * (m3 * m3)) {
      ts[0] = -m1 / m3;
      ns[0] = fmul(-rinv, ba);
      valid[0] = 1;
    }
  } else if (m2 > 0.0f) {
    const v3 q = sub(fmul(m3, ob), fmul(m2, rd));
    if (dot(q, q) < (rb * rb) * (m3 * m3)) {
      ts[0] = -m2 / m3;
      ns[0] = fmul(rinv, ba);
      valid[0] = 1;
    }
  }
  const float m4 = dot(rd, oa), m5 = dot(oa, oa);
  const float rr = ra - rb;
  const float hy = m0 + rr * rr;
  const float m00 = m0 * m0;
  const float k2 = m00 - (m3 * m3) * hy;
  const float k1 = (m00 * m4 - (m1 * m3) * hy) + (m0 * ra) * (rr * m3);
  const float k0 = (m00 * m5 - (m1 * m1) * hy) + (m0 * ra) * ((rr * m1) * 2.0f - m0 * ra);
  const float h = k1 * k1 - k2 * k0;
  if (h >= 0.0f) {
    const float t = (-k1 - sqrtf(h)) / k2;
    const float y = m1 + t * m3;
    if (y > 0.0f && y < m0) {
      const v3 pt = add(oa, fmul(t, rd));
      ts[1] = t;
      ns[1] = sub(fmul(m0, add(fmul(m0, pt), fmul(rr * ra, ba))), fmul(hy * y, ba));
      valid[1] = 1;
    }
  }
  return valid[0] | valid[1];
}

int orc_intersect_cone(const float o[3], const float d[3], const float p0[3], const float p1[3],
    float ra, float rb, float *ts, float *ns, int *valid)
{
  v3 n[2];
  int r = isect_cone(ld3(o), ld3(d), ld3(p0), ld3(p1), ra, rb, ts, n, valid);
  for (int i = 0; i < 2; i++) {
    if (valid[i]) {
      ns[3 * i] = n[i].x; ns[3 * i + 1] = n[i].y; ns[3 * i + 2] = n[i].z;
    }
  }
  return r;
}

/* Round linear curve segment (Curve.cpp:83): the convex hull of the spheres
 * (pa, ra) and (pb, rb).  The ray enters it on the tangent cone (rounded-cone
 * closed form, I. Quilez) when the entry lies between the tangent circles,
 * else on the nearer end sphere; one sphere inside the other leaves that
 * sphere.  Same operation order as the device. */
static inline int isect_round_cone(v3 ro, v3 rd, v3 pa, v3 pb, float ra, float rb, float *t, v3 *n)
{
  const v3 ba = sub(pb, pa), oa = sub(ro, pa), ob = sub(ro, pb);
  const float rr = ra - rb;
  const float m0 = dot(ba, ba), m1 = dot(ba, oa), m2 = dot(ba, rd), m3 = dot(rd, oa);
  const float m5 = dot(oa, oa), m6 = dot(ob, rd), m7 = dot(ob, ob);
  const float d2 = m0 - rr * rr;
  if (d2 > 0.0f) {
    const float k2 = d2 - m2 * m2;
    const float k1 = (d2 * m3 - m1 * m2) + (m2 * rr) * ra;
    const float k0 = ((d2 * m5 - m1 * m1) + ((m1 * rr) * ra) * 2.0f) - (m0 * ra) * ra;
    const float h = k1 * k1 - k0 * k2;
    if (h < 0.0f)
      return 0;
    const float tc = (-sqrtf(h) - k1) / k2;
    const float y = (m1 - ra * rr) + tc * m2;
    if (y > 0.0f && y < d2) {
      *t = tc;
      *n = sub(fmul(d2, add(oa, fmul(tc, rd))), fmul(y, ba));
      return 1;
    }
  }
  const float h1 = (m3 * m3 - m5) + ra * ra;
  const float h2 = (m6 * m6 - m7) + rb * rb;
  const int use1 = d2 > 0.0f || ra >= rb, use2 = d2 > 0.0f || rb > ra;
  int hit = 0;
  if (use1 && h1 > 0.0f) {
    *t = -m3 - sqrtf(h1);
    *n = add(oa, fmul(*t, rd));
    hit = 1;
  }
  if (use2 && h2 > 0.0f) {
    const float t2 = -m6 - sqrtf(h2);
    if (!hit || t2 < *t) {
      *t = t2;
      *n = add(ob, fmul(t2, rd));
      hit = 1;
    }
  }
  return hit;
}

int orc_intersect_round_cone(const float o[3], const float d[3], const float pa[3],
    const float pb[3], float ra, float rb, float *t, float n[3])
{
  v3 nn;
  int r = isect_round_cone(ld3(o), ld3(d), ld3(pa), ld3(pb), ra, rb, t, &nn);
  if (r) {
    n[0] = nn.x; n[1] = nn.y; n[2] = nn.z;
  }
  return r;
}

int orc_intersect_cylinder(const float o[3], const float d[3], const float p0[3],
    const float p1[3], float radius, int caps, float *ts, float *ns, int *valid)
{
  v3 n[3];
  int r = isect_cylinder(ld3(o), ld3(d), ld3(p0), ld3(p1), fabsf(radius), caps, ts, n, valid);
  for (int i = 0; i < 3; i++) {
    if (valid[i]) {
      ns[3 * i] = n[i].x; ns[3 * i + 1] = n[i].y; ns[3 * i + 2] = n[i].z;
    }
  }
  return r;
}

/* Watertight ray/triangle test (Woop, Benthin, Wald; JCGT 2(1), 2013). */
typedef struct { int kx, ky, kz; float Sx, Sy, Sz; } wt_ray;

static inline wt_ray wt_setup(v3 d)
{
  wt_ray w;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = 0;
  if (ay > ax)
    kz = 1;
  if (az > (kz == 0 ? ax : ay))
    kz = 2;
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  if (comp(d, kz) < 0.0f) {
    int tmp = kx;
    kx = ky;
    ky = tmp;
  }
  w.kx = kx; w.ky = ky; w.kz = kz;
  w.Sx = comp(d, kx) / comp(d, kz);
  w.Sy = comp(d, ky) / comp(d, kz);
  w.Sz = 1.0f / comp(d, kz);
  return w;
}

static inline int isect_tri_wt(const wt_ray *w, v3 o, v3 v0, v3 v1, v3 v2, float *t, float *bu, float *bv)
{
  v3 A = sub(v0, o), B = sub(v1, o), C = sub(v2, o);
  float Akz = comp(A, w->kz), Bkz = comp(B, w->kz), Ckz = comp(C, w->kz);
  float Ax = comp(A, w->kx) - w->Sx * Akz;
  float Ay = comp(A, w->ky) - w->Sy * Akz;
  float Bx = comp(B, w->kx) - w->Sx * Bkz;
  float By = comp(B, w->ky) - w->Sy * Bkz;
  float Cx = comp(C, w->kx) - w->Sx * Ckz;
  float Cy = comp(C, w->ky) - w->Sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
    double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
    double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
    U = (float)(CxBy - CyBx);
    V = (float)(AxCy - AyCx);
    W = (float)(BxAy - ByAx);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
    return 0;
  float det = (U + V) + W;
  if (det == 0.0f)
    return 0;
  float Az = w->Sz * Akz, Bz = w->Sz * Bkz, Cz = w->Sz * Ckz;
  float T = (U * Az + V * Bz) + W * Cz;
  float rcp = 1.0f / det;
  *t = T * rcp;
  *bu = V * rcp;
  *bv = W * rcp;
  return 1;
}

int orc_intersect_triangle(const float o[3], const float d[3], const float v0[3],
    const float v1[3], const float v2[3], float *t, float *u, float *v)
{
  wt_ray w = wt_setup(ld3(d));
  return isect_tri_wt(&w, ld3(o), ld3(v0), ld3(v1), ld3(v2), t, u, v);
}

/* Moller-Trumbore in FP64, used by tests to bound the fp32 watertight test. */
double orc_intersect_triangle_f64(const double o[3], const double d[3], const double v0[3],
    const double v1[3], const double v2[3])
{
  double e1[3], e2[3], p[3], q[3], s[3];
  for (int k = 0; k < 3; k++) {
    e1[k] = v1[k] - v0[k];
    e2[k] = v2[k] - v0[k];
    s[k] = o[k] - v0[k];
  }
  p[0] = d[1] * e2[2] - d[2] * e2[1];
  p[1] = d[2] * e2[0] - d[0] * e2[2];
  p[2] = d[0] * e2[1] - d[1] * e2[0];
  double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
  if (det == 0.0)
    return NAN;
  double inv = 1.0 / det;
  double u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
  if (u < 0.0 || u > 1.0)
    return NAN;
  q[0] = s[1] * e1[2] - s[2] * e1[1];
  q[1] = s[2] * e1[0] - s[0] * e1[2];
  q[2] = s[0] * e1[1] - s[1] * e1[0];
  double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
  if (v < 0.0 || u + v > 1.0)
    return NAN;
  return (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
}

/* ------------------------------------------------------------------------- */
/* Instances: 3x4 column-major transforms, inverse computed in FP64          */
/* ------------------------------------------------------------------------- */
static int is_identity(const float *m)
{
  static const float id[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  for (int i = 0; i < 12; i++)
    if (m[i] != id[i])
      return 0;
  return 1;
}

static void invert_xfm(const float *m, float *inv)
{
  double a = m[0], b = m[3], c = m[6]; /* row 0: a b c */
  double d = m[1], e = m[4], f = m[7]; /* row 1 */
  double g = m[2], h = m[5], i = m[8]; /* row 2 */
  double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
  double det = a * A + b * B + c * C;
  double r = 1.0 / det;
  double im[9];
  /* inverse rows */
  im[0] = A * r; im[3] = -(b * i - c * h) * r; im[6] = (b * f - c * e) * r;
  im[1] = B * r; im[4] = (a * i - c * g) * r;  im[7] = -(a * f - c * d) * r;
  im[2] = C * r; im[5] = -(a * h - b * g) * r; im[8] = (a * e - b * d) * r;
  double tx = m[9], ty = m[10], tz = m[11];
  for (int k = 0; k < 9; k++)
    inv[k] = (float)im[k];
  inv[9] = (float)(-(im[0] * tx + im[3] * ty + im[6] * tz));
  inv[10] = (float)(-(im[1] * tx + im[4] * ty + im[7] * tz));
  inv[11] = (float)(-(im[2] * tx + im[5] * ty + im[8] * tz));
}

static inline v3 xfm_point(const float *m, v3 p)
{
  return mk(((m[0] * p.x + m[3] * p.y) + m[6] * p.z) + m[9],
      ((m[1] * p.x + m[4] * p.y) + m[7] * p.z) + m[10],
      ((m[2] * p.x + m[5] * p.y) + m[8] * p.z) + m[11]);
}

static inline v3 xfm_vector(const float *m, v3 p)
{
  return mk((m[0] * p.x + m[3] * p.y) + m[6] * p.z, (m[1] * p.x + m[4] * p.y) + m[7] * p.z,
      (m[2] * p.x + m[5] * p.y) + m[8] * p.z);
}

/* normal: transpose(inverse) * n */
static inline v3 xfm_normal(const float *inv, v3 n)
{
  return mk((inv[0] * n.x + inv[1] * n.y) + inv[2] * n.z, (inv[3] * n.x + inv[4] * n.y) + inv[5] * n.z,
      (inv[6] * n.x + inv[7] * n.y) + inv[8] * n.z);
}

/* ------------------------------------------------------------------------- */
/* Scene                                                                      */
/* ------------------------------------------------------------------------- */
int orc_scene_build(orc_scene *s)
{
  s->bvhs = (orc_bvh *)calloc(s->num_geoms ? s->num_geoms : 1, sizeof(orc_bvh));
  s->tlas = NULL;
  for (uint32_t g = 0; g < s->num_geoms; g++) {
    const orc_geom *G = &s->geoms[g];
    if (orc_bvh_build(G->type, G->num_prims, G->position, G->index, G->radius,
            G->global_radius, &s->bvhs[g]))
      return -1;
  }
  if (s->num_entries > 1) {
    /* An acceleration structure over the entries, for the oracle's own speed
       on scenes of many instances: world boxes of each entry's BLAS root box
       (8 corners transformed in FP64, padded outwards), an LBVH over them.
       It only prunes entries; every hit is still found by the exact leaf
       tests, and the closest-hit order key is independent of it. */
    float *boxes = (float *)malloc(sizeof(float) * 6 * (size_t)s->num_entries);
    if (!boxes)
      return -1;
    for (uint32_t e = 0; e < s->num_entries; e++) {
      const orc_entry *E = &s->entries[e];
      const orc_bvh *B = &s->bvhs[E->geom];
      double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (int c = 0; c < 8 && B->n > 0; c++) {
        double p[3] = {B->root_box[(c & 1) ? 3 : 0], B->root_box[(c & 2) ? 4 : 1],
            B->root_box[(c & 4) ? 5 : 2]};
        for (int k = 0; k < 3; k++) {
          double w = (double)E->xfm[k] * p[0] + (double)E->xfm[3 + k] * p[1] +
              (double)E->xfm[6 + k] * p[2] + (double)E->xfm[9 + k];
          lo[k] = w < lo[k] ? w : lo[k];
          hi[k] = w > hi[k] ? w : hi[k];
        }
      }
      float *b = boxes + 6 * (size_t)e;
      for (int k = 0; k < 3; k++) {
        if (B->n == 0) {
          b[k] = b[3 + k] = 0.0f; /* empty BLAS: its traversal returns at once */
          continue;
        }
        double pad = 1e-5 * (hi[k] - lo[k]) + 1e-5 * (fabs(lo[k]) + fabs(hi[k])) + 1e-30;
        b[k] = (float)(lo[k] - pad);
        b[3 + k] = (float)(hi[k] + pad);
      }
    }
    orc_bvh *t = (orc_bvh *)calloc(1, sizeof(orc_bvh));
    if (!t || orc_bvh_build(ORC_BOXES, s->num_entries, boxes, NULL, NULL, 0.0f, t)) {
      free(boxes);
      free(t);
      return -1;
    }
    free(boxes);
    s->tlas = t;
  }
  return 0;
}

void orc_scene_free(orc_scene *s)
{
  if (!s->bvhs)
    return;
  for (uint32_t g = 0; g < s->num_geoms; g++)
    orc_bvh_free(&s->bvhs[g]);
  free(s->bvhs);
  s->bvhs = NULL;
  if (s->tlas) {
    orc_bvh_free((orc_bvh *)s->tlas);
    free(s->tlas);
    s->tlas = NULL;
  }
}

/* ray box slab test, shared rule with the device traversal */
static inline int slab(const float *b, v3 o, v3 inv, float tmin, float tmax, float *tn_out)
{
  float t0x = (b[0] - o.x) * inv.x, t1x = (b[3] - o.x) * inv.x;
  float t0y = (b[1] - o.y) * inv.y, t1y = (b[4] - o.y) * inv.y;
  float t0z = (b[2] - o.z) * inv.z, t1z = (b[5] - o.z) * inv.z;
  float tnx = fminf(t0x, t1x), tfx = fmaxf(t0x, t1x);
  float tny = fminf(t0y, t1y), tfy = fmaxf(t0y, t1y);
  float tnz = fminf(t0z, t1z), tfz = fmaxf(t0z, t1z);
  float tn = fmaxf(fmaxf(tnx, tny), tnz);
  float tf = fminf(fminf(tfx, tfy), tfz);
  float tnear = fmaxf(tn * BOX_LO, tmin);
  float tfar = fminf(tf * BOX_HI, tmax);
  *tn_out = tnear;
  return tnear <= tfar;
}

typedef struct {
  float t;
  uint64_t key;
  int found;
  float acc; /* any-hit: accumulated opacity */
  uint32_t entry, prim;
  float a0, a1, a2; /* tri: u, v ; cylinder: normal */
} best_t;

static inline int better(const best_t *b, float t, uint64_t key)
{
  if (!b->found)
    return 1;
  return t < b->t || (t == b->t && key < b->key);
}

typedef struct {
  const orc_scene *s;
  uint32_t e;
  const orc_geom *G;
  const orc_bvh *B;
  v3 o, d, inv;
  wt_ray wt;
  int cull;
  int any; /* any-hit mode */
} tctx;

/* Any-hit of shadow and AO rays (AmbientOcclusion_ptx.cu:53-69,
   DirectLight_ptx.cu:255-280): every reported intersection adds its
   surface's opacity, o += a (1 - o); the ray ends once o reaches 0.99. */
static inline int any_hit(tctx *c, best_t *best)
{
  best->acc = best->acc + c->G->opacity * (1.0f - best->acc);
  return !(best->acc < 0.99f);
}

/* test primitive p (original index).  returns 1 if any-hit terminated. */
static inline int test_prim(tctx *c, uint32_t p, float tmin, float *tmax, best_t *best)
{
  const orc_geom *G = c->G;
  uint64_t key = ((uint64_t)c->e << 32) | p;
  if (G->type == ORC_SPHERE) {
    uint32_t vi = G->index ? G->index[p] : p;
    float r = G->radius ? G->radius[vi] : G->global_radius;
    float t;
    if (isect_sphere(c->o, c->d, ld3(G->position + 3 * (size_t)vi), r, &t)) {
      if (t >= tmin && t <= *tmax) {
        if (c->any)
          return any_hit(c, best);
        if (better(best, t, key)) {
          best->found = 1; best->t = t; best->key = key; best->entry = c->e; best->prim = p;
          *tmax = t;
        }
      }
    }
  } else if (G->type == ORC_CYLINDER) {
    uint32_t a = G->index ? G->index[2 * (size_t)p] : 2 * p;
    uint32_t b = G->index ? G->index[2 * (size_t)p + 1] : 2 * p + 1;
    float r = fabsf(G->radius ? G->radius[p] : G->global_radius);
    float ts[3];
    v3 ns[3];
    int valid[3];
    if (isect_cylinder(c->o, c->d, ld3(G->position + 3 * (size_t)a),
            ld3(G->position + 3 * (size_t)b), r, G->caps, ts, ns, valid)) {
      for (int k = 0; k < 3; k++) {
        if (!valid[k])
          continue;
        float t = ts[k];
        if (t >= tmin && t <= *tmax) {
          if (c->any) {
            if (any_hit(c, best))
              return 1;
            continue;
          }
          if (better(best, t, key)) {
            best->found = 1; best->t = t; best->key = key; best->entry = c->e; best->prim = p;
            best->a0 = ns[k].x; best->a1 = ns[k].y; best->a2 = ns[k].z;
            *tmax = t;
          }
        }
      }
    }
  } else if (G->type == ORC_CURVE) {
    uint32_t a = G->index[p];
    float r0 = G->radius ? G->radius[a] : G->global_radius;
    float r1 = G->radius ? G->radius[a + 1] : G->global_radius;
    float t;
    v3 nn;
    if (isect_round_cone(c->o, c->d, ld3(G->position + 3 * (size_t)a),
            ld3(G->position + 3 * (size_t)a + 3), r0, r1, &t, &nn)
        && t >= tmin && t <= *tmax) {
      if (c->any)
        return any_hit(c, best);
      if (better(best, t, key)) {
        best->found = 1; best->t = t; best->key = key; best->entry = c->e; best->prim = p;
        best->a0 = nn.x; best->a1 = nn.y; best->a2 = nn.z;
        *tmax = t;
      }
    }
  } else if (G->type == ORC_CONE) {
    uint32_t a = G->index ? G->index[2 * (size_t)p] : 2 * p;
    uint32_t b = G->index ? G->index[2 * (size_t)p + 1] : 2 * p + 1;
    float ts[2];
    v3 ns[2];
    int valid[2];
    if (isect_cone(c->o, c->d, ld3(G->position + 3 * (size_t)a), ld3(G->position + 3 * (size_t)b),
            G->radius[a], G->radius[b], ts, ns, valid)) {
      for (int k = 0; k < 2; k++) {
        if (!valid[k])
          continue;
        float t = ts[k];
        if (t >= tmin && t <= *tmax) {
          if (c->any) {
            if (any_hit(c, best))
              return 1;
            continue;
          }
          if (better(best, t, key)) {
            best->found = 1; best->t = t; best->key = key; best->entry = c->e; best->prim = p;
            best->a0 = ns[k].x; best->a1 = ns[k].y; best->a2 = ns[k].z;
            *tmax = t;
          }
        }
      }
    }
  } else {
    uint32_t ia = G->index ? G->index[3 * (size_t)p] : 3 * p;
    uint32_t ib = G->index ? G->index[3 * (size_t)p + 1] : 3 * p + 1;
    uint32_t ic = G->index ? G->index[3 * (size_t)p + 2] : 3 * p + 2;
    v3 v0 = ld3(G->position + 3 * (size_t)ia), v1 = ld3(G->position + 3 * (size_t)ib),
       v2 = ld3(G->position + 3 * (size_t)ic);
    float t, u, v;
    if (isect_tri_wt(&c->wt, c->o, v0, v1, v2, &t, &u, &v)) {
      if (t >= tmin && t <= *tmax) {
        if (c->cull) {
          v3 ng = cross(sub(v1, v0), sub(v2, v0));
          if (!(dot(ng, c->d) < 0.0f))
            return 0;
        }
        if (c->any)
          return any_hit(c, best);
        if (better(best, t, key)) {
          best->found = 1; best->t = t; best->key = key; best->entry = c->e; best->prim = p;
          best->a0 = u; best->a1 = v;
          *tmax = t;
        }
      }
    }
  }
  return 0;
}

/* BVH2 traversal of one entry, near-first.  returns 1 if any-hit found. */
static int traverse_entry(tctx *c, float tmin, float *tmax, best_t *best, uint64_t *stats)
{
  const orc_bvh *B = c->B;
  if (B->n == 0)
    return 0;
  float tn;
  if (!slab(B->root_box, c->o, c->inv, tmin, *tmax, &tn))
    return 0;
  if (B->n == 1) {
    if (stats)
      stats[1]++;
    return test_prim(c, B->order[0], tmin, tmax, best);
  }
  uint32_t stack[256];
  int sp = 0;
  uint32_t cur = 0;
  for (;;) {
    if (!(cur & LEAF_BIT)) {
      if (stats)
        stats[0]++;
      const float *nb = B->node_boxes + 12 * (size_t)cur;
      float tn0, tn1;
      int h0 = slab(nb, c->o, c->inv, tmin, *tmax, &tn0);
      int h1 = slab(nb + 6, c->o, c->inv, tmin, *tmax, &tn1);
      uint32_t c0 = B->child[2 * (size_t)cur], c1 = B->child[2 * (size_t)cur + 1];
      if (h0 && h1) {
        if (tn1 < tn0) {
          uint32_t tmp = c0;
          c0 = c1;
          c1 = tmp;
        }
        stack[sp++] = c1;
        cur = c0;
        continue;
      } else if (h0) {
        cur = c0;
        continue;
      } else if (h1) {
        cur = c1;
        continue;
      }
    } else {
      if (stats)
        stats[1]++;
      if (test_prim(c, B->order[cur & ~LEAF_BIT], tmin, tmax, best))
        return 1;
    }
    if (sp == 0)
      break;
    cur = stack[--sp];
  }
  return 0;
}

static void setup_entry(tctx *c, const orc_scene *s, uint32_t e, v3 wo, v3 wd, float *inv_out,
    int *ident)
{
  const orc_entry *E = &s->entries[e];
  c->s = s;
  c->e = e;
  c->G = &s->geoms[E->geom];
  c->B = &s->bvhs[E->geom];
  *ident = is_identity(E->xfm);
  if (*ident) {
    c->o = wo;
    c->d = wd;
  } else {
    invert_xfm(E->xfm, inv_out);
    c->o = xfm_point(inv_out, wo);
    c->d = xfm_vector(inv_out, wd);
  }
  c->inv = mk(1.0f / c->d.x, 1.0f / c->d.y, 1.0f / c->d.z);
  c->wt = wt_setup(c->d);
}

static int trace_impl(const orc_scene *s, v3 o, v3 d, float tmin, float tmax, int cull_all,
    int any, best_t *best, uint64_t *stats)
{
  memset(best, 0, sizeof(*best));
  float tcur = tmax;
  const orc_bvh *T = (const orc_bvh *)s->tlas;
  if (!T) {
    for (uint32_t e = 0; e < s->num_entries; e++) {
      tctx c;
      float inv[12];
      int ident;
      setup_entry(&c, s, e, o, d, inv, &ident);
      c.cull = !any && (cull_all || c.G->cull_backfaces);
      c.any = any;
      if (traverse_entry(&c, tmin, &tcur, best, stats))
        return 1;
    }
    return any ? 0 : best->found;
  }
  /* entries through the oracle's own instance-level BVH (near first) */
  const v3 winv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  uint32_t stack[256];
  int sp = 0;
  uint32_t cur = T->n == 1 ? LEAF_BIT : 0u;
  float tn;
  if (!slab(T->root_box, o, winv, tmin, tcur, &tn))
    return 0;
  for (;;) {
    if (!(cur & LEAF_BIT)) {
      const float *nb = T->node_boxes + 12 * (size_t)cur;
      float tn0, tn1;
      int h0 = slab(nb, o, winv, tmin, tcur, &tn0);
      int h1 = slab(nb + 6, o, winv, tmin, tcur, &tn1);
      uint32_t c0 = T->child[2 * (size_t)cur], c1 = T->child[2 * (size_t)cur + 1];
      if (h0 && h1) {
        if (tn1 < tn0) {
          uint32_t tmp = c0;
          c0 = c1;
          c1 = tmp;
        }
        stack[sp++] = c1;
        cur = c0;
        continue;
      } else if (h0) {
        cur = c0;
        continue;
      } else if (h1) {
        cur = c1;
        continue;
      }
    } else {
      uint32_t e = T->order[cur & ~LEAF_BIT];
      tctx c;
      float inv[12];
      int ident;
      setup_entry(&c, s, e, o, d, inv, &ident);
      c.cull = !any && (cull_all || c.G->cull_backfaces);
      c.any = any;
      if (traverse_entry(&c, tmin, &tcur, best, stats))
        return 1;
    }
    if (sp == 0)
      break;
    cur = stack[--sp];
  }
  return any ? 0 : best->found;
}

/* Fill world-space hit attributes (populateHit.h:150-368). */
static void hit_attributes(const orc_scene *s, const best_t *b, v3 wo, v3 wd, orc_hit *h)
{
  const orc_entry *E = &s->entries[b->entry];
  const orc_geom *G = &s->geoms[E->geom];
  int ident = is_identity(E->xfm);
  float inv[12];
  v3 o = wo, d = wd;
  if (!ident) {
    invert_xfm(E->xfm, inv);
    o = xfm_point(inv, wo);
    d = xfm_vector(inv, wd);
  }
  v3 ng, ns;
  h->u = h->v = 0.0f;
  if (G->type == ORC_TRIANGLE) {
    uint32_t p = b->prim;
    uint32_t ia = G->index ? G->index[3 * (size_t)p] : 3 * p;
    uint32_t ib = G->index ? G->index[3 * (size_t)p + 1] : 3 * p + 1;
    uint32_t ic = G->index ? G->index[3 * (size_t)p + 2] : 3 * p + 2;
    v3 v0 = ld3(G->position + 3 * (size_t)ia), v1 = ld3(G->position + 3 * (size_t)ib),
       v2 = ld3(G->position + 3 * (size_t)ic);
    v3 raw = cross(sub(v1, v0), sub(v2, v0));
    int front = dot(raw, d) < 0.0f;
    ng = normalize(raw);
    if (!front)
      ng = neg(ng);
    h->u = b->a0;
    h->v = b->a1;
    if (G->vertex_normal) {
      float bx = (1.0f - b->a0) - b->a1, by = b->a0, bz = b->a1;
      v3 n0 = ld3(G->vertex_normal + 3 * (size_t)ia), n1 = ld3(G->vertex_normal + 3 * (size_t)ib),
         n2 = ld3(G->vertex_normal + 3 * (size_t)ic);
      ns = add(add(fmul(bx, n0), fmul(by, n1)), fmul(bz, n2));
    } else
      ns = ng;
    ns = normalize(ns);
    if (dot(ng, ns) < 0.0f)
      ns = neg(ns);
  } else if (G->type == ORC_SPHERE) {
    uint32_t vi = G->index ? G->index[b->prim] : b->prim;
    v3 c = ld3(G->position + 3 * (size_t)vi);
    v3 hp = add(o, fmul(b->t, d));
    ng = ns = sub(hp, c);
  } else {
    ng = ns = mk(b->a0, b->a1, b->a2);
  }
  if (!ident) {
    ng = xfm_normal(inv, ng);
    ns = xfm_normal(inv, ns);
  }
  ng = normalize(ng);
  ns = normalize(ns);
  h->t = b->t;
  h->entry = b->entry;
  h->prim = b->prim;
  h->ng[0] = ng.x; h->ng[1] = ng.y; h->ng[2] = ng.z;
  h->ns[0] = ns.x; h->ns[1] = ns.y; h->ns[2] = ns.z;
}

int orc_trace_closest(const orc_scene *s, const float org[3], const float dir[3], float tmin,
    float tmax, int cull, orc_hit *hit, uint64_t *stats)
{
  best_t b;
  v3 o = ld3(org), d = ld3(dir);
  if (!trace_impl(s, o, d, tmin, tmax, cull, 0, &b, stats))
    return 0;
  hit_attributes(s, &b, o, d, hit);
  return 1;
}

int orc_trace_any(const orc_scene *s, const float org[3], const float dir[3], float tmin,
    float tmax, uint64_t *stats)
{
  best_t b;
  return trace_impl(s, ld3(org), ld3(dir), tmin, tmax, 0, 1, &b, stats);
}

static float surface_atten(const orc_scene *s, v3 o, v3 d, float tmin, float tmax,
    uint64_t *stats)
{
  best_t b;
  if (trace_impl(s, o, d, tmin, tmax, 0, 1, &b, stats))
    return 1.0f; /* reached 0.99: fully blocked */
  return b.acc;
}

float orc_surface_attenuation(const orc_scene *s, const float org[3], const float dir[3],
    float tmin, float tmax)
{
  return surface_atten(s, ld3(org), ld3(dir), tmin, tmax, NULL);
}

int orc_trace_brute(const orc_scene *s, const float org[3], const float dir[3], float tmin,
    float tmax, int cull, orc_hit *hit)
{
  best_t best;
  memset(&best, 0, sizeof(best));
  v3 o = ld3(org), d = ld3(dir);
  float tcur = tmax;
  for (uint32_t e = 0; e < s->num_entries; e++) {
    tctx c;
    float inv[12];
    int ident;
    setup_entry(&c, s, e, o, d, inv, &ident);
    c.cull = cull || c.G->cull_backfaces;
    c.any = 0;
    for (uint32_t p = 0; p < c.G->num_prims; p++)
      test_prim(&c, p, tmin, &tcur, &best);
  }
  if (!best.found)
    return 0;
  hit_attributes(s, &best, o, d, hit);
  return 1;
}

void orc_trace_rays(const orc_scene *s, uint32_t n, const float *org, const float *dir,
    float tmin, float tmax, float *t_out, uint32_t *entry_out, uint32_t *prim_out, uint64_t *stats)
{
  uint64_t nodes = 0, prims = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : nodes, prims)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    best_t b;
    uint64_t st[2] = {0, 0};
    int hit = trace_impl(s, ld3(org + 3 * i), ld3(dir + 3 * i), tmin, tmax, 0, 0, &b, st);
    nodes += st[0];
    prims += st[1];
    t_out[i] = hit ? b.t : INFINITY;
    entry_out[i] = hit ? b.entry : ~0u;
    prim_out[i] = hit ? b.prim : ~0u;
  }
  if (stats) {
    stats[0] += nodes;
    stats[1] += prims;
  }
}

/* ------------------------------------------------------------------------- */
/* Camera (Perspective.cpp:40-72, cameraCreateRay.h:38-81)                    */
/* ------------------------------------------------------------------------- */
typedef struct { v3 pos, d00, du, dv; float region[4]; } cam_x;

static cam_x cam_setup(const orc_camera *c)
{
  cam_x x;
  v3 dir = normalize(ld3(c->dir));
  v3 up = normalize(ld3(c->up));
  float h = 2.0f * tanf(0.5f * c->fovy);
  float w = h * c->aspect;
  x.du = mulf(normalize(cross(dir, up)), w);
  x.dv = mulf(normalize(cross(x.du, dir)), h);
  x.d00 = sub(sub(dir, fmul(0.5f, x.du)), fmul(0.5f, x.dv));
  x.pos = ld3(c->pos);
  memcpy(x.region, c->region, sizeof(x.region));
  return x;
}

static inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

static void cam_ray(const cam_x *c, float sx, float sy, v3 *o, v3 *d)
{
  sx = mixf(c->region[0], c->region[2], sx);
  sy = mixf(c->region[1], c->region[3], sy);
  *o = c->pos;
  *d = normalize(add(add(c->d00, fmul(sx, c->du)), fmul(sy, c->dv)));
}

void orc_camera_ray(const orc_camera *c, uint32_t w, uint32_t h, float px, float py,
    float org[3], float dir[3])
{
  cam_x x = cam_setup(c);
  float iw = 1.0f / (float)w, ih = 1.0f / (float)h;
  v3 o, d;
  cam_ray(&x, px * iw, py * ih, &o, &d);
  org[0] = o.x; org[1] = o.y; org[2] = o.z;
  dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}

/* ------------------------------------------------------------------------- */
/* Volume (FP64)                                                              */
/* ------------------------------------------------------------------------- */
static inline double clampd(double x, double a, double b) { return x < a ? a : (x > b ? b : x); }

static double field_at(const uint32_t dims[3], const float *data, const double *ddata,
    int64_t x, int64_t y, int64_t z)
{
  size_t i = (size_t)x + (size_t)dims[0] * ((size_t)y + (size_t)dims[1] * (size_t)z);
  return ddata ? ddata[i] : (double)data[i];
}

static double trilinear(const uint32_t dims[3], const float origin[3], const float spacing[3],
    const float *data, const double *ddata, const double p[3], int64_t cell[3], double fr[3])
{
  int64_t i0[3];
  double f[3];
  for (int k = 0; k < 3; k++) {
    double g = (p[k] - (double)origin[k]) / (double)spacing[k];
    double gmax = (double)dims[k] - 1.0;
    g = clampd(g, 0.0, gmax);
    int64_t i = (int64_t)floor(g);
    if (dims[k] >= 2 && i > (int64_t)dims[k] - 2)
      i = (int64_t)dims[k] - 2;
    if (dims[k] < 2)
      i = 0;
    i0[k] = i;
    f[k] = dims[k] < 2 ? 0.0 : g - (double)i;
  }
  int64_t x1 = dims[0] >= 2 ? i0[0] + 1 : i0[0];
  int64_t y1 = dims[1] >= 2 ? i0[1] + 1 : i0[1];
  int64_t z1 = dims[2] >= 2 ? i0[2] + 1 : i0[2];
  double c000 = field_at(dims, data, ddata, i0[0], i0[1], i0[2]);
  double c100 = field_at(dims, data, ddata, x1, i0[1], i0[2]);
  double c010 = field_at(dims, data, ddata, i0[0], y1, i0[2]);
  double c110 = field_at(dims, data, ddata, x1, y1, i0[2]);
  double c001 = field_at(dims, data, ddata, i0[0], i0[1], z1);
  double c101 = field_at(dims, data, ddata, x1, i0[1], z1);
  double c011 = field_at(dims, data, ddata, i0[0], y1, z1);
  double c111 = field_at(dims, data, ddata, x1, y1, z1);
  double c00 = c000 + (c100 - c000) * f[0];
  double c10 = c010 + (c110 - c010) * f[0];
  double c01 = c001 + (c101 - c001) * f[0];
  double c11 = c011 + (c111 - c011) * f[0];
  double c0 = c00 + (c10 - c00) * f[1];
  double c1 = c01 + (c11 - c01) * f[1];
  if (cell) {
    for (int k = 0; k < 3; k++) {
      cell[k] = i0[k];
      fr[k] = f[k];
    }
  }
  return c0 + (c1 - c0) * f[2];
}

double orc_sample_field(const orc_volume *v, const double p[3])
{
  return trilinear(v->dims, v->origin, v->spacing, v->data, NULL, p, NULL, NULL);
}

/* classifySample (volumeIntegration.h:46-62) with a linear, clamped 256 LUT */
static void classify_lut(const double *lutd, const float *lutf, const float vr[2], double s,
    double rgba[4], int *i0_out, double *f_out, double *du_ds)
{
  double lo = vr[0], hi = vr[1];
  /* position() (gpu_math.h:185-189): glm::clamp(s, lo, hi) = min(max(s, lo),
     hi), then (s - lo) / (hi - lo).  A reversed range (lo > hi) clamps every
     sample to hi, so u = 1 (the last texel); an empty range (lo == hi) gives
     0/0, a NaN texture coordinate in the reference, read here as texel 0
     (the device does the same; parity unpinned for that case). */
  double sc = fmin(fmax(s, lo), hi);
  double u = (sc - lo) / (hi - lo);
  double x = u * 256.0 - 0.5;
  int clamped = 0;
  if (!(x >= 0.0)) {
    x = 0.0;
    clamped = 1;
  }
  if (x > 255.0) {
    x = 255.0;
    clamped = 1;
  }
  int i = (int)floor(x);
  if (i > 254)
    i = 254;
  double f = x - (double)i;
  for (int k = 0; k < 4; k++) {
    double a = lutd ? lutd[4 * i + k] : (double)lutf[4 * i + k];
    double b = lutd ? lutd[4 * (i + 1) + k] : (double)lutf[4 * (i + 1) + k];
    rgba[k] = a + (b - a) * f;
  }
  if (i0_out)
    *i0_out = i;
  if (f_out)
    *f_out = f;
  if (du_ds) {
    int in_range = s > lo && s < hi;
    *du_ds = (clamped || !in_range) ? 0.0 : 256.0 / (hi - lo);
  }
}

void orc_classify(const orc_volume *v, double s, double rgba[4])
{
  classify_lut(NULL, v->lut, v->value_range, s, rgba, NULL, NULL, NULL);
}

/* Decision margins for fp32-vs-FP64 comparisons: a pixel whose sample count
   would change under fp32 rounding (entry/exit distances within a few ulp of
   a sample position, or opacity within 1e-5 of the 0.99 early-termination
   threshold) is flagged ambiguous.  The opacity margin is 20x the largest
   |O_fp32 - O_fp64| measured over 6,912 rays of a fp32 emulation of the
   march (4.7e-7, 0.5-voxel steps, LUT opacity up to 0.3), so an fp32 march
   cannot take its early-termination decision differently on an unflagged
   pixel. */
#define AMB_T_REL 2e-6
#define AMB_OPACITY 1e-5
/* Which decision is in doubt, as bits of the flag: a segment end or slab
   tie (the sample count of that segment may differ by one) and an
   early-termination tie (the march may stop one or more samples earlier or
   later while the opacity stays within AMB_OPACITY of 0.99). */
#define AMB_TIE_T 1
#define AMB_TIE_STOP 2

static inline int near_rel(double a, double b)
{
  if (isinf(a) || isinf(b))
    return 0; /* an unbounded ray end is never a tie */
  double m = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
  return fabs(a - b) <= AMB_T_REL * (m > 1.0 ? m : 1.0);
}

static int vol_slab_amb(const orc_volume *v, const double o[3], const double d[3], double tmin,
    double tmax, double *t0, double *t1, int *amb);

static int vol_slab(const orc_volume *v, const double o[3], const double d[3], double tmin,
    double tmax, double *t0, double *t1)
{
  return vol_slab_amb(v, o, d, tmin, tmax, t0, t1, NULL);
}

static int vol_slab_amb(const orc_volume *v, const double o[3], const double d[3], double tmin,
    double tmax, double *t0, double *t1, int *amb)
{
  double tn = -INFINITY, tf = INFINITY;
  for (int k = 0; k < 3; k++) {
    double lo = v->origin[k];
    double hi = (double)v->origin[k] + ((double)v->dims[k] - 1.0) * (double)v->spacing[k];
    double a = (lo - o[k]) / d[k], b = (hi - o[k]) / d[k];
    double n = a < b ? a : b, f = a < b ? b : a;
    if (n != n)
      n = -INFINITY;
    if (f != f)
      f = INFINITY;
    tn = n > tn ? n : tn;
    tf = f < tf ? f : tf;
  }
  if (amb && (near_rel(tn, tf) || near_rel(tf, tmin) || near_rel(tn, tmax)))
    *amb |= AMB_TIE_T;
  if (!(tn < tf))
    return 0;
  if (tf < tmin || tn > tmax)
    return 0;
  *t0 = clampd(tn, tmin, tmax);
  *t1 = clampd(tf, tmin, tmax);
  return 1;
}

static double march_amb(const orc_volume *v, const double org[3], const double dir[3],
    double tmin, double tmax, double inv_sampling_rate, int early_stop, double color[3],
    double *opacity, uint64_t *nsamples, int *amb);

double orc_volume_march(const orc_volume *v, const double org[3], const double dir[3],
    double tmin, double tmax, double inv_sampling_rate, int early_stop, double color[3],
    double *opacity, uint64_t *nsamples)
{
  return march_amb(v, org, dir, tmin, tmax, inv_sampling_rate, early_stop, color, opacity,
      nsamples, NULL);
}

static double march_amb(const orc_volume *v, const double org[3], const double dir[3],
    double tmin, double tmax, double inv_sampling_rate, int early_stop, double color[3],
    double *opacity, uint64_t *nsamples, int *amb)
{
  double t0, t1;
  if (!vol_slab_amb(v, org, dir, tmin, tmax, &t0, &t1, amb))
    return tmax;
  float ss = fminf(fminf(v->spacing[0] / 2.0f, v->spacing[1] / 2.0f), v->spacing[2] / 2.0f);
  double step = (double)ss * inv_sampling_rate;
  double expo = step * (1.0 / (double)v->unit_distance);
  double T = 1.0;
  if (!(ss > 0.0f))
    return t0;
  for (uint64_t k = 0; k < MAX_SAMPLES_PER_SEGMENT; k++) {
    double t = t0 + (double)k * step;
    if (amb && near_rel(t, t1))
      *amb |= AMB_TIE_T;
    if (!(t <= t1))
      break;
    if (amb && early_stop && fabs(*opacity - 0.99) < AMB_OPACITY)
      *amb |= AMB_TIE_STOP;
    if (early_stop && !(*opacity < 0.99))
      break;
    double p[3] = {org[0] + dir[0] * t, org[1] + dir[1] * t, org[2] + dir[2] * t};
    double s = trilinear(v->dims, v->origin, v->spacing, v->data, NULL, p, NULL, NULL);
    if (nsamples)
      (*nsamples)++;
    if (s != s)
      continue;
    double co[4];
    classify_lut(NULL, v->lut, v->value_range, s, co, NULL, NULL, NULL);
    double st = pow(1.0 - co[3], expo);
    double w = T * (1.0 - st);
    color[0] += w * co[0];
    color[1] += w * co[1];
    color[2] += w * co[2];
    *opacity += w;
    T *= st;
  }
  return t0;
}

void orc_volume_image(const orc_volume *v, const orc_camera *cam, uint32_t w, uint32_t h,
    double inv_sampling_rate, int early_stop, double *rgba_out, uint64_t *nsamples, int32_t threads,
    uint8_t *ambiguous)
{
  cam_x cx = cam_setup(cam);
  float iw = 1.0f / (float)w, ih = 1.0f / (float)h;
  uint64_t total = 0;
#ifdef _OPENMP
  if (threads > 0)
    omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total)
  for (int64_t y = 0; y < (int64_t)h; y++) {
    for (uint32_t x = 0; x < w; x++) {
      v3 o, d;
      cam_ray(&cx, (float)x * iw, (float)y * ih, &o, &d);
      double od[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
      double c[3] = {0, 0, 0}, op = 0.0;
      uint64_t ns = 0;
      int amb = 0;
      march_amb(v, od, dd, 0.0, (double)FLT_MAX, inv_sampling_rate, early_stop, c, &op, &ns, &amb);
      total += ns;
      if (ambiguous)
        ambiguous[(size_t)y * w + x] = (uint8_t)amb;
      double *out = rgba_out + 4 * ((size_t)y * w + x);
      out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = op;
    }
  }
  if (nsamples)
    *nsamples = total;
}

/* ------------------------------------------------------------------------- */
/* TF discretisation (TransferFunction1D.cpp:101-150, colorMapHelpers.h)     */
/* ------------------------------------------------------------------------- */
static inline float posf(float v, float lo, float hi)
{
  float c = v < lo ? lo : (v > hi ? hi : v);
  return (c - lo) * (1.0f / (hi - lo));
}

static void interp(const float *vals, uint32_t ncomp, uint32_t n, const float *positions,
    float lo, float hi, float p, float *out)
{
  for (uint32_t i = 0; i + 1 < n; i++) {
    float r0 = posf(positions[i], lo, hi), r1 = posf(positions[i + 1], lo, hi);
    if (p >= r0 && p <= r1) {
      float t = posf(p, r0, r1);
      for (uint32_t k = 0; k < ncomp; k++)
        out[k] = vals[ncomp * i + k] * (1.0f - t) + vals[ncomp * (i + 1) + k] * t;
      return;
    }
  }
  uint32_t j = p <= posf(positions[0], lo, hi) ? 0 : n - 1;
  for (uint32_t k = 0; k < ncomp; k++)
    out[k] = vals[ncomp * j + k];
}

static void linear_positions(uint32_t n, float lo, float hi, float *pos)
{
  pos[0] = 0.0f;
  pos[n - 1] = 1.0f;
  float w = 1.0f / (float)(n - 1);
  for (uint32_t i = 1; i + 1 < n; i++)
    pos[i] = pos[i - 1] + w;
  for (uint32_t i = 0; i < n; i++)
    pos[i] = pos[i] * (hi - lo) + lo;
}

void orc_tf_discretize(const float *color, uint32_t ncolor, uint32_t ccomp, const float *opacity,
    uint32_t nopacity, const float uniform_color[4], float uniform_opacity,
    const float value_range[2], float *lut)
{
  float lo = value_range[0], hi = value_range[1];
  float *cpos = color ? (float *)malloc(sizeof(float) * ncolor) : NULL;
  float *opos = opacity ? (float *)malloc(sizeof(float) * nopacity) : NULL;
  if (cpos)
    linear_positions(ncolor, lo, hi, cpos);
  if (opos)
    linear_positions(nopacity, lo, hi, opos);
  for (uint32_t i = 0; i < 256; i++) {
    float p = (float)i / 255.0f;
    float c[4] = {uniform_color[0], uniform_color[1], uniform_color[2], uniform_color[3]};
    if (color) {
      float tmp[4] = {0, 0, 0, 1};
      interp(color, ccomp, ncolor, cpos, lo, hi, p, tmp);
      c[0] = tmp[0]; c[1] = tmp[1]; c[2] = tmp[2];
      c[3] = ccomp == 4 ? tmp[3] : 1.0f;
    }
    float o = uniform_opacity;
    if (opacity)
      interp(opacity, 1, nopacity, opos, lo, hi, p, &o);
    lut[4 * i + 0] = c[0];
    lut[4 * i + 1] = c[1];
    lut[4 * i + 2] = c[2];
    lut[4 * i + 3] = c[3] * o;
  }
  free(cpos);
  free(opos);
}

/* ------------------------------------------------------------------------- */
/* Channel packing (glm func_packing.inl:67-83, color_space.inl:7-55)        */
/* ------------------------------------------------------------------------- */
static inline float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

uint32_t orc_pack_unorm4x8(const float v[4])
{
  uint32_t r = 0;
  for (int k = 0; k < 4; k++) {
    float q = roundf(clamp01(v[k]) * 255.0f);
    r |= ((uint32_t)(unsigned char)q) << (8 * k);
  }
  return r;
}

void orc_linear_to_srgb(const float in[4], float out[4])
{
  for (int k = 0; k < 3; k++) {
    float c = clamp01(in[k]);
    out[k] = c < 0.0031308f ? c * 12.92f : powf(c, 0.41666f) * 1.055f - 0.055f;
  }
  out[3] = in[3];
}

/* ------------------------------------------------------------------------- */
/* Renderers                                                                  */
/* ------------------------------------------------------------------------- */
static inline float cmax3(v3 v) { return fmaxf(fmaxf(v.x, v.y), v.z); }

static float eps_from(v3 P, v3 dir, float t)
{
  float a = fmaxf(fmaxf(fabsf(P.x), fabsf(P.y)), fabsf(P.z));
  float m = fmaxf(fmaxf(fabsf(dir.x), fabsf(dir.y)), fabsf(dir.z)) * t;
  return fmaxf(a, m) * 0x1.fp-21f;
}

/* rayMarchAllVolumes (volumeIntegration.h:317-350), marching in FP64. */
/* The first sample of a segment (volumeIntegration.h:83,120): the entry t0
   plus stepSize * u1 (the field's step, rayMarchVolume) plus the sampling
   step * u2 (_rayMarchVolume), u1 and u2 two draws of the pixel's PCG32
   sequence, when jit is given; the offsets are formed in fp32 as the device
   forms them. */
static double jittered_start(const orc_volume *v, double t0, float inv_rate, orc_pcg *jit)
{
  if (!jit)
    return t0;
  const float ss = fminf(fminf(v->spacing[0] / 2.0f, v->spacing[1] / 2.0f), v->spacing[2] / 2.0f);
  const float u1 = orc_pcg_float(jit), u2 = orc_pcg_float(jit);
  return t0 + (double)(ss * u1) + (double)((ss * inv_rate) * u2);
}

static float march_all(const orc_scene *s, v3 o, v3 d, float tlower, float tfar, int early,
    float inv_rate, v3 *color, float *opacity, uint32_t *obj, uint32_t *inst, uint64_t *ns,
    int *amb, orc_pcg *jit)
{
  float depth = tfar;
  int first = 1;
  int last = -1;
  double c[3] = {color->x, color->y, color->z};
  double op = *opacity;
  double od[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  double tl = tlower;
  for (int seg = 0; seg < MAX_VOLUME_SEGMENTS; seg++) {
    int best = -1;
    double bt0 = 0, bt1 = 0;
    for (uint32_t i = 0; i < s->num_volumes; i++) {
      if ((int)i == last)
        continue;
      double t0, t1;
      if (vol_slab_amb(&s->volumes[i], od, dd, tl, (double)tfar, &t0, &t1, amb)) {
        if (best < 0 || t0 < bt0) {
          best = (int)i;
          bt0 = t0;
          bt1 = t1;
        }
      }
    }
    if (best < 0)
      break;
    if (first) {
      *obj = s->volumes[best].id;
      *inst = s->volumes[best].inst_id;
      first = 0;
    }
    if ((float)bt0 < depth)
      depth = (float)bt0;
    march_amb(&s->volumes[best], od, dd, jittered_start(&s->volumes[best], bt0, inv_rate, jit),
        bt1, (double)inv_rate, early, c, &op, ns, amb);
    tl = bt1 + 1e-3;
    last = best;
    if (!(op < 0.99))
      break;
  }
  *color = mk((float)c[0], (float)c[1], (float)c[2]);
  *opacity = (float)op;
  return depth;
}

typedef struct {
  int found;
  float t, eps;
  v3 hp, ng, ns;
  uint32_t prim_out, obj, inst;
  const orc_geom *G;
} shit;

static int surface_hit(const orc_scene *s, v3 o, v3 d, float tmin, float tmax, int cull, shit *h,
    uint64_t *stats)
{
  best_t b;
  h->found = 0;
  if (!trace_impl(s, o, d, tmin, tmax, cull, 0, &b, stats))
    return 0;
  orc_hit oh;
  hit_attributes(s, &b, o, d, &oh);
  const orc_entry *E = &s->entries[b.entry];
  const orc_geom *G = &s->geoms[E->geom];
  h->found = 1;
  h->t = b.t;
  h->hp = add(o, fmul(b.t, d));
  h->eps = eps_from(h->hp, d, b.t);
  h->ng = ld3(oh.ng);
  h->ns = ld3(oh.ns);
  h->prim_out = G->primitive_id ? G->primitive_id[b.prim] : b.prim;
  h->obj = G->surface_id;
  h->inst = E->id;
  h->G = G;
  return 1;
}

static inline void acc3(v3 *a, v3 b, float i)
{
  float w = 1.0f - i;
  a->x = a->x + b.x * w;
  a->y = a->y + b.y * w;
  a->z = a->z + b.z * w;
}

static inline void acc1(float *a, float b, float i) { *a = *a + b * (1.0f - i); }

typedef struct { v3 c; float a; } v4;

static inline v4 tonemap4(v4 v)
{
  float s = 1.0f + fmaxf(0.0f, cmax3(v.c));
  v4 r = {mk(v.c.x / s, v.c.y / s, v.c.z / s), v.a};
  return r;
}

static inline v4 inv_tonemap4(v4 v)
{
  float s = fmaxf(1e-12f, 1.0f - cmax3(v.c));
  v4 r = {mk(v.c.x / s, v.c.y / s, v.c.z / s), v.a};
  return r;
}

/* computeAO (gpu/computeAO.h:38-60): n cosine-weighted hemisphere rays from
   the geometric-normal offset point, four PCG draws each, each weighted by
   the surface attenuation along it. */
static float ao_factor(const orc_scene *s, const shit *h, float dist, int n, orc_pcg *rs,
    uint64_t *rays, uint64_t *st)
{
  float weights = 0.0f, hits = 0.0f;
  for (int k = 0; k < n; k++) {
    v3 ao_o = add(h->hp, fmul(h->eps, h->ng));
    float q0 = orc_pcg_float(rs), q1 = orc_pcg_float(rs), q2 = orc_pcg_float(rs);
    (void)orc_pcg_float(rs);
    v3 dir = normalize(mk(2.0f * q0 - 1.0f, 2.0f * q1 - 1.0f, 2.0f * q2 - 1.0f));
    if (!(dot(dir, h->ns) > 0.0f))
      dir = neg(dir);
    float wgt = fmaxf(0.0f, dot(dir, h->ns));
    weights += wgt;
    if (wgt != 0.0f) {
      (*rays)++;
      hits += wgt * surface_atten(s, ao_o, dir, 0.0f, dist, st);
    }
  }
  return weights > 0.0f ? 1.0f - hits / weights : 0.0f;
}

/* volumeAttenuation (DirectLight_ptx.cu:62-69, any-hit :270-279): every
   volume the shadow ray crosses on [0, tmax] marched into one opacity
   (FP64, no colour), volumes in index order, stopping at 0.99. */
static float volume_atten(const orc_scene *s, v3 o, v3 d, float tmax, float inv_rate,
    uint64_t *ns, int *amb, orc_pcg *jit)
{
  double op = 0.0, c[3] = {0.0, 0.0, 0.0};
  double od[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
  for (uint32_t i = 0; i < s->num_volumes && op < 0.99; i++) {
    double t0, t1;
    if (!vol_slab_amb(&s->volumes[i], od, dd, 0.0, (double)tmax, &t0, &t1, amb))
      continue;
    march_amb(&s->volumes[i], od, dd, jittered_start(&s->volumes[i], t0, inv_rate, jit), t1,
        (double)inv_rate, 1, c, &op, ns, amb);
  }
  return (float)op;
}

/* sampleLight (gpu/sampleLight.h:58-100), lights in world space. */
typedef struct {
  v3 radiance, dir;
  float dist;
} lsample;

static lsample sample_light(const orc_light *L, v3 hp)
{
  lsample ls;
  v3 rad = ld3(L->radiance);
  if (L->type == ORC_LIGHT_DIRECTIONAL) {
    ls.dir = ld3(L->dir);
    ls.dist = INFINITY;
    ls.radiance = rad;
  } else if (L->type == ORC_LIGHT_POINT) {
    v3 v = sub(ld3(L->pos), hp);
    ls.dist = sqrtf(dot(v, v));
    ls.dir = normalize(v);
    ls.radiance = rad;
  } else {
    /* spot: the distance is the length of the normalised direction, as the
       reference computes it (sampleLight.h:81-82) */
    ls.dir = normalize(sub(ld3(L->pos), hp));
    ls.dist = sqrtf(dot(ls.dir, ls.dir));
    float spot = dot(normalize(ld3(L->dir)), neg(ls.dir));
    if (spot < L->cos_outer)
      spot = 0.0f;
    else if (spot > L->cos_inner)
      spot = 1.0f;
    else {
      spot = (spot - L->cos_outer) / (L->cos_inner - L->cos_outer);
      spot = spot * spot * (3.0f - 2.0f * spot);
    }
    ls.radiance = mulf(rad, spot);
  }
  return ls;
}

/* shadeSurface of the directLight renderer (DirectLight_ptx.cu:71-225) for
   matte surfaces (MatteShader_ptx.cu:71-79): AO factor, ambient times tint,
   and per light base/pi * max(0, Ns.L) * radiance times the surface and
   volume transmission of the shadow ray from hp + eps*Ns.  Matte surfaces
   spawn no bounce ray. */
static v4 shade_direct(const orc_scene *s, const orc_renderer *r, const shit *h, int aos,
    float inv_rate, orc_pcg *rs, uint64_t *rays, uint64_t *vs, int *amb, uint64_t *st)
{
  v3 sp = add(h->hp, fmul(h->eps, h->ns));
  float aof = aos > 0 ? ao_factor(s, h, r->ao_distance, aos, rs, rays, st) : 1.0f;
  v3 base = ld3(h->G->color);
  v3 contrib = mk(0, 0, 0);
  if (r->ambient_radiance > 0.0f)
    contrib = mulv(mulf(ld3(r->ambient_color), r->ambient_radiance), base);
  v3 kd = mulf(base, (float)0.318309886183790671538); /* 1/pi */
  for (uint32_t l = 0; l < s->num_lights; l++) {
    lsample ls = sample_light(&s->lights[l], h->hp);
    float ndl = fmaxf(0.0f, dot(h->ns, ls.dir));
    v3 c = mulv(mulf(kd, ndl), ls.radiance);
    if (isnan(c.x) || isnan(c.y) || isnan(c.z))
      continue;
    if (c.x == 0.0f && c.y == 0.0f && c.z == 0.0f)
      continue; /* adds exactly zero whatever the shadow ray finds */
    (*rays)++;
    float so = 1.0f - surface_atten(s, sp, ls.dir, 0.0f, ls.dist, st);
    float vo = 1.0f;
    if (s->num_volumes)
      vo = 1.0f - volume_atten(s, sp, ls.dir, ls.dist, inv_rate, vs, amb,
                          r->volume_jitter ? rs : NULL);
    contrib = add(contrib, mulf(c, so * vo));
  }
  contrib = mulf(contrib, aof);
  v4 out = {contrib, h->G->opacity};
  return out;
}

int orc_render(const orc_scene *s, const orc_camera *cam, const orc_renderer *r, orc_frame *f)
{
  cam_x cx = cam_setup(cam);
  const uint32_t W = f->width, H = f->height;
  float iw = 1.0f / (float)W, ih = 1.0f / (float)H;
  const int sampled = r->kind == ORC_AO || r->kind == ORC_DIRECT;
  int spp = sampled ? (r->pixel_samples > 1 ? r->pixel_samples : 1) : 1;
  int aos = r->ambient_samples < 0 ? 0 : (r->ambient_samples > 256 ? 256 : r->ambient_samples);
  float inv_rate = 1.0f / r->volume_sampling_rate;
  uint64_t s_nodes = 0, s_prims = 0, s_rays = 0, s_vs = 0;
#ifdef _OPENMP
  if (f->threads > 0)
    omp_set_num_threads(f->threads);
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : s_nodes, s_prims, s_rays, s_vs)
  for (int64_t y = 0; y < (int64_t)H; y++) {
    for (uint32_t x = 0; x < W; x++) {
      size_t idx = (size_t)x + (size_t)y * W;
      orc_pcg rs;
      orc_pcg_seed(&rs, (uint64_t)idx, (uint64_t)f->frame_id);
      v4 accum = {mk(0, 0, 0), 0.0f};
      float dbuf = FLT_MAX;
      uint32_t pid = 0, oid = 0, iid = 0;
      int amb = 0;
      uint64_t st[2] = {0, 0};
      for (int i = 0; i < spp; i++) {
        v3 o, d;
        if (sampled) {
          float r0 = orc_pcg_float(&rs), r1 = orc_pcg_float(&rs);
          (void)orc_pcg_float(&rs);
          (void)orc_pcg_float(&rs);
          cam_ray(&cx, ((float)x + r0) * iw, ((float)y + r1) * ih, &o, &d);
        } else {
          cam_ray(&cx, (float)x * iw, (float)y * ih, &o, &d);
        }
        float tlower = 0.0f;
        const float tmax = FLT_MAX;
        v3 outc = mk(0, 0, 0);
        float outo = 0.0f, depth = 1e30f;
        uint32_t prim = ~0u, obj = ~0u, inst = ~0u;
        int first = 1;
        for (int layer = 0; layer < MAX_SURFACE_LAYERS && outo < 0.99f; layer++) {
          shit h;
          s_rays++;
          surface_hit(s, o, d, tlower, tmax, r->cull_backfaces, &h, st);
          v3 color = mk(0, 0, 0);
          float opacity = 0.0f;
          if (h.found) {
            uint32_t vo = ~0u, vi = ~0u;
            float vdepth = march_all(s, o, d, tlower, h.t, 1, inv_rate, &color, &opacity, &vo,
                &vi, &s_vs, &amb, r->volume_jitter ? &rs : NULL);
            if (first) {
              if (vdepth < h.t) {
                depth = vdepth;
                prim = 0;
                obj = vo;
                inst = vi;
              } else {
                depth = h.t;
                prim = h.prim_out;
                obj = h.obj;
                inst = h.inst;
              }
              first = 0;
            }
            if (r->kind == ORC_DIRECT) {
              uint64_t nr = 0;
              v4 sr = shade_direct(s, r, &h, aos, inv_rate, &rs, &nr, &s_vs, &amb, st);
              s_rays += nr;
              /* DirectLight_ptx.cu:372-381: the shading result replaces the
                 volume colour/opacity in front of the surface and is then
                 accumulated onto itself */
              if (isnan(sr.c.x) || isnan(sr.c.y) || isnan(sr.c.z)) {
                color = mk(0, 0, 0);
                opacity = 0.0f;
              } else {
                color = sr.c;
                opacity = sr.a;
              }
              acc3(&color, sr.c, opacity);
              acc1(&opacity, sr.a, opacity);
            } else {
              v3 base = ld3(h.G->color);
              float mop = h.G->opacity;
              v3 lighting;
              if (r->kind == ORC_AO) {
                uint64_t nr = 0;
                float ao = aos > 0 ? ao_factor(s, &h, r->ao_distance, aos, &rs, &nr, st) : 1.0f;
                s_rays += nr;
                float sc = ao * r->ambient_radiance;
                lighting = fmul(sc, ld3(r->ambient_color));
              } else {
                float sc = fabsf(dot(d, h.ns));
                lighting = fmul(sc, ld3(r->ambient_color));
              }
              acc3(&color, mulv(base, lighting), opacity);
              acc1(&opacity, mop, opacity);
            }
            color = mulf(color, opacity);
            acc3(&outc, color, outo);
            acc1(&outo, opacity, outo);
            tlower = h.t + h.eps;
          } else {
            uint32_t vo = ~0u, vi = ~0u;
            float vdepth = march_all(s, o, d, tlower, tmax, 1, inv_rate, &color, &opacity, &vo,
                &vi, &s_vs, &amb, r->volume_jitter ? &rs : NULL);
            if (first) {
              depth = fminf(depth, vdepth);
              prim = 0;
              obj = vo;
              inst = vi;
            }
            color = mulf(color, opacity);
            acc3(&color, ld3(r->background), opacity);
            acc1(&opacity, r->background[3], opacity);
            acc3(&outc, color, outo);
            acc1(&outo, opacity, outo);
            break;
          }
        }
        /* accumResults (gpu_util.h:393-443) */
        v4 cv = {outc, outo};
        v4 tm = tonemap4(cv);
        accum.c = add(accum.c, tm.c);
        accum.a = accum.a + tm.a;
        if (depth < dbuf) {
          dbuf = depth;
          pid = prim;
          oid = obj;
          iid = inst;
        }
        float div = (float)(f->frame_id + (uint32_t)i + 1u);
        v4 avg = {mk(accum.c.x / div, accum.c.y / div, accum.c.z / div), accum.a / div};
        v4 outv = inv_tonemap4(avg);
        float ov[4] = {outv.c.x, outv.c.y, outv.c.z, outv.a};
        if (i == spp - 1) {
          if (f->color_f32) {
            float *cp = f->color_f32 + 4 * idx;
            cp[0] = ov[0]; cp[1] = ov[1]; cp[2] = ov[2]; cp[3] = ov[3];
          }
          if (f->color_u32) {
            if (f->color_format == ORC_COLOR_SRGB8) {
              float sv[4];
              orc_linear_to_srgb(ov, sv);
              f->color_u32[idx] = orc_pack_unorm4x8(sv);
            } else
              f->color_u32[idx] = orc_pack_unorm4x8(ov);
          }
        }
      }
      if (f->ambiguous)
        f->ambiguous[idx] = (uint8_t)amb;
      if (f->depth)
        f->depth[idx] = dbuf;
      if (f->prim_id)
        f->prim_id[idx] = pid;
      if (f->obj_id)
        f->obj_id[idx] = oid;
      if (f->inst_id)
        f->inst_id[idx] = iid;
      s_nodes += st[0];
      s_prims += st[1];
    }
  }
  if (f->stats) {
    f->stats[0] += s_nodes;
    f->stats[1] += s_prims;
    f->stats[2] += s_rays;
    f->stats[3] += s_vs;
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Differentiable volume (FP64 forward used for central finite differences)  */
/* ------------------------------------------------------------------------- */
static void diff_march(const uint32_t dims[3], const float origin[3], const float spacing[3],
    const double *density, const double *lut, const double vr[2], double unit_distance,
    double step, const double o[3], const double d[3], int early_stop, double out[4], int *amb)
{
  orc_volume v;
  memset(&v, 0, sizeof(v));
  memcpy(v.dims, dims, sizeof(v.dims));
  memcpy(v.origin, origin, sizeof(v.origin));
  memcpy(v.spacing, spacing, sizeof(v.spacing));
  double t0, t1;
  out[0] = out[1] = out[2] = out[3] = 0.0;
  if (!vol_slab_amb(&v, o, d, 0.0, (double)FLT_MAX, &t0, &t1, amb))
    return;
  float vrf[2] = {(float)vr[0], (float)vr[1]};
  double expo = step / unit_distance;
  double T = 1.0;
  for (uint64_t k = 0;; k++) {
    double t = t0 + (double)k * step;
    if (amb && near_rel(t, t1))
      *amb |= AMB_TIE_T;
    if (!(t <= t1))
      break;
    if (amb && early_stop && fabs(out[3] - 0.99) < AMB_OPACITY)
      *amb |= AMB_TIE_STOP;
    if (early_stop && !(out[3] < 0.99))
      break;
    double p[3] = {o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t};
    double s = trilinear(dims, origin, spacing, NULL, density, p, NULL, NULL);
    double co[4];
    classify_lut(lut, NULL, vrf, s, co, NULL, NULL, NULL);
    double st = pow(1.0 - co[3], expo);
    double w = T * (1.0 - st);
    out[0] += w * co[0];
    out[1] += w * co[1];
    out[2] += w * co[2];
    out[3] += w;
    T *= st;
  }
}

void orc_diffvol_forward(const uint32_t dims[3], const float origin[3], const float spacing[3],
    const double *density, const double *lut, const double value_range[2], double unit_distance,
    double step, uint32_t n, const float *org, const float *dir, double *rgba_out, int early_stop,
    int32_t threads, uint8_t *ambiguous)
{
#ifdef _OPENMP
  if (threads > 0)
    omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    double o[3] = {org[3 * i], org[3 * i + 1], org[3 * i + 2]};
    double d[3] = {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
    int amb = 0;
    diff_march(dims, origin, spacing, density, lut, value_range, unit_distance, step, o, d,
        early_stop, rgba_out + 4 * i, &amb);
    if (ambiguous)
      ambiguous[i] = (uint8_t)amb;
  }
}

double orc_diffvol_loss(const uint32_t dims[3], const float origin[3], const float spacing[3],
    const double *density, const double *lut, const double value_range[2], double unit_distance,
    double step, uint32_t n, const float *org, const float *dir, const double *w, int early_stop,
    int32_t threads)
{
  double total = 0.0;
#ifdef _OPENMP
  if (threads > 0)
    omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    double o[3] = {org[3 * i], org[3 * i + 1], org[3 * i + 2]};
    double d[3] = {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
    double out[4];
    diff_march(dims, origin, spacing, density, lut, value_range, unit_distance, step, o, d,
        early_stop, out, NULL);
    total += w[4 * i] * out[0] + w[4 * i + 1] * out[1] + w[4 * i + 2] * out[2] + w[4 * i + 3] * out[3];
  }
  return total;
}

/* ------------------------------------------------------------------------- */
/* Macro-cell majorants and Woodcock tracking                                 */
/* ------------------------------------------------------------------------- */
/* Majorant grid of 16^3-voxel cells (scene/volume/space_skipping/
   UniformGrid.cu:44-258 in layout; the value range is conservative here):
   min / max over each cell's voxels plus a one-voxel halo (NaN ignored), then
   the largest LUT alpha over the entries that range classifies to, with the
   device's fp32 clamp / scale / floor. */
void orc_majorant_grid(const orc_volume *v, float *out)
{
  const int m[3] = {((int)v->dims[0] + 15) / 16, ((int)v->dims[1] + 15) / 16,
      ((int)v->dims[2] + 15) / 16};
  for (int cz = 0; cz < m[2]; cz++)
    for (int cy = 0; cy < m[1]; cy++)
      for (int cx = 0; cx < m[0]; cx++) {
        const int c[3] = {cx, cy, cz};
        int b0[3], b1[3];
        for (int a = 0; a < 3; a++) {
          b0[a] = 16 * c[a] - 1 < 0 ? 0 : 16 * c[a] - 1;
          b1[a] = 16 * c[a] + 17 > (int)v->dims[a] - 1 ? (int)v->dims[a] - 1 : 16 * c[a] + 17;
        }
        float lo = INFINITY, hi = -INFINITY;
        for (int z = b0[2]; z <= b1[2]; z++)
          for (int y = b0[1]; y <= b1[1]; y++)
            for (int x = b0[0]; x <= b1[0]; x++) {
              float val = v->data[(size_t)x + (size_t)v->dims[0] * ((size_t)y +
                  (size_t)v->dims[1] * (size_t)z)];
              lo = fminf(lo, val);
              hi = fmaxf(hi, val);
            }
        float mj = 0.0f;
        if (lo <= hi) {
          const float vlo = v->value_range[0], vhi = v->value_range[1];
          const float inv = 1.0f / (vhi - vlo);
          int e[2];
          const float ends[2] = {lo, hi};
          for (int q = 0; q < 2; q++) {
            float cl = fminf(fmaxf(ends[q], vlo), vhi);
            float x = fminf(fmaxf((cl - vlo) * inv * 256.0f - 0.5f, 0.0f), 255.0f);
            e[q] = (int)x < 254 ? (int)x : 254;
          }
          int i0 = e[0] < e[1] ? e[0] : e[1], i1 = (e[0] > e[1] ? e[0] : e[1]) + 1;
          for (int i = i0; i <= i1; i++)
            mj = fmaxf(mj, v->lut[4 * i + 3]);
        }
        out[((size_t)cz * m[1] + cy) * m[0] + cx] = mj;
      }
}

/* Woodcock tracking (volumeIntegration.h:167-228 _sampleDistance over the
   cells of dda.h:43-122): in a cell of majorant m, tentative collisions at
   t -= log(1 - u) / m * step; real when the LUT alpha there >= u' m.  The
   stepping is fp32 as on the device; field samples and classification are
   FP64.  PCG32 keyed (ray index, seed). */
void orc_sample_distance(const orc_volume *v, const float *majorant, uint32_t n,
    const float *org, const float *dir, const float *tmm, uint64_t seed, float *t_out,
    float *albedo, float *tr_out)
{
  const int m[3] = {((int)v->dims[0] + 15) / 16, ((int)v->dims[1] + 15) / 16,
      ((int)v->dims[2] + 15) / 16};
  float sp2 = v->spacing[0] / 2.0f;
  if (v->spacing[1] / 2.0f < sp2)
    sp2 = v->spacing[1] / 2.0f;
  if (v->spacing[2] / 2.0f < sp2)
    sp2 = v->spacing[2] / 2.0f;
  const float step = sp2 * 1.0f;
  double lutd[1024];
  for (int i = 0; i < 1024; i++)
    lutd[i] = v->lut[i];
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const float *o = org + 3 * i, *d = dir + 3 * i;
    float tout = tmm[2 * i + 1], tr = 1.0f, alb[3] = {0.0f, 0.0f, 0.0f};
    /* the interval clip in fp32, as the device's vol_slab */
    float tn = -INFINITY, tf = INFINITY;
    for (int a = 0; a < 3; a++) {
      float lo = v->origin[a], hi = v->origin[a] + ((float)v->dims[a] - 1.0f) * v->spacing[a];
      float inv = 1.0f / d[a];
      float ta = (lo - o[a]) * inv, tb = (hi - o[a]) * inv;
      tn = fmaxf(tn, fminf(ta, tb));
      tf = fminf(tf, fmaxf(ta, tb));
    }
    const float tmin = tmm[2 * i], tmax = tmm[2 * i + 1];
    if (tn < tf && !(tf < tmin || tn > tmax)) {
      const float t0 = fminf(fmaxf(tn, tmin), tmax), t1 = fminf(fmaxf(tf, tmin), tmax);
      float g0[3], gd[3];
      for (int a = 0; a < 3; a++) {
        float inv_sp = 1.0f / v->spacing[a];
        g0[a] = (o[a] - v->origin[a]) * inv_sp;
        gd[a] = d[a] * inv_sp;
      }
      orc_pcg rs;
      orc_pcg_seed(&rs, (uint64_t)i, seed);
      float t = t0;
      const int max_cells = 3 * (m[0] + m[1] + m[2]) + 8;
      int hit = 0;
      for (int cell = 0; cell < max_cells && t < t1 && !hit; cell++) {
        float tm = fminf(t + 1e-3f * step, 0.5f * (t + t1));
        int ci[3];
        for (int a = 0; a < 3; a++) {
          int q = (int)floorf((g0[a] + gd[a] * tm) * 0.0625f);
          ci[a] = q < 0 ? 0 : (q > m[a] - 1 ? m[a] - 1 : q);
        }
        float te = t1;
        for (int a = 0; a < 3; a++) {
          if (gd[a] > 0.0f)
            te = fminf(te, (16.0f * (float)(ci[a] + 1) - g0[a]) / gd[a]);
          else if (gd[a] < 0.0f)
            te = fminf(te, (16.0f * (float)ci[a] - g0[a]) / gd[a]);
        }
        if (!(te > t))
          te = tm;
        const float mj = majorant[((size_t)ci[2] * m[1] + ci[1]) * m[0] + ci[0]];
        if (mj > 0.0f) {
          for (int k = 0; k < 4096; k++) {
            t -= logf(1.0f - orc_pcg_float(&rs)) / mj * step;
            if (t >= te)
              break;
            double p[3];
            for (int a = 0; a < 3; a++)
              p[a] = (double)v->origin[a] + ((double)g0[a] + (double)gd[a] * (double)t) *
                  (double)v->spacing[a];
            double s = orc_sample_field(v, p);
            if (s == s) {
              double rgba[4];
              classify_lut(lutd, v->lut, v->value_range, s, rgba, NULL, NULL, NULL);
              alb[0] = (float)rgba[0];
              alb[1] = (float)rgba[1];
              alb[2] = (float)rgba[2];
              float u = orc_pcg_float(&rs);
              if ((float)rgba[3] >= u * mj) {
                tr = 0.0f;
                tout = t;
                hit = 1;
                break;
              }
            }
          }
        }
        t = te;
      }
    }
    t_out[i] = tout;
    tr_out[i] = tr;
    albedo[3 * i] = alb[0];
    albedo[3 * i + 1] = alb[1];
    albedo[3 * i + 2] = alb[2];
  }
}
