// vrx_trace.h -- device-side ray traversal, intersection and hit population.
//
// Replaces what the reference delegates to OptiX and its programs:
//   - optixTrace + RT-core traversal (gpu/intersectRay.h:39-111) -> stackful
//     BVH2 traversal, near child first, short stack in LDS with spill to HBM
//   - OptiX builtin triangle test -> watertight test (Woop, Benthin, Wald,
//     JCGT 2013)
//   - custom intersectors scene/Intersectors_ptx.cu:74-190 (sphere, cylinder)
//   - closest-hit population gpu/populateHit.h:150-368 and
//     computeGeometryPrimId gpu/gpu_util.h:309-315
//   - back-face culling gpu/populateHit.h:320-332 (+ the renderer's
//     cullTriangleBackfaces ray flag, gpu/intersectRay.h:79-83)
// Closest hit = lexicographic minimum of (t, entry index, primitive index)
// over every reported intersection in [tmin, tmax]: independent of the
// traversal order, so it is the same hit the CPU brute force finds.
#pragma once

#include "vrx_lbvh.h"
#include "vrx_math.h"

namespace vrx {

constexpr uint32_t kMarker = 0xFFFFFFFFu; // "leave instance" stack marker

// ---- short stack: top `sd` entries in LDS, older ones spilled to HBM ------
struct Stack {
  uint32_t *lds;     // this lane's column: lds[slot * stride]
  uint32_t stride;
  uint32_t sd_mask;  // sd - 1 (sd a power of two)
  uint32_t sd;
  uint32_t *ovf;     // this thread's spill column: ovf[pos * ovf_stride]
  uint32_t ovf_stride;
  uint32_t ovf_cap;  // spill entries available
  uint32_t cnt;
  uint32_t *err;     // set to 1 if the spill area would overflow

  __device__ __forceinline__ void push(uint32_t x)
  {
    uint32_t slot = cnt & sd_mask;
    if (cnt >= sd) {
      uint32_t pos = cnt - sd;
      if (pos < ovf_cap)
        gst(ovf, (size_t)pos * ovf_stride, lds[slot * stride]);
      else
        gst(err, 0, 1u);
    }
    lds[slot * stride] = x;
    cnt++;
  }
  __device__ __forceinline__ uint32_t pop()
  {
    cnt--;
    uint32_t slot = cnt & sd_mask;
    uint32_t x = lds[slot * stride];
    if (cnt >= sd) {
      uint32_t pos = cnt - sd;
      lds[slot * stride] = pos < ovf_cap ? gld(ovf, (size_t)pos * ovf_stride) : kMarker;
    }
    return x;
  }
};

struct WT {
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

__device__ __forceinline__ WT wt_setup(v3 d)
{
  WT w;
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
  w.kx = kx;
  w.ky = ky;
  w.kz = kz;
  float dz = comp(d, kz);
  w.Sx = comp(d, kx) / dz;
  w.Sy = comp(d, ky) / dz;
  w.Sz = 1.0f / dz;
  return w;
}

__device__ __forceinline__ bool isect_tri(
    const WT &w, v3 o, v3 v0, v3 v1, v3 v2, float &t, float &bu, float &bv)
{
  v3 A = sub(v0, o), B = sub(v1, o), C = sub(v2, o);
  float Akz = comp(A, w.kz), Bkz = comp(B, w.kz), Ckz = comp(C, w.kz);
  float Ax = comp(A, w.kx) - w.Sx * Akz;
  float Ay = comp(A, w.ky) - w.Sy * Akz;
  float Bx = comp(B, w.kx) - w.Sx * Bkz;
  float By = comp(B, w.ky) - w.Sy * Bkz;
  float Cx = comp(C, w.kx) - w.Sx * Ckz;
  float Cy = comp(C, w.ky) - w.Sy * Ckz;
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
    return false;
  float det = (U + V) + W;
  if (det == 0.0f)
    return false;
  float Az = w.Sz * Akz, Bz = w.Sz * Bkz, Cz = w.Sz * Ckz;
  float T = (U * Az + V * Bz) + W * Cz;
  float rcp = 1.0f / det;
  t = T * rcp;
  bu = V * rcp;
  bv = W * rcp;
  return true;
}

// Intersectors_ptx.cu:74-99 (near root only)
__device__ __forceinline__ bool isect_sphere(v3 o, v3 d, v3 c, float r, float &t)
{
  float rd2 = 1.0f / dot(d, d);
  v3 CO = sub(c, o);
  float projCO = dot(CO, d) * rd2;
  v3 perp = sub(CO, fmul(projCO, d));
  float l2 = dot(perp, perp);
  float r2 = r * r;
  if (l2 > r2)
    return false;
  float td = sqrtf((r2 - l2) * rd2);
  t = projCO - td;
  return true;
}

// Intersectors_ptx.cu:101-190: up to three reports (tin, tout, cap), in order.
__device__ __forceinline__ int isect_cylinder(
    v3 ro, v3 rd, v3 p0, v3 p1, float radius, bool caps, float ts[3], v3 ns[3])
{
  int valid = 0;
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
      ts[0] = tin;
      ns[0] = N_in;
      valid |= 1;
    }
    if (tout > cap_tin && tout < cap_tout) {
      ts[1] = tout;
      ns[1] = N_out;
      valid |= 2;
    }
  } else if (sf <= 0.0f && sf >= -s2) {
    ts[0] = tin;
    ns[0] = N_in;
    ts[1] = tout;
    ns[1] = N_out;
    valid |= 3;
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
        valid |= 4;
      }
    }
  }
  return valid;
}

// Slab test shared with the CPU path: conservative (kBoxLo/kBoxHi slack).
__device__ __forceinline__ bool slab(
    float lx, float ly, float lz, float hx, float hy, float hz, v3 o, v3 inv, float tmin, float tmax,
    float &tnear)
{
  float t0x = (lx - o.x) * inv.x, t1x = (hx - o.x) * inv.x;
  float t0y = (ly - o.y) * inv.y, t1y = (hy - o.y) * inv.y;
  float t0z = (lz - o.z) * inv.z, t1z = (hz - o.z) * inv.z;
  float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
  float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
  tnear = fmaxf(tn * kBoxLo, tmin);
  float tfar = fminf(tf * kBoxHi, tmax);
  return tnear <= tfar;
}

struct Hit {
  float t;
  uint32_t entry, prim; // prim = original primitive index
  float a0, a1, a2;     // triangle: u, v ; cylinder: object-space normal
  bool found;
};

struct BlasCtx {
  const Node *nodes;
  const Node4 *wnodes; // 4-wide nodes (traversal of width 4 only)
  const float4 *leaves;
  const uint32_t *ids;
  int type;
  bool cull;
  bool caps;
  uint32_t entry;
  float opacity; // surface opacity (any-hit rays only)
};

__device__ __forceinline__ void load_blas(
    const vrx_scene &sc, uint32_t e, v3 wo, v3 wd, bool cull_all, bool any, BlasCtx &b, v3 &o,
    v3 &d, uint32_t &root, uint32_t &n)
{
  const vrx_entry *Ep = sc.d_entries + e;
  const uint32_t gi = gld(&Ep->geom);
  const vrx_geom *G = sc.d_geoms + gi;
  b.nodes = (const Node *)gld(&G->bvh.d_nodes);
  b.wnodes = (const Node4 *)gld(&G->bvh.d_wide);
  b.leaves = (const float4 *)gld(&G->bvh.d_leaves);
  b.ids = gld(&G->bvh.d_leaf_ids);
  b.type = gld(&G->type);
  b.cull = !any && (cull_all || gld(&G->cull_backfaces) != 0);
  b.caps = gld(&G->caps) != 0;
  b.entry = e;
  b.opacity = any ? gld(&G->opacity) : 1.0f;
  n = gld(&G->n);
  root = n == 1 ? kLeafBit : 0u;
  if (gld(&Ep->identity)) {
    o = wo;
    d = wd;
  } else {
    float inv[12];
#pragma unroll
    for (int q = 0; q < 12; q++)
      inv[q] = gld(&Ep->inv[q]);
    o = xfm_point(inv, wo);
    d = xfm_vector(inv, wd);
  }
}

// Capped cone between p0 (radius ra) and p1 (radius rb), as the reference's
// custom intersector tests it (Intersectors_ptx.cu:192-246): the cap disc
// at the end the ray origin lies beyond (its plane side by the projection
// onto the axis), then the lateral surface -- the published closed form for
// a ray against a capped cone (I. Quilez), with the conic's quadratic
// k2 t^2 + 2 k1 t + k0 = 0 taken at its smaller root.  Up to two candidate
// hits (ts/ns: t and unnormalised normal; bit q of the result marks q).
__device__ __forceinline__ int isect_cone(
    v3 ro, v3 rd, v3 p0, v3 p1, float ra, float rb, float ts[2], v3 ns[2])
{
  int valid = 0;
  const v3 ba = sub(p1, p0), oa = sub(ro, p0), ob = sub(ro, p1);
  const float m0 = dot(ba, ba), m1 = dot(oa, ba), m2 = dot(ob, ba), m3 = dot(rd, ba);
  const float rinv = 1.0f / sqrtf(m0);
  if (m1 < 0.0f) {
    const v3 q = sub(fmul(m3, oa), fmul(m1, rd));
    if (dot(q, q) < (ra * ra) * (m3 * m3)) {
      ts[0] = -m1 / m3;
      ns[0] = fmul(-rinv, ba);
      valid |= 1;
    }
  } else if (m2 > 0.0f) {
    const v3 q = sub(fmul(m3, ob), fmul(m2, rd));
    if (dot(q, q) < (rb * rb) * (m3 * m3)) {
      ts[0] = -m2 / m3;
      ns[0] = fmul(rinv, ba);
      valid |= 1;
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
      valid |= 2;
    }
  }
  return valid;
}

// Round linear curve segment: the sphere swept from (pa, ra) to (pb, rb),
// the reference's curve primitive type (Curve.cpp:83, OptiX round linear
// curves).  The surface is the convex hull of the two end spheres, so the
// ray enters it once: on the tangent cone (the published closed form for a
// rounded cone, I. Quilez), when the entry lies between the two tangent
// circles, else on the nearer end sphere.  When one end sphere contains the
// other the segment is that sphere.  Returns hit; t is the entry distance,
// n an unnormalised outward normal.
__device__ __forceinline__ bool isect_round_cone(
    v3 ro, v3 rd, v3 pa, v3 pb, float ra, float rb, float &t, v3 &n)
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
      return false; // the ray misses the tangent cone, so both spheres
    const float tc = (-sqrtf(h) - k1) / k2;
    const float y = (m1 - ra * rr) + tc * m2;
    if (y > 0.0f && y < d2) {
      t = tc;
      n = sub(fmul(d2, add(oa, fmul(tc, rd))), fmul(y, ba));
      return true;
    }
  }
  // end spheres (for d2 <= 0 only the larger one bounds the segment)
  const float h1 = (m3 * m3 - m5) + ra * ra;
  const float h2 = (m6 * m6 - m7) + rb * rb;
  const bool use1 = d2 > 0.0f || ra >= rb, use2 = d2 > 0.0f || rb > ra;
  bool hit = false;
  if (use1 && h1 > 0.0f) {
    t = -m3 - sqrtf(h1);
    n = add(oa, fmul(t, rd));
    hit = true;
  }
  if (use2 && h2 > 0.0f) {
    const float t2 = -m6 - sqrtf(h2);
    if (!hit || t2 < t) {
      t = t2;
      n = add(ob, fmul(t2, rd));
      hit = true;
    }
  }
  return hit;
}

template <int PT>
__device__ __forceinline__ int blas_type(const BlasCtx &b)
{
  if constexpr (PT >= 0)
    return PT;
  else
    return b.type;
}

// Accept a candidate t (already known to lie in [tmin, tmax]).
__device__ __forceinline__ bool take(Hit &h, float t, float tmax, uint32_t entry, uint32_t prim)
{
  if (!h.found || t < tmax)
    return true;
  // t == tmax == h.t : lowest (entry, prim) wins
  return entry < h.entry || (entry == h.entry && prim < h.prim);
}

// Any-hit of shadow and AO rays (DirectLight_ptx.cu:255-280,
// AmbientOcclusion_ptx.cu:53-69): every reported intersection in the ray
// interval adds its surface's opacity, o += a (1 - o), and the ray ends once
// o reaches 0.99.  h.a0 carries o (trace() starts it at 0); an opaque surface
// ends the ray at its first hit.
__device__ __forceinline__ bool any_hit(const BlasCtx &b, Hit &h)
{
  h.a0 = h.a0 + b.opacity * (1.0f - h.a0);
  return !(h.a0 < 0.99f);
}

// Test leaf k of the current BLAS.  Returns true for an any-hit termination.
template <int PT, bool ANY>
__device__ __forceinline__ bool test_leaf(
    const BlasCtx &b, uint32_t k, v3 o, v3 d, const WT &wt, float tmin, float &tmax, Hit &h)
{
  const int type = blas_type<PT>(b);
  if (type == VRX_SPHERE) {
    float4 L = gld(b.leaves, k);
    float t;
    if (isect_sphere(o, d, mk(L.x, L.y, L.z), L.w, t) && t >= tmin && t <= tmax) {
      if (ANY)
        return any_hit(b, h);
      uint32_t prim = gld(b.ids, k);
      if (take(h, t, tmax, b.entry, prim)) {
        h.found = true;
        h.t = t;
        h.entry = b.entry;
        h.prim = prim;
        tmax = t;
      }
    }
  } else if (type == VRX_TRIANGLE) {
    const float4 *L = b.leaves + 3 * (size_t)k;
    float4 A = gld(L, 0), B = gld(L, 1), C = gld(L, 2);
    v3 v0 = mk(A.x, A.y, A.z), v1 = mk(B.x, B.y, B.z), v2 = mk(C.x, C.y, C.z);
    float t, u, v;
    if (isect_tri(wt, o, v0, v1, v2, t, u, v) && t >= tmin && t <= tmax) {
      if (b.cull) {
        v3 ng = cross(sub(v1, v0), sub(v2, v0));
        if (!(dot(ng, d) < 0.0f))
          return false;
      }
      if (ANY)
        return any_hit(b, h);
      uint32_t prim = gld(b.ids, k);
      if (take(h, t, tmax, b.entry, prim)) {
        h.found = true;
        h.t = t;
        h.entry = b.entry;
        h.prim = prim;
        h.a0 = u;
        h.a1 = v;
        tmax = t;
      }
    }
  } else if (type == VRX_CYLINDER) {
    const float4 *L = b.leaves + 2 * (size_t)k;
    float4 P0 = gld(L, 0), P1 = gld(L, 1);
    float ts[3];
    v3 ns[3];
    int vm = isect_cylinder(o, d, mk(P0.x, P0.y, P0.z), mk(P1.x, P1.y, P1.z), P0.w,
        b.caps, ts, ns);
#pragma unroll
    for (int q = 0; q < 3; q++) {
      if (!(vm & (1 << q)))
        continue;
      float t = ts[q];
      if (t >= tmin && t <= tmax) {
        if (ANY) {
          if (any_hit(b, h))
            return true;
          continue;
        }
        uint32_t prim = gld(b.ids, k);
        if (take(h, t, tmax, b.entry, prim)) {
          h.found = true;
          h.t = t;
          h.entry = b.entry;
          h.prim = prim;
          h.a0 = ns[q].x;
          h.a1 = ns[q].y;
          h.a2 = ns[q].z;
          tmax = t;
        }
      }
    }
  } else if (type == VRX_CURVE) {
    const float4 *L = b.leaves + 2 * (size_t)k;
    float4 P0 = gld(L, 0), P1 = gld(L, 1);
    float t;
    v3 nn;
    if (isect_round_cone(o, d, mk(P0.x, P0.y, P0.z), mk(P1.x, P1.y, P1.z), P0.w, P1.w, t, nn)
        && t >= tmin && t <= tmax) {
      if (ANY)
        return any_hit(b, h);
      uint32_t prim = gld(b.ids, k);
      if (take(h, t, tmax, b.entry, prim)) {
        h.found = true;
        h.t = t;
        h.entry = b.entry;
        h.prim = prim;
        h.a0 = nn.x;
        h.a1 = nn.y;
        h.a2 = nn.z;
        tmax = t;
      }
    }
  } else if (type == VRX_CONE) {
    const float4 *L = b.leaves + 2 * (size_t)k;
    float4 P0 = gld(L, 0), P1 = gld(L, 1);
    float ts[2];
    v3 ns[2];
    int vm = isect_cone(o, d, mk(P0.x, P0.y, P0.z), mk(P1.x, P1.y, P1.z), P0.w, P1.w, ts, ns);
#pragma unroll
    for (int q = 0; q < 2; q++) {
      if (!(vm & (1 << q)))
        continue;
      float t = ts[q];
      if (t >= tmin && t <= tmax) {
        if (ANY) {
          if (any_hit(b, h))
            return true;
          continue;
        }
        uint32_t prim = gld(b.ids, k);
        if (take(h, t, tmax, b.entry, prim)) {
          h.found = true;
          h.t = t;
          h.entry = b.entry;
          h.prim = prim;
          h.a0 = ns[q].x;
          h.a1 = ns[q].y;
          h.a2 = ns[q].z;
          tmax = t;
        }
      }
    }
  }
  return false;
}

// Full traversal.  TWO_LEVEL: start in the instance BVH (sc.tlas) over
// entries; otherwise trace entry 0 only.  Returns hit found (closest) or
// occluded (ANY).
// order (t, link) pairs by t: a 5-comparator sorting network
__device__ __forceinline__ void cswap(float &ta, uint32_t &la, float &tb, uint32_t &lb)
{
  if (tb < ta) {
    const float t = ta;
    ta = tb;
    tb = t;
    const uint32_t l = la;
    la = lb;
    lb = l;
  }
}

// Visit a 4-wide node: test the child boxes, continue with the nearest hit
// child and push the others far-to-near.  Returns false when no child is hit.
__device__ __forceinline__ bool visit4(const Node4 &nd, v3 o, v3 inv, float tmin, float tmax,
    Stack &st, uint32_t &node)
{
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
    const bool hit = (uint32_t)c < nd.meta.x && slab(lx[c], ly[c], lz[c], hx[c], hy[c], hz[c], o,
                                                     inv, tmin, tmax, tn);
    t[c] = hit ? tn : INFINITY;
  }
  cswap(t[0], l[0], t[1], l[1]);
  cswap(t[2], l[2], t[3], l[3]);
  cswap(t[0], l[0], t[2], l[2]);
  cswap(t[1], l[1], t[3], l[3]);
  cswap(t[1], l[1], t[2], l[2]);
  if (!(t[0] < INFINITY))
    return false;
  if (t[3] < INFINITY)
    st.push(l[3]);
  if (t[2] < INFINITY)
    st.push(l[2]);
  if (t[1] < INFINITY)
    st.push(l[1]);
  node = l[0];
  return true;
}

// BUDGET: the walk gives up after `budget` loop iterations and sets
// *deferred (the caller finishes that ray another way); h is then partial.
template <int PT, bool TWO_LEVEL, bool ANY, bool WIDE = false, bool BUDGET = false>
__device__ bool trace(const vrx_scene &sc, v3 wo, v3 wd, float tmin, float tmax, bool cull_all,
    Stack &st, Hit &h, uint32_t budget = 0, bool *deferred = nullptr)
{
  h.found = false;
  h.t = tmax;
  h.entry = 0;
  h.prim = 0;
  h.a0 = h.a1 = h.a2 = 0.0f;
  if (sc.num_entries == 0)
    return false;
  st.cnt = 0;
  const v3 winv = mk(1.0f / wd.x, 1.0f / wd.y, 1.0f / wd.z);
  v3 o = wo, d = wd, inv = winv;
  WT wt;
  BlasCtx b;
  bool in_blas;
  uint32_t node;
  const Node *nodes;
  const Node4 *wnodes = nullptr;
  const uint32_t *tlas_ids = sc.tlas.d_leaf_ids;
  const Node *tlas_nodes = (const Node *)sc.tlas.d_nodes;
  const Node4 *tlas_wnodes = (const Node4 *)sc.tlas.d_wide;
  if (TWO_LEVEL) {
    in_blas = false;
    nodes = tlas_nodes;
    wnodes = tlas_wnodes;
    node = sc.num_entries == 1 ? kLeafBit : 0u;
  } else {
    uint32_t n, root;
    load_blas(sc, 0u, wo, wd, cull_all, ANY, b, o, d, root, n);
    if (n == 0)
      return false;
    inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    if (blas_type<PT>(b) == VRX_TRIANGLE)
      wt = wt_setup(d);
    in_blas = true;
    nodes = b.nodes;
    wnodes = b.wnodes;
    node = root;
  }
  [[maybe_unused]] uint32_t iters = 0;
  for (;;) {
    if constexpr (BUDGET) {
      if (++iters > budget) {
        *deferred = true;
        return false;
      }
    }
    bool do_pop = true;
    if (WIDE && !(node & kLeafBit)) {
      const Node4 nd = gld_node4(wnodes, node);
      do_pop = !visit4(nd, o, inv, tmin, h.t, st, node);
    } else if (!(node & kLeafBit)) {
      const Node nd = gld_node(nodes, node);
      float tn0, tn1;
      bool h0 = slab(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, o, inv, tmin, h.t, tn0);
      bool h1 = slab(nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, o, inv, tmin, h.t, tn1);
      if (h0 && h1) {
        uint32_t nearc = nd.d.x, farc = nd.d.y;
        if (tn1 < tn0) {
          nearc = nd.d.y;
          farc = nd.d.x;
        }
        st.push(farc);
        node = nearc;
        do_pop = false;
      } else if (h0) {
        node = nd.d.x;
        do_pop = false;
      } else if (h1) {
        node = nd.d.y;
        do_pop = false;
      }
    } else if (in_blas) {
      float tcur = h.t;
      if (test_leaf<PT, ANY>(b, node & ~kLeafBit, o, d, wt, tmin, tcur, h))
        return true;
    } else {
      // instance leaf: enter the entry's BLAS in object space
      uint32_t e = gld(tlas_ids, node & ~kLeafBit);
      uint32_t n, root;
      load_blas(sc, e, wo, wd, cull_all, ANY, b, o, d, root, n);
      if (n != 0) {
        inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        if (blas_type<PT>(b) == VRX_TRIANGLE)
          wt = wt_setup(d);
        st.push(kMarker);
        in_blas = true;
        nodes = b.nodes;
        wnodes = b.wnodes;
        node = root;
        do_pop = false;
      } else {
        o = wo;
        d = wd;
      }
    }
    if (do_pop) {
      for (;;) {
        if (st.cnt == 0)
          return ANY ? false : h.found;
        node = st.pop();
        if (TWO_LEVEL && node == kMarker) {
          in_blas = false;
          nodes = tlas_nodes;
          wnodes = tlas_wnodes;
          o = wo;
          d = wd;
          inv = winv;
          continue;
        }
        break;
      }
    }
  }
}

// Surface hit in world space (populateHit.h:150-368).
struct SurfHit {
  float t, eps;
  v3 hp, ng, ns;
  uint32_t prim_out, obj, inst;
  v3 color;
  float opacity;
};

__device__ __forceinline__ void populate(
    const vrx_scene &sc, const Hit &h, v3 wo, v3 wd, SurfHit &s)
{
  const vrx_entry *Ep = sc.d_entries + h.entry;
  const vrx_geom *Gp = sc.d_geoms + gld(&Ep->geom);
  const int gtype = gld(&Gp->type);
  const float *pos = gld(&Gp->d_position);
  const uint32_t *index = gld(&Gp->d_index);
  v3 o = wo, d = wd;
  const bool ident = gld(&Ep->identity) != 0;
  float inv[12];
  if (!ident) {
#pragma unroll
    for (int q = 0; q < 12; q++)
      inv[q] = gld(&Ep->inv[q]);
    o = xfm_point(inv, wo);
    d = xfm_vector(inv, wd);
  }
  v3 ng, ns;
  if (gtype == VRX_TRIANGLE) {
    uint32_t p = h.prim;
    uint32_t ia = index ? gld(index, 3 * (size_t)p) : 3 * p;
    uint32_t ib = index ? gld(index, 3 * (size_t)p + 1) : 3 * p + 1;
    uint32_t ic = index ? gld(index, 3 * (size_t)p + 2) : 3 * p + 2;
    v3 v0 = gld3(pos, ia), v1 = gld3(pos, ib), v2 = gld3(pos, ic);
    v3 raw = cross(sub(v1, v0), sub(v2, v0));
    bool front = dot(raw, d) < 0.0f;
    ng = normalize(raw);
    if (!front)
      ng = neg(ng);
    const float *vn = gld(&Gp->d_vertex_normal);
    if (vn) {
      float bx = (1.0f - h.a0) - h.a1, by = h.a0, bz = h.a1;
      v3 n0 = gld3(vn, ia), n1 = gld3(vn, ib), n2 = gld3(vn, ic);
      ns = add(add(fmul(bx, n0), fmul(by, n1)), fmul(bz, n2));
    } else {
      ns = ng;
    }
    ns = normalize(ns);
    if (dot(ng, ns) < 0.0f)
      ns = neg(ns);
  } else if (gtype == VRX_SPHERE) {
    uint32_t vi = index ? gld(index, h.prim) : h.prim;
    v3 c = gld3(pos, vi);
    v3 hp = add(o, fmul(h.t, d));
    ng = ns = sub(hp, c);
  } else {
    ng = ns = mk(h.a0, h.a1, h.a2);
  }
  if (!ident) {
    ng = xfm_normal(inv, ng);
    ns = xfm_normal(inv, ns);
  }
  s.ng = normalize(ng);
  s.ns = normalize(ns);
  s.t = h.t;
  s.hp = add(wo, fmul(h.t, wd));
  s.eps = eps_from(s.hp, wd, h.t);
  const uint32_t *pid = gld(&Gp->d_primitive_id);
  s.prim_out = pid ? gld(pid, h.prim) : h.prim;
  s.obj = gld(&Gp->surface_id);
  s.inst = gld(&Ep->inst_id);
  s.color = mk(gld(&Gp->color[0]), gld(&Gp->color[1]), gld(&Gp->color[2]));
  s.opacity = gld(&Gp->opacity);
}

} // namespace vrx
