// objects.cpp -- ANARI objects of the visrtx_amd device (host C++).
//
// Each class cites the reference object it mirrors (paths relative to the
// VisRTX tree).  GPU resources are created through include/vrx_hip.h only.
#include "objects.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>

namespace vrxa {

// ---------------------------------------------------------------------------
// type helpers
// ---------------------------------------------------------------------------
size_t sizeOf(ANARIDataType t)
{
  switch (t) {
  case ANARI_BOOL:
    return 4;
  case ANARI_INT8:
  case ANARI_UINT8:
  case ANARI_UFIXED8:
  case ANARI_UFIXED8_R_SRGB:
    return 1;
  case ANARI_INT16:
  case ANARI_UINT16:
    return 2;
  case ANARI_INT32:
  case ANARI_UINT32:
  case ANARI_FLOAT32:
  case ANARI_UFIXED8_VEC4:
  case ANARI_UFIXED8_RGBA_SRGB:
  case ANARI_DATA_TYPE:
    return 4;
  case ANARI_INT32_VEC2:
  case ANARI_UINT32_VEC2:
  case ANARI_FLOAT32_VEC2:
  case ANARI_FLOAT32_BOX1:
  case ANARI_INT64:
  case ANARI_UINT64:
  case ANARI_FLOAT64:
    return 8;
  case ANARI_INT32_VEC3:
  case ANARI_UINT32_VEC3:
  case ANARI_FLOAT32_VEC3:
    return 12;
  case ANARI_INT32_VEC4:
  case ANARI_UINT32_VEC4:
  case ANARI_FLOAT32_VEC4:
  case ANARI_FLOAT32_BOX2:
  case ANARI_FLOAT64_BOX1:
    return 16;
  case ANARI_FLOAT32_BOX3:
    return 24;
  case ANARI_FLOAT32_BOX4:
    return 32;
  case ANARI_FLOAT32_MAT3x4:
    return 48;
  case ANARI_FLOAT32_MAT4:
    return 64;
  case ANARI_VOID_POINTER:
  case ANARI_FUNCTION_POINTER:
  case ANARI_MEMORY_DELETER:
  case ANARI_STATUS_CALLBACK:
  case ANARI_FRAME_COMPLETION_CALLBACK:
    return sizeof(void *);
  default:
    if (isObjectType(t))
      return sizeof(void *);
    return 0;
  }
}

bool isObjectType(ANARIDataType t) { return t >= ANARI_OBJECT && t <= ANARI_WORLD; }

const char *typeName(ANARIDataType t)
{
  switch (t) {
  case ANARI_ARRAY1D:
    return "ANARI_ARRAY1D";
  case ANARI_ARRAY2D:
    return "ANARI_ARRAY2D";
  case ANARI_ARRAY3D:
    return "ANARI_ARRAY3D";
  case ANARI_CAMERA:
    return "ANARI_CAMERA";
  case ANARI_FRAME:
    return "ANARI_FRAME";
  case ANARI_GEOMETRY:
    return "ANARI_GEOMETRY";
  case ANARI_GROUP:
    return "ANARI_GROUP";
  case ANARI_INSTANCE:
    return "ANARI_INSTANCE";
  case ANARI_LIGHT:
    return "ANARI_LIGHT";
  case ANARI_MATERIAL:
    return "ANARI_MATERIAL";
  case ANARI_RENDERER:
    return "ANARI_RENDERER";
  case ANARI_SURFACE:
    return "ANARI_SURFACE";
  case ANARI_SAMPLER:
    return "ANARI_SAMPLER";
  case ANARI_SPATIAL_FIELD:
    return "ANARI_SPATIAL_FIELD";
  case ANARI_VOLUME:
    return "ANARI_VOLUME";
  case ANARI_WORLD:
    return "ANARI_WORLD";
  case ANARI_FLOAT32:
    return "ANARI_FLOAT32";
  case ANARI_FLOAT32_VEC3:
    return "ANARI_FLOAT32_VEC3";
  case ANARI_UINT32:
    return "ANARI_UINT32";
  default:
    return "ANARI_TYPE";
  }
}

// ---------------------------------------------------------------------------
// Object base (Object.cpp, RegisteredObject.h)
// ---------------------------------------------------------------------------
Object::Object(Device *d, ANARIDataType t, std::string sub) : dev(d), type(t), subtype(std::move(sub))
{}

static void dropParams(std::map<std::string, Param> &m)
{
  for (auto &kv : m)
    if (kv.second.obj)
      releaseInternal(kv.second.obj);
  m.clear();
}

Object::~Object()
{
  observeParams(false);
  dropParams(staged);
  dropParams(params);
}

static bool isArrayType(ANARIDataType t)
{
  return t == ANARI_ARRAY1D || t == ANARI_ARRAY2D || t == ANARI_ARRAY3D;
}

void Object::observeParams(bool on)
{
  for (auto &kv : params) {
    Object *o = kv.second.obj;
    if (!o || !isArrayType(o->type) || o == this)
      continue;
    auto &v = ((Array *)o)->observers;
    if (on) {
      v.push_back(this);
    } else {
      auto it = std::find(v.begin(), v.end(), this);
      if (it != v.end())
        v.erase(it);
    }
  }
}

bool Object::getProperty(const std::string &name, ANARIDataType t, void *mem, uint64_t size, uint32_t)
{
  if (name == "valid" && t == ANARI_BOOL && size >= 4) {
    int v = isValid() ? 1 : 0;
    memcpy(mem, &v, 4);
    return true;
  }
  return false;
}

const Param *Object::find(const char *name) const
{
  auto it = params.find(name);
  return it == params.end() ? nullptr : &it->second;
}

bool Object::getRaw(const char *name, ANARIDataType t, void *out) const
{
  const Param *p = find(name);
  if (!p || p->type != t || p->bytes.empty())
    return false;
  memcpy(out, p->bytes.data(), p->bytes.size());
  return true;
}

std::string Object::getString(const char *name, const std::string &def) const
{
  const Param *p = find(name);
  if (!p || p->type != ANARI_STRING)
    return def;
  return p->str;
}

Object *Object::getObject(const char *name, ANARIDataType t) const
{
  const Param *p = find(name);
  if (!p || !p->obj)
    return nullptr;
  if (t != ANARI_OBJECT && p->obj->type != t)
    return nullptr;
  return p->obj;
}

void Object::report(int severity, int code, const char *fmt, ...) const
{
  va_list ap;
  va_start(ap, fmt);
  dev->report(const_cast<Object *>(this), severity, code, fmt, ap);
  va_end(ap);
}

void retainInternal(Object *o)
{
  if (o)
    o->intRef++;
}

void releaseInternal(Object *o)
{
  if (!o)
    return;
  o->intRef--;
  if (o->intRef <= 0 && o->pubRef <= 0)
    delete o;
}

// ---------------------------------------------------------------------------
// Array (array/Array.cpp:38-197,248-297; Array1D.cpp:43-91; MemoryAllocation.cpp)
// ---------------------------------------------------------------------------
Array::Array(Device *d, ANARIDataType arrType, const void *appMem, ANARIMemoryDeleter del,
    const void *userData, ANARIDataType e, uint64_t n1, uint64_t n2, uint64_t n3)
    : Object(d, arrType, ""), elem(e), dims{n1, n2, n3}, elemSize(sizeOf(e)), app(appMem),
      deleter(del), deleterUser(userData)
{
  end = capacity();
  if (!appMem) {
    owned.assign(capacityBytes() ? capacityBytes() : 1, 0); // managed: zeroed, app fills via map
    app = nullptr;
  } else if (vrx_pointer_kind(appMem, &space) != VRX_OK) {
    space = VRX_PTR_HOST;
  }
  if (isObjectType(elem))
    refreshObjects();
}

Array::~Array()
{
  for (Object *o : objects)
    releaseInternal(o);
  objects.clear();
  if (devOwned)
    vrx_free(const_cast<void *>(app));
  else if (app && deleter)
    deleter(deleterUser, app);
  if (dptr)
    vrx_free(dptr);
}

const void *Array::host()
{
  const uint8_t *base = nullptr;
  if (space == VRX_PTR_DEVICE || space == VRX_PTR_FOREIGN) {
    // device memory: a host copy made on demand (commit-time reads only)
    if (mirrorVersion != version || mirror.size() != capacityBytes()) {
      mirror.resize(capacityBytes());
      if (!mirror.empty()
          && !dev->check(vrx_memcpy_sync(mirror.data(), app, mirror.size(), VRX_D2H), this,
              "array read-back"))
        return nullptr;
      mirrorVersion = version;
    }
    base = mirror.data();
  } else {
    base = owned.empty() ? (const uint8_t *)app : owned.data();
  }
  return base ? base + begin * elemSize : nullptr;
}

void *Array::map()
{
  if (mapped)
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_OPERATION,
        "array mapped again without being previously unmapped");
  mapped = true;
  // the whole allocation, where the application's data lives (device
  // memory for device-resident arrays)
  return owned.empty() ? const_cast<void *>(app) : owned.data();
}

void Array::unmap()
{
  if (!mapped) {
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_OPERATION,
        "array unmapped again without being previously mapped");
    return;
  }
  mapped = false;
  version++;
  if (isObjectType(elem))
    refreshObjects();
  notifyObservers();
}

void Array::commit()
{
  if (type != ANARI_ARRAY1D)
    return;
  // Array1D::commitParameters / finalize (Array1D.cpp:43-67)
  const uint64_t cap = capacity();
  auto read = [&](const char *name, uint64_t def) -> uint64_t {
    const Param *p = find(name);
    if (!p || p->bytes.empty())
      return def;
    if (p->type == ANARI_UINT64 || p->type == ANARI_INT64) {
      uint64_t v;
      memcpy(&v, p->bytes.data(), 8);
      return p->type == ANARI_INT64 && (int64_t)v < 0 ? 0 : v;
    }
    if (p->type == ANARI_UINT32) {
      uint32_t v;
      memcpy(&v, p->bytes.data(), 4);
      return v;
    }
    if (p->type == ANARI_INT32) {
      int32_t v;
      memcpy(&v, p->bytes.data(), 4);
      return v < 0 ? 0 : (uint64_t)v;
    }
    return def;
  };
  if (cap == 0) {
    report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT, "array size must be greater than zero");
    begin = end = 0;
    return;
  }
  uint64_t b = std::min(read("begin", 0), cap - 1);
  uint64_t e = std::clamp(read("end", cap), (uint64_t)1, cap);
  if (b > e) {
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "array 'begin' is not less than 'end', swapping values");
    std::swap(b, e);
  }
  begin = b;
  end = e;
  version++;
  if (isObjectType(elem))
    refreshObjects();
  notifyObservers();
}

void Array::notifyObservers()
{
  for (Object *o : observers)
    dev->queueCommit(o);
}

void Array::refreshObjects()
{
  std::vector<Object *> fresh;
  const Object *const *h = (const Object *const *)host();
  for (uint64_t i = 0; h && i < count(); i++) {
    Object *o = const_cast<Object *>(h[i]);
    retainInternal(o);
    fresh.push_back(o);
  }
  for (Object *o : objects)
    releaseInternal(o);
  objects.swap(fresh);
}

void Array::onPublicReleaseToZero()
{
  // Array.cpp:164-182 (privatize): a shared array still referenced
  // internally is copied, host memory to host memory and device memory to
  // a device allocation of its own
  if (!app || deleter || devOwned || intRef <= 0)
    return;
  if (space == VRX_PTR_HOST || space == VRX_PTR_MANAGED) {
    owned.assign((const uint8_t *)app, (const uint8_t *)app + capacityBytes());
    app = nullptr;
    space = VRX_PTR_HOST;
    return;
  }
  report(ANARI_SEVERITY_PERFORMANCE_WARNING, ANARI_STATUS_NO_ERROR,
      "making private copy of shared GPU array");
  void *copy = nullptr;
  if (!dev->check(vrx_malloc(&copy, capacityBytes()), this, "array privatization"))
    return;
  if (!dev->check(vrx_memcpy(copy, app, capacityBytes(), VRX_D2D, dev->stream), this,
          "array privatization")
      || !dev->check(vrx_stream_sync(dev->stream), this, "array privatization")) {
    vrx_free(copy);
    return;
  }
  app = copy;
  devOwned = true;
  space = VRX_PTR_DEVICE;
}

const void *Array::device()
{
  const size_t off = (size_t)begin * elemSize;
  if (deviceResident())
    return app ? (const uint8_t *)app + off : nullptr;
  if (space == VRX_PTR_FOREIGN) {
    report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
        "array memory belongs to another device");
    return nullptr;
  }
  if (dptr && dversion == version)
    return (const uint8_t *)dptr + off;
  size_t b = capacityBytes();
  if (b == 0)
    return nullptr;
  if (!dptr || dbytes < b) {
    if (dptr)
      vrx_free(dptr);
    dptr = nullptr;
    if (!dev->check(vrx_malloc(&dptr, b), this, "array allocation"))
      return nullptr;
    dbytes = b;
  }
  const void *src = owned.empty() ? app : owned.data();
  if (!dev->check(vrx_memcpy(dptr, src, b, VRX_H2D, dev->stream), this, "array upload"))
    return nullptr;
  dversion = version;
  return (const uint8_t *)dptr + off;
}

// ---------------------------------------------------------------------------
// Geometry (scene/surface/geometry/{Geometry,Triangle,Quad,Sphere,Cylinder,Cone,Curve})
// ---------------------------------------------------------------------------
Geometry::Geometry(Device *d, const std::string &sub) : Object(d, ANARI_GEOMETRY, sub)
{
  if (sub == "triangle")
    ptype = VRX_TRIANGLE;
  else if (sub == "quad") {
    ptype = VRX_TRIANGLE;
    quad = true;
  } else if (sub == "sphere")
    ptype = VRX_SPHERE;
  else if (sub == "cylinder")
    ptype = VRX_CYLINDER;
  else if (sub == "cone")
    ptype = VRX_CONE;
  else if (sub == "curve")
    ptype = VRX_CURVE;
  if (ptype == VRX_CYLINDER || ptype == VRX_CURVE)
    globalRadius = 1.0f; // Cylinder.cpp:49, Curve.cpp:49
}

Geometry::~Geometry()
{
  freeBVH();
  vrx_free(dGenIdx);
}

const uint32_t *Geometry::deviceIndex()
{
  if (!useGenIdx)
    return index ? (const uint32_t *)index->device() : nullptr;
  if (genIdxDirty) {
    const size_t bytes = genIdx.size() * sizeof(uint32_t);
    if (bytes > capGenIdx) {
      vrx_free(dGenIdx);
      dGenIdx = nullptr;
      capGenIdx = 0;
      if (!dev->check(vrx_malloc(&dGenIdx, bytes), this, "quad index allocation"))
        return nullptr;
      capGenIdx = bytes;
    }
    if (bytes && !dev->check(vrx_memcpy_sync(dGenIdx, genIdx.data(), bytes, VRX_H2D), this,
                     "quad index upload"))
      return nullptr;
    genIdxDirty = false;
  }
  return (const uint32_t *)dGenIdx;
}

void Geometry::freeBVH()
{
  vrx_free(dNodes);
  vrx_free(dLeaves);
  vrx_free(dIds);
  vrx_free(dHeader);
  vrx_free(dWide);
  dNodes = dLeaves = dIds = dHeader = dWide = nullptr;
  capNodes = capLeaves = capIds = capWide = 0;
  wideValid = false;
  bvh.d_wide = nullptr;
}

static Array *arrayParam(const Object *o, const char *name)
{
  return (Array *)o->getObject(name, ANARI_ARRAY1D);
}

void Geometry::commit()
{
  valid = false;
  bvhDirty = true;
  if (ptype < 0) {
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "unsupported geometry subtype '%s'", subtype.c_str());
    return;
  }
  pos = arrayParam(this, "vertex.position");
  index = arrayParam(this, "primitive.index");
  primId = arrayParam(this, "primitive.id");
  normal = ptype == VRX_TRIANGLE ? arrayParam(this, "vertex.normal") : nullptr;
  genIdx.clear();
  genIdxDirty = true;
  useGenIdx = quad;
  cull = get<int>("cullBackfaces", ANARI_BOOL, 0) != 0;
  if (ptype == VRX_SPHERE) {
    radius = arrayParam(this, "vertex.radius");
    globalRadius = get<float>("radius", ANARI_FLOAT32, 0.01f); // Sphere.cu:52
  } else if (ptype == VRX_CYLINDER) {
    radius = arrayParam(this, "primitive.radius");
    globalRadius = get<float>("radius", ANARI_FLOAT32, 1.0f);
    caps = getString("caps", "none") != "none"; // Cylinder.cpp:47
  } else if (ptype == VRX_CONE) {
    // Cone.cpp:42-62: per-vertex radii are required.  "caps" is read there
    // but the intersector always tests the end discs (Intersectors_ptx.cu:
    // 215-227), so it changes nothing here either.
    radius = arrayParam(this, "vertex.radius");
    caps = getString("caps", "none") != "none";
    if (!radius) {
      report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
          "missing required parameter 'vertex.radius' on cone geometry");
      return;
    }
  } else if (ptype == VRX_CURVE) {
    // Curve.cpp:45-50: per-vertex radii, else "radius" (default 1)
    radius = arrayParam(this, "vertex.radius");
    globalRadius = get<float>("radius", ANARI_FLOAT32, 1.0f);
  } else {
    radius = nullptr;
  }
  if (!pos) {
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "missing required parameter 'vertex.position' on %s geometry", subtype.c_str());
    return;
  }
  if (pos->elem != ANARI_FLOAT32_VEC3) {
    report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
        "'vertex.position' must be ANARI_FLOAT32_VEC3");
    return;
  }
  const uint64_t nv = pos->count();
  if (quad) {
    // Quad.cpp:57-63 (soup must be a multiple of 4), :124-145 (the split)
    if (index && index->elem != ANARI_UINT32_VEC4) {
      report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
          "'primitive.index' must be ANARI_UINT32_VEC4 on quad geometry");
      return;
    }
    if (!index && nv % 4 != 0) {
      report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
          "'vertex.position' on quad geometry is a non-multiple of 4 without 'primitive.index' "
          "present");
      return;
    }
    const uint64_t nq = index ? index->count() : nv / 4;
    if (2 * nq >= 0x7FFFFFFFull) {
      report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT, "too many quads");
      return;
    }
    const uint32_t *ix = index ? (const uint32_t *)index->host() : nullptr;
    uint32_t mx = 0;
    genIdx.resize(6 * nq);
    for (uint64_t q = 0; q < nq; q++) {
      uint32_t x, y, z, w;
      if (ix) {
        x = ix[4 * q];
        y = ix[4 * q + 1];
        z = ix[4 * q + 2];
        w = ix[4 * q + 3];
      } else {
        x = (uint32_t)(4 * q);
        y = x + 1;
        z = x + 2;
        w = x + 3;
      }
      mx = std::max(std::max(mx, std::max(x, y)), std::max(z, w));
      const uint32_t t[6] = {x, y, w, z, w, y};
      for (int k = 0; k < 6; k++)
        genIdx[6 * q + k] = t[k];
    }
    if (nq && (uint64_t)mx >= nv) {
      report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
          "'primitive.index' refers to vertex %u but only %llu vertices exist", mx,
          (unsigned long long)nv);
      return;
    }
    n = (uint32_t)(2 * nq);
  } else if (ptype == VRX_CURVE) {
    // segment i joins vertices index[i] and index[i] + 1 (round linear
    // curves); without an index, segments start at 0 .. nv/2 - 1
    // (Curve.cpp:107-118)
    if (index && index->elem != ANARI_UINT32) {
      report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
          "'primitive.index' must be ANARI_UINT32 on curve geometry");
      return;
    }
    const uint64_t ns = index ? index->count() : nv / 2;
    const uint32_t *ix = index ? (const uint32_t *)index->host() : nullptr;
    uint32_t mx = 0;
    for (uint64_t i = 0; ix && i < ns; i++)
      mx = std::max(mx, ix[i]);
    if (ns && (uint64_t)(ix ? mx : (uint32_t)(ns - 1)) + 1 >= nv) {
      report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
          "curve segment starting at vertex %u needs vertex %u but only %llu vertices exist",
          ix ? mx : (uint32_t)(ns - 1), (ix ? mx : (uint32_t)(ns - 1)) + 1, (unsigned long long)nv);
      return;
    }
    if (!ix) {
      genIdx.resize(ns);
      for (uint64_t i = 0; i < ns; i++)
        genIdx[i] = (uint32_t)i;
      useGenIdx = true;
    }
    n = (uint32_t)ns;
  } else if (ptype == VRX_TRIANGLE) {
    if (index) {
      if (index->elem != ANARI_UINT32_VEC3) {
        report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
            "'primitive.index' must be ANARI_UINT32_VEC3 on triangle geometry");
        return;
      }
      n = (uint32_t)index->count();
    } else {
      if (nv % 3 != 0) {
        // Triangle.cpp:56-69
        report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
            "triangle soup 'vertex.position' size is not a multiple of 3");
        return;
      }
      n = (uint32_t)(nv / 3);
    }
  } else if (ptype == VRX_SPHERE) {
    if (index && index->elem != ANARI_UINT32) {
      report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
          "'primitive.index' must be ANARI_UINT32 on sphere geometry");
      return;
    }
    n = (uint32_t)(index ? index->count() : nv);
  } else {
    if (index && index->elem != ANARI_UINT32_VEC2) {
      report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
          "'primitive.index' must be ANARI_UINT32_VEC2 on %s geometry", subtype.c_str());
      return;
    }
    n = (uint32_t)(index ? index->count() : nv / 2);
  }
  if (index && !quad && ptype != VRX_CURVE) {
    // every index must address an existing vertex (checked once per commit)
    const uint32_t *ix = (const uint32_t *)index->host();
    const size_t ni = index->count()
        * (ptype == VRX_TRIANGLE ? 3 : (ptype == VRX_CYLINDER || ptype == VRX_CONE ? 2 : 1));
    uint32_t mx = 0;
    for (size_t i = 0; i < ni; i++)
      mx = ix[i] > mx ? ix[i] : mx;
    if (ni && (uint64_t)mx >= nv) {
      report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
          "'primitive.index' refers to vertex %u but only %llu vertices exist", mx,
          (unsigned long long)nv);
      return;
    }
  }
  if (radius) {
    const uint64_t need = ptype == VRX_SPHERE || ptype == VRX_CONE || ptype == VRX_CURVE ? nv : n;
    if (radius->count() < need) {
      report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
          "radius array has %llu entries, %llu needed", (unsigned long long)radius->count(),
          (unsigned long long)need);
      return;
    }
  }
  if (radius && radius->elem != ANARI_FLOAT32) {
    report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT, "radius array must be ANARI_FLOAT32");
    return;
  }
  if (primId && (primId->elem != ANARI_UINT32 || primId->count() < n)) {
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "ignoring 'primitive.id' (needs ANARI_UINT32 with one entry per primitive)");
    primId = nullptr;
  }
  if (normal && (normal->elem != ANARI_FLOAT32_VEC3 || normal->count() < nv))
    normal = nullptr;
  valid = true;
}

int Geometry::buildBVH(vrx_stream s)
{
  vrx_bvh_sizes sz;
  vrx_bvh_query_sizes(ptype, n, &sz);
  auto grow = [&](void *&p, size_t &cap, size_t need) -> bool {
    if (p && cap >= need)
      return true;
    vrx_free(p);
    p = nullptr;
    cap = 0;
    if (!dev->check(vrx_malloc(&p, need), this, "BVH allocation"))
      return false;
    cap = need;
    return true;
  };
  if (!grow(dNodes, capNodes, sz.nodes) || !grow(dLeaves, capLeaves, sz.leaves)
      || !grow(dIds, capIds, sz.leaf_ids))
    return -1;
  if (!dHeader && !dev->check(vrx_malloc(&dHeader, sz.header), this, "BVH header"))
    return -1;
  vrx_prims p{};
  p.type = ptype;
  p.n = n;
  p.d_position = (const float *)pos->device();
  p.d_index = deviceIndex();
  if ((useGenIdx || index) && !p.d_index && n)
    return -1;
  p.d_radius = radius ? (const float *)radius->device() : nullptr;
  p.global_radius = globalRadius;
  if (primId)
    primId->device();
  if (normal)
    normal->device();
  bvh.d_nodes = dNodes;
  bvh.d_leaves = dLeaves;
  bvh.d_leaf_ids = (uint32_t *)dIds;
  bvh.d_header = dHeader;
  void *ws = dev->workspace(sz.workspace);
  if (!ws)
    return -1;
  wideValid = false;
  bvh.d_wide = nullptr;
  if (!dev->check(vrx_bvh_build(&p, &bvh, ws, sz.workspace, s), this, "BVH build"))
    return -1;
  bvhDirty = false;
  return 0;
}

int Geometry::allocWide()
{
  const size_t need = vrx_bvh_wide_bytes(n);
  if (need > capWide) {
    vrx_free(dWide);
    dWide = nullptr;
    capWide = 0;
    if (!dev->check(vrx_malloc(&dWide, need), this, "4-wide node allocation"))
      return -1;
    capWide = need;
  }
  return 0;
}

int Geometry::buildWide(vrx_stream s)
{
  if (wideValid)
    return 0;
  if (allocWide() != 0)
    return -1;
  bvh.d_wide = dWide;
  if (!dev->check(vrx_bvh_widen(&bvh, n, s), this, "4-wide nodes"))
    return -1;
  wideValid = true;
  return 0;
}

// ---------------------------------------------------------------------------
// Material: matte (scene/surface/material/Matte.cpp:39-74,
// gpu/evalMaterialParameters.h:392-403 alpha modes)
// ---------------------------------------------------------------------------
void Material::commit()
{
  float c4[4] = {0.8f, 0.8f, 0.8f, 1.0f};
  if (!getRaw("color", ANARI_FLOAT32_VEC4, c4)) {
    float c3[3];
    if (getRaw("color", ANARI_FLOAT32_VEC3, c3)) {
      c4[0] = c3[0];
      c4[1] = c3[1];
      c4[2] = c3[2];
    }
  }
  color[0] = c4[0];
  color[1] = c4[1];
  color[2] = c4[2];
  float op = get<float>("opacity", ANARI_FLOAT32, 1.0f);
  std::string mode = getString("alphaMode", "opaque");
  float cutoff = get<float>("alphaCutoff", ANARI_FLOAT32, 0.5f);
  float in = c4[3] * op;
  if (mode == "blend")
    opacity = in;
  else if (mode == "mask")
    opacity = in < cutoff ? 0.0f : 1.0f;
  else
    opacity = 1.0f;
  if (subtype != "matte")
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "material subtype '%s' is not supported; only 'matte'", subtype.c_str());
}

// ---------------------------------------------------------------------------
// Surface (scene/surface/Surface.cpp:42-118)
// ---------------------------------------------------------------------------
void Surface::commit()
{
  geom = (Geometry *)getObject("geometry", ANARI_GEOMETRY);
  mat = (Material *)getObject("material", ANARI_MATERIAL);
  id = get<uint32_t>("id", ANARI_UINT32, ~0u);
  if (!geom)
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "missing 'geometry' on ANARISurface");
}

bool Surface::isValid() const { return geom && geom->isValid(); }

// ---------------------------------------------------------------------------
// SpatialField: structuredRegular (StructuredRegularField.cpp:90-212)
// ---------------------------------------------------------------------------
SpatialField::~SpatialField() { vrx_free(dData); }

void SpatialField::commit()
{
  valid = false;
  float o[3] = {0, 0, 0}, s[3] = {1, 1, 1};
  getRaw("origin", ANARI_FLOAT32_VEC3, o);
  getRaw("spacing", ANARI_FLOAT32_VEC3, s);
  memcpy(origin, o, 12);
  memcpy(spacing, s, 12);
  data = (Array *)getObject("data", ANARI_ARRAY3D);
  uploadedVersion = 0;
  if (subtype != "structuredRegular") {
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "spatial field subtype '%s' is not supported", subtype.c_str());
    return;
  }
  if (!data) {
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "missing required parameter 'data' on structuredRegular spatial field");
    return;
  }
  if (data->elem != ANARI_FLOAT32 && data->elem != ANARI_UFIXED8 && data->elem != ANARI_UINT8) {
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "invalid data array type encountered in structuredRegular spatial field(%s)",
        typeName(data->elem));
    return;
  }
  for (int k = 0; k < 3; k++)
    dims[k] = (uint32_t)data->dims[k];
  valid = true;
}

const float *SpatialField::deviceData()
{
  if (!valid)
    return nullptr;
  if (dData && uploadedVersion == data->version)
    return (const float *)dData;
  size_t n = data->count();
  if (data->elem == ANARI_FLOAT32) {
    // reuse the array's own device copy (no second 512 MiB copy)
    uploadedVersion = data->version;
    return (const float *)data->device();
  }
  std::vector<float> conv(n);
  const uint8_t *src = (const uint8_t *)data->host();
  for (size_t i = 0; i < n; i++)
    conv[i] = data->elem == ANARI_UFIXED8 ? (float)src[i] / 255.0f : (float)src[i];
  if (!dData && !dev->check(vrx_malloc(&dData, n * 4), this, "field allocation"))
    return nullptr;
  if (!dev->check(vrx_memcpy_sync(dData, conv.data(), n * 4, VRX_H2D), this, "field upload"))
    return nullptr;
  uploadedVersion = data->version;
  return (const float *)dData;
}

void SpatialField::bounds(float lo[3], float hi[3]) const
{
  for (int k = 0; k < 3; k++) {
    lo[k] = origin[k];
    hi[k] = origin[k] + ((float)dims[k] - 1.0f) * spacing[k];
  }
}

// ---------------------------------------------------------------------------
// Volume: transferFunction1D (scene/volume/TransferFunction1D.cpp:46-150,
// utility/colorMapHelpers.h:43-73)
// ---------------------------------------------------------------------------
Volume::~Volume()
{
  vrx_free(dLut);
  vrx_free(dGrid);
  vrx_free(dPairs);
}

static float posf(float v, float lo, float hi)
{
  float c = v < lo ? lo : (v > hi ? hi : v);
  return (c - lo) * (1.0f / (hi - lo));
}

static std::vector<float> linearPositions(size_t n, float lo, float hi)
{
  std::vector<float> p(n);
  p.front() = 0.0f;
  p.back() = 1.0f;
  float w = 1.0f / (float)(n - 1);
  for (size_t i = 1; i + 1 < n; i++)
    p[i] = p[i - 1] + w;
  for (auto &v : p)
    v = v * (hi - lo) + lo;
  return p;
}

static void interpolated(const float *vals, int ncomp, const std::vector<float> &positions, float lo,
    float hi, float pos, float *out)
{
  size_t n = positions.size();
  for (size_t i = 0; i + 1 < n; i++) {
    float r0 = posf(positions[i], lo, hi), r1 = posf(positions[i + 1], lo, hi);
    if (pos >= r0 && pos <= r1) {
      float t = posf(pos, r0, r1);
      for (int k = 0; k < ncomp; k++)
        out[k] = vals[ncomp * i + k] * (1.0f - t) + vals[ncomp * (i + 1) + k] * t;
      return;
    }
  }
  size_t j = pos <= posf(positions[0], lo, hi) ? 0 : n - 1;
  for (int k = 0; k < ncomp; k++)
    out[k] = vals[ncomp * j + k];
}

void Volume::commit()
{
  id = get<uint32_t>("id", ANARI_UINT32, ~0u); // Volume.cpp:45-48
  field = (SpatialField *)getObject("value", ANARI_SPATIAL_FIELD);
  if (subtype != "transferFunction1D" && subtype != "scivis") {
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "volume subtype '%s' is not supported", subtype.c_str());
    field = nullptr;
    return;
  }
  if (!field) {
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "missing parameter 'value' on transferFunction1D ANARIVolume");
    return;
  }
  Array *colorArr = arrayParam(this, "color");
  Array *opacityArr = arrayParam(this, "opacity");
  float uc[4] = {1, 1, 1, 1};
  float c3[3];
  if (getRaw("color", ANARI_FLOAT32_VEC3, c3)) {
    uc[0] = c3[0];
    uc[1] = c3[1];
    uc[2] = c3[2];
  }
  getRaw("color", ANARI_FLOAT32_VEC4, uc);
  float uo = get<float>("opacity", ANARI_FLOAT32, 1.0f) * uc[3];
  unitDistance = get<float>("unitDistance", ANARI_FLOAT32, 1.0f);
  float vr[2] = {0.0f, 1.0f};
  getRaw("valueRange", ANARI_FLOAT32_VEC2, vr);
  getRaw("valueRange", ANARI_FLOAT32_BOX1, vr);
  double vrd[2];
  if (getRaw("valueRange", ANARI_FLOAT64_BOX1, vrd)) {
    vr[0] = (float)vrd[0];
    vr[1] = (float)vrd[1];
  }
  valueRange[0] = vr[0];
  valueRange[1] = vr[1];
  // discritizeTFData (TransferFunction1D.cpp:101-150): 256 entries
  std::vector<float> cpos, opos;
  int ccomp = 0;
  if (colorArr) {
    if (colorArr->elem == ANARI_FLOAT32_VEC3)
      ccomp = 3;
    else if (colorArr->elem == ANARI_FLOAT32_VEC4)
      ccomp = 4;
    else
      report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
          "unusable tf1D color array type set (%s)", typeName(colorArr->elem));
    if (ccomp)
      cpos = linearPositions(colorArr->count(), vr[0], vr[1]);
  }
  if (opacityArr && opacityArr->elem == ANARI_FLOAT32)
    opos = linearPositions(opacityArr->count(), vr[0], vr[1]);
  for (int i = 0; i < 256; i++) {
    float p = (float)i / 255.0f;
    float c[4] = {uc[0], uc[1], uc[2], uc[3]};
    if (ccomp) {
      float tmp[4] = {0, 0, 0, 1};
      interpolated((const float *)colorArr->host(), ccomp, cpos, vr[0], vr[1], p, tmp);
      c[0] = tmp[0];
      c[1] = tmp[1];
      c[2] = tmp[2];
      c[3] = ccomp == 4 ? tmp[3] : 1.0f;
    }
    float o = uo;
    if (!opos.empty())
      interpolated((const float *)opacityArr->host(), 1, opos, vr[0], vr[1], p, &o);
    lut[4 * i + 0] = c[0];
    lut[4 * i + 1] = c[1];
    lut[4 * i + 2] = c[2];
    lut[4 * i + 3] = c[3] * o;
  }
  if (!dLut && !dev->check(vrx_malloc(&dLut, sizeof(lut)), this, "LUT allocation"))
    return;
  dev->check(vrx_memcpy(dLut, lut, sizeof(lut), VRX_H2D, dev->stream), this, "LUT upload");
}

// ---------------------------------------------------------------------------
// Group / Instance / World (scene/Group.cpp, Instance.cpp, World.cpp)
// ---------------------------------------------------------------------------
template <typename T>
static std::vector<T *> objectList(const Object *o, const char *name, ANARIDataType t)
{
  std::vector<T *> out;
  Array *a = arrayParam(o, name);
  if (!a || a->elem != t)
    return out;
  for (Object *x : a->objects)
    if (x)
      out.push_back((T *)x);
  return out;
}

// replaces an owning list: the new entries are retained before the old ones
// are released (an entry in both lists survives)
template <typename T>
static void holdList(std::vector<T *> &dst, std::vector<T *> src)
{
  for (T *o : src)
    retainInternal(o);
  for (T *o : dst)
    releaseInternal(o);
  dst.swap(src);
}

Group::~Group()
{
  holdList(surfaces, {});
  holdList(volumes, {});
  holdList(lights, {});
}

void Group::commit()
{
  holdList(surfaces, objectList<Surface>(this, "surface", ANARI_SURFACE));
  holdList(volumes, objectList<Volume>(this, "volume", ANARI_VOLUME));
  holdList(lights, objectList<Light>(this, "light", ANARI_LIGHT)); // Group.cpp:92
}

// Group.cpp:65-86: "bounds" (object space) is the union of the group's
// surface BVH boxes (built here if they are not yet) and volume boxes.
bool Group::getProperty(
    const std::string &name, ANARIDataType t, void *mem, uint64_t size, uint32_t mask)
{
  if (name == "bounds" && t == ANARI_FLOAT32_BOX3 && size >= 24) {
    dev->flushCommits();
    vrx_stream s = dev->stream;
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    auto extend = [&](const float *lo, const float *hi) {
      for (int k = 0; k < 3; k++) {
        b[k] = std::min(b[k], lo[k]);
        b[3 + k] = std::max(b[3 + k], hi[k]);
      }
    };
    for (Surface *sf : surfaces) {
      if (!sf || !sf->isValid())
        continue;
      Geometry *g = sf->geom;
      if (g->bvhDirty && g->buildBVH(s) != 0)
        return false;
      if (!dev->check(vrx_memcpy(&g->header, g->dHeader, sizeof(vrx_bvh_header), VRX_D2H, s),
              this, "header readback")
          || !dev->check(vrx_stream_sync(s), this, "stream sync"))
        return false;
      if (g->header.n > 0)
        extend(g->header.lo, g->header.hi);
    }
    for (Volume *v : volumes) {
      if (!v || !v->isValid())
        continue;
      float lo[3], hi[3];
      v->field->bounds(lo, hi);
      extend(lo, hi);
    }
    memcpy(mem, b, 24);
    return true;
  }
  return Object::getProperty(name, t, mem, size, mask);
}

// ---------------------------------------------------------------------------
// Light (scene/light/Light.cpp:47-54, Directional.cpp:40-47, Point.cpp:40-48,
// Spot.cpp:40-68)
// ---------------------------------------------------------------------------
void Light::commit()
{
  type = -1;
  float c[3] = {1, 1, 1};
  getRaw("color", ANARI_FLOAT32_VEC3, c);
  memcpy(color, c, 12);
  const float fmax = std::numeric_limits<float>::max();
  if (subtype == "directional") {
    type = VRX_LIGHT_DIRECTIONAL;
    float dv[3] = {0, 0, -1};
    getRaw("direction", ANARI_FLOAT32_VEC3, dv);
    const float inv = 1.0f / std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    for (int k = 0; k < 3; k++)
      direction[k] = dv[k] * inv;
    irradiance = std::clamp(get<float>("irradiance", ANARI_FLOAT32, 1.0f), 0.0f, fmax);
  } else if (subtype == "point" || subtype == "spot") {
    type = subtype == "point" ? VRX_LIGHT_POINT : VRX_LIGHT_SPOT;
    float pv[3] = {0, 0, 0};
    getRaw("position", ANARI_FLOAT32_VEC3, pv);
    memcpy(position, pv, 12);
    // "intensity", else "power", else 1
    const float power = get<float>("power", ANARI_FLOAT32, 1.0f);
    intensity = std::clamp(get<float>("intensity", ANARI_FLOAT32, power), 0.0f, fmax);
    if (type == VRX_LIGHT_SPOT) {
      float dv[3] = {0, 0, -1};
      getRaw("direction", ANARI_FLOAT32_VEC3, dv);
      memcpy(direction, dv, 12);
      const float opening = get<float>("openingAngle", ANARI_FLOAT32, 3.14159265358979f);
      const float falloff = get<float>("falloffAngle", ANARI_FLOAT32, 0.1f);
      const float inner = opening - 2.0f * falloff;
      if (inner < 0.0f)
        report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
            "falloffAngle should be smaller than half of openingAngle");
      cosOuter = std::cos(opening);
      cosInner = std::cos(inner);
    }
  } else {
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "light subtype '%s' is not supported (directional, point, spot)", subtype.c_str());
  }
}

vrx_light Light::record(const float *m) const
{
  static const float id3[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  if (!m)
    m = id3;
  vrx_light l{};
  l.type = type;
  // gpu/sampleLight.h:58-100: xfmVec / xfmPoint of the instance's matrix,
  // summed column by column
  auto vec = [&](const float *v, float *out) {
    for (int i = 0; i < 3; i++)
      out[i] = m[i] * v[0] + m[3 + i] * v[1] + m[6 + i] * v[2];
  };
  auto point = [&](const float *v, float *out) {
    for (int i = 0; i < 3; i++)
      out[i] = m[i] * v[0] + m[3 + i] * v[1] + m[6 + i] * v[2] + m[9 + i];
  };
  if (type == VRX_LIGHT_DIRECTIONAL) {
    const float nd[3] = {-direction[0], -direction[1], -direction[2]};
    vec(nd, l.dir);
    for (int k = 0; k < 3; k++)
      l.radiance[k] = color[k] * irradiance;
  } else {
    point(position, l.pos);
    for (int k = 0; k < 3; k++)
      l.radiance[k] = color[k] * intensity;
    if (type == VRX_LIGHT_SPOT) {
      memcpy(l.dir, direction, 12); // not transformed by the reference either
      l.cos_outer = cosOuter;
      l.cos_inner = cosInner;
    }
  }
  return l;
}

void Instance::commit()
{
  group = (Group *)getObject("group", ANARI_GROUP);
  id = get<uint32_t>("id", ANARI_UINT32, ~0u); // Instance.cpp:42-43
  float m4[16];
  float m34[12];
  if (getRaw("transform", ANARI_FLOAT32_MAT4, m4)) {
    // column-major 4x4 -> 3x4 (drop the last row)
    for (int c = 0; c < 4; c++)
      for (int r = 0; r < 3; r++)
        xfm[3 * c + r] = m4[4 * c + r];
  } else if (getRaw("transform", ANARI_FLOAT32_MAT3x4, m34)) {
    memcpy(xfm, m34, sizeof(xfm));
  } else {
    const float id3[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
    memcpy(xfm, id3, sizeof(xfm));
  }
  xfms.clear();
  ids.clear();
  // transform / id arrays (Instance.cpp:40-47, 80-97, 115-128)
  if (Array *xa = arrayParam(this, "transform")) {
    if (xa->elem != ANARI_FLOAT32_MAT4 || !xa->host()) {
      report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
          "'transform' array elements must be FLOAT32_MAT4; array ignored");
    } else {
      const float *m = (const float *)xa->host();
      const size_t n = xa->count();
      xfms.resize(12 * n);
      for (size_t i = 0; i < n; i++)
        for (int c = 0; c < 4; c++)
          for (int r = 0; r < 3; r++)
            xfms[12 * i + 3 * c + r] = m[16 * i + 4 * c + r];
      if (Array *ia = arrayParam(this, "id")) {
        if (ia->elem != ANARI_UINT32 || ia->count() < n || !ia->host())
          report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
              "'id' array must hold one UINT32 per transform; array ignored");
        else {
          const uint32_t *v = (const uint32_t *)ia->host();
          ids.assign(v, v + n);
        }
      }
    }
  }
  if (!group)
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT, "missing 'group' on ANARIInstance");
}

World::~World()
{
  freeScene();
  holdList(instances, {});
  holdList(surfaces, {});
  holdList(volumes, {});
  holdList(lights, {});
}

void World::freeScene()
{
  vrx_free(dGeoms);
  vrx_free(dEntries);
  vrx_free(dVolumes);
  vrx_free(dLights);
  dLights = nullptr;
  vrx_free(dTlasNodes);
  vrx_free(dTlasLeaves);
  vrx_free(dTlasIds);
  vrx_free(dTlasHeader);
  vrx_free(dTlasWide);
  dGeoms = dEntries = dVolumes = dTlasNodes = dTlasLeaves = dTlasIds = dTlasHeader = nullptr;
  dTlasWide = nullptr;
}

void World::commit()
{
  holdList(instances, objectList<Instance>(this, "instance", ANARI_INSTANCE));
  holdList(surfaces, objectList<Surface>(this, "surface", ANARI_SURFACE));
  holdList(volumes, objectList<Volume>(this, "volume", ANARI_VOLUME));
  holdList(lights, objectList<Light>(this, "light", ANARI_LIGHT)); // World.cpp:124
  id = get<uint32_t>("id", ANARI_UINT32, ~0u); // World.cpp:159 (zero instance id)
  sceneDirty = true;
}

bool World::needsRebuild() const { return sceneDirty || builtAt < dev->lastSceneFinalize; }

static bool isIdentity(const float *m)
{
  static const float id3[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  return memcmp(m, id3, sizeof(id3)) == 0;
}

// inverse of a 3x4 column-major transform, computed in FP64
static void invertXfm(const float *m, float *inv)
{
  double a = m[0], b = m[3], c = m[6];
  double d = m[1], e = m[4], f = m[7];
  double g = m[2], h = m[5], i = m[8];
  double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
  double det = a * A + b * B + c * C;
  double r = 1.0 / det;
  double im[9];
  im[0] = A * r;
  im[3] = -(b * i - c * h) * r;
  im[6] = (b * f - c * e) * r;
  im[1] = B * r;
  im[4] = (a * i - c * g) * r;
  im[7] = -(a * f - c * d) * r;
  im[2] = C * r;
  im[5] = -(a * h - b * g) * r;
  im[8] = (a * e - b * d) * r;
  double tx = m[9], ty = m[10], tz = m[11];
  for (int k = 0; k < 9; k++)
    inv[k] = (float)im[k];
  inv[9] = (float)(-(im[0] * tx + im[3] * ty + im[6] * tz));
  inv[10] = (float)(-(im[1] * tx + im[4] * ty + im[7] * tz));
  inv[11] = (float)(-(im[2] * tx + im[5] * ty + im[8] * tz));
}

static void xfmBox(const float *m, const float lo[3], const float hi[3], float out[6])
{
  out[0] = out[1] = out[2] = INFINITY;
  out[3] = out[4] = out[5] = -INFINITY;
  for (int c = 0; c < 8; c++) {
    float p[3] = {(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2]};
    for (int r = 0; r < 3; r++) {
      float q = ((m[r] * p[0] + m[3 + r] * p[1]) + m[6 + r] * p[2]) + m[9 + r];
      out[r] = std::min(out[r], q);
      out[3 + r] = std::max(out[3 + r], q);
    }
  }
}

// World::rebuildWorld (World.cpp:202-341): BLAS per geometry, instance entries,
// instance-level BVH, volume list, world bounds.
int World::rebuild()
{
  vrx_stream s = dev->stream;
  struct Occ {
    Surface *surf;
    const float *xfm;
    uint32_t inst;
  };
  std::vector<Occ> occ;
  for (Surface *sf : surfaces)
    if (sf && sf->isValid())
      occ.push_back({sf, nullptr, id});
  for (Instance *in : instances) {
    if (!in || !in->group)
      continue;
    for (size_t t = 0; t < in->numTransforms(); t++)
      for (Surface *sf : in->group->surfaces)
        if (sf && sf->isValid())
          occ.push_back({sf, in->xfmAt(t), in->idAt(t)});
  }
  std::vector<Volume *> vols;
  std::vector<uint32_t> volInst;
  for (Volume *v : volumes)
    if (v && v->isValid()) {
      vols.push_back(v);
      volInst.push_back(id);
    }
  for (Instance *in : instances) {
    if (!in || !in->group)
      continue;
    for (size_t t = 0; t < in->numTransforms(); t++)
      for (Volume *v : in->group->volumes) {
        if (!v || !v->isValid())
          continue;
        if (!isIdentity(in->xfmAt(t))) {
          report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_OPERATION,
              "instanced volumes with a non-identity transform are not supported; skipped");
          continue;
        }
        vols.push_back(v);
        volInst.push_back(in->idAt(t));
      }
  }

  vrx_event e0 = nullptr, e1 = nullptr;
  vrx_event_create(&e0);
  vrx_event_create(&e1);
  vrx_event_record(e0, s);
  // BLAS builds for dirty geometries
  std::vector<Geometry *> geoms;
  for (auto &o : occ)
    if (std::find(geoms.begin(), geoms.end(), o.surf->geom) == geoms.end())
      geoms.push_back(o.surf->geom);
  for (Geometry *g : geoms)
    if (g->bvhDirty && g->buildBVH(s) != 0)
      return -1;
  vrx_event_record(e1, s);
  // read back headers (heights, boxes)
  for (Geometry *g : geoms)
    if (!dev->check(vrx_memcpy(&g->header, g->dHeader, sizeof(vrx_bvh_header), VRX_D2H, s), this,
            "header readback"))
      return -1;
  dev->check(vrx_stream_sync(s), this, "stream sync");
  float ms = 0.0f;
  vrx_event_elapsed_ms(e0, e1, &ms);
  lastBuildMs = ms;

  // vrx_geom per surface occurrence (deduplicated by surface)
  std::vector<vrx_geom> hg;
  std::vector<Surface *> gsurf;
  std::vector<Geometry *> geomOf;
  std::vector<vrx_entry> he;
  uint32_t maxH = 0;
  int single = -2;
  float wb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (auto &o : occ) {
    auto it = std::find(gsurf.begin(), gsurf.end(), o.surf);
    uint32_t gi;
    if (it == gsurf.end()) {
      Geometry *g = o.surf->geom;
      vrx_geom vg{};
      vg.type = g->ptype;
      vg.n = g->n;
      vg.bvh = g->bvh;
      vg.bvh.d_wide = g->wideValid ? g->dWide : nullptr;
      vg.d_position = (const float *)g->pos->device();
      vg.d_index = g->deviceIndex();
      vg.d_radius = g->radius ? (const float *)g->radius->device() : nullptr;
      vg.global_radius = g->globalRadius;
      vg.d_primitive_id = g->primId ? (const uint32_t *)g->primId->device() : nullptr;
      vg.d_vertex_normal = g->normal ? (const float *)g->normal->device() : nullptr;
      vg.caps = g->caps;
      vg.cull_backfaces = g->cull;
      Material *m = o.surf->mat;
      vg.color[0] = m ? m->color[0] : 0.8f; // evalShading.h:61-81 defaults
      vg.color[1] = m ? m->color[1] : 0.8f;
      vg.color[2] = m ? m->color[2] : 0.8f;
      vg.opacity = m ? m->opacity : 1.0f;
      vg.surface_id = o.surf->id;
      gi = (uint32_t)hg.size();
      hg.push_back(vg);
      gsurf.push_back(o.surf);
      geomOf.push_back(g);
      maxH = std::max(maxH, g->header.height);
      single = single == -2 ? g->ptype : (single == g->ptype ? single : -1);
    } else {
      gi = (uint32_t)(it - gsurf.begin());
    }
    vrx_entry e{};
    static const float id3[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
    memcpy(e.xfm, o.xfm ? o.xfm : id3, sizeof(e.xfm));
    e.identity = isIdentity(e.xfm) ? 1u : 0u;
    if (e.identity)
      memcpy(e.inv, id3, sizeof(e.inv));
    else
      invertXfm(e.xfm, e.inv);
    e.geom = gi;
    e.inst_id = o.inst;
    he.push_back(e);
    const vrx_bvh_header &H = o.surf->geom->header;
    if (H.n > 0) {
      float bx[6];
      if (e.identity) {
        memcpy(bx, H.lo, 12);
        memcpy(bx + 3, H.hi, 12);
      } else
        xfmBox(e.xfm, H.lo, H.hi, bx);
      for (int k = 0; k < 3; k++) {
        wb[k] = std::min(wb[k], bx[k]);
        wb[3 + k] = std::max(wb[3 + k], bx[3 + k]);
      }
    }
  }
  std::vector<vrx_volume> hv;
  uint64_t maxVox = 0;
  uint32_t nGrids = 0;
  for (size_t i = 0; i < vols.size(); i++) {
    Volume *v = vols[i];
    vrx_volume vv{};
    vv.d_data = v->field->deviceData();
    memcpy(vv.dims, v->field->dims, 12);
    memcpy(vv.origin, v->field->origin, 12);
    memcpy(vv.spacing, v->field->spacing, 12);
    vv.d_lut = (const float *)v->dLut;
    vv.value_range[0] = v->valueRange[0];
    vv.value_range[1] = v->valueRange[1];
    vv.unit_distance = v->unitDistance;
    vv.id = v->id;
    vv.inst_id = volInst[i];
    vv.d_majorant = nullptr;
    if (vv.d_data && vv.d_lut) {
      const size_t cells = vrx_volume_grid_cells(&vv);
      if (v->gridCells != cells) {
        vrx_free(v->dGrid);
        v->dGrid = nullptr;
        v->gridCells = 0;
        if (dev->check(vrx_malloc(&v->dGrid, cells * sizeof(float)), this, "majorant grid"))
          v->gridCells = cells;
      }
      std::vector<float> mg(cells);
      if (v->dGrid && dev->check(vrx_volume_majorant_grid(&vv, (float *)v->dGrid, s), this,
                                 "majorant grid")
          && dev->check(vrx_memcpy(mg.data(), v->dGrid, cells * sizeof(float), VRX_D2H, s), this,
                 "majorant readback")
          && dev->check(vrx_stream_sync(s), this, "stream sync")) {
        const size_t empty = (size_t)std::count(mg.begin(), mg.end(), 0.0f);
        v->emptyCellFraction = cells ? (float)empty / (float)cells : 0.0f;
        if (8 * empty >= cells && empty > 0)
          vv.d_majorant = (const float *)v->dGrid;
      }
    }
    // a dense field alone in its world is marched by the four-lane kernel,
    // which reads the y-pair copy of the field (8 B per voxel)
    vv.d_ypairs = nullptr;
    const uint64_t nvox = (uint64_t)vv.dims[0] * vv.dims[1] * vv.dims[2];
    if (vv.d_data && !vv.d_majorant && vols.size() == 1 && he.empty() && vv.dims[0] >= 2
        && nvox < (1ull << 29)) {
      if (v->pairBytes != nvox * 8) {
        vrx_free(v->dPairs);
        v->dPairs = nullptr;
        v->pairBytes = 0;
        if (dev->check(vrx_malloc(&v->dPairs, nvox * 8), this, "volume y-pairs"))
          v->pairBytes = nvox * 8;
      }
      if (v->dPairs && dev->check(vrx_volume_build_ypairs(&vv, (float *)v->dPairs, s), this,
                          "volume y-pairs"))
        vv.d_ypairs = (const float *)v->dPairs;
    }
    nGrids += vv.d_majorant ? 1u : 0u;
    hv.push_back(vv);
    maxVox = std::max(maxVox, (uint64_t)vv.dims[0] * vv.dims[1] * vv.dims[2]);
    float lo[3], hi[3];
    v->field->bounds(lo, hi);
    for (int k = 0; k < 3; k++) {
      wb[k] = std::min(wb[k], lo[k]);
      wb[3 + k] = std::max(wb[3 + k], hi[k]);
    }
  }
  // light instances (World.cpp:462-486): the world's own lights, then each
  // instance's group lights under that instance's transform
  std::vector<vrx_light> hl;
  for (Light *l : lights)
    if (l && l->isValid())
      hl.push_back(l->record(nullptr));
  for (Instance *in : instances) {
    if (!in || !in->group)
      continue;
    for (size_t t = 0; t < in->numTransforms(); t++)
      for (Light *l : in->group->lights)
        if (l && l->isValid())
          hl.push_back(l->record(in->xfmAt(t)));
  }
  freeScene();
  auto upload = [&](void *&dst, const void *src, size_t bytes) -> bool {
    if (!bytes)
      return true;
    if (!dev->check(vrx_malloc(&dst, bytes), this, "scene allocation"))
      return false;
    return dev->check(vrx_memcpy(dst, src, bytes, VRX_H2D, s), this, "scene upload");
  };
  if (!upload(dGeoms, hg.data(), hg.size() * sizeof(vrx_geom))
      || !upload(dEntries, he.data(), he.size() * sizeof(vrx_entry))
      || !upload(dVolumes, hv.data(), hv.size() * sizeof(vrx_volume))
      || !upload(dLights, hl.data(), hl.size() * sizeof(vrx_light)))
    return -1;
  memset(&scene, 0, sizeof(scene));
  scene.d_geoms = (const vrx_geom *)dGeoms;
  scene.num_geoms = (uint32_t)hg.size();
  scene.d_entries = (const vrx_entry *)dEntries;
  scene.num_entries = (uint32_t)he.size();
  scene.d_volumes = (const vrx_volume *)dVolumes;
  scene.num_volumes = (uint32_t)hv.size();
  scene.d_lights = (const vrx_light *)dLights;
  scene.num_lights = (uint32_t)hl.size();
  scene.num_majorant_grids = nGrids;
  scene.max_blas_height = maxH;
  scene.single_type = single >= 0 ? single : -1;
  scene.max_volume_voxels = maxVox;
  if (he.size() > 1) {
    vrx_bvh_sizes sz;
    vrx_tlas_query_sizes((uint32_t)he.size(), &sz);
    if (!dev->check(vrx_malloc(&dTlasNodes, sz.nodes), this, "tlas")
        || !dev->check(vrx_malloc(&dTlasLeaves, sz.leaves), this, "tlas")
        || !dev->check(vrx_malloc(&dTlasIds, sz.leaf_ids), this, "tlas")
        || !dev->check(vrx_malloc(&dTlasHeader, sz.header), this, "tlas"))
      return -1;
    scene.tlas.d_nodes = dTlasNodes;
    scene.tlas.d_leaves = dTlasLeaves;
    scene.tlas.d_leaf_ids = (uint32_t *)dTlasIds;
    scene.tlas.d_header = dTlasHeader;
    void *ws = dev->workspace(sz.workspace);
    if (!ws)
      return -1;
    vrx_event_record(e0, s);
    if (!dev->check(vrx_tlas_build(&scene, &scene.tlas, ws, sz.workspace, s), this, "tlas build"))
      return -1;
    vrx_event_record(e1, s);
    vrx_bvh_header th;
    dev->check(vrx_memcpy(&th, dTlasHeader, sizeof(th), VRX_D2H, s), this, "tlas header");
    dev->check(vrx_stream_sync(s), this, "stream sync");
    float ms2 = 0.0f;
    vrx_event_elapsed_ms(e0, e1, &ms2);
    lastBuildMs += ms2;
    scene.tlas_height = th.height;
  }
  vrx_event_destroy(e0);
  vrx_event_destroy(e1);
  memcpy(bounds, wb, sizeof(bounds));
  hostGeoms = std::move(hg);
  hostGeomOf = std::move(geomOf);
  sceneDirty = false;
  builtAt = dev->tick();
  return 0;
}

int World::ensureWide()
{
  if (scene.wide || scene.num_entries == 0)
    return 0;
  vrx_stream s = dev->stream;
  // allocate first: the timed region covers only the widen kernels
  for (Geometry *g : hostGeomOf)
    if (!g->wideValid && g->allocWide() != 0)
      return -1;
  if (scene.num_entries > 1 && !dTlasWide
      && !dev->check(vrx_malloc(&dTlasWide, vrx_bvh_wide_bytes(scene.num_entries)), this,
          "4-wide instance nodes"))
    return -1;
  vrx_event e0 = nullptr, e1 = nullptr;
  vrx_event_create(&e0);
  vrx_event_create(&e1);
  vrx_event_record(e0, s);
  for (size_t i = 0; i < hostGeoms.size(); i++) {
    Geometry *g = hostGeomOf[i];
    if (g->buildWide(s) != 0)
      return -1;
    hostGeoms[i].bvh.d_wide = g->dWide;
  }
  if (scene.num_entries > 1) {
    scene.tlas.d_wide = dTlasWide;
    if (!dev->check(vrx_bvh_widen(&scene.tlas, scene.num_entries, s), this, "4-wide instance nodes"))
      return -1;
  }
  vrx_event_record(e1, s);
  if (!hostGeoms.empty()
      && !dev->check(vrx_memcpy(dGeoms, hostGeoms.data(), hostGeoms.size() * sizeof(vrx_geom),
                         VRX_H2D, s),
          this, "scene upload"))
    return -1;
  dev->check(vrx_stream_sync(s), this, "stream sync");
  float ms = 0.0f;
  vrx_event_elapsed_ms(e0, e1, &ms);
  lastWideMs = ms;
  vrx_event_destroy(e0);
  vrx_event_destroy(e1);
  scene.wide = 1;
  return 0;
}

bool World::getProperty(
    const std::string &name, ANARIDataType t, void *mem, uint64_t size, uint32_t mask)
{
  if (name == "bounds" && t == ANARI_FLOAT32_BOX3 && size >= 24) {
    // World.cpp:100-118: flush commits and rebuild before answering
    dev->flushCommits();
    if (needsRebuild() && rebuild() != 0)
      return false;
    memcpy(mem, bounds, 24);
    return true;
  }
  if (name == "bvhBuildDuration" && t == ANARI_FLOAT32 && size >= 4) {
    float sec = lastBuildMs / 1000.0f;
    memcpy(mem, &sec, 4);
    return true;
  }
  // extension: the device-side scene description (vrx_scene of vrx_hip.h),
  // valid until the world is next rebuilt; lets a caller drive the batched
  // multi-view entry points on a scene built through this API
  if (name == "bvhWidenDuration" && t == ANARI_FLOAT32 && size >= 4) {
    float sec = lastWideMs / 1000.0f;
    memcpy(mem, &sec, 4);
    return true;
  }
  // "vrxSceneWide" also builds the 4-wide nodes (node width 4)
  if ((name == "vrxScene" || name == "vrxSceneWide") && t == ANARI_VOID_POINTER
      && size >= sizeof(void *)) {
    dev->flushCommits();
    if (needsRebuild() && rebuild() != 0)
      return false;
    if (name == "vrxSceneWide" && ensureWide() != 0)
      return false;
    const void *p = &scene;
    memcpy(mem, &p, sizeof(p));
    return true;
  }
  return Object::getProperty(name, t, mem, size, mask);
}

// ---------------------------------------------------------------------------
// Camera: perspective (camera/Perspective.cpp:40-72, Camera.cpp:68-76)
// ---------------------------------------------------------------------------
void Camera::commit()
{
  float region[4] = {0, 0, 1, 1};
  getRaw("imageRegion", ANARI_FLOAT32_BOX2, region);
  float pos[3] = {0, 0, 0}, dir[3] = {0, 0, 1}, up[3] = {0, 1, 0};
  getRaw("position", ANARI_FLOAT32_VEC3, pos);
  getRaw("direction", ANARI_FLOAT32_VEC3, dir);
  getRaw("up", ANARI_FLOAT32_VEC3, up);
  float fovy = get<float>("fovy", ANARI_FLOAT32, 1.04719755f); // radians(60)
  float aspect = get<float>("aspect", ANARI_FLOAT32, 1.0f);
  if (subtype != "perspective")
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "camera subtype '%s' is not supported", subtype.c_str());
  vrx_camera_perspective(pos, dir, up, fovy, aspect, region, &cam);
}

// ---------------------------------------------------------------------------
// Renderer (renderer/Renderer.cpp:97-207, Raycast.cpp:45-50, AmbientOcclusion.cpp:44-56,
// DirectLight.cpp:46-64)
// ---------------------------------------------------------------------------
Renderer::Renderer(Device *d, const std::string &sub) : Object(d, ANARI_RENDERER, sub)
{
  // Renderer.cpp:114-121: "default" is the directLight renderer
  if (sub == "raycast")
    kind = VRX_RENDER_RAYCAST;
  else if (sub == "ao")
    kind = VRX_RENDER_AO;
  else if (sub == "directLight" || sub == "default")
    kind = VRX_RENDER_DIRECT;
}

void Renderer::commit()
{
  memset(&r, 0, sizeof(r));
  r.kind = kind < 0 ? VRX_RENDER_RAYCAST : kind;
  float bg[4] = {0, 0, 0, 1};
  getRaw("background", ANARI_FLOAT32_VEC4, bg);
  memcpy(r.background, bg, 16);
  float ac[3] = {1, 1, 1};
  getRaw("ambientColor", ANARI_FLOAT32_VEC3, ac);
  memcpy(r.ambient_color, ac, 12);
  // default ambient radiance: 1 for ao, 0 for directLight and raycast
  // (AmbientOcclusion.cpp:44, DirectLight.cpp:46, Renderer.h:56)
  r.ambient_radiance =
      get<float>("ambientRadiance", ANARI_FLOAT32, kind == VRX_RENDER_AO ? 1.0f : 0.0f);
  // directLight's lightFalloff is clamped to [0, 1] and not used in shading
  lightFalloff = std::clamp(get<float>("lightFalloff", ANARI_FLOAT32, 1.0f), 0.0f, 1.0f);
  r.pixel_samples = get<int>("pixelSamples", ANARI_INT32, 1);
  r.ambient_samples = std::clamp(get<int>("ambientSamples", ANARI_INT32, 1), 0, 256);
  r.ao_distance = get<float>("ambientOcclusionDistance", ANARI_FLOAT32, 1e20f);
  r.cull_backfaces = get<int>("cullTriangleBackfaces", ANARI_BOOL, 0);
  float rate = std::clamp(get<float>("volumeSamplingRate", ANARI_FLOAT32, 0.125f), 1e-3f, 10.0f);
  r.inv_volume_sampling_rate = 1.0f / rate;
  r.stack_lds = get<int>("stackDepthLDS", ANARI_INT32, 0);
  r.vol_unroll = get<int>("volumeUnroll", ANARI_INT32, 0);
  // 4-wide nodes by default: on MI355X they trace 6-22% faster than BVH2
  // (the world builds them once, lazily, from the BVH2)
  r.node_width = get<int>("nodeWidth", ANARI_INT32, 4) == 2 ? 2 : 4;
  // extension: skip empty macro cells in volume marches (same image)
  r.skip_empty = get<int>("emptySpaceSkipping", ANARI_BOOL, 1) != 0 ? 1 : 0;
  // extension: the reference always jitters volume samples
  // (volumeIntegration.h:83,120); here it is opt-in, off by default
  r.volume_jitter = get<int>("volumeJitter", ANARI_BOOL, 0) != 0 ? 1 : 0;
  // extension: raycast primary walks longer than this many steps finish in a
  // second launch, 16 lanes per ray (same pixels); unset: tuned per frame
  r.defer_steps = std::max(get<int>("longRaySteps", ANARI_INT32, 0), 0);
  tuneWidth = find("nodeWidth") == nullptr;
  tuneStack = find("stackDepthLDS") == nullptr;
  tuneDefer = find("longRaySteps") == nullptr && r.kind == VRX_RENDER_RAYCAST;
  sampleLimit = get<int>("sampleLimit", ANARI_INT32, 128);
  if (kind == VRX_RENDER_RAYCAST)
    sampleLimit = 1; // single-shot renderer
  if (kind < 0)
    report(ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "renderer subtype '%s' is not supported (raycast, ao, directLight)", subtype.c_str());
}

bool Renderer::getProperty(
    const std::string &name, ANARIDataType t, void *mem, uint64_t size, uint32_t mask)
{
  // extension: the committed vrx_renderer (vrx_hip.h), valid until the next
  // commit of this renderer
  if (name == "vrxRenderer" && t == ANARI_VOID_POINTER && size >= sizeof(void *)) {
    dev->flushCommits();
    const void *p = &r;
    memcpy(mem, &p, sizeof(p));
    return true;
  }
  return Object::getProperty(name, t, mem, size, mask);
}

// ---------------------------------------------------------------------------
// Frame (frame/Frame.cu:75-310, 312-443, 559-660)
// ---------------------------------------------------------------------------
Frame::Frame(Device *d) : Object(d, ANARI_FRAME, "")
{
  vrx_event_create(&evStart);
  vrx_event_create(&evEnd);
  vrx_event_create(&evDone);
  void *p = nullptr;
  // rays traced, volume samples and the traversal stack-overflow word,
  // written by vrx_frame_collect_counters after each frame
  if (vrx_host_malloc(&p, 32) == 0) {
    hReadback = (uint64_t *)p;
    memset(hReadback, 0, 32);
    void *d = nullptr;
    if (vrx_host_device_pointer(p, &d) == 0)
      dReadback = (uint64_t *)d;
  }
}

Frame::~Frame()
{
  if (rendered)
    vrx_event_sync(evDone);
  freeBuffers();
  vrx_event_destroy(evStart);
  vrx_event_destroy(evEnd);
  vrx_event_destroy(evDone);
  vrx_host_free(hReadback);
}

void Frame::freeBuffers()
{
  for (void **p : {&dAccum, &dColor, &dDepth, &dPrim, &dObj, &dInst, &dScratch, &dCounters}) {
    vrx_free(*p);
    *p = nullptr;
  }
  scratchBytes = 0;
  allocated = false;
}

bool Frame::isValid() const
{
  return renderer && renderer->isValid() && camera && camera->isValid() && world && world->isValid();
}

void Frame::commit()
{
  renderer = (Renderer *)getObject("renderer", ANARI_RENDERER);
  camera = (Camera *)getObject("camera", ANARI_CAMERA);
  world = (World *)getObject("world", ANARI_WORLD);
  const Param *cb = find("frameCompletionCallback");
  callback = cb && !cb->bytes.empty() ? *(ANARIFrameCompletionCallback *)cb->bytes.data() : nullptr;
  const Param *cbu = find("frameCompletionCallbackUserData");
  callbackUser = cbu && !cbu->bytes.empty() ? *(const void **)cbu->bytes.data() : nullptr;
  uint32_t sz[2] = {10, 10};
  getRaw("size", ANARI_UINT32_VEC2, sz);
  size[0] = sz[0];
  size[1] = sz[1];
  colorType = get<ANARIDataType>("channel.color", ANARI_DATA_TYPE, ANARI_UFIXED8_RGBA_SRGB);
  depthType = get<ANARIDataType>("channel.depth", ANARI_DATA_TYPE, ANARI_UNKNOWN);
  primType = get<ANARIDataType>("channel.primitiveId", ANARI_DATA_TYPE, ANARI_UNKNOWN);
  objType = get<ANARIDataType>("channel.objectId", ANARI_DATA_TYPE, ANARI_UNKNOWN);
  instType = get<ANARIDataType>("channel.instanceId", ANARI_DATA_TYPE, ANARI_UNKNOWN);
  // Frame.cu:119-122: any id channel implies depth
  if (primType == ANARI_UINT32 || objType == ANARI_UINT32 || instType == ANARI_UINT32)
    depthType = ANARI_FLOAT32;
  freeBuffers();
  nextReset = true;
}

void Frame::allocate()
{
  size_t np = (size_t)size[0] * size[1];
  bool f32 = colorType == ANARI_FLOAT32_VEC4;
  bool ok = dev->check(vrx_malloc(&dAccum, np * 16), this, "frame buffer")
      && dev->check(vrx_malloc(&dColor, np * (f32 ? 16 : 4)), this, "frame buffer")
      && dev->check(vrx_malloc(&dCounters, VRX_COUNTER_BYTES), this, "frame counters")
      && dev->check(vrx_fill_u32(dCounters, 0u, VRX_COUNTER_BYTES / 4, dev->stream), this,
          "frame counters");
  if (ok && depthType == ANARI_FLOAT32)
    ok = dev->check(vrx_malloc(&dDepth, np * 4), this, "depth buffer");
  if (ok && primType == ANARI_UINT32)
    ok = dev->check(vrx_malloc(&dPrim, np * 4), this, "primitiveId buffer");
  if (ok && objType == ANARI_UINT32)
    ok = dev->check(vrx_malloc(&dObj, np * 4), this, "objectId buffer");
  if (ok && instType == ANARI_UINT32)
    ok = dev->check(vrx_malloc(&dInst, np * 4), this, "instanceId buffer");
  allocated = ok;
}

void Frame::wait()
{
  if (!rendered)
    return;
  vrx_event_sync(evDone);
  // the traversal stacks are sized from the BVH heights, so this flag stays
  // clear; it is checked so that a violated bound can never pass silently
  if (hReadback && (uint32_t)hReadback[2] != 0 && !stackReported) {
    report(ANARI_SEVERITY_ERROR, ANARI_STATUS_UNKNOWN_ERROR,
        "traversal stack overflow in the last frame: results are incomplete");
    stackReported = true;
  }
}

void Frame::Sync::waitFor(uint64_t gen)
{
  std::unique_lock<std::mutex> lk(m);
  cv.wait(lk, [&] { return done >= gen; });
}

bool Frame::callbacksDone()
{
  std::lock_guard<std::mutex> lk(sync->m);
  return sync->done >= callbacksIssued;
}

bool Frame::ready()
{
  if (!rendered)
    return true;
  return vrx_event_query(evDone) == 1 && callbacksDone();
}

// step budget of the long-ray pass the frame autotuner tries (raycast of
// one-level scenes with one primitive type)
constexpr int kLongRaySteps = 32;

void Frame::pickLaunch(vrx_renderer &rr)
{
  uint64_t key = (uint64_t)(uintptr_t)world ^ (world->builtAt << 20) ^ (uint64_t)(uintptr_t)renderer
      ^ (renderer->lastFinalized << 40) ^ ((uint64_t)size[0] << 8) ^ ((uint64_t)size[1] << 28);
  if (key != tune.key) {
    tune = Tune{};
    tune.key = key;
    const int widths[2] = {4, 2}, sds[2] = {16, 32};
    const int nw = renderer->tuneWidth ? 2 : 1, ns = renderer->tuneStack ? 2 : 1;
    // the long-ray pass applies to one-level scenes of one primitive type;
    // it is tried once, on the fastest node width and stack depth
    const int st = world->scene.single_type;
    tune.deferPending = renderer->tuneDefer && world->scene.num_entries == 1
        && world->scene.num_volumes == 0
        && (st == VRX_TRIANGLE || st == VRX_SPHERE || st == VRX_CYLINDER);
    for (int a = 0; a < ns; a++)
      for (int b = 0; b < nw; b++) {
        tune.width[tune.n] = renderer->tuneWidth ? widths[b] : rr.node_width;
        tune.sd[tune.n] = renderer->tuneStack ? sds[a] : rr.stack_lds;
        tune.defer[tune.n] = tune.deferPending ? 0 : rr.defer_steps;
        tune.n++;
      }
  } else if (tune.best < 0 && tune.last >= 0) {
    // the previous render used candidate `last`; wait() has made its events
    // complete
    float ms = 0.0f;
    if (vrx_event_elapsed_ms(evStart, evEnd, &ms) == 0)
      tune.ms[tune.last] = ms;
    if (tune.next >= tune.n) {
      int b = 0;
      for (int i = 1; i < tune.n; i++)
        if (tune.ms[i] < tune.ms[b])
          b = i;
      if (tune.deferPending) {
        // one more candidate: the fastest launch with the long-ray pass
        tune.deferPending = false;
        tune.width[tune.n] = tune.width[b];
        tune.sd[tune.n] = tune.sd[b];
        tune.defer[tune.n] = kLongRaySteps;
        tune.n++;
      } else {
        tune.best = b;
      }
    }
  }
  const int use = tune.best >= 0 ? tune.best : tune.next;
  if (tune.best < 0)
    tune.next++;
  tune.last = tune.best >= 0 ? -1 : use;
  rr.node_width = tune.width[use];
  rr.stack_lds = tune.sd[use];
  rr.defer_steps = tune.defer[use];
}

void Frame::render()
{
  wait();
  dev->flushCommits();
  if (world && world->needsRebuild()) {
    if (world->rebuild() != 0)
      return;
  }
  if (!isValid()) {
    const char *why = !renderer ? "missing ANARIRenderer"
        : !renderer->isValid()  ? "invalid ANARIRenderer"
        : !camera               ? "missing ANARICamera"
        : !camera->isValid()    ? "invalid ANARICamera"
        : !world                ? "missing ANARIWorld"
                                : "invalid ANARIWorld";
    report(ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_OPERATION,
        "skipping render of incomplete or invalid frame object -- issue: %s", why);
    return;
  }
  if (!allocated)
    allocate();
  if (!allocated)
    return;
  // Frame.cu:574-588 accumulation reset on any finalize since the last frame
  if (dev->lastFinalize > lastSeenFinalize) {
    lastSeenFinalize = dev->lastFinalize;
    nextReset = true;
  }
  const int spp = renderer->spp();
  if (!nextReset && renderer->sampleLimit > 0 && frameID >= (uint32_t)renderer->sampleLimit)
    return; // Frame.cu:251-253
  if (nextReset) {
    frameID = 0;
    nextReset = false;
  } else {
    frameID += (uint32_t)spp;
  }
  vrx_renderer rr = renderer->r;
  pickLaunch(rr);
  if (rr.node_width == 4 && world->ensureWide() != 0)
    return;
  vrx_frame f{};
  f.width = size[0];
  f.height = size[1];
  f.color_format = colorType == ANARI_FLOAT32_VEC4 ? VRX_COLOR_F32
      : colorType == ANARI_UFIXED8_RGBA_SRGB        ? VRX_COLOR_SRGB8
                                                    : VRX_COLOR_UNORM8;
  f.frame_id = frameID;
  f.d_accum = (float *)dAccum;
  f.d_color_f32 = colorType == ANARI_FLOAT32_VEC4 ? (float *)dColor : nullptr;
  f.d_color_u32 = colorType == ANARI_FLOAT32_VEC4 ? nullptr : (uint32_t *)dColor;
  f.d_depth = (float *)dDepth;
  f.d_prim_id = (uint32_t *)dPrim;
  f.d_obj_id = (uint32_t *)dObj;
  f.d_inst_id = (uint32_t *)dInst;
  size_t need = vrx_render_scratch_bytes(&world->scene, &rr, f.width, f.height);
  if (need > scratchBytes) {
    vrx_free(dScratch);
    dScratch = nullptr;
    scratchBytes = 0;
    if (!dev->check(vrx_malloc(&dScratch, need), this, "render scratch"))
      return;
    scratchBytes = need;
    vrx_fill_u32(dScratch, 0u, 64, dev->stream);
  }
  f.d_scratch = dScratch;
  f.scratch_bytes = scratchBytes;
  f.d_counters = (uint64_t *)dCounters;
  vrx_event_record(evStart, dev->stream);
  dev->check(vrx_render(&world->scene, &camera->cam, &rr, &f, dev->stream), this, "render");
  vrx_event_record(evEnd, dev->stream);
  // totals to the pinned read-back words, slots cleared for the next frame
  if (dReadback)
    dev->check(vrx_frame_collect_counters(&f, dReadback, dev->stream), this, "frame counters");
  if (callback) {
    // Frame.cu:295-304 calls the callback from cudaLaunchHostFunc; here the
    // host function only hands the job to the device's callback thread
    dev->startCallbackThread();
    CallbackJob *job = new CallbackJob{dev, callback, callbackUser, this, sync, callbacksIssued + 1};
    if (dev->check(vrx_launch_host_func(
                       dev->stream, [](void *p) { ((CallbackJob *)p)->dev->postCallback((CallbackJob *)p); },
                       job),
            this, "completion callback"))
      callbacksIssued++;
    else
      delete job;
  }
  vrx_event_record(evDone, dev->stream);
  rendered = true;
}

const void *Frame::map(const std::string &channel, uint32_t *w, uint32_t *h, ANARIDataType *pt)
{
  wait();
  *pt = ANARI_UNKNOWN;
  const size_t np = (size_t)size[0] * size[1];
  auto fetch = [&](void *d, std::vector<uint8_t> &hb, size_t bytes) -> const void * {
    if (!d)
      return nullptr;
    hb.resize(bytes);
    if (!dev->check(vrx_memcpy_sync(hb.data(), d, bytes, VRX_D2H), this, "frame download"))
      return nullptr;
    return hb.data();
  };
  const bool f32 = colorType == ANARI_FLOAT32_VEC4;
  const void *ret = nullptr;
  ANARIDataType t = ANARI_UNKNOWN;
  bool gpu = false;
  std::string ch = channel;
  for (const char *suf : {"GPU", "CUDA", "HIP"}) {
    size_t L = strlen(suf);
    if (ch.size() > L && ch.compare(ch.size() - L, L, suf) == 0) {
      ch = ch.substr(0, ch.size() - L);
      gpu = true;
      break;
    }
  }
  if (!allocated)
    return nullptr;
  if (ch == "channel.color") {
    t = colorType;
    ret = gpu ? dColor : fetch(dColor, hColor, np * (f32 ? 16 : 4));
  } else if (ch == "channel.depth" && dDepth) {
    t = ANARI_FLOAT32;
    ret = gpu ? dDepth : fetch(dDepth, hDepth, np * 4);
  } else if (ch == "channel.primitiveId" && dPrim) {
    t = ANARI_UINT32;
    ret = gpu ? dPrim : fetch(dPrim, hPrim, np * 4);
  } else if (ch == "channel.objectId" && dObj) {
    t = ANARI_UINT32;
    ret = gpu ? dObj : fetch(dObj, hObj, np * 4);
  } else if (ch == "channel.instanceId" && dInst) {
    t = ANARI_UINT32;
    ret = gpu ? dInst : fetch(dInst, hInst, np * 4);
  }
  if (ret) {
    *w = size[0];
    *h = size[1];
    *pt = t;
  }
  return ret;
}

bool Frame::getProperty(const std::string &name, ANARIDataType t, void *mem, uint64_t sz, uint32_t mask)
{
  if (name == "duration" && t == ANARI_FLOAT32 && sz >= 4) {
    if (mask & ANARI_WAIT)
      wait();
    float ms = 0.0f;
    if (rendered)
      vrx_event_elapsed_ms(evStart, evEnd, &ms);
    duration = ms / 1000.0f;
    memcpy(mem, &duration, 4);
    return true;
  }
  if (name == "numSamples" && t == ANARI_INT32 && sz >= 4) {
    if (mask & ANARI_WAIT)
      wait();
    int v = (int)frameID;
    memcpy(mem, &v, 4);
    return true;
  }
  if (name == "nextFrameReset" && t == ANARI_BOOL && sz >= 4) {
    if (mask & ANARI_WAIT)
      wait();
    int v = (nextReset || dev->lastFinalize > lastSeenFinalize) ? 1 : 0;
    memcpy(mem, &v, 4);
    return true;
  }
  if ((name == "tunedNodeWidth" || name == "tunedStackDepth" || name == "tunedLongRaySteps")
      && t == ANARI_INT32 && sz >= 4) {
    // the launch configuration the autotuner settled on (0 while tuning)
    int v = 0;
    if (tune.best >= 0)
      v = name == "tunedNodeWidth" ? tune.width[tune.best]
          : name == "tunedStackDepth" ? tune.sd[tune.best] : tune.defer[tune.best];
    memcpy(mem, &v, 4);
    return true;
  }
  if (name == "raysTraced" && t == ANARI_UINT64 && sz >= 8) {
    wait();
    const uint64_t v = hReadback ? hReadback[0] : 0;
    memcpy(mem, &v, 8);
    return true;
  }
  if (name == "volumeSamples" && t == ANARI_UINT64 && sz >= 8) {
    wait();
    const uint64_t v = hReadback ? hReadback[1] : 0;
    memcpy(mem, &v, 8);
    return true;
  }
  return Object::getProperty(name, t, mem, sz, mask);
}

// ---------------------------------------------------------------------------
// Device (VisRTXDevice.cpp:102-762)
// ---------------------------------------------------------------------------
Device::Device(ANARIStatusCallback cb, const void *user) : statusCB(cb), statusUser(user) {}

Device::~Device()
{
  if (stream)
    vrx_stream_sync(stream);
  // every host function has run: the queue holds all outstanding callbacks,
  // which the thread runs before it ends
  if (cbThread.joinable()) {
    {
      std::lock_guard<std::mutex> lk(cbMutex);
      cbStop = true;
    }
    cbCv.notify_all();
    if (cbThread.get_id() == std::this_thread::get_id())
      cbThread.detach(); // released from inside a callback
    else
      cbThread.join();
  }
  for (Object *o : commitQueue)
    releaseInternal(o);
  commitQueue.clear();
  vrx_free(dWorkspace);
  if (stream)
    vrx_stream_destroy(stream);
}

void Device::startCallbackThread()
{
  if (!cbThread.joinable())
    cbThread = std::thread([this] { callbackLoop(); });
}

void Device::postCallback(CallbackJob *j)
{
  {
    std::lock_guard<std::mutex> lk(cbMutex);
    cbQueue.push_back(j);
  }
  cbCv.notify_one();
}

void Device::callbackLoop()
{
  for (;;) {
    CallbackJob *j = nullptr;
    {
      std::unique_lock<std::mutex> lk(cbMutex);
      cbCv.wait(lk, [&] { return cbStop || !cbQueue.empty(); });
      if (cbQueue.empty())
        return;
      j = cbQueue.front();
      cbQueue.pop_front();
    }
    j->cb(j->user, (ANARIDevice)j->dev, (ANARIFrame)j->frame);
    {
      std::lock_guard<std::mutex> lk(j->sync->m);
      j->sync->done = std::max(j->sync->done, j->gen);
    }
    j->sync->cv.notify_all();
    delete j;
  }
}

// initDevice (VisRTXDevice.cpp:435-458): lazy, latched on failure
bool Device::init()
{
  if (initialized)
    return true;
  if (initFailed)
    return false;
  int count = 0;
  if (vrx_device_count(&count) != 0 || count == 0) {
    reportf(nullptr, ANARI_SEVERITY_FATAL_ERROR, ANARI_STATUS_UNSUPPORTED_DEVICE,
        "no HIP device found: %s", vrx_last_error());
    initFailed = true;
    return false;
  }
  int want = 0;
  for (const char *k : {"hipDevice", "cudaDevice"}) {
    auto it = params.find(k);
    if (it != params.end() && it->second.bytes.size() >= 4)
      memcpy(&want, it->second.bytes.data(), 4);
  }
  if (want < 0 || want >= count) {
    // VisRTXDevice.cpp:570-577
    reportf(nullptr, ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
        "device index %d out of range (%d devices), using device 0", want, count);
    want = 0;
  }
  hipDevice = want;
  if (vrx_set_device(want) != 0 || vrx_stream_create(&stream) != 0) {
    reportf(nullptr, ANARI_SEVERITY_FATAL_ERROR, ANARI_STATUS_UNKNOWN_ERROR,
        "device initialisation failed: %s", vrx_last_error());
    initFailed = true;
    return false;
  }
  initialized = true;
  return true;
}

void *Device::workspace(size_t bytes)
{
  if (dWorkspace && workspaceBytes >= bytes)
    return dWorkspace;
  if (dWorkspace) {
    vrx_stream_sync(stream);
    vrx_free(dWorkspace);
  }
  dWorkspace = nullptr;
  workspaceBytes = 0;
  if (!check(vrx_malloc(&dWorkspace, bytes), nullptr, "workspace allocation"))
    return nullptr;
  workspaceBytes = bytes;
  return dWorkspace;
}

void Device::queueCommit(Object *o)
{
  if (o->queued)
    return;
  o->queued = true;
  retainInternal(o);
  commitQueue.push_back(o);
}

// commitBuffer.flush (RegisteredObject.h:66-72, Frame.cu:201-208): arrays
// and leaf objects first so that parents see finalized children.
static int commitOrder(ANARIDataType t)
{
  switch (t) {
  case ANARI_ARRAY1D:
  case ANARI_ARRAY2D:
  case ANARI_ARRAY3D:
    return 0;
  case ANARI_SPATIAL_FIELD:
  case ANARI_GEOMETRY:
  case ANARI_MATERIAL:
  case ANARI_CAMERA:
  case ANARI_RENDERER:
    return 1;
  case ANARI_SURFACE:
  case ANARI_VOLUME:
  case ANARI_LIGHT:
    return 2;
  case ANARI_GROUP:
    return 3;
  case ANARI_INSTANCE:
    return 4;
  case ANARI_WORLD:
    return 5;
  default:
    return 6;
  }
}

void Device::flushCommits()
{
  // commits can queue more commits (an array's observers), so repeat until
  // the queue stays empty
  while (!commitQueue.empty())
    flushOnce();
}

void Device::flushOnce()
{
  std::vector<Object *> q;
  q.swap(commitQueue);
  std::stable_sort(q.begin(), q.end(),
      [](Object *a, Object *b) { return commitOrder(a->type) < commitOrder(b->type); });
  for (Object *o : q) {
    o->queued = false;
    o->commit();
    o->lastFinalized = tick();
    lastFinalize = o->lastFinalized;
    if (o->type != ANARI_CAMERA && o->type != ANARI_RENDERER && o->type != ANARI_FRAME)
      lastSceneFinalize = o->lastFinalized;
  }
  // a changed child dirties the worlds that reference it
  for (Object *o : q) {
    if (o->type == ANARI_WORLD)
      ((World *)o)->sceneDirty = true;
  }
  for (Object *o : q)
    releaseInternal(o);
}

void Device::report(Object *src, int severity, int code, const char *fmt, va_list ap)
{
  if (!statusCB)
    return;
  char buf[1024];
  vsnprintf(buf, sizeof(buf), fmt, ap);
  statusCB(statusUser, (ANARIDevice)this, (ANARIObject)src, src ? src->type : ANARI_DEVICE,
      severity, code, buf);
}

void Device::reportf(Object *src, int severity, int code, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  report(src, severity, code, fmt, ap);
  va_end(ap);
}

bool Device::check(int st, Object *src, const char *what)
{
  if (st == VRX_OK)
    return true;
  reportf(src, st == VRX_ERR_NOMEM ? ANARI_SEVERITY_FATAL_ERROR : ANARI_SEVERITY_ERROR,
      st == VRX_ERR_NOMEM ? ANARI_STATUS_OUT_OF_MEMORY : ANARI_STATUS_UNKNOWN_ERROR, "%s: %s", what,
      vrx_last_error());
  return false;
}

} // namespace vrxa
