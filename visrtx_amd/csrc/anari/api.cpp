// api.cpp -- the exported ANARI C API of the visrtx_amd device.
//
// The reference exposes the ANARI API through the SDK loader and
// helium::BaseDevice overrides (devices/rtx/VisRTXDevice.cpp:102-382,
// VisRTXLibrary.cpp:40-80).  This library implements the same entry points
// directly, so an application can link it without the SDK, and also exports
// makeVisRTXAMDDevice (VisRTXLibrary.cpp:76-80 counterpart).  Every call takes
// the device lock (object creation is thread-safe, as exercised by
// tests/api/TestMultiThreadedObjectCreation.cpp:57-95) and never throws.
// Waits for a frame (frameReady WAIT, mapFrame, blocking frame properties)
// release the lock while they block.
#include <string.h>

#include <new>

#include "anari/ext/visrtx_amd/visrtx_amd.h"
#include "objects.h"

using namespace vrxa;

namespace {

struct Library {
  ANARIStatusCallback cb;
  const void *user;
};

inline Device *D(ANARIDevice d) { return (Device *)d; }
inline Object *O(ANARIObject o) { return (Object *)o; }

#define GUARD(dev, ret)                                          \
  if (!(dev))                                                    \
    return ret;                                                  \
  std::lock_guard<std::recursive_mutex> lock_(D(dev)->mutex);    \
  if (!D(dev)->init())                                           \
    return ret;

const char *g_deviceSubtypes[] = {"default", "visrtx_amd", nullptr};
const char *g_geomSubtypes[] = {"triangle", "quad", "sphere", "cylinder", "cone", "curve", nullptr};
const char *g_rendererSubtypes[] = {"default", "raycast", "ao", "directLight", nullptr};
const char *g_lightSubtypes[] = {"directional", "point", "spot", nullptr};
const char *g_cameraSubtypes[] = {"perspective", nullptr};
const char *g_materialSubtypes[] = {"matte", nullptr};
const char *g_fieldSubtypes[] = {"structuredRegular", nullptr};
const char *g_volumeSubtypes[] = {"transferFunction1D", "scivis", nullptr};
const char *g_instanceSubtypes[] = {"transform", nullptr};
const char *g_none[] = {nullptr};

template <typename T, typename... A>
ANARIObject create(ANARIDevice d, A &&...args)
{
  try {
    return (ANARIObject) new T(D(d), std::forward<A>(args)...);
  } catch (const std::bad_alloc &) {
    D(d)->reportf(nullptr, ANARI_SEVERITY_FATAL_ERROR, ANARI_STATUS_OUT_OF_MEMORY,
        "out of host memory");
    return nullptr;
  }
}

void freeParam(Param &p)
{
  if (p.obj)
    releaseInternal(p.obj);
  p.obj = nullptr;
}

// Waits for a frame's last render and its completion callback WITHOUT the
// device lock, so that other threads, and a callback that calls the API,
// proceed meanwhile.  The frame is held by an internal reference while the
// lock is released.  Returns false for a handle that is not a frame.
bool waitFrameUnlocked(Device *dev, Object *o)
{
  vrx_event ev = nullptr;
  std::shared_ptr<Frame::Sync> sync;
  uint64_t gen = 0;
  {
    std::lock_guard<std::recursive_mutex> lock(dev->mutex);
    if (!dev->init() || !o || o->type != ANARI_FRAME)
      return false;
    Frame *fr = (Frame *)o;
    if (!fr->rendered)
      return true;
    retainInternal(fr);
    ev = fr->evDone;
    sync = fr->sync;
    gen = fr->callbacksIssued;
  }
  vrx_event_sync(ev);
  sync->waitFor(gen);
  std::lock_guard<std::recursive_mutex> lock(dev->mutex);
  ((Frame *)o)->wait(); // complete now: reports a traversal stack overflow
  releaseInternal(o);
  return true;
}

} // namespace

namespace vrx_anari {
const void *objectInfo(ANARIDataType type, const char *subtype, const char *info, ANARIDataType infoType);
const void *parameterInfo(ANARIDataType type, const char *subtype, const char *param,
    ANARIDataType paramType, const char *info, ANARIDataType infoType);
const char *const *deviceExtensions();
} // namespace vrx_anari

extern "C" {

// ---------------------------------------------------------------------------
ANARIDevice makeVisRTXAMDDevice(ANARIStatusCallback cb, const void *userPtr)
{
  try {
    return (ANARIDevice) new Device(cb, userPtr);
  } catch (...) {
    return nullptr;
  }
}

ANARIDevice visrtx_amd_library_new_device(const char *, ANARIStatusCallback cb, const void *userPtr)
{
  return makeVisRTXAMDDevice(cb, userPtr);
}

ANARILibrary anariLoadLibrary(const char *name, ANARIStatusCallback cb, const void *user)
{
  if (name && strcmp(name, "visrtx_amd") != 0 && strcmp(name, "visrtx") != 0
      && strcmp(name, "default") != 0 && strcmp(name, "environment") != 0) {
    if (cb)
      cb(user, nullptr, nullptr, ANARI_LIBRARY, ANARI_SEVERITY_ERROR,
          ANARI_STATUS_INVALID_ARGUMENT, "unknown ANARI library name");
    return nullptr;
  }
  return (ANARILibrary) new Library{cb, user};
}

void anariUnloadLibrary(ANARILibrary l) { delete (Library *)l; }

const char **anariGetDeviceSubtypes(ANARILibrary) { return g_deviceSubtypes; }

ANARIDevice anariNewDevice(ANARILibrary l, const char *type)
{
  if (!l)
    return nullptr;
  Library *lib = (Library *)l;
  return visrtx_amd_library_new_device(type, lib->cb, lib->user);
}

// ---------------------------------------------------------------------------
static ANARIArray newArray(ANARIDevice d, ANARIDataType at, const void *mem, ANARIMemoryDeleter del,
    const void *user, ANARIDataType t, uint64_t n1, uint64_t n2, uint64_t n3)
{
  GUARD(d, nullptr);
  if (sizeOf(t) == 0) {
    D(d)->reportf(nullptr, ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_ARGUMENT,
        "unsupported array element type %d", t);
    return nullptr;
  }
  int space = VRX_PTR_HOST;
  if (isObjectType(t) && mem && vrx_pointer_kind(mem, &space) == VRX_OK && space != VRX_PTR_HOST) {
    // Array.cpp:79-82
    D(d)->reportf(nullptr, ANARI_SEVERITY_ERROR, ANARI_STATUS_INVALID_OPERATION,
        "illegal operation: cannot create object arrays from GPU memory");
    return nullptr;
  }
  return create<Array>(d, at, mem, del, user, t, n1, n2, n3);
}

ANARIArray1D anariNewArray1D(ANARIDevice d, const void *mem, ANARIMemoryDeleter del,
    const void *user, ANARIDataType t, uint64_t n1)
{
  return newArray(d, ANARI_ARRAY1D, mem, del, user, t, n1, 1, 1);
}

ANARIArray2D anariNewArray2D(ANARIDevice d, const void *mem, ANARIMemoryDeleter del,
    const void *user, ANARIDataType t, uint64_t n1, uint64_t n2)
{
  return newArray(d, ANARI_ARRAY2D, mem, del, user, t, n1, n2, 1);
}

ANARIArray3D anariNewArray3D(ANARIDevice d, const void *mem, ANARIMemoryDeleter del,
    const void *user, ANARIDataType t, uint64_t n1, uint64_t n2, uint64_t n3)
{
  return newArray(d, ANARI_ARRAY3D, mem, del, user, t, n1, n2, n3);
}

void *anariMapArray(ANARIDevice d, ANARIArray a)
{
  GUARD(d, nullptr);
  if (!a)
    return nullptr;
  return ((Array *)a)->map();
}

void anariUnmapArray(ANARIDevice d, ANARIArray a)
{
  GUARD(d, );
  if (a)
    ((Array *)a)->unmap();
}

// ---------------------------------------------------------------------------
ANARILight anariNewLight(ANARIDevice d, const char *type)
{
  GUARD(d, nullptr);
  return create<Light>(d, std::string(type ? type : ""));
}

ANARICamera anariNewCamera(ANARIDevice d, const char *type)
{
  GUARD(d, nullptr);
  return create<Camera>(d, std::string(type ? type : ""));
}

ANARIGeometry anariNewGeometry(ANARIDevice d, const char *type)
{
  GUARD(d, nullptr);
  return create<Geometry>(d, std::string(type ? type : ""));
}

ANARISpatialField anariNewSpatialField(ANARIDevice d, const char *type)
{
  GUARD(d, nullptr);
  return create<SpatialField>(d, std::string(type ? type : ""));
}

ANARIVolume anariNewVolume(ANARIDevice d, const char *type)
{
  GUARD(d, nullptr);
  return create<Volume>(d, std::string(type ? type : ""));
}

ANARISurface anariNewSurface(ANARIDevice d)
{
  GUARD(d, nullptr);
  return create<Surface>(d);
}

ANARIMaterial anariNewMaterial(ANARIDevice d, const char *type)
{
  GUARD(d, nullptr);
  return create<Material>(d, std::string(type ? type : ""));
}

ANARISampler anariNewSampler(ANARIDevice d, const char *type)
{
  GUARD(d, nullptr);
  return create<Object>(d, ANARI_SAMPLER, std::string(type ? type : ""));
}

ANARIGroup anariNewGroup(ANARIDevice d)
{
  GUARD(d, nullptr);
  return create<Group>(d);
}

ANARIInstance anariNewInstance(ANARIDevice d, const char *type)
{
  GUARD(d, nullptr);
  return create<Instance>(d, std::string(type ? type : "transform"));
}

ANARIWorld anariNewWorld(ANARIDevice d)
{
  GUARD(d, nullptr);
  return create<World>(d);
}

ANARIRenderer anariNewRenderer(ANARIDevice d, const char *type)
{
  GUARD(d, nullptr);
  return create<Renderer>(d, std::string(type ? type : "default"));
}

ANARIFrame anariNewFrame(ANARIDevice d)
{
  GUARD(d, nullptr);
  return create<Frame>(d);
}

// ---------------------------------------------------------------------------
void anariSetParameter(ANARIDevice d, ANARIObject o, const char *name, ANARIDataType t, const void *mem)
{
  // a null object (a failed creation) is ignored; the device handle itself
  // addresses device parameters
  if (!d || !o || !name)
    return;
  std::lock_guard<std::recursive_mutex> lock(D(d)->mutex);
  Param p;
  p.type = t;
  if (t == ANARI_STRING) {
    p.str = mem ? (const char *)mem : "";
  } else if (isObjectType(t)) {
    p.obj = mem ? *(Object *const *)mem : nullptr;
    retainInternal(p.obj);
  } else {
    size_t sz = sizeOf(t);
    if (sz == 0 || !mem) {
      D(d)->reportf(O(o), ANARI_SEVERITY_WARNING, ANARI_STATUS_INVALID_ARGUMENT,
          "ignoring parameter '%s' of unsupported type %d", name, t);
      return;
    }
    p.bytes.assign((const uint8_t *)mem, (const uint8_t *)mem + sz);
  }
  auto &m = (void *)o == (void *)d ? D(d)->params : O(o)->staged;
  auto it = m.find(name);
  if (it != m.end())
    freeParam(it->second);
  m[name] = std::move(p);
}

void anariUnsetParameter(ANARIDevice d, ANARIObject o, const char *name)
{
  if (!d || !o || !name)
    return;
  std::lock_guard<std::recursive_mutex> lock(D(d)->mutex);
  auto &m = (void *)o == (void *)d ? D(d)->params : O(o)->staged;
  auto it = m.find(name);
  if (it != m.end()) {
    freeParam(it->second);
    m.erase(it);
  }
}

void anariUnsetAllParameters(ANARIDevice d, ANARIObject o)
{
  if (!d || !o)
    return;
  std::lock_guard<std::recursive_mutex> lock(D(d)->mutex);
  auto &m = (void *)o == (void *)d ? D(d)->params : O(o)->staged;
  for (auto &kv : m)
    freeParam(kv.second);
  m.clear();
}

void anariCommitParameters(ANARIDevice d, ANARIObject o)
{
  if (!d || !o)
    return;
  std::lock_guard<std::recursive_mutex> lock(D(d)->mutex);
  if ((void *)o == (void *)d) {
    // device parameters (cudaDevice / hipDevice, forceInit): VisRTXDevice.cpp:460-518
    auto it = D(d)->params.find("forceInit");
    if (it != D(d)->params.end())
      D(d)->init();
    return;
  }
  if (!D(d)->init())
    return;
  Object *obj = O(o);
  // snapshot staged parameters (internal references are duplicated); the
  // object observes the arrays of its new snapshot only
  obj->observeParams(false);
  for (auto &kv : obj->params)
    freeParam(kv.second);
  obj->params.clear();
  for (auto &kv : obj->staged) {
    Param p = kv.second;
    retainInternal(p.obj);
    obj->params[kv.first] = std::move(p);
  }
  obj->observeParams(true);
  D(d)->queueCommit(obj);
}

void anariRelease(ANARIDevice d, ANARIObject o)
{
  // a null object is a no-op (a failed creation's handle); only the device
  // handle itself releases the device
  if (!d || !o)
    return;
  if ((void *)o == (void *)d) {
    Device *dev = D(d);
    bool last;
    {
      std::lock_guard<std::recursive_mutex> lock(dev->mutex);
      last = --dev->refCount <= 0;
    }
    if (last)
      delete dev;
    return;
  }
  std::lock_guard<std::recursive_mutex> lock(D(d)->mutex);
  Object *obj = O(o);
  obj->pubRef--;
  if (obj->pubRef <= 0) {
    obj->onPublicReleaseToZero();
    if (obj->intRef <= 0)
      delete obj;
  }
}

void anariRetain(ANARIDevice d, ANARIObject o)
{
  if (!d || !o)
    return;
  std::lock_guard<std::recursive_mutex> lock(D(d)->mutex);
  if ((void *)o == (void *)d)
    D(d)->refCount++;
  else
    O(o)->pubRef++;
}

// ---------------------------------------------------------------------------
const char **anariGetObjectSubtypes(ANARIDevice, ANARIDataType t)
{
  switch (t) {
  case ANARI_GEOMETRY:
    return g_geomSubtypes;
  case ANARI_RENDERER:
    return g_rendererSubtypes;
  case ANARI_CAMERA:
    return g_cameraSubtypes;
  case ANARI_MATERIAL:
    return g_materialSubtypes;
  case ANARI_SPATIAL_FIELD:
    return g_fieldSubtypes;
  case ANARI_VOLUME:
    return g_volumeSubtypes;
  case ANARI_INSTANCE:
    return g_instanceSubtypes;
  case ANARI_LIGHT:
    return g_lightSubtypes;
  default:
    return g_none;
  }
}

const void *anariGetObjectInfo(ANARIDevice, ANARIDataType objectType, const char *subtype,
    const char *infoName, ANARIDataType infoType)
{
  return vrx_anari::objectInfo(objectType, subtype, infoName, infoType);
}

const void *anariGetParameterInfo(ANARIDevice, ANARIDataType objectType, const char *subtype,
    const char *paramName, ANARIDataType paramType, const char *infoName, ANARIDataType infoType)
{
  return vrx_anari::parameterInfo(objectType, subtype, paramName, paramType, infoName, infoType);
}

int anariGetProperty(ANARIDevice d, ANARIObject o, const char *name, ANARIDataType t, void *mem,
    uint64_t size, ANARIWaitMask mask)
{
  if (!d || !name || !mem)
    return 0;
  if (o && (void *)o != (void *)d && (mask & ANARI_WAIT)) {
    // a blocking query waits for the object's frame outside the lock
    bool isFrame;
    {
      std::lock_guard<std::recursive_mutex> lock(D(d)->mutex);
      isFrame = O(o)->type == ANARI_FRAME;
    }
    if (isFrame)
      waitFrameUnlocked(D(d), O(o));
  }
  std::lock_guard<std::recursive_mutex> lock(D(d)->mutex);
  if (!o || (void *)o == (void *)d) {
    // device properties (VisRTXDevice.cpp:520-552)
    std::string n = name;
    if (n == "version" && t == ANARI_INT32 && size >= 4) {
      int v = vrx_version();
      memcpy(mem, &v, 4);
      return 1;
    }
    if (n == "version.major" && t == ANARI_INT32 && size >= 4) {
      int v = 0;
      memcpy(mem, &v, 4);
      return 1;
    }
    if (n == "version.minor" && t == ANARI_INT32 && size >= 4) {
      int v = 1;
      memcpy(mem, &v, 4);
      return 1;
    }
    if (n == "version.patch" && t == ANARI_INT32 && size >= 4) {
      int v = 0;
      memcpy(mem, &v, 4);
      return 1;
    }
    // the extension list and the renderer subtypes as string lists (a
    // pointer to a null-terminated array of C strings)
    if (n == "extension" && t == ANARI_STRING_LIST && size >= sizeof(void *)) {
      const char *const *v = vrx_anari::deviceExtensions();
      memcpy(mem, &v, sizeof(v));
      return 1;
    }
    if (n == "subtypes.renderer" && t == ANARI_STRING_LIST && size >= sizeof(void *)) {
      const char *const *v = g_rendererSubtypes;
      memcpy(mem, &v, sizeof(v));
      return 1;
    }
    if (n == "hipDevice" && t == ANARI_INT32 && size >= 4) {
      if (!D(d)->init())
        return 0;
      memcpy(mem, &D(d)->hipDevice, 4);
      return 1;
    }
    return 0;
  }
  if (!D(d)->init())
    return 0;
  return O(o)->getProperty(name, t, mem, size, (uint32_t)mask) ? 1 : 0;
}

// ---------------------------------------------------------------------------
const void *anariMapFrame(ANARIDevice d, ANARIFrame f, const char *channel, uint32_t *width,
    uint32_t *height, ANARIDataType *pixelType)
{
  uint32_t w0 = 0, h0 = 0;
  ANARIDataType t0 = ANARI_UNKNOWN;
  if (!width)
    width = &w0;
  if (!height)
    height = &h0;
  if (!pixelType)
    pixelType = &t0;
  *width = *height = 0;
  *pixelType = ANARI_UNKNOWN;
  if (!d || !f || !channel || !waitFrameUnlocked(D(d), O(f)))
    return nullptr;
  GUARD(d, nullptr);
  return ((Frame *)f)->map(channel, width, height, pixelType);
}

void anariUnmapFrame(ANARIDevice, ANARIFrame, const char *) {}

void anariRenderFrame(ANARIDevice d, ANARIFrame f)
{
  GUARD(d, );
  if (!f || O(f)->type != ANARI_FRAME)
    return;
  ((Frame *)f)->render();
}

int anariFrameReady(ANARIDevice d, ANARIFrame f, ANARIWaitMask m)
{
  if (!d || !f)
    return 0;
  if (m == ANARI_NO_WAIT) {
    GUARD(d, 0);
    if (O(f)->type != ANARI_FRAME)
      return 0;
    return ((Frame *)f)->ready() ? 1 : 0;
  }
  // Frame.cu:430-438: WAIT returns once the frame, and its completion
  // callback, are done
  return waitFrameUnlocked(D(d), O(f)) ? 1 : 0;
}

void anariDiscardFrame(ANARIDevice, ANARIFrame) {}

} // extern "C"
