// objects.h -- ANARI object model of the visrtx_amd device (host C++).
//
// Mirrors the reference device's object layer (devices/rtx/Object.{h,cpp},
// RegisteredObject.h:39-92, array/, camera/, scene/, renderer/, frame/) with
// the helium semantics the reference inherits from the ANARI-SDK:
// parameters are staged until anariCommitParameters, commits are deferred
// and flushed at the next renderFrame / blocking property query
// (frame/Frame.cu:201-208), objects are reference counted (public + internal
// references), and no exception crosses the C API -- errors go to the status
// callback.  All GPU work goes through the C ABI in include/vrx_hip.h.
#pragma once

#include <stdarg.h>
#include <stdint.h>

#include <array>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "anari/anari.h"
#include "vrx_hip.h"

namespace vrxa {

struct Device;
struct Object;

size_t sizeOf(ANARIDataType t);
bool isObjectType(ANARIDataType t);
const char *typeName(ANARIDataType t);

struct Param {
  ANARIDataType type{ANARI_UNKNOWN};
  std::vector<uint8_t> bytes;
  std::string str;
  Object *obj{nullptr};
};

struct Object {
  Object(Device *d, ANARIDataType t, std::string sub);
  virtual ~Object();

  Device *dev;
  ANARIDataType type;
  std::string subtype;
  int pubRef{1};
  int intRef{0};
  std::map<std::string, Param> staged;
  std::map<std::string, Param> params; // committed snapshot
  bool queued{false};
  uint64_t lastFinalized{0};

  // called by the device when the deferred commit is flushed
  virtual void commit() {}
  // register (on) or unregister (off) this object as an observer of the
  // arrays among its committed parameters: unmapping or re-committing such an
  // array queues this object's commit again (helium ChangeObserverPtr)
  void observeParams(bool on);
  virtual bool isValid() const { return true; }
  virtual bool getProperty(const std::string &name, ANARIDataType type, void *mem, uint64_t size,
      uint32_t mask);
  virtual void onPublicReleaseToZero() {}

  // parameter access on the committed snapshot
  const Param *find(const char *name) const;
  bool getRaw(const char *name, ANARIDataType t, void *out) const;
  template <typename T>
  T get(const char *name, ANARIDataType t, T def) const
  {
    T v = def;
    getRaw(name, t, &v);
    return v;
  }
  std::string getString(const char *name, const std::string &def) const;
  Object *getObject(const char *name, ANARIDataType t) const;
  bool has(const char *name) const { return find(name) != nullptr; }

  void report(int severity, int code, const char *fmt, ...) const;
};

void retainInternal(Object *o);
void releaseInternal(Object *o);

// ---------------------------------------------------------------------------
struct Array : Object {
  Array(Device *d, ANARIDataType arrType, const void *appMem, ANARIMemoryDeleter del,
      const void *userData, ANARIDataType elem, uint64_t n1, uint64_t n2, uint64_t n3);
  ~Array() override;

  ANARIDataType elem;
  uint64_t dims[3];
  size_t elemSize;
  const void *app{nullptr};
  ANARIMemoryDeleter deleter{nullptr};
  const void *deleterUser{nullptr};
  std::vector<uint8_t> owned; // managed or privatized host storage
  // where `app` lives (vrx_pointer_kind): shared and captured arrays may hold
  // device or managed memory, which is used in place (Array.cpp:38-65)
  int space{VRX_PTR_HOST};
  bool devOwned{false}; // `app` is a device copy made by privatization
  bool mapped{false};
  uint64_t version{1};
  // Array1D region [begin, end) of the capacity (Array1D.cpp:43-91)
  uint64_t begin{0}, end{0};
  // device copy of host storage
  void *dptr{nullptr};
  size_t dbytes{0};
  uint64_t dversion{0};
  // host copy of device storage, for commit-time reads (index checks, LUTs)
  std::vector<uint8_t> mirror;
  uint64_t mirrorVersion{0};
  std::vector<Object *> objects; // object arrays: internal references
  std::vector<Object *> observers; // objects whose committed parameters hold this array

  uint64_t capacity() const { return dims[0] * dims[1] * dims[2]; }
  // elements in use: the region for Array1D, else the capacity
  uint64_t count() const { return type == ANARI_ARRAY1D ? end - begin : capacity(); }
  size_t capacityBytes() const { return capacity() * elemSize; }
  size_t bytes() const { return count() * elemSize; }
  bool deviceResident() const { return space == VRX_PTR_DEVICE || space == VRX_PTR_MANAGED; }
  const void *host(); // first element of the region, host readable
  void *map();
  void unmap();
  void commit() override; // Array1D "begin" / "end"
  const void *device(); // first element of the region on the device (uploads host data)
  void onPublicReleaseToZero() override; // privatize shared memory
  void refreshObjects();
  void notifyObservers();
};

// ---------------------------------------------------------------------------
struct Geometry : Object {
  Geometry(Device *d, const std::string &sub);
  ~Geometry() override;
  void commit() override;
  bool isValid() const override { return valid; }

  int ptype{-1};
  bool quad{false}; // "quad": built and traced as two triangles per quad
  bool valid{false};
  Array *pos{nullptr}, *index{nullptr}, *radius{nullptr}, *primId{nullptr}, *normal{nullptr};
  float globalRadius{0.01f};
  bool caps{false};
  bool cull{false};
  uint32_t n{0};
  // BVH
  bool bvhDirty{true};
  vrx_bvh bvh{};
  vrx_bvh_header header{};
  void *dNodes{nullptr}, *dLeaves{nullptr}, *dIds{nullptr}, *dHeader{nullptr};
  size_t capNodes{0}, capLeaves{0}, capIds{0};
  // 4-wide nodes of the current BVH, built on demand (node width 4)
  void *dWide{nullptr};
  size_t capWide{0};
  bool wideValid{false};
  // generated indices: the triangles of quads (Quad.cpp:124-145), the
  // segments of a curve soup (Curve.cpp:107-118)
  bool useGenIdx{false};
  std::vector<uint32_t> genIdx;
  void *dGenIdx{nullptr};
  size_t capGenIdx{0};
  bool genIdxDirty{true};
  // device triangle/segment/centre indices (the generated ones for quads)
  const uint32_t *deviceIndex();
  int buildBVH(vrx_stream s);
  int allocWide();
  int buildWide(vrx_stream s);
  void freeBVH();
};

struct Material : Object {
  Material(Device *d, const std::string &sub) : Object(d, ANARI_MATERIAL, sub) {}
  void commit() override;
  bool isValid() const override { return subtype == "matte"; }
  float color[3]{0.8f, 0.8f, 0.8f};
  float opacity{1.0f};
};

struct Surface : Object {
  Surface(Device *d) : Object(d, ANARI_SURFACE, "") {}
  void commit() override;
  bool isValid() const override;
  Geometry *geom{nullptr};
  Material *mat{nullptr};
  uint32_t id{~0u};
};

struct SpatialField : Object {
  SpatialField(Device *d, const std::string &sub) : Object(d, ANARI_SPATIAL_FIELD, sub) {}
  ~SpatialField() override;
  void commit() override;
  bool isValid() const override { return valid; }
  bool valid{false};
  Array *data{nullptr};
  float origin[3]{0, 0, 0};
  float spacing[3]{1, 1, 1};
  uint32_t dims[3]{0, 0, 0};
  void *dData{nullptr}; // float32 copy on the device
  uint64_t uploadedVersion{0};
  const float *deviceData();
  void bounds(float lo[3], float hi[3]) const;
};

struct Volume : Object {
  Volume(Device *d, const std::string &sub) : Object(d, ANARI_VOLUME, sub) {}
  ~Volume() override;
  void commit() override;
  bool isValid() const override { return field && field->isValid(); }
  SpatialField *field{nullptr};
  uint32_t id{~0u};
  float valueRange[2]{0.0f, 1.0f};
  float unitDistance{1.0f};
  float lut[256 * 4];
  void *dLut{nullptr};
  // macro-cell majorant grid (rebuilt with the world; attached to the
  // device volume when at least 1/8 of its cells are empty)
  void *dGrid{nullptr};
  size_t gridCells{0};
  void *dPairs{nullptr}; // the y-pair copy of the field for the dense march
  size_t pairBytes{0};
  float emptyCellFraction{0.0f};
};

// directional / point / spot lights (scene/light/*.cpp); the light record
// in world space is made by the world from these and the instance transform
struct Light : Object {
  Light(Device *d, const std::string &sub) : Object(d, ANARI_LIGHT, sub) {}
  void commit() override;
  bool isValid() const override { return type >= 0; }
  int type{-1};
  float color[3]{1, 1, 1};
  float direction[3]{0, 0, -1};
  float position[3]{0, 0, 0};
  float irradiance{1.0f}, intensity{1.0f};
  float cosOuter{-1.0f}, cosInner{-1.0f};
  // the world-space record under the 3x4 column-major transform xfm (or identity)
  vrx_light record(const float *xfm) const;
};

struct Group : Object {
  Group(Device *d) : Object(d, ANARI_GROUP, "") {}
  ~Group() override;
  void commit() override;
  bool getProperty(const std::string &name, ANARIDataType type, void *mem, uint64_t size,
      uint32_t mask) override;
  // internal references, taken at commit and dropped at the next commit or
  // destruction (the object array may change or die in between)
  std::vector<Surface *> surfaces;
  std::vector<Volume *> volumes;
  std::vector<Light *> lights;
};

struct Instance : Object {
  Instance(Device *d, const std::string &sub) : Object(d, ANARI_INSTANCE, sub) {}
  void commit() override;
  Group *group{nullptr};
  float xfm[12]{1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
  uint32_t id{~0u};
  // a `transform` array instances the group once per element, each with the
  // matching element of an `id` array if one is given (Instance.cpp:115-128)
  std::vector<float> xfms; // 12 floats (column-major 3x4) per transform
  std::vector<uint32_t> ids;
  size_t numTransforms() const { return xfms.empty() ? 1 : xfms.size() / 12; }
  const float *xfmAt(size_t i) const { return xfms.empty() ? xfm : xfms.data() + 12 * i; }
  uint32_t idAt(size_t i) const { return !xfms.empty() && !ids.empty() ? ids[i] : id; }
};

struct World : Object {
  World(Device *d) : Object(d, ANARI_WORLD, "") {}
  ~World() override;
  void commit() override;
  bool getProperty(const std::string &name, ANARIDataType type, void *mem, uint64_t size,
      uint32_t mask) override;
  // internal references, as Group's
  std::vector<Instance *> instances;
  std::vector<Surface *> surfaces;
  std::vector<Volume *> volumes;
  std::vector<Light *> lights;
  uint32_t id{~0u};

  // device scene
  bool sceneDirty{true};
  uint64_t builtAt{0};
  vrx_scene scene{};
  void *dGeoms{nullptr}, *dEntries{nullptr}, *dVolumes{nullptr}, *dLights{nullptr};
  void *dTlasNodes{nullptr}, *dTlasLeaves{nullptr}, *dTlasIds{nullptr}, *dTlasHeader{nullptr};
  void *dTlasWide{nullptr};
  float bounds[6]{0, 0, 0, 0, 0, 0};
  float lastBuildMs{0.0f}, lastWideMs{0.0f};
  // host copies of the scene tables (for re-upload when 4-wide nodes appear)
  std::vector<vrx_geom> hostGeoms;
  std::vector<Geometry *> hostGeomOf;
  int rebuild();
  int ensureWide(); // 4-wide nodes for every BLAS and the TLAS; sets scene.wide
  void freeScene();
  bool needsRebuild() const;
};

struct Camera : Object {
  Camera(Device *d, const std::string &sub) : Object(d, ANARI_CAMERA, sub) {}
  void commit() override;
  bool isValid() const override { return subtype == "perspective"; }
  vrx_camera cam{};
};

struct Renderer : Object {
  Renderer(Device *d, const std::string &sub);
  void commit() override;
  bool isValid() const override { return kind >= 0; }
  bool getProperty(const std::string &name, ANARIDataType type, void *mem, uint64_t size,
      uint32_t mask) override;
  int kind{-1};
  vrx_renderer r{};
  int sampleLimit{128};
  float lightFalloff{1.0f};
  // launch parameters left to the frame's autotuner (not set explicitly)
  bool tuneWidth{true}, tuneStack{true}, tuneDefer{false};
  int spp() const
  {
    return r.kind == VRX_RENDER_RAYCAST ? 1 : (r.pixel_samples > 1 ? r.pixel_samples : 1);
  }
};

struct Frame : Object {
  Frame(Device *d);
  ~Frame() override;
  void commit() override;
  bool isValid() const override;
  bool getProperty(const std::string &name, ANARIDataType type, void *mem, uint64_t size,
      uint32_t mask) override;

  Renderer *renderer{nullptr};
  Camera *camera{nullptr};
  World *world{nullptr};
  ANARIFrameCompletionCallback callback{nullptr};
  const void *callbackUser{nullptr};
  uint32_t size[2]{10, 10};
  ANARIDataType colorType{ANARI_UFIXED8_RGBA_SRGB};
  ANARIDataType depthType{ANARI_UNKNOWN}, primType{ANARI_UNKNOWN}, objType{ANARI_UNKNOWN},
      instType{ANARI_UNKNOWN};

  // device buffers
  void *dAccum{nullptr}, *dColor{nullptr}, *dDepth{nullptr}, *dPrim{nullptr}, *dObj{nullptr},
      *dInst{nullptr}, *dScratch{nullptr}, *dCounters{nullptr};
  size_t scratchBytes{0};
  std::vector<uint8_t> hColor, hDepth, hPrim, hObj, hInst;
  vrx_event evStart{nullptr}, evEnd{nullptr};
  vrx_event evDone{nullptr}; // after the result read-backs that follow evEnd
  bool rendered{false};
  bool stackReported{false};
  // Online launch autotuning: the first frames of a (world, renderer, size)
  // configuration each use one candidate (node width x LDS stack depth), timed
  // by the frame's events; then the fastest is kept.  Every candidate gives
  // bit-identical channels, so tuning only changes the time.
  struct Tune {
    uint64_t key{0};
    int n{0}, next{0}, best{-1}, last{-1};
    bool deferPending{false}; // the long-ray candidate is still to be added
    int width[8]{}, sd[8]{}, defer[8]{};
    float ms[8]{};
  } tune;
  void pickLaunch(vrx_renderer &rr);
  // completion callbacks: each render with a callback posts one job; the
  // device's callback thread runs it and advances `sync`
  struct Sync {
    std::mutex m;
    std::condition_variable cv;
    uint64_t done{0};
    void waitFor(uint64_t gen);
  };
  std::shared_ptr<Sync> sync{std::make_shared<Sync>()};
  uint64_t callbacksIssued{0};
  bool callbacksDone(); // every posted callback has returned
  bool nextReset{true};
  uint64_t lastSeenFinalize{0};
  uint32_t frameID{0};
  float duration{0.0f};
  // pinned read-back slots: [0] rays traced, [1] volume samples, [2] the
  // stack overflow flag (pinned, so the copies stay asynchronous)
  uint64_t *hReadback{nullptr}; // pinned: [0] rays, [1] volume samples, [2] stack flag
  uint64_t *dReadback{nullptr}; // its device-side address
  bool allocated{false};

  void allocate();
  void freeBuffers();
  void render();
  void wait(); // GPU work of the last render (not its callback)
  bool ready(); // GPU work done and its callback returned
  const void *map(const std::string &channel, uint32_t *w, uint32_t *h, ANARIDataType *pt);
};

// A frame completion callback waiting to run (posted by the HIP host function
// of Frame::render, run by the device's callback thread).
struct CallbackJob {
  Device *dev;
  ANARIFrameCompletionCallback cb;
  const void *user;
  Frame *frame;
  std::shared_ptr<Frame::Sync> sync;
  uint64_t gen;
};

// ---------------------------------------------------------------------------
struct Device {
  Device(ANARIStatusCallback cb, const void *user);
  ~Device();

  std::recursive_mutex mutex;
  ANARIStatusCallback statusCB{nullptr};
  const void *statusUser{nullptr};
  int refCount{1};
  int hipDevice{0};
  bool initialized{false};
  bool initFailed{false};
  vrx_stream stream{nullptr};
  std::vector<Object *> commitQueue;
  uint64_t finalizeClock{1};
  uint64_t lastFinalize{0};
  uint64_t lastSceneFinalize{0}; // last finalize of an object a world depends on
  // BVH build workspace (grow only)
  void *dWorkspace{nullptr};
  size_t workspaceBytes{0};
  std::map<std::string, Param> params;
  // Frame completion callbacks run on this thread, never on the HIP runtime's
  // host-function thread: the stream then never waits for application code,
  // and a callback may call the API (it takes the device lock like any
  // caller) while another thread waits for the frame without the lock.
  std::thread cbThread;
  std::mutex cbMutex;
  std::condition_variable cbCv;
  std::deque<CallbackJob *> cbQueue;
  bool cbStop{false};
  void postCallback(CallbackJob *j); // from the HIP host function: enqueue only
  void startCallbackThread();
  void callbackLoop();

  bool init();
  void *workspace(size_t bytes);
  void flushCommits();
  void flushOnce();
  void queueCommit(Object *o);
  uint64_t tick() { return ++finalizeClock; }
  void report(Object *src, int severity, int code, const char *fmt, va_list ap);
  void reportf(Object *src, int severity, int code, const char *fmt, ...);
  bool check(int vrxStatus, Object *src, const char *what);
};

} // namespace vrxa
