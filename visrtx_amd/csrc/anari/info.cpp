// info.cpp -- ANARI object and parameter introspection (anariGetObjectInfo /
// anariGetParameterInfo) for the objects this device implements.
//
// The reference answers these queries from tables generated out of
// devices/rtx/visrtx_device.json by the ANARI-SDK code generator
// (devices/rtx/CMakeLists.txt:385-391).  Here the tables are written out
// directly for the subtypes on this device's path, with the defaults and
// clamps the device applies when it commits each object (objects.cpp).
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <anari/anari.h>

namespace vrx_anari {

namespace {

struct ParamDef {
  const char *name;
  ANARIDataType type;
  bool required;
  const void *def; // default value of `type`, or null
  const void *min;
  const void *max;
  const char *desc;
};

struct ObjDef {
  ANARIDataType type;
  const char *subtype; // null: the object type has no subtypes
  const char *desc;
  std::vector<ParamDef> params;
  std::vector<ANARIParameter> list; // null-terminated, built on first use
};

const float kZero3[3] = {0.0f, 0.0f, 0.0f};
const float kOne3[3] = {1.0f, 1.0f, 1.0f};
const float kGrey3[3] = {0.8f, 0.8f, 0.8f};
const float kDirZ[3] = {0.0f, 0.0f, 1.0f};
const float kUpY[3] = {0.0f, 1.0f, 0.0f};
const float kBlack4[4] = {0.0f, 0.0f, 0.0f, 1.0f};
const float kRegion[4] = {0.0f, 0.0f, 1.0f, 1.0f};
const float kRange01[2] = {0.0f, 1.0f};
const float kOne = 1.0f, kZero = 0.0f;
const float kPi = 3.14159265358979f, kFalloff = 0.1f;
const float kDirNegZ[3] = {0.0f, 0.0f, -1.0f};
const float kSphereRadius = 0.01f;
const float kFovy = 1.04719755f; // 60 degrees
const float kRate = 0.125f, kRateMin = 1e-3f, kRateMax = 10.0f;
const float kAoDist = 1e20f;
const float kHalf = 0.5f;
const int32_t kInt1 = 1, kInt0 = 0, kInt2 = 2, kInt4 = 4, kInt5 = 5, kInt256 = 256, kLimit = 128;
const int32_t kFalse = 0, kTrue1 = 1;
const uint32_t kNoId = 0xFFFFFFFFu;
const ANARIDataType kSrgb = ANARI_UFIXED8_RGBA_SRGB;
const float kIdentity4[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};

std::vector<ObjDef> &tables()
{
  static std::vector<ObjDef> t = {
      {ANARI_GEOMETRY, "sphere", "spheres given by centres and radii (Sphere.cu)",
          {{"vertex.position", ANARI_ARRAY1D, true, nullptr, nullptr, nullptr, "FLOAT32_VEC3 centres"},
              {"vertex.radius", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "FLOAT32 per-centre radii"},
              {"radius", ANARI_FLOAT32, false, &kSphereRadius, nullptr, nullptr, "radius when no vertex.radius"},
              {"primitive.index", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "UINT32 centre indices"},
              {"primitive.id", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "UINT32 primitiveId values"}}},
      {ANARI_GEOMETRY, "cylinder", "cylinders between vertex pairs (Cylinder.cpp)",
          {{"vertex.position", ANARI_ARRAY1D, true, nullptr, nullptr, nullptr, "FLOAT32_VEC3 end points"},
              {"primitive.index", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "UINT32_VEC2 end point pairs"},
              {"primitive.radius", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "FLOAT32 per-cylinder radii"},
              {"radius", ANARI_FLOAT32, false, &kOne, nullptr, nullptr, "radius when no primitive.radius"},
              {"caps", ANARI_STRING, false, "none", nullptr, nullptr, "\"none\" or \"both\""},
              {"primitive.id", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "UINT32 primitiveId values"}}},
      {ANARI_GEOMETRY, "triangle", "triangle soup or indexed mesh (Triangle.cpp)",
          {{"vertex.position", ANARI_ARRAY1D, true, nullptr, nullptr, nullptr, "FLOAT32_VEC3 vertices"},
              {"primitive.index", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "UINT32_VEC3 triangles"},
              {"vertex.normal", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "FLOAT32_VEC3 shading normals"},
              {"primitive.id", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "UINT32 primitiveId values"},
              {"cullBackfaces", ANARI_BOOL, false, &kFalse, nullptr, nullptr, "skip back-facing hits"}}},
      {ANARI_GEOMETRY, "cone", "capped cones between vertex pairs (Cone.cpp)",
          {{"vertex.position", ANARI_ARRAY1D, true, nullptr, nullptr, nullptr, "FLOAT32_VEC3 end points"},
              {"vertex.radius", ANARI_ARRAY1D, true, nullptr, nullptr, nullptr, "FLOAT32 per-vertex radii"},
              {"primitive.index", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "UINT32_VEC2 end point pairs"},
              {"caps", ANARI_STRING, false, "none", nullptr, nullptr, "read, as the reference does; end discs are always tested"},
              {"primitive.id", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "UINT32 primitiveId values"}}},
      {ANARI_GEOMETRY, "curve", "round linear curve segments (Curve.cpp)",
          {{"vertex.position", ANARI_ARRAY1D, true, nullptr, nullptr, nullptr, "FLOAT32_VEC3 control points"},
              {"vertex.radius", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "FLOAT32 per-vertex radii"},
              {"radius", ANARI_FLOAT32, false, &kOne, nullptr, nullptr, "radius when no vertex.radius"},
              {"primitive.index", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr,
                  "UINT32 first vertex of each segment"},
              {"primitive.id", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "UINT32 primitiveId values"}}},
      {ANARI_GEOMETRY, "quad", "quads, each traced as two triangles (Quad.cpp)",
          {{"vertex.position", ANARI_ARRAY1D, true, nullptr, nullptr, nullptr, "FLOAT32_VEC3 vertices"},
              {"primitive.index", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "UINT32_VEC4 quads"},
              {"vertex.normal", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "FLOAT32_VEC3 shading normals"},
              {"primitive.id", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr,
                  "UINT32 primitiveId values, indexed by triangle (2 per quad)"},
              {"cullBackfaces", ANARI_BOOL, false, &kFalse, nullptr, nullptr, "skip back-facing hits"}}},
      {ANARI_MATERIAL, "matte", "constant-colour matte material (Matte.cpp)",
          {{"color", ANARI_FLOAT32_VEC3, false, kGrey3, nullptr, nullptr, "base colour"},
              {"opacity", ANARI_FLOAT32, false, &kOne, nullptr, nullptr, "opacity"},
              {"alphaMode", ANARI_STRING, false, "opaque", nullptr, nullptr, "opaque | blend | mask"},
              {"alphaCutoff", ANARI_FLOAT32, false, &kHalf, nullptr, nullptr, "mask threshold"}}},
      {ANARI_SURFACE, nullptr, "a geometry with a material",
          {{"geometry", ANARI_GEOMETRY, true, nullptr, nullptr, nullptr, "the geometry"},
              {"material", ANARI_MATERIAL, false, nullptr, nullptr, nullptr, "the material"},
              {"id", ANARI_UINT32, false, &kNoId, nullptr, nullptr, "objectId value"}}},
      {ANARI_SPATIAL_FIELD, "structuredRegular", "regular grid of samples (StructuredRegularField.cpp)",
          {{"data", ANARI_ARRAY3D, true, nullptr, nullptr, nullptr, "FLOAT32 / UINT8 / UFIXED8 samples, x fastest"},
              {"origin", ANARI_FLOAT32_VEC3, false, kZero3, nullptr, nullptr, "position of sample (0,0,0)"},
              {"spacing", ANARI_FLOAT32_VEC3, false, kOne3, nullptr, nullptr, "sample spacing"},
              {"filter", ANARI_STRING, false, "linear", nullptr, nullptr, "reconstruction filter"}}},
      {ANARI_VOLUME, "transferFunction1D", "emission-absorption volume (TransferFunction1D.cpp)",
          {{"value", ANARI_SPATIAL_FIELD, true, nullptr, nullptr, nullptr, "the field"},
              {"color", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "FLOAT32_VEC3/VEC4 colour map"},
              {"opacity", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "FLOAT32 opacity map"},
              {"valueRange", ANARI_FLOAT32_BOX1, false, kRange01, nullptr, nullptr, "field values mapped to the maps"},
              {"unitDistance", ANARI_FLOAT32, false, &kOne, nullptr, nullptr, "distance of one opacity unit"},
              {"id", ANARI_UINT32, false, &kNoId, nullptr, nullptr, "objectId value"}}},
      {ANARI_LIGHT, "directional", "light from one direction (Directional.cpp)",
          {{"color", ANARI_FLOAT32_VEC3, false, kOne3, nullptr, nullptr, "light colour"},
              {"direction", ANARI_FLOAT32_VEC3, false, kDirNegZ, nullptr, nullptr, "direction the light travels"},
              {"irradiance", ANARI_FLOAT32, false, &kOne, &kZero, nullptr, "irradiance"}}},
      {ANARI_LIGHT, "point", "point light (Point.cpp)",
          {{"color", ANARI_FLOAT32_VEC3, false, kOne3, nullptr, nullptr, "light colour"},
              {"position", ANARI_FLOAT32_VEC3, false, kZero3, nullptr, nullptr, "light position"},
              {"intensity", ANARI_FLOAT32, false, &kOne, &kZero, nullptr, "intensity (else power)"},
              {"power", ANARI_FLOAT32, false, &kOne, &kZero, nullptr, "used when no intensity is set"}}},
      {ANARI_LIGHT, "spot", "spot light with a smooth edge (Spot.cpp)",
          {{"color", ANARI_FLOAT32_VEC3, false, kOne3, nullptr, nullptr, "light colour"},
              {"position", ANARI_FLOAT32_VEC3, false, kZero3, nullptr, nullptr, "light position"},
              {"direction", ANARI_FLOAT32_VEC3, false, kDirNegZ, nullptr, nullptr, "cone axis"},
              {"openingAngle", ANARI_FLOAT32, false, &kPi, nullptr, nullptr, "full cone angle (radians)"},
              {"falloffAngle", ANARI_FLOAT32, false, &kFalloff, nullptr, nullptr, "width of the smooth edge (radians)"},
              {"intensity", ANARI_FLOAT32, false, &kOne, &kZero, nullptr, "intensity (else power)"},
              {"power", ANARI_FLOAT32, false, &kOne, &kZero, nullptr, "used when no intensity is set"}}},
      {ANARI_GROUP, nullptr, "surfaces, volumes and lights instanced together",
          {{"surface", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "SURFACE array"},
              {"volume", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "VOLUME array"},
              {"light", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "LIGHT array"}}},
      {ANARI_INSTANCE, "transform", "a group under an affine transform (Instance.cpp)",
          {{"group", ANARI_GROUP, true, nullptr, nullptr, nullptr, "the group"},
              {"transform", ANARI_FLOAT32_MAT4, false, kIdentity4, nullptr, nullptr, "object to world"},
              {"transform", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr,
                  "FLOAT32_MAT4 array: one instance of the group per element"},
              {"id", ANARI_UINT32, false, &kNoId, nullptr, nullptr, "instanceId value"},
              {"id", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr,
                  "UINT32 array: instanceId per transform-array element"}}},
      {ANARI_WORLD, nullptr, "the scene",
          {{"surface", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "SURFACE array"},
              {"volume", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "VOLUME array"},
              {"instance", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "INSTANCE array"},
              {"light", ANARI_ARRAY1D, false, nullptr, nullptr, nullptr, "LIGHT array"},
              {"id", ANARI_UINT32, false, &kNoId, nullptr, nullptr, "instanceId of world-level objects"}}},
      {ANARI_CAMERA, "perspective", "pinhole camera (Perspective.cpp)",
          {{"position", ANARI_FLOAT32_VEC3, false, kZero3, nullptr, nullptr, "eye position"},
              {"direction", ANARI_FLOAT32_VEC3, false, kDirZ, nullptr, nullptr, "view direction"},
              {"up", ANARI_FLOAT32_VEC3, false, kUpY, nullptr, nullptr, "up vector"},
              {"fovy", ANARI_FLOAT32, false, &kFovy, nullptr, nullptr, "vertical field of view (radians)"},
              {"aspect", ANARI_FLOAT32, false, &kOne, nullptr, nullptr, "width / height"},
              {"imageRegion", ANARI_FLOAT32_BOX2, false, kRegion, nullptr, nullptr, "rendered part of the image plane"}}},
      {ANARI_FRAME, nullptr, "frame buffer and render target",
          {{"world", ANARI_WORLD, true, nullptr, nullptr, nullptr, "the scene"},
              {"camera", ANARI_CAMERA, true, nullptr, nullptr, nullptr, "the camera"},
              {"renderer", ANARI_RENDERER, true, nullptr, nullptr, nullptr, "the renderer"},
              {"size", ANARI_UINT32_VEC2, true, nullptr, nullptr, nullptr, "width, height"},
              {"channel.color", ANARI_DATA_TYPE, false, &kSrgb, nullptr, nullptr, "UFIXED8_RGBA_SRGB | UFIXED8_VEC4 | FLOAT32_VEC4"},
              {"channel.depth", ANARI_DATA_TYPE, false, nullptr, nullptr, nullptr, "FLOAT32"},
              {"channel.primitiveId", ANARI_DATA_TYPE, false, nullptr, nullptr, nullptr, "UINT32"},
              {"channel.objectId", ANARI_DATA_TYPE, false, nullptr, nullptr, nullptr, "UINT32"},
              {"channel.instanceId", ANARI_DATA_TYPE, false, nullptr, nullptr, nullptr, "UINT32"},
              {"frameCompletionCallback", ANARI_FRAME_COMPLETION_CALLBACK, false, nullptr, nullptr, nullptr, "called when a frame completes"},
              {"frameCompletionCallbackUserData", ANARI_VOID_POINTER, false, nullptr, nullptr, nullptr, "callback argument"}}},
  };
  return t;
}

enum { kRaycast = 0, kAO = 1, kDirect = 2 };

std::vector<ParamDef> rendererParams(int kind)
{
  const bool ao = kind != kRaycast;
  std::vector<ParamDef> p = {
      {"background", ANARI_FLOAT32_VEC4, false, kBlack4, nullptr, nullptr, "colour of rays that miss"},
      {"ambientColor", ANARI_FLOAT32_VEC3, false, kOne3, nullptr, nullptr, "ambient light colour"},
      {"ambientRadiance", ANARI_FLOAT32, false, kind == kAO ? &kOne : &kZero, nullptr, nullptr,
          "ambient light intensity"},
      {"cullTriangleBackfaces", ANARI_BOOL, false, &kFalse, nullptr, nullptr, "skip back-facing triangles"},
      {"volumeSamplingRate", ANARI_FLOAT32, false, &kRate, &kRateMin, &kRateMax, "samples per half voxel"},
      {"sampleLimit", ANARI_INT32, false, &kLimit, &kInt0, nullptr, "accumulation limit (raycast: 1)"},
      {"stackDepthLDS", ANARI_INT32, false, &kInt0, &kInt0, nullptr, "traversal stack entries kept in LDS (0: default)"},
      {"volumeUnroll", ANARI_INT32, false, &kInt0, &kInt0, nullptr, "volume samples gathered per step (0: default)"},
      {"nodeWidth", ANARI_INT32, false, &kInt4, &kInt2, &kInt4, "BVH node width traversed: 2 or 4"},
      {"emptySpaceSkipping", ANARI_BOOL, false, &kTrue1, nullptr, nullptr,
          "skip empty 16^3 macro cells in volume marches (same image)"},
      {"volumeJitter", ANARI_BOOL, false, &kFalse, nullptr, nullptr,
          "start each volume segment at a random offset (per-pixel PCG32)"},
      {"longRaySteps", ANARI_INT32, false, &kInt0, &kInt0, nullptr,
          "raycast: primary walks past this many steps finish 16 lanes per ray (0: off; unset: tuned)"}};
  if (ao) {
    p.push_back({"pixelSamples", ANARI_INT32, false, &kInt1, &kInt1, nullptr, "samples per pixel per frame"});
    p.push_back({"ambientSamples", ANARI_INT32, false, &kInt1, &kInt0, &kInt256, "occlusion rays per hit"});
    p.push_back({"ambientOcclusionDistance", ANARI_FLOAT32, false, &kAoDist, nullptr, nullptr, "occlusion ray length"});
  }
  if (kind == kDirect) {
    p.push_back({"lightFalloff", ANARI_FLOAT32, false, &kOne, &kZero, &kOne, "accepted; not used in shading"});
    // matte surfaces spawn no bounce ray, so the bounce loop ends after the
    // first hit whatever the depth
    p.push_back({"maxRayDepth", ANARI_INT32, false, &kInt5, &kInt0, nullptr, "accepted; matte surfaces do not bounce"});
  }
  return p;
}

ObjDef *find(ANARIDataType type, const char *subtype)
{
  static std::vector<ObjDef> renderers;
  if (type == ANARI_RENDERER) {
    if (renderers.empty()) {
      renderers.push_back({ANARI_RENDERER, "raycast", "first-hit shading (Raycast_ptx.cu)", rendererParams(kRaycast), {}});
      renderers.push_back({ANARI_RENDERER, "ao", "ambient occlusion (AmbientOcclusion_ptx.cu)", rendererParams(kAO), {}});
      renderers.push_back({ANARI_RENDERER, "directLight",
          "lights with shadow rays and ambient occlusion (DirectLight_ptx.cu)", rendererParams(kDirect), {}});
      renderers.push_back({ANARI_RENDERER, "default", "the directLight renderer", rendererParams(kDirect), {}});
    }
    for (auto &o : renderers)
      if (subtype && std::strcmp(subtype, o.subtype) == 0)
        return &o;
    return nullptr;
  }
  for (auto &o : tables()) {
    if (o.type != type)
      continue;
    if (!o.subtype || (subtype && std::strcmp(subtype, o.subtype) == 0))
      return &o;
  }
  return nullptr;
}

} // namespace

// The extensions this device implements, named as ANARI's query tables name
// them ("ANARI_" + the extension's identifier in upper case); the reference
// lists its set in visrtx_device.json "dependencies", and
// VisRTXFeatureUtility.cpp:39-61 reads the same strings back.  CUDA arrays
// and output buffers, HDRI lights, physically based materials, background
// images, samplers, NanoVDB fields and the orthographic / depth-of-field
// cameras are not built here.
const char *const kDeviceExtensions[] = {"ANARI_KHR_ARRAY1D_REGION", "ANARI_KHR_CAMERA_PERSPECTIVE",
    "ANARI_KHR_DEVICE_SYNCHRONIZATION", "ANARI_KHR_FRAME_ACCUMULATION",
    "ANARI_KHR_FRAME_CHANNEL_PRIMITIVE_ID", "ANARI_KHR_FRAME_CHANNEL_OBJECT_ID",
    "ANARI_KHR_FRAME_CHANNEL_INSTANCE_ID", "ANARI_KHR_FRAME_COMPLETION_CALLBACK",
    "ANARI_KHR_GEOMETRY_CONE", "ANARI_KHR_GEOMETRY_CURVE", "ANARI_KHR_GEOMETRY_CYLINDER",
    "ANARI_KHR_GEOMETRY_QUAD", "ANARI_KHR_GEOMETRY_SPHERE", "ANARI_KHR_GEOMETRY_TRIANGLE",
    "ANARI_KHR_INSTANCE_TRANSFORM", "ANARI_KHR_INSTANCE_TRANSFORM_ARRAY",
    "ANARI_KHR_LIGHT_DIRECTIONAL", "ANARI_KHR_LIGHT_POINT", "ANARI_KHR_LIGHT_SPOT",
    "ANARI_KHR_MATERIAL_MATTE", "ANARI_KHR_RENDERER_AMBIENT_LIGHT",
    "ANARI_KHR_RENDERER_BACKGROUND_COLOR", "ANARI_KHR_SPATIAL_FIELD_STRUCTURED_REGULAR",
    "ANARI_KHR_VOLUME_TRANSFER_FUNCTION1D", "ANARI_VISRTX_TRIANGLE_BACK_FACE_CULLING", nullptr};

const char *const *deviceExtensions()
{
  return kDeviceExtensions;
}

namespace {

// anariGetObjectInfo(..., "feature", ANARI_STRING_LIST): the extensions an
// object subtype belongs to (what visrtxGetObjectExtensions,
// VisRTXFeatureUtility.cpp:64-77, reads)
const char *const *features(ANARIDataType type, const char *sub)
{
  static const char *const kTri[] = {"ANARI_KHR_GEOMETRY_TRIANGLE", "ANARI_VISRTX_TRIANGLE_BACK_FACE_CULLING", nullptr};
  static const char *const kQuad[] = {"ANARI_KHR_GEOMETRY_QUAD", nullptr};
  static const char *const kSphere[] = {"ANARI_KHR_GEOMETRY_SPHERE", nullptr};
  static const char *const kCyl[] = {"ANARI_KHR_GEOMETRY_CYLINDER", nullptr};
  static const char *const kCone[] = {"ANARI_KHR_GEOMETRY_CONE", nullptr};
  static const char *const kCurve[] = {"ANARI_KHR_GEOMETRY_CURVE", nullptr};
  static const char *const kRen[] = {"ANARI_KHR_RENDERER_AMBIENT_LIGHT", "ANARI_KHR_RENDERER_BACKGROUND_COLOR",
      "ANARI_VISRTX_TRIANGLE_BACK_FACE_CULLING", nullptr};
  static const char *const kDir[] = {"ANARI_KHR_LIGHT_DIRECTIONAL", nullptr};
  static const char *const kPoint[] = {"ANARI_KHR_LIGHT_POINT", nullptr};
  static const char *const kSpot[] = {"ANARI_KHR_LIGHT_SPOT", nullptr};
  static const char *const kCam[] = {"ANARI_KHR_CAMERA_PERSPECTIVE", nullptr};
  static const char *const kMatte[] = {"ANARI_KHR_MATERIAL_MATTE", nullptr};
  static const char *const kField[] = {"ANARI_KHR_SPATIAL_FIELD_STRUCTURED_REGULAR", nullptr};
  static const char *const kVol[] = {"ANARI_KHR_VOLUME_TRANSFER_FUNCTION1D", nullptr};
  static const char *const kInst[] = {"ANARI_KHR_INSTANCE_TRANSFORM", "ANARI_KHR_INSTANCE_TRANSFORM_ARRAY", nullptr};
  static const char *const kFrame[] = {"ANARI_KHR_FRAME_ACCUMULATION", "ANARI_KHR_FRAME_CHANNEL_PRIMITIVE_ID",
      "ANARI_KHR_FRAME_CHANNEL_OBJECT_ID", "ANARI_KHR_FRAME_CHANNEL_INSTANCE_ID",
      "ANARI_KHR_FRAME_COMPLETION_CALLBACK", nullptr};
  static const char *const kNone[] = {nullptr};
  const auto is = [&](const char *n) { return sub && std::strcmp(sub, n) == 0; };
  switch (type) {
  case ANARI_GEOMETRY:
    return is("triangle") ? kTri : is("quad") ? kQuad : is("sphere") ? kSphere
        : is("cylinder") ? kCyl : is("cone") ? kCone : is("curve") ? kCurve : nullptr;
  case ANARI_RENDERER:
    return kRen;
  case ANARI_LIGHT:
    return is("directional") ? kDir : is("point") ? kPoint : is("spot") ? kSpot : nullptr;
  case ANARI_CAMERA:
    return is("perspective") ? kCam : nullptr;
  case ANARI_MATERIAL:
    return is("matte") ? kMatte : nullptr;
  case ANARI_SPATIAL_FIELD:
    return is("structuredRegular") ? kField : nullptr;
  case ANARI_VOLUME:
    return is("transferFunction1D") ? kVol : nullptr;
  case ANARI_INSTANCE:
    return kInst;
  case ANARI_FRAME:
    return kFrame;
  default:
    return kNone;
  }
}

} // namespace

const void *objectInfo(ANARIDataType type, const char *subtype, const char *info, ANARIDataType infoType)
{
  ObjDef *o = find(type, subtype);
  if (!o || !info)
    return nullptr;
  if (std::strcmp(info, "feature") == 0 && infoType == ANARI_STRING_LIST)
    return features(type, subtype);
  if (std::strcmp(info, "description") == 0 && infoType == ANARI_STRING)
    return o->desc;
  if (std::strcmp(info, "parameter") == 0 && infoType == ANARI_PARAMETER_LIST) {
    if (o->list.empty()) {
      for (const ParamDef &p : o->params)
        o->list.push_back({p.name, p.type});
      o->list.push_back({nullptr, ANARI_UNKNOWN});
    }
    return o->list.data();
  }
  return nullptr;
}

const void *parameterInfo(ANARIDataType type, const char *subtype, const char *param,
    ANARIDataType paramType, const char *info, ANARIDataType infoType)
{
  static const int32_t kTrue = 1;
  ObjDef *o = find(type, subtype);
  if (!o || !param || !info)
    return nullptr;
  for (const ParamDef &p : o->params) {
    if (std::strcmp(p.name, param) != 0 || p.type != paramType)
      continue;
    if (std::strcmp(info, "required") == 0 && infoType == ANARI_BOOL)
      return p.required ? &kTrue : &kFalse;
    if (std::strcmp(info, "description") == 0 && infoType == ANARI_STRING)
      return p.desc;
    if (std::strcmp(info, "default") == 0 && infoType == p.type)
      return p.def;
    if (std::strcmp(info, "minimum") == 0 && infoType == p.type)
      return p.min;
    if (std::strcmp(info, "maximum") == 0 && infoType == p.type)
      return p.max;
    return nullptr;
  }
  return nullptr;
}

} // namespace vrx_anari
