/*
 * anari.h -- the subset of the ANARI 1.0 C API that the visrtx_amd device
 * implements, written by this project (the ANARI-SDK is not vendored in the
 * VisRTX tree and is not installed here).
 *
 * Function names, argument order and meaning follow the Khronos ANARI 1.0
 * specification, which is what VisRTX's device implements through the SDK's
 * helium::BaseDevice (devices/rtx/VisRTXDevice.h:56-121).
 *
 * Source of the numeric values.  Neither the VisRTX tree nor this image holds
 * an ANARI-SDK header (anari_enums.h) or any file with the numbers, so none
 * of them is checked against one.  Each group below is marked:
 *   [spec]        restated from the ANARI 1.0 specification's enumerations
 *                 (type list order, severity and status tables) from memory;
 *   [unverified]  as [spec], but the position in the type list is uncertain
 *                 (the box, region and matrix types after the sRGB types).
 * An application compiled against the real SDK headers must pass the same
 * values; a mismatch shows as "unsupported type" warnings or ignored
 * parameters, never as memory corruption, because every parameter is
 * checked against the type it is read as.
 */
#ifndef VRX_ANARI_H
#define VRX_ANARI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ANARI_INTERFACE __attribute__((visibility("default")))

#ifdef __cplusplus
/* as in the SDK's C++ view of the handles: a device handle converts to an
   object handle (anariSetParameter(device, device, ...)) */
struct _ANARIObject {};
struct _ANARIDevice : _ANARIObject {};
#endif
typedef struct _ANARILibrary *ANARILibrary;
typedef struct _ANARIDevice *ANARIDevice;
typedef struct _ANARIObject *ANARIObject;
typedef ANARIObject ANARIArray;
typedef ANARIObject ANARIArray1D;
typedef ANARIObject ANARIArray2D;
typedef ANARIObject ANARIArray3D;
typedef ANARIObject ANARICamera;
typedef ANARIObject ANARIFrame;
typedef ANARIObject ANARIGeometry;
typedef ANARIObject ANARIGroup;
typedef ANARIObject ANARIInstance;
typedef ANARIObject ANARILight;
typedef ANARIObject ANARIMaterial;
typedef ANARIObject ANARIRenderer;
typedef ANARIObject ANARISurface;
typedef ANARIObject ANARISampler;
typedef ANARIObject ANARISpatialField;
typedef ANARIObject ANARIVolume;
typedef ANARIObject ANARIWorld;

typedef int ANARIDataType;
/* one entry of an ANARI_PARAMETER_LIST (anariGetObjectInfo "parameter"),
   terminated by {NULL, ANARI_UNKNOWN} */
typedef struct {
  const char *name;
  ANARIDataType type;
} ANARIParameter;
typedef int ANARIStatusSeverity;
typedef int ANARIStatusCode;
typedef int ANARIWaitMask;

/* data types [spec] */
#define ANARI_UNKNOWN 0
#define ANARI_DATA_TYPE 100
#define ANARI_STRING 101
#define ANARI_VOID_POINTER 102
#define ANARI_BOOL 103
#define ANARI_STRING_LIST 150
#define ANARI_TYPE_LIST 151
#define ANARI_PARAMETER_LIST 152
#define ANARI_FUNCTION_POINTER 200
#define ANARI_MEMORY_DELETER 201
#define ANARI_STATUS_CALLBACK 202
#define ANARI_FRAME_COMPLETION_CALLBACK 203
#define ANARI_LIBRARY 500
#define ANARI_DEVICE 501
#define ANARI_OBJECT 502
#define ANARI_ARRAY 503
#define ANARI_ARRAY1D 504
#define ANARI_ARRAY2D 505
#define ANARI_ARRAY3D 506
#define ANARI_CAMERA 507
#define ANARI_FRAME 508
#define ANARI_GEOMETRY 509
#define ANARI_GROUP 510
#define ANARI_INSTANCE 511
#define ANARI_LIGHT 512
#define ANARI_MATERIAL 513
#define ANARI_RENDERER 514
#define ANARI_SURFACE 515
#define ANARI_SAMPLER 516
#define ANARI_SPATIAL_FIELD 517
#define ANARI_VOLUME 518
#define ANARI_WORLD 519
#define ANARI_INT8 1000
#define ANARI_UINT8 1004
#define ANARI_INT16 1008
#define ANARI_UINT16 1012
#define ANARI_INT32 1016
#define ANARI_INT32_VEC2 1017
#define ANARI_INT32_VEC3 1018
#define ANARI_INT32_VEC4 1019
#define ANARI_UINT32 1020
#define ANARI_UINT32_VEC2 1021
#define ANARI_UINT32_VEC3 1022
#define ANARI_UINT32_VEC4 1023
#define ANARI_INT64 1024
#define ANARI_UINT64 1028
#define ANARI_UFIXED8 1036
#define ANARI_UFIXED8_VEC4 1039
#define ANARI_FLOAT32 1068
#define ANARI_FLOAT32_VEC2 1069
#define ANARI_FLOAT32_VEC3 1070
#define ANARI_FLOAT32_VEC4 1071
#define ANARI_FLOAT64 1072
#define ANARI_UFIXED8_R_SRGB 2000
#define ANARI_UFIXED8_RA_SRGB 2001
#define ANARI_UFIXED8_RGB_SRGB 2002
#define ANARI_UFIXED8_RGBA_SRGB 2003
/* [unverified] box and matrix types */
#define ANARI_FLOAT32_BOX1 2004
#define ANARI_FLOAT32_BOX2 2005
#define ANARI_FLOAT32_BOX3 2006
#define ANARI_FLOAT32_BOX4 2007
#define ANARI_FLOAT64_BOX1 2008
#define ANARI_FLOAT32_MAT4 2026
#define ANARI_FLOAT32_MAT3x4 2028

/* status callback severities [spec]: numbered from 1 (fatal) to 6 (debug) */
#define ANARI_SEVERITY_FATAL_ERROR 1
#define ANARI_SEVERITY_ERROR 2
#define ANARI_SEVERITY_WARNING 3
#define ANARI_SEVERITY_PERFORMANCE_WARNING 4
#define ANARI_SEVERITY_INFO 5
#define ANARI_SEVERITY_DEBUG 6

/* status codes [spec] */
#define ANARI_STATUS_NO_ERROR 0
#define ANARI_STATUS_UNKNOWN_ERROR 1
#define ANARI_STATUS_INVALID_ARGUMENT 2
#define ANARI_STATUS_INVALID_OPERATION 3
#define ANARI_STATUS_OUT_OF_MEMORY 4
#define ANARI_STATUS_UNSUPPORTED_DEVICE 5
#define ANARI_STATUS_VERSION_MISMATCH 6

/* wait masks [spec] */
#define ANARI_NO_WAIT 0
#define ANARI_WAIT 1

typedef void (*ANARIStatusCallback)(const void *userData, ANARIDevice device,
    ANARIObject source, ANARIDataType sourceType, ANARIStatusSeverity severity,
    ANARIStatusCode code, const char *message);
typedef void (*ANARIMemoryDeleter)(const void *userData, const void *appMemory);
typedef void (*ANARIFrameCompletionCallback)(const void *userData, ANARIDevice device,
    ANARIFrame frame);

/* library / device */
ANARI_INTERFACE ANARILibrary anariLoadLibrary(const char *name, ANARIStatusCallback statusCallback,
    const void *statusCallbackUserData);
ANARI_INTERFACE void anariUnloadLibrary(ANARILibrary);
ANARI_INTERFACE const char **anariGetDeviceSubtypes(ANARILibrary);
ANARI_INTERFACE ANARIDevice anariNewDevice(ANARILibrary, const char *type);

/* arrays */
ANARI_INTERFACE ANARIArray1D anariNewArray1D(ANARIDevice, const void *appMemory,
    ANARIMemoryDeleter deleter, const void *userData, ANARIDataType dataType, uint64_t numItems1);
ANARI_INTERFACE ANARIArray2D anariNewArray2D(ANARIDevice, const void *appMemory,
    ANARIMemoryDeleter deleter, const void *userData, ANARIDataType dataType, uint64_t numItems1,
    uint64_t numItems2);
ANARI_INTERFACE ANARIArray3D anariNewArray3D(ANARIDevice, const void *appMemory,
    ANARIMemoryDeleter deleter, const void *userData, ANARIDataType dataType, uint64_t numItems1,
    uint64_t numItems2, uint64_t numItems3);
ANARI_INTERFACE void *anariMapArray(ANARIDevice, ANARIArray);
ANARI_INTERFACE void anariUnmapArray(ANARIDevice, ANARIArray);

/* objects */
ANARI_INTERFACE ANARILight anariNewLight(ANARIDevice, const char *type);
ANARI_INTERFACE ANARICamera anariNewCamera(ANARIDevice, const char *type);
ANARI_INTERFACE ANARIGeometry anariNewGeometry(ANARIDevice, const char *type);
ANARI_INTERFACE ANARISpatialField anariNewSpatialField(ANARIDevice, const char *type);
ANARI_INTERFACE ANARIVolume anariNewVolume(ANARIDevice, const char *type);
ANARI_INTERFACE ANARISurface anariNewSurface(ANARIDevice);
ANARI_INTERFACE ANARIMaterial anariNewMaterial(ANARIDevice, const char *type);
ANARI_INTERFACE ANARISampler anariNewSampler(ANARIDevice, const char *type);
ANARI_INTERFACE ANARIGroup anariNewGroup(ANARIDevice);
ANARI_INTERFACE ANARIInstance anariNewInstance(ANARIDevice, const char *type);
ANARI_INTERFACE ANARIWorld anariNewWorld(ANARIDevice);
ANARI_INTERFACE ANARIRenderer anariNewRenderer(ANARIDevice, const char *type);
ANARI_INTERFACE ANARIFrame anariNewFrame(ANARIDevice);

/* parameters and lifetime */
ANARI_INTERFACE void anariSetParameter(ANARIDevice, ANARIObject, const char *name,
    ANARIDataType dataType, const void *mem);
ANARI_INTERFACE void anariUnsetParameter(ANARIDevice, ANARIObject, const char *name);
ANARI_INTERFACE void anariUnsetAllParameters(ANARIDevice, ANARIObject);
ANARI_INTERFACE void anariCommitParameters(ANARIDevice, ANARIObject);
ANARI_INTERFACE void anariRelease(ANARIDevice, ANARIObject);
ANARI_INTERFACE void anariRetain(ANARIDevice, ANARIObject);

/* introspection */
ANARI_INTERFACE const char **anariGetObjectSubtypes(ANARIDevice, ANARIDataType objectType);
ANARI_INTERFACE const void *anariGetObjectInfo(ANARIDevice, ANARIDataType objectType,
    const char *objectSubtype, const char *infoName, ANARIDataType infoType);
ANARI_INTERFACE const void *anariGetParameterInfo(ANARIDevice, ANARIDataType objectType,
    const char *objectSubtype, const char *parameterName, ANARIDataType parameterType,
    const char *infoName, ANARIDataType infoType);
ANARI_INTERFACE int anariGetProperty(ANARIDevice, ANARIObject, const char *name,
    ANARIDataType type, void *mem, uint64_t size, ANARIWaitMask mask);

/* frames */
ANARI_INTERFACE const void *anariMapFrame(ANARIDevice, ANARIFrame, const char *channel,
    uint32_t *width, uint32_t *height, ANARIDataType *pixelType);
ANARI_INTERFACE void anariUnmapFrame(ANARIDevice, ANARIFrame, const char *channel);
ANARI_INTERFACE void anariRenderFrame(ANARIDevice, ANARIFrame);
ANARI_INTERFACE int anariFrameReady(ANARIDevice, ANARIFrame, ANARIWaitMask);
ANARI_INTERFACE void anariDiscardFrame(ANARIDevice, ANARIFrame);

#ifdef __cplusplus
}
#endif
#endif
