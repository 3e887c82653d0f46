/*
 * visrtx_amd.h -- direct construction of the MI355X ANARI device.
 *
 * Counterpart of include/anari/ext/visrtx/visrtx.h:49-51 and
 * VisRTXLibrary.cpp:76-80 (makeVisRTXDevice) in the VisRTX tree: an
 * application that links this library instead of going through the ANARI
 * loader creates the device with this call.
 */
#ifndef VISRTX_AMD_H
#define VISRTX_AMD_H

#include "../../anari.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Same shape as makeVisRTXDevice(ANARIStatusCallback, const void *). */
ANARI_INTERFACE ANARIDevice makeVisRTXAMDDevice(ANARIStatusCallback cb, const void *userPtr);

/* The device behind this library's own anariNewDevice(library, subtype):
 * subtype is ignored, as in VisRTXLibrary::newDevice (VisRTXLibrary.cpp:
 * 56-59).  It is not the ANARI-SDK's loader entry point: that one
 * (ANARI_DEFINE_LIBRARY_ENTRYPOINT, VisRTXLibrary.cpp:70-74) returns a
 * library object whose devices implement the SDK's anari::DeviceImpl, which
 * this library does not provide. */
ANARI_INTERFACE ANARIDevice visrtx_amd_library_new_device(const char *subtype,
    ANARIStatusCallback cb, const void *userPtr);

#ifdef __cplusplus
}
#endif
#endif
