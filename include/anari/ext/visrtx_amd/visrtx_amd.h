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

/* Loader-style entry (VisRTXLibrary.cpp:70-74 ANARI_DEFINE_LIBRARY_ENTRYPOINT):
 * returns a device for anariNewDevice(library, subtype); subtype is ignored
 * as in VisRTXLibrary::newDevice (VisRTXLibrary.cpp:56-59). */
ANARI_INTERFACE ANARIDevice visrtx_amd_library_new_device(const char *subtype,
    ANARIStatusCallback cb, const void *userPtr);

#ifdef __cplusplus
}
#endif
#endif
