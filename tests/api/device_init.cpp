// device_init.cpp -- the reference's tests/api/TestDeviceInit.cpp:37-44 for
// this device: create the device, set "forceInit" and commit it on the device
// handle (device initialisation happens there, VisRTXDevice.cpp:460-518),
// release.  Exits 0 when no error was reported and the device answered
// after initialisation.
#include <anari/anari.h>
#include <anari/ext/visrtx_amd/visrtx_amd.h>

#include <cstdio>

static int g_errors = 0;

static void status(const void *, ANARIDevice, ANARIObject, ANARIDataType,
    ANARIStatusSeverity sev, ANARIStatusCode, const char *msg)
{
  if (sev <= ANARI_SEVERITY_ERROR) {
    std::fprintf(stderr, "[anari] error: %s\n", msg);
    g_errors++;
  }
}

int main()
{
  ANARIDevice device = makeVisRTXAMDDevice(status, nullptr);
  if (!device)
    return 2;
  const int yes = 1;
  anariSetParameter(device, device, "forceInit", ANARI_BOOL, &yes);
  anariCommitParameters(device, device);
  int hipDevice = -1;
  const int ok = anariGetProperty(
      device, device, "hipDevice", ANARI_INT32, &hipDevice, sizeof(hipDevice), ANARI_WAIT);
  anariRelease(device, device);
  std::printf("errors=%d initialised=%d hipDevice=%d\n", g_errors, ok, hipDevice);
  return g_errors == 0 && ok && hipDevice >= 0 ? 0 : 1;
}
