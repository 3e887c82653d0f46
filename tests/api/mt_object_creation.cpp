// mt_object_creation.cpp -- the reference's
// tests/api/TestMultiThreadedObjectCreation.cpp:57-95 for this device: four
// threads released together each create 100 x {directional light,
// perspective camera, triangle geometry, world, frame} and release them.
// Exits 0 when every creation succeeded and no error was reported.
#include <anari/anari.h>
#include <anari/ext/visrtx_amd/visrtx_amd.h>

#include <atomic>
#include <cstdio>
#include <thread>
#include <vector>

static std::atomic<int> g_errors{0};

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
  std::atomic<bool> gate{false};
  std::atomic<int> nullHandles{0};
  std::atomic<int> created{0};
  std::vector<std::thread> threads;
  for (int i = 0; i < 4; i++) {
    threads.emplace_back([&]() {
      while (!gate)
        std::this_thread::yield();
      std::vector<ANARIObject> objects;
      for (int j = 0; j < 100; j++) {
        objects.push_back(anariNewLight(device, "directional"));
        objects.push_back(anariNewCamera(device, "perspective"));
        objects.push_back(anariNewGeometry(device, "triangle"));
        objects.push_back(anariNewWorld(device));
        objects.push_back(anariNewFrame(device));
      }
      for (ANARIObject o : objects) {
        if (!o)
          nullHandles++;
        else
          created++;
        anariRelease(device, o);
      }
    });
  }
  gate = true;
  for (auto &t : threads)
    t.join();
  anariRelease(device, device);
  std::printf("errors=%d created=%d null=%d\n", g_errors.load(), created.load(), nullHandles.load());
  return g_errors == 0 && nullHandles == 0 && created == 2000 ? 0 : 1;
}
