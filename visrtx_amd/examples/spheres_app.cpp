// spheres_app.cpp -- an ANARI client in C++ that renders random spheres with
// the MI355X device, following the object graph of the reference's
// examples/simple/testApp_spheres.cpp:167-263 (geometry -> surface -> world,
// perspective camera, raycast renderer, frame with colour, depth and id
// channels).  Only the library name differs from a visrtx client.
//
//   spheres_app [num_spheres] [width] [height] [out.ppm]
//
// Prints one line of key=value results (frame duration, hit pixels, a
// primitiveId checksum) and writes the colour channel as a binary PPM.
#include <anari/anari.h>
#include <anari/ext/visrtx_amd/visrtx_amd.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

int g_errors = 0;

void status(const void *, ANARIDevice, ANARIObject, ANARIDataType, ANARIStatusSeverity sev,
    ANARIStatusCode, const char *msg)
{
  if (sev <= ANARI_SEVERITY_ERROR) {
    std::fprintf(stderr, "[anari] error: %s\n", msg);
    g_errors++;
  } else if (sev == ANARI_SEVERITY_WARNING) {
    std::fprintf(stderr, "[anari] warning: %s\n", msg);
  }
}

// xorshift64*: a deterministic, dependency-free scene generator
struct Rng {
  uint64_t s;
  float next()
  {
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return (float)((s * 2685821657736338717ull) >> 40) * (1.0f / 16777216.0f);
  }
};

void setF3(ANARIDevice d, ANARIObject o, const char *name, float x, float y, float z)
{
  float v[3] = {x, y, z};
  anariSetParameter(d, o, name, ANARI_FLOAT32_VEC3, v);
}

} // namespace

int main(int argc, char **argv)
{
  const uint32_t n = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 10) : 100000u;
  const uint32_t W = argc > 2 ? (uint32_t)std::strtoul(argv[2], nullptr, 10) : 800u;
  const uint32_t H = argc > 3 ? (uint32_t)std::strtoul(argv[3], nullptr, 10) : 600u;
  const char *out = argc > 4 ? argv[4] : "spheres.ppm";
  // renderer subtype: "raycast" (default) or "directLight" (then the world
  // gets a directional light and the renderer a little ambient light)
  const char *rtype = argc > 5 ? argv[5] : "raycast";
  const bool lit = std::strcmp(rtype, "directLight") == 0;

  ANARILibrary lib = anariLoadLibrary("visrtx_amd", status, nullptr);
  if (!lib) {
    std::fprintf(stderr, "could not load the visrtx_amd library\n");
    return 1;
  }
  ANARIDevice dev = anariNewDevice(lib, "default");
  if (!dev) {
    std::fprintf(stderr, "could not create the device\n");
    return 1;
  }

  // scene data in application memory (shared arrays, as in testApp_spheres)
  std::vector<float> pos(3 * (size_t)n), rad(n);
  Rng rng{0x9E3779B97F4A7C15ull};
  for (uint32_t i = 0; i < n; i++) {
    pos[3 * i] = rng.next();
    pos[3 * i + 1] = rng.next();
    pos[3 * i + 2] = rng.next();
    rad[i] = 0.002f + 0.01f * rng.next();
  }

  ANARIGeometry geom = anariNewGeometry(dev, "sphere");
  if (!geom) { // the device initialises lazily: a failed start shows here
    std::fprintf(stderr, "device initialisation failed\n");
    anariRelease(dev, (ANARIObject)dev);
    anariUnloadLibrary(lib);
    return 1;
  }
  ANARIArray1D aPos = anariNewArray1D(dev, pos.data(), nullptr, nullptr, ANARI_FLOAT32_VEC3, n);
  ANARIArray1D aRad = anariNewArray1D(dev, rad.data(), nullptr, nullptr, ANARI_FLOAT32, n);
  anariSetParameter(dev, geom, "vertex.position", ANARI_ARRAY1D, &aPos);
  anariSetParameter(dev, geom, "vertex.radius", ANARI_ARRAY1D, &aRad);
  anariCommitParameters(dev, geom);
  anariRelease(dev, aPos);
  anariRelease(dev, aRad);

  ANARIMaterial mat = anariNewMaterial(dev, "matte");
  setF3(dev, mat, "color", 0.9f, 0.6f, 0.3f);
  anariCommitParameters(dev, mat);

  ANARISurface surf = anariNewSurface(dev);
  anariSetParameter(dev, surf, "geometry", ANARI_GEOMETRY, &geom);
  anariSetParameter(dev, surf, "material", ANARI_MATERIAL, &mat);
  uint32_t sid = 7;
  anariSetParameter(dev, surf, "id", ANARI_UINT32, &sid);
  anariCommitParameters(dev, surf);

  ANARIWorld world = anariNewWorld(dev);
  ANARIArray1D aSurf = anariNewArray1D(dev, &surf, nullptr, nullptr, ANARI_SURFACE, 1);
  anariSetParameter(dev, world, "surface", ANARI_ARRAY1D, &aSurf);
  ANARILight light = nullptr;
  if (lit) {
    light = anariNewLight(dev, "directional");
    setF3(dev, light, "direction", 0.3f, -1.0f, 0.5f);
    float irradiance = 2.0f;
    anariSetParameter(dev, light, "irradiance", ANARI_FLOAT32, &irradiance);
    anariCommitParameters(dev, light);
    ANARIArray1D aLight = anariNewArray1D(dev, &light, nullptr, nullptr, ANARI_LIGHT, 1);
    anariSetParameter(dev, world, "light", ANARI_ARRAY1D, &aLight);
    anariRelease(dev, aLight);
  }
  anariCommitParameters(dev, world);
  anariRelease(dev, aSurf);

  ANARICamera cam = anariNewCamera(dev, "perspective");
  setF3(dev, cam, "position", 0.5f, 0.5f, -1.2f);
  setF3(dev, cam, "direction", 0.0f, 0.0f, 1.0f);
  setF3(dev, cam, "up", 0.0f, 1.0f, 0.0f);
  float aspect = (float)W / (float)H;
  anariSetParameter(dev, cam, "aspect", ANARI_FLOAT32, &aspect);
  anariCommitParameters(dev, cam);

  ANARIRenderer ren = anariNewRenderer(dev, rtype);
  float bg[4] = {0.1f, 0.1f, 0.1f, 1.0f};
  anariSetParameter(dev, ren, "background", ANARI_FLOAT32_VEC4, bg);
  if (lit) {
    float ambient = 0.2f;
    anariSetParameter(dev, ren, "ambientRadiance", ANARI_FLOAT32, &ambient);
  }
  anariCommitParameters(dev, ren);

  ANARIFrame frame = anariNewFrame(dev);
  uint32_t size[2] = {W, H};
  ANARIDataType colorType = ANARI_UFIXED8_RGBA_SRGB, depthType = ANARI_FLOAT32,
                idType = ANARI_UINT32;
  anariSetParameter(dev, frame, "size", ANARI_UINT32_VEC2, size);
  anariSetParameter(dev, frame, "channel.color", ANARI_DATA_TYPE, &colorType);
  anariSetParameter(dev, frame, "channel.depth", ANARI_DATA_TYPE, &depthType);
  anariSetParameter(dev, frame, "channel.primitiveId", ANARI_DATA_TYPE, &idType);
  anariSetParameter(dev, frame, "world", ANARI_WORLD, &world);
  anariSetParameter(dev, frame, "camera", ANARI_CAMERA, &cam);
  anariSetParameter(dev, frame, "renderer", ANARI_RENDERER, &ren);
  anariCommitParameters(dev, frame);

  anariRenderFrame(dev, frame);
  anariFrameReady(dev, frame, ANARI_WAIT);
  float duration = 0.0f;
  anariGetProperty(dev, frame, "duration", ANARI_FLOAT32, &duration, sizeof(duration), ANARI_WAIT);

  uint32_t fw = 0, fh = 0;
  ANARIDataType ft = ANARI_UNKNOWN;
  const uint32_t *px = (const uint32_t *)anariMapFrame(dev, frame, "channel.color", &fw, &fh, &ft);
  const uint32_t *ids =
      (const uint32_t *)anariMapFrame(dev, frame, "channel.primitiveId", &fw, &fh, &ft);
  const float *depth = (const float *)anariMapFrame(dev, frame, "channel.depth", &fw, &fh, &ft);
  uint64_t hits = 0, checksum = 0;
  if (px && ids && depth) {
    for (size_t i = 0; i < (size_t)fw * fh; i++) {
      if (depth[i] < 1e29f) {
        hits++;
        checksum = checksum * 1099511628211ull + ids[i];
      }
    }
    if (FILE *f = std::fopen(out, "wb")) {
      std::fprintf(f, "P6\n%u %u\n255\n", fw, fh);
      // row 0 of the frame is the bottom row: write top-down
      for (int64_t y = (int64_t)fh - 1; y >= 0; y--)
        for (uint32_t x = 0; x < fw; x++) {
          uint32_t c = px[(size_t)y * fw + x];
          unsigned char rgb[3] = {(unsigned char)(c & 0xff), (unsigned char)((c >> 8) & 0xff),
              (unsigned char)((c >> 16) & 0xff)};
          std::fwrite(rgb, 1, 3, f);
        }
      std::fclose(f);
    }
  }
  anariUnmapFrame(dev, frame, "channel.color");
  anariUnmapFrame(dev, frame, "channel.primitiveId");
  anariUnmapFrame(dev, frame, "channel.depth");

  std::printf("spheres=%u size=%ux%u duration_ms=%.4f hits=%llu checksum=%016llx errors=%d\n", n,
      fw, fh, duration * 1000.0f, (unsigned long long)hits, (unsigned long long)checksum,
      g_errors);

  for (ANARIObject o : {(ANARIObject)frame, (ANARIObject)ren, (ANARIObject)cam,
           (ANARIObject)world, (ANARIObject)surf, (ANARIObject)mat, (ANARIObject)geom})
    anariRelease(dev, o);
  if (light)
    anariRelease(dev, light);
  anariRelease(dev, (ANARIObject)dev);
  anariUnloadLibrary(lib);
  return (px && hits > 0 && g_errors == 0) ? 0 : 2;
}
