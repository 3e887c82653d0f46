// scenegen.cpp -- seeded synthetic scenes (host, OpenMP).
//
// Generators for the benchmark and test scenes: PCG32 (O'Neill 2014) for
// random primitives, hashed-lattice value noise with smoothstep
// interpolation and octave sums (frequency x2, amplitude x1/2, normalised to
// [0,1)) for heightfields, volumes and displaced meshes.  These produce
// inputs only; nothing here is on the rendering path.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/vrx_hip.h"

namespace {

inline uint32_t fmix32(uint32_t h)
{
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

inline float lattice(int x, int y, int z, uint32_t seed)
{
  uint32_t h = seed * 0x9e3779b9u;
  h ^= (uint32_t)x * 0x8da6b343u;
  h = fmix32(h);
  h ^= (uint32_t)y * 0xd8163841u;
  h = fmix32(h);
  h ^= (uint32_t)z * 0xcb1ab31fu;
  h = fmix32(h);
  return (float)(h >> 8) * 0x1p-24f;
}

inline float smooth(float t) { return t * t * (3.0f - 2.0f * t); }
inline float lerp(float a, float b, float t) { return a + (b - a) * t; }

float value_noise(uint32_t seed, float x, float y, float z)
{
  float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  float tx = smooth(x - fx), ty = smooth(y - fy), tz = smooth(z - fz);
  float c000 = lattice(ix, iy, iz, seed), c100 = lattice(ix + 1, iy, iz, seed);
  float c010 = lattice(ix, iy + 1, iz, seed), c110 = lattice(ix + 1, iy + 1, iz, seed);
  float c001 = lattice(ix, iy, iz + 1, seed), c101 = lattice(ix + 1, iy, iz + 1, seed);
  float c011 = lattice(ix, iy + 1, iz + 1, seed), c111 = lattice(ix + 1, iy + 1, iz + 1, seed);
  float a = lerp(lerp(c000, c100, tx), lerp(c010, c110, tx), ty);
  float b = lerp(lerp(c001, c101, tx), lerp(c011, c111, tx), ty);
  return lerp(a, b, tz);
}

float fractal(uint32_t seed, int octaves, float x, float y, float z)
{
  float sum = 0.0f, amp = 1.0f, norm = 0.0f, f = 1.0f;
  for (int o = 0; o < octaves; o++) {
    sum += amp * value_noise(seed + 1013u * (uint32_t)o, x * f, y * f, z * f);
    norm += amp;
    amp *= 0.5f;
    f *= 2.0f;
  }
  return sum / norm;
}

struct Pcg32 {
  uint64_t state, inc;
  uint32_t next()
  {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
  }
  void seed(uint64_t initstate, uint64_t initseq)
  {
    state = 0u;
    inc = (initseq << 1u) | 1u;
    next();
    state += initstate;
    next();
  }
  float uniform() { return (float)(next() >> 8) * 0x1p-24f; }
};

} // namespace

extern "C" {

uint32_t vrx_gen_pcg32(uint64_t *state, uint64_t inc)
{
  Pcg32 r{*state, inc | 1u};
  uint32_t v = r.next();
  *state = r.state;
  return v;
}

float vrx_gen_value_noise3(uint64_t seed, float x, float y, float z)
{
  return value_noise((uint32_t)seed, x, y, z);
}

void vrx_gen_spheres(uint64_t seed, uint32_t n, float rmin, float rmax, float *centers, float *radii)
{
  // one PCG32 stream per 64k-sphere chunk so the result does not depend on
  // the thread count
  const uint32_t chunk = 65536;
  const uint32_t nchunks = (n + chunk - 1) / chunk;
  const double lr = log((double)rmax / (double)rmin);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < (int64_t)nchunks; c++) {
    Pcg32 r;
    r.seed(seed, (uint64_t)c);
    uint32_t beg = (uint32_t)c * chunk, end = beg + chunk < n ? beg + chunk : n;
    for (uint32_t i = beg; i < end; i++) {
      centers[3 * (size_t)i + 0] = r.uniform();
      centers[3 * (size_t)i + 1] = r.uniform();
      centers[3 * (size_t)i + 2] = r.uniform();
      float u = r.uniform();
      radii[i] = (float)((double)rmin * exp((double)u * lr));
    }
  }
}

void vrx_gen_heightfield(uint64_t seed, uint32_t nx, uint32_t ny, float amplitude, float *verts,
    uint32_t *indices)
{
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < (int64_t)ny; j++) {
    for (uint32_t i = 0; i < nx; i++) {
      float x = (float)i / (float)(nx - 1), y = (float)j / (float)(ny - 1);
      float z = amplitude * fractal((uint32_t)seed, 4, x * 4.0f, y * 4.0f, 0.5f);
      float *v = verts + 3 * ((size_t)j * nx + i);
      v[0] = x;
      v[1] = y;
      v[2] = z;
    }
  }
  if (!indices)
    return;
  size_t t = 0;
  for (uint32_t j = 0; j + 1 < ny; j++) {
    for (uint32_t i = 0; i + 1 < nx; i++) {
      uint32_t a = j * nx + i, b = a + 1, c = a + nx, d = c + 1;
      indices[t++] = a;
      indices[t++] = b;
      indices[t++] = d;
      indices[t++] = a;
      indices[t++] = d;
      indices[t++] = c;
    }
  }
}

void vrx_gen_noise_volume(uint64_t seed, uint32_t nx, uint32_t ny, uint32_t nz, int octaves,
    float base_freq, float *out)
{
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)nz; k++) {
    for (uint32_t j = 0; j < ny; j++) {
      for (uint32_t i = 0; i < nx; i++) {
        float x = nx > 1 ? (float)i / (float)(nx - 1) : 0.0f;
        float y = ny > 1 ? (float)j / (float)(ny - 1) : 0.0f;
        float z = nz > 1 ? (float)k / (float)(nz - 1) : 0.0f;
        out[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k)] =
            fractal((uint32_t)seed, octaves, x * base_freq, y * base_freq, z * base_freq);
      }
    }
  }
}

void vrx_gen_sphere_mesh_counts(uint32_t target_tris, uint32_t *nverts, uint32_t *ntris)
{
  uint32_t nv = (uint32_t)(sqrt((double)target_tris / 4.0) + 0.5);
  if (nv < 3)
    nv = 3;
  uint32_t nu = 2 * nv;
  // rings 1..nv-1 (nu verts each) + 2 poles; caps 2*nu tris, body 2*nu*(nv-2)
  *nverts = nu * (nv - 1) + 2;
  *ntris = 2 * nu + 2 * nu * (nv - 2);
}

void vrx_gen_sphere_mesh(uint64_t seed, uint32_t target_tris, float displacement, float *verts,
    uint32_t *indices)
{
  uint32_t nv = (uint32_t)(sqrt((double)target_tris / 4.0) + 0.5);
  if (nv < 3)
    nv = 3;
  uint32_t nu = 2 * nv;
  const float pi = 3.14159265358979f;
  auto put = [&](size_t vi, float x, float y, float z) {
    float n = fractal((uint32_t)seed, 3, x * 2.0f + 7.0f, y * 2.0f + 7.0f, z * 2.0f + 7.0f);
    float r = 1.0f + displacement * (n - 0.5f);
    verts[3 * vi] = x * r;
    verts[3 * vi + 1] = y * r;
    verts[3 * vi + 2] = z * r;
  };
  put(0, 0.0f, 0.0f, 1.0f);
  for (uint32_t j = 1; j < nv; j++) {
    float th = pi * (float)j / (float)nv;
    for (uint32_t i = 0; i < nu; i++) {
      float ph = 2.0f * pi * (float)i / (float)nu;
      put(1 + (size_t)(j - 1) * nu + i, sinf(th) * cosf(ph), sinf(th) * sinf(ph), cosf(th));
    }
  }
  const uint32_t south = 1 + nu * (nv - 1);
  put(south, 0.0f, 0.0f, -1.0f);
  size_t t = 0;
  for (uint32_t i = 0; i < nu; i++) {
    uint32_t a = 1 + i, b = 1 + (i + 1) % nu;
    indices[t++] = 0;
    indices[t++] = a;
    indices[t++] = b;
  }
  for (uint32_t j = 1; j + 1 < nv; j++) {
    uint32_t r0 = 1 + (j - 1) * nu, r1 = 1 + j * nu;
    for (uint32_t i = 0; i < nu; i++) {
      uint32_t i1 = (i + 1) % nu;
      indices[t++] = r0 + i;
      indices[t++] = r1 + i;
      indices[t++] = r1 + i1;
      indices[t++] = r0 + i;
      indices[t++] = r1 + i1;
      indices[t++] = r0 + i1;
    }
  }
  uint32_t rl = 1 + (nv - 2) * nu;
  for (uint32_t i = 0; i < nu; i++) {
    uint32_t a = rl + i, b = rl + (i + 1) % nu;
    indices[t++] = south;
    indices[t++] = b;
    indices[t++] = a;
  }
}

} // extern "C"
