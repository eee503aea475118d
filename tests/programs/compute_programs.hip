// User compute programs for the gprtComputeLaunch tests. Built into a gfx950 code object
// (tests/programs/Makefile: hipcc --genco) and loaded at run time with gprtModuleCreate, as an
// application loads its own device code. Each program is an ordinary HIP kernel with C linkage;
// its parameter list is what the host passes to gprtComputeLaunch, argument by argument.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gprt_hip_device.h"

struct Float3 {
  float x, y, z;
};

// The geometry record of the reference's compute-bounds sample
// (samples/s5-computePrograms/s5-0-computeAABBs/sharedCode.h:27-34).
struct AABBGeomData {
  Float3 *vertex;
  float *radius;
  Float3 *aabbs;
};

// One sphere's box per thread along x, as the sample's ComputeBounds
// (samples/s5-computePrograms/s5-0-computeAABBs/deviceCode.slang:58-69), which is launched with
// one thread per group; `count` bounds the grid when groups hold more threads.
extern "C" __global__ void ComputeBounds(AABBGeomData record, uint32_t count) {
  const uint32_t prim = blockIdx.x * blockDim.x + threadIdx.x;
  if (prim >= count) return;
  const Float3 p = record.vertex[prim];
  const float r = record.radius[prim];
  record.aabbs[2 * prim] = Float3{p.x - r, p.y - r, p.z - r};
  record.aabbs[2 * prim + 1] = Float3{p.x + r, p.y + r, p.z + r};
}

// Rays of the GPRT ray record (origin, tmin, direction, tmax) on a w x h grid over the unit
// square at z = zOrigin, pointing along +z.
struct Ray {
  float ox, oy, oz, tmin, dx, dy, dz, tmax;
};

extern "C" __global__ void OrthoRays(Ray *rays, uint32_t w, uint32_t h, float zOrigin) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= w || y >= h) return;
  Ray r;
  r.ox = ((float)x + 0.5f) / (float)w;
  r.oy = ((float)y + 0.5f) / (float)h;
  r.oz = zOrigin;
  r.tmin = 0.0f;
  r.dx = 0.0f;
  r.dy = 0.0f;
  r.dz = 1.0f;
  r.tmax = 1e30f;
  rays[y * w + x] = r;
}

// Arguments of mixed sizes: each sits at its natural alignment in the kernel's argument block.
extern "C" __global__ void Mixed(uint32_t n, uint64_t *out, uint16_t add, double scale, uint8_t flag) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = (uint64_t)((double)i * scale) + add + ((uint64_t)flag << 40);
}

// Grid and group shape: every thread writes its (group, thread) coordinates to its linear slot.
extern "C" __global__ void Grid3D(uint32_t *out) {
  const uint32_t g = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  const uint32_t t = (threadIdx.z * blockDim.y + threadIdx.y) * blockDim.x + threadIdx.x;
  const uint32_t slot = g * (blockDim.x * blockDim.y * blockDim.z) + t;
  out[2 * slot] = blockIdx.x | (blockIdx.y << 16) | (blockIdx.z << 24);
  out[2 * slot + 1] = threadIdx.x | (threadIdx.y << 10) | (threadIdx.z << 20);
}

// A full 256-byte argument block: the output pointer and 62 words.
struct Full {
  uint32_t *out;
  uint32_t words[62];
};

extern "C" __global__ void Echo(Full f) {
  const uint32_t i = threadIdx.x;
  if (i < 62) f.out[i] = f.words[i];
}

// Closest hits of a ray batch traced inside the program against the accel whose
// gprtAccelGetDeviceAddress is `accel` (a triangle or AABB accel).
extern "C" __global__ void TraceRays(const void *accel, const GPRTRay *rays, GPRTHit *hits, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  hits[i] = gprtTraceClosestHit(accel, rays[i]);
}

// The ray-generation program of the reference's compute-bounds sample
// (samples/s5-computePrograms/s5-0-computeAABBs/deviceCode.slang:30-55): a pinhole camera
// {pos, dir_00, dir_du, dir_dv} (sharedCode.h), one ray per pixel through the pixel centre,
// tmin 0, tmax 10000, traced against record.world. The sample turns the hit into a colour; this
// one stores the hit and the ray it traced.
struct Camera {
  Float3 pos, dir_00, dir_du, dir_dv;
};
struct RayGenData {
  const void *world;
  GPRTHit *frameBuffer;
  GPRTRay *rays;
};

extern "C" __global__ void SimpleRayGen(RayGenData record, Camera cam, uint32_t w, uint32_t h) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= w || y >= h) return;
  const float sx = ((float)x + 0.5f) / (float)w, sy = ((float)y + 0.5f) / (float)h;
  float d[3] = {cam.dir_00.x + sx * cam.dir_du.x + sy * cam.dir_dv.x,
                cam.dir_00.y + sx * cam.dir_du.y + sy * cam.dir_dv.y,
                cam.dir_00.z + sx * cam.dir_du.z + sy * cam.dir_dv.z};
  const float inv = 1.0f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  GPRTRay ray;
  ray.ox = cam.pos.x;
  ray.oy = cam.pos.y;
  ray.oz = cam.pos.z;
  ray.tmin = 0.0f;
  ray.dx = d[0] * inv;
  ray.dy = d[1] * inv;
  ray.dz = d[2] * inv;
  ray.tmax = 10000.0f;
  const uint32_t px = y * w + x;
  record.rays[px] = ray;
  record.frameBuffer[px] = gprtTraceClosestHit(record.world, ray);
}
