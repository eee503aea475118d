// User compute programs: modules (gfx950 code objects loaded with hipModuleLoadData), compute
// handles (one kernel of a module, found by symbol) and launches on the context's stream.
// Replaces GPRT's SPIR-V module and compute-pipeline path (gprt/gprt.cpp:2030-2145 Compute,
// :9416-9479 _gprtComputeLaunch): there a launch binds the pipeline and pushes a 256-byte
// push-constant block; here the packed arguments are the kernel's argument block.
#include <cstring>

#include "internal.hpp"

namespace gprt {

// A module stays allocated while compute handles refer to it, so a launch through a handle of
// a destroyed module finds `mod == nullptr` and reports it instead of using a freed module.
struct Module {
  Context *ctx = nullptr;
  hipModule_t mod = nullptr;
  uint32_t refs = 1;  // the module itself + one per compute handle
};

struct Compute {
  Context *ctx = nullptr;
  Module *module = nullptr;
  hipFunction_t fn = nullptr;
  std::string name;
};

static void module_release(Module *m) {
  if (--m->refs == 0) delete m;
}

// The argument block handed to the runtime: the packed arguments, zero padded. The runtime
// copies each declared kernel parameter from its offset in this block, so a kernel that
// declares more parameters than the caller packed reads zeros from here, never past it.
constexpr size_t kArgBlock = 4096;

}  // namespace gprt

using namespace gprt;

GPRT_API GPRTModule gprtModuleCreate(GPRTContext _ctx, const void *code, size_t bytes) {
  Context *ctx = (Context *)_ctx;
  if (!ctx || !code) {
    fail("gprtModuleCreate: null argument");
    return nullptr;
  }
  // An ELF code object or a clang offload bundle; anything else is not device code for HIP.
  static const char kElf[4] = {0x7f, 'E', 'L', 'F'};
  static const char kBundle[] = "__CLANG_OFFLOAD_BUNDLE__";
  const bool elf = bytes >= 64 && std::memcmp(code, kElf, 4) == 0;
  const bool bundle = bytes >= 64 && std::memcmp(code, kBundle, sizeof(kBundle) - 1) == 0;
  if (!elf && !bundle) {
    fail("gprtModuleCreate: %zu bytes are neither a code object nor an offload bundle (build the device code "
         "with hipcc --genco --offload-arch=gfx950)",
         bytes);
    return nullptr;
  }
  if (!hip_ok(hipSetDevice(ctx->device), "hipSetDevice")) return nullptr;
  hipModule_t mod = nullptr;
  if (!hip_ok(hipModuleLoadData(&mod, code), "gprtModuleCreate: hipModuleLoadData")) return nullptr;
  Module *m = new Module();
  m->ctx = ctx;
  m->mod = mod;
  return (GPRTModule)m;
}

GPRT_API void gprtModuleDestroy(GPRTModule _m) {
  Module *m = (Module *)_m;
  if (!m || !m->mod) return;
  hipStreamSynchronize(m->ctx->stream);  // launches of its kernels may still be in flight
  hipModuleUnload(m->mod);
  m->mod = nullptr;
  module_release(m);
}

GPRT_API GPRTCompute gprtComputeCreate(GPRTContext _ctx, GPRTModule _m, const char *entrypoint) {
  Context *ctx = (Context *)_ctx;
  Module *m = (Module *)_m;
  if (!ctx || !m || !entrypoint) {
    fail("gprtComputeCreate: null argument");
    return nullptr;
  }
  if (!m->mod) {
    fail("gprtComputeCreate: the module was destroyed");
    return nullptr;
  }
  if (m->ctx != ctx) {
    fail("gprtComputeCreate: the module belongs to another context");
    return nullptr;
  }
  hipFunction_t fn = nullptr;
  if (hipModuleGetFunction(&fn, m->mod, entrypoint) != hipSuccess) {
    hipGetLastError();
    fail("gprtComputeCreate: the module has no program named \"%s\" (declare it extern \"C\" __global__)",
         entrypoint);
    return nullptr;
  }
  Compute *c = new Compute();
  c->ctx = ctx;
  c->module = m;
  c->fn = fn;
  c->name = entrypoint;
  ++m->refs;
  return (GPRTCompute)c;
}

GPRT_API void gprtComputeDestroy(GPRTCompute _c) {
  Compute *c = (Compute *)_c;
  if (!c) return;
  module_release(c->module);
  delete c;
}

GPRT_API uint64_t gprtComputeLaunchArgs(GPRTCompute _c, const uint32_t numGroups[3], const uint32_t groupSize[3],
                                        const void *args, size_t bytes) {
  const StageRange range("gprtComputeLaunchArgs");
  Compute *c = (Compute *)_c;
  if (!c || !numGroups || !groupSize || (bytes && !args)) {
    fail("gprtComputeLaunch: null argument");
    return 0;
  }
  if (bytes > GPRT_PUSH_CONSTANTS_LIMIT) {
    fail("gprtComputeLaunch: %zu bytes of arguments exceed GPRT_PUSH_CONSTANTS_LIMIT (%d)", bytes,
         GPRT_PUSH_CONSTANTS_LIMIT);
    return 0;
  }
  // The reference's launch asserts (gprt/gprt_host.h:2549-2554).
  for (int i = 0; i < 3; ++i) {
    if (numGroups[i] > GPRT_WORKGROUP_LIMIT) {
      fail("gprtComputeLaunch: Workgroup count exceeds WORKGROUP_LIMIT (%u groups in dimension %d, limit %d). "
           "Increase numthreads and reduce the number of workgroups.",
           numGroups[i], i, GPRT_WORKGROUP_LIMIT);
      return 0;
    }
  }
  const uint64_t threads = (uint64_t)groupSize[0] * groupSize[1] * groupSize[2];
  if (threads == 0 || threads > GPRT_THREADGROUP_LIMIT) {
    fail("gprtComputeLaunch: Thread count per group (%llu) must be 1..THREADGROUP_LIMIT (%d).",
         (unsigned long long)threads, GPRT_THREADGROUP_LIMIT);
    return 0;
  }
  if (!c->module->mod) {
    fail("gprtComputeLaunch: the module of program \"%s\" was destroyed", c->name.c_str());
    return 0;
  }
  if (numGroups[0] == 0 || numGroups[1] == 0 || numGroups[2] == 0) return 0;  // an empty dispatch
  Context *ctx = c->ctx;
  alignas(16) unsigned char block[kArgBlock] = {};
  if (bytes) std::memcpy(block, args, bytes);
  size_t blockBytes = kArgBlock;
  void *extra[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, block, HIP_LAUNCH_PARAM_BUFFER_SIZE, &blockBytes,
                   HIP_LAUNCH_PARAM_END};
  hipEvent_t a = nullptr, b = nullptr;
  const bool timed = ctx->timer.enabled;
  if (timed) {
    a = ctx->timer.get();
    b = ctx->timer.get();
    hipEventRecord(a, ctx->stream);
  }
  const hipError_t e = hipModuleLaunchKernel(c->fn, numGroups[0], numGroups[1], numGroups[2], groupSize[0],
                                             groupSize[1], groupSize[2], 0, ctx->stream, nullptr, extra);
  if (timed) {
    hipEventRecord(b, ctx->stream);
    ctx->timer.pending.push_back({c->name, a, b});
  }
  hip_ok(e, "gprtComputeLaunch: hipModuleLaunchKernel");
  return 0;
}
