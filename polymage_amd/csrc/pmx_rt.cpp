// libpmx_rt: native runtime of the MI355X PolyMage backend.
// See include/pmx_rt.h for the contract of every entry point.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "pmx_rt.h"

namespace {
thread_local std::string g_err;

int fail(const char* where, int code, const char* what) {
  char buf[512];
  snprintf(buf, sizeof buf, "%s: %s (code %d)", where, what ? what : "", code);
  g_err = buf;
  return code ? -std::abs(code) : -1;
}

int hip_check(hipError_t e, const char* where) {
  if (e == hipSuccess) return 0;
  return fail(where, (int)e, hipGetErrorString(e));
}

struct Slot {
  void* ptr = nullptr;
  size_t bytes = 0;
};
std::mutex g_arena_mu;
std::map<std::pair<int, std::string>, Slot> g_arena;
}  // namespace

extern "C" {

const char* pmx_last_error(void) { return g_err.c_str(); }
void pmx_set_error(const char* msg) { g_err = msg ? msg : ""; }

int pmx_compile_hip(const char* src, const char* name, const char* arch,
                    const char* const* opts, int nopts, const char* out_path,
                    char* log, size_t log_cap) {
  if (log && log_cap) log[0] = 0;
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, src, name, 0, nullptr, nullptr);
  if (r != HIPRTC_SUCCESS) return fail("hiprtcCreateProgram", (int)r, hiprtcGetErrorString(r));
  std::vector<std::string> all;
  all.push_back(std::string("--offload-arch=") + (arch ? arch : "gfx950"));
  for (int i = 0; i < nopts; ++i) all.push_back(opts[i]);
  std::vector<const char*> cargs;
  for (auto& s : all) cargs.push_back(s.c_str());
  r = hiprtcCompileProgram(prog, (int)cargs.size(), cargs.data());
  size_t lsz = 0;
  hiprtcGetProgramLogSize(prog, &lsz);
  if (log && log_cap && lsz > 1) {
    std::vector<char> tmp(lsz + 1, 0);
    hiprtcGetProgramLog(prog, tmp.data());
    size_t n = std::min(log_cap - 1, lsz);
    memcpy(log, tmp.data(), n);
    log[n] = 0;
  }
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return fail("hiprtcCompileProgram", (int)r, hiprtcGetErrorString(r));
  }
  size_t csz = 0;
  hiprtcGetCodeSize(prog, &csz);
  std::vector<char> code(csz);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  std::string tmp_path = std::string(out_path) + ".tmp";
  FILE* f = fopen(tmp_path.c_str(), "wb");
  if (!f) return fail("pmx_compile_hip", 1, "cannot open output");
  size_t w = fwrite(code.data(), 1, csz, f);
  fclose(f);
  if (w != csz) return fail("pmx_compile_hip", 1, "short write");
  if (rename(tmp_path.c_str(), out_path) != 0) return fail("pmx_compile_hip", 1, "rename");
  return 0;
}

int pmx_device_count(int* n) { return hip_check(hipGetDeviceCount(n), "hipGetDeviceCount"); }

int pmx_device_name(int dev, char* buf, size_t cap) {
  hipDeviceProp_t p;
  int rc = hip_check(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
  if (rc) return rc;
  snprintf(buf, cap, "%s", p.gcnArchName);
  return 0;
}

int pmx_arena_get(const char* slot, size_t bytes, void** ptr) {
  int dev = 0;
  int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_arena_mu);
  Slot& s = g_arena[{dev, std::string(slot)}];
  if (s.bytes < bytes) {
    if (s.ptr) (void)hipFree(s.ptr);
    s.ptr = nullptr;
    s.bytes = 0;
    size_t want = (bytes + 4095) & ~size_t(4095);
    rc = hip_check(hipMalloc(&s.ptr, want), "hipMalloc(arena)");
    if (rc) return rc;
    s.bytes = want;
  }
  *ptr = s.ptr;
  return 0;
}

int pmx_arena_get_zeroed(const char* slot, size_t bytes, void** ptr) {
  int dev = 0;
  int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  bool fresh = false;
  {
    std::lock_guard<std::mutex> lk(g_arena_mu);
    auto it = g_arena.find({dev, std::string(slot)});
    fresh = (it == g_arena.end() || it->second.bytes < bytes);
  }
  rc = pmx_arena_get(slot, bytes, ptr);
  if (rc) return rc;
  if (fresh) {
    rc = hip_check(hipMemset(*ptr, 0, bytes), "hipMemset(arena)");
    if (rc) return rc;
    rc = hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  }
  return rc;
}

int pmx_arena_release(void) {
  std::lock_guard<std::mutex> lk(g_arena_mu);
  for (auto& kv : g_arena)
    if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  g_arena.clear();
  return 0;
}

size_t pmx_arena_bytes(void) {
  std::lock_guard<std::mutex> lk(g_arena_mu);
  size_t n = 0;
  for (auto& kv : g_arena) n += kv.second.bytes;
  return n;
}

int pmx_module_load_data(const void* image, void** module) {
  hipModule_t m;
  int rc = hip_check(hipModuleLoadData(&m, image), "hipModuleLoadData");
  if (rc) return rc;
  *module = (void*)m;
  return 0;
}

int pmx_module_function(void* module, const char* name, void** function) {
  hipFunction_t f;
  int rc = hip_check(hipModuleGetFunction(&f, (hipModule_t)module, name), name);
  if (rc) return rc;
  *function = (void*)f;
  return 0;
}

int pmx_module_unload(void* module) {
  return hip_check(hipModuleUnload((hipModule_t)module), "hipModuleUnload");
}

int pmx_launch(void* function, unsigned gx, unsigned gy, unsigned gz,
               unsigned bx, unsigned by, unsigned bz, unsigned shmem,
               void* stream, void** args) {
  if (gx == 0 || gy == 0 || gz == 0) return 0;
  return hip_check(hipModuleLaunchKernel((hipFunction_t)function, gx, gy, gz, bx, by, bz,
                                         shmem, (hipStream_t)stream, args, nullptr),
                   "hipModuleLaunchKernel");
}

int pmx_occupancy(void* function, int threads, size_t dyn_lds, int* blocks_per_cu) {
  int n = 0;
  int rc = hip_check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(
                         &n, (hipFunction_t)function, threads, dyn_lds),
                     "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
  if (rc) return rc;
  *blocks_per_cu = n;
  return 0;
}

int pmx_device_cus(int dev, int* n) {
  return hip_check(hipDeviceGetAttribute(n, hipDeviceAttributeMultiprocessorCount, dev),
                   "hipDeviceGetAttribute");
}

int pmx_stream_create(void** stream) {
  hipStream_t s;
  int rc = hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
  if (rc) return rc;
  *stream = (void*)s;
  return 0;
}
int pmx_stream_destroy(void* stream) {
  return hip_check(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
}

int pmx_stream_sync(void* stream) {
  return hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

int pmx_event_create(void** ev) {
  hipEvent_t e;
  int rc = hip_check(hipEventCreate(&e), "hipEventCreate");
  if (rc) return rc;
  *ev = (void*)e;
  return 0;
}
int pmx_event_destroy(void* ev) { return hip_check(hipEventDestroy((hipEvent_t)ev), "hipEventDestroy"); }
int pmx_event_record(void* ev, void* stream) {
  return hip_check(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream), "hipEventRecord");
}
int pmx_event_elapsed_ms(void* start, void* stop, float* ms) {
  int rc = hip_check(hipEventSynchronize((hipEvent_t)stop), "hipEventSynchronize");
  if (rc) return rc;
  return hip_check(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop),
                   "hipEventElapsedTime");
}

int pmx_graph_begin(void* stream) {
  return hip_check(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal),
                   "hipStreamBeginCapture");
}
int pmx_graph_end(void* stream, void** graph_exec) {
  hipGraph_t g;
  int rc = hip_check(hipStreamEndCapture((hipStream_t)stream, &g), "hipStreamEndCapture");
  if (rc) return rc;
  hipGraphExec_t ex;
  rc = hip_check(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  (void)hipGraphDestroy(g);
  if (rc) return rc;
  *graph_exec = (void*)ex;
  return 0;
}
int pmx_graph_launch(void* graph_exec, void* stream) {
  return hip_check(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream), "hipGraphLaunch");
}
int pmx_graph_destroy(void* graph_exec) {
  return hip_check(hipGraphExecDestroy((hipGraphExec_t)graph_exec), "hipGraphExecDestroy");
}

int pmx_malloc(void** ptr, size_t bytes) { return hip_check(hipMalloc(ptr, bytes ? bytes : 1), "hipMalloc"); }
int pmx_free(void* ptr) { return hip_check(hipFree(ptr), "hipFree"); }
int pmx_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  return hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream),
                   "hipMemcpyAsync(H2D)");
}
int pmx_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  return hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream),
                   "hipMemcpyAsync(D2H)");
}
int pmx_memset_d8(void* dst, int value, size_t bytes, void* stream) {
  return hip_check(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream), "hipMemsetAsync");
}

void pool_init(void) {}
void pool_destroy(void) { pmx_arena_release(); }

}  // extern "C"
