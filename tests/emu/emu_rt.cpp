// CPU emulation of the libpmx_rt C ABI used by generated host stubs.
//
// Test infrastructure only: it runs a generated pipeline's kernels on the
// CPU so that index arithmetic, region boxes and LDS layouts can be
// checked (under AddressSanitizer) and compared with the oracle before
// any of it runs on a GPU.  "Device" memory is host memory; a launch runs
// every workgroup of the grid in turn on `blockDim.x` host threads that
// meet at a barrier for __syncthreads(); workgroup-shared arrays are
// function statics, shared by those threads.
#include <pthread.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>
#include "pmx_rt.h"
#include "emu_prelude.h"

thread_local pmx_dim3 threadIdx;
thread_local pmx_dim3 blockIdx;
pmx_dim3 blockDim;
pmx_dim3 gridDim;

static thread_local std::string g_err;
static pthread_barrier_t* g_bar = nullptr;

extern "C" const char* pmx_last_error(void) { return g_err.c_str(); }
extern "C" void pmx_set_error(const char* msg) { g_err = msg ? msg : ""; }
static int fail(const char* m) { pmx_set_error(m); return -1; }

void __syncthreads() { pthread_barrier_wait(g_bar); }

// lane exchange of a wavefront (__shfl): 64 threads share an exchange row
// and meet at their own barrier
unsigned long long* pmx_emu_xchg = nullptr;
extern "C" { int pmx_emu_misaligned = 0; }
static std::vector<pthread_barrier_t>* g_wbar = nullptr;
void pmx_emu_wave_sync() { pthread_barrier_wait(&(*g_wbar)[threadIdx.x / 64]); }

// ---- kernel registry ------------------------------------------------------
typedef void (*pmx_tramp)(void**);
static std::map<std::string, pmx_tramp>& registry() {
  static std::map<std::string, pmx_tramp> r;
  return r;
}
extern "C" void pmx_emu_register(const char* name, pmx_tramp t) { registry()[name] = t; }

extern "C" int pmx_compile_hip(const char*, const char*, const char*, const char* const*, int,
                               const char*, char*, size_t) { return fail("emulator: no compiler"); }
extern "C" int pmx_device_count(int* n) { *n = 1; return 0; }
extern "C" int pmx_device_name(int, char* buf, size_t cap) { snprintf(buf, cap, "cpu-emulator"); return 0; }

// ---- arena ------------------------------------------------------------------
static std::map<std::string, std::vector<unsigned char>*>& arena() {
  static std::map<std::string, std::vector<unsigned char>*> a;
  return a;
}
static int arena_get(const char* slot, size_t bytes, void** ptr, bool zero) {
  auto& a = arena();
  auto it = a.find(slot);
  if (it == a.end() || it->second->size() < bytes) {
    if (it != a.end()) delete it->second;
    auto* v = new std::vector<unsigned char>(bytes ? bytes : 1);
    if (!zero) memset(v->data(), 0xA5, v->size());   // poison: uninitialised reads show up
    a[slot] = v;
  }
  *ptr = a[slot]->data();
  return 0;
}
extern "C" int pmx_arena_get(const char* s, size_t b, void** p) { return arena_get(s, b, p, false); }
extern "C" int pmx_arena_get_zeroed(const char* s, size_t b, void** p) { return arena_get(s, b, p, true); }
extern "C" int pmx_arena_release(void) {
  for (auto& kv : arena()) delete kv.second;
  arena().clear();
  return 0;
}
extern "C" size_t pmx_arena_bytes(void) {
  size_t n = 0;
  for (auto& kv : arena()) n += kv.second->size();
  return n;
}

// ---- modules and launches -----------------------------------------------------
extern "C" int pmx_module_load_data(const void*, void** module) {
  static int token;
  *module = &token;
  return 0;
}
extern "C" int pmx_module_function(void*, const char* name, void** function) {
  auto it = registry().find(name);
  if (it == registry().end()) return fail("emulator: kernel not registered");
  *function = (void*)it->second;
  return 0;
}
extern "C" int pmx_module_unload(void*) { return 0; }
extern "C" int pmx_occupancy(void*, int, size_t, int* n) { *n = 2; return 0; }
extern "C" int pmx_device_cus(int, int* n) { *n = 4; return 0; }

struct Launch { pmx_tramp t; void** args; unsigned grid; unsigned tid; };
static void* worker(void* p) {
  Launch* L = (Launch*)p;
  threadIdx.x = L->tid; threadIdx.y = 0; threadIdx.z = 0;
  for (unsigned b = 0; b < L->grid; ++b) {
    blockIdx.x = b; blockIdx.y = 0; blockIdx.z = 0;
    L->t(L->args);
    pthread_barrier_wait(g_bar);     // workgroup done before the next reuses shared arrays
  }
  return nullptr;
}
extern "C" int pmx_launch(void* function, unsigned gx, unsigned gy, unsigned gz, unsigned bx,
                          unsigned by, unsigned bz, unsigned shmem, void* stream, void** args) {
  (void)shmem; (void)stream;
  if (gy != 1 || gz != 1 || by != 1 || bz != 1) return fail("emulator: 1-D launches only");
  if (bx == 0 || bx > 1024) return fail("emulator: block size");
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  blockDim = {bx, 1, 1};
  gridDim = {gx, 1, 1};
  pthread_barrier_t bar;
  pthread_barrier_init(&bar, nullptr, bx);
  g_bar = &bar;
  // one barrier per 64-thread wavefront (the whole block when smaller)
  std::vector<pthread_barrier_t> wbar((bx + 63) / 64);
  for (unsigned w = 0; w < wbar.size(); ++w)
    pthread_barrier_init(&wbar[w], nullptr, bx % 64 && w + 1 == wbar.size() ? bx % 64 : 64);
  g_wbar = &wbar;
  std::vector<unsigned long long> xchg(bx);
  pmx_emu_xchg = xchg.data();
  std::vector<pthread_t> th(bx);
  std::vector<Launch> L(bx);
  pthread_attr_t attr;
  pthread_attr_init(&attr);
  pthread_attr_setstacksize(&attr, 1 << 20);
  for (unsigned i = 0; i < bx; ++i) {
    L[i] = Launch{(pmx_tramp)function, args, gx, i};
    if (pthread_create(&th[i], &attr, worker, &L[i])) { fprintf(stderr, "emu: pthread_create\n"); abort(); }
  }
  for (unsigned i = 0; i < bx; ++i) pthread_join(th[i], nullptr);
  pthread_attr_destroy(&attr);
  pthread_barrier_destroy(&bar);
  for (auto& b : wbar) pthread_barrier_destroy(&b);
  g_bar = nullptr;
  g_wbar = nullptr;
  pmx_emu_xchg = nullptr;
  return 0;
}

// ---- streams, events, graphs ----------------------------------------------------------
extern "C" int pmx_stream_create(void** s) { static int token; *s = &token; return 0; }
extern "C" int pmx_stream_destroy(void*) { return 0; }
extern "C" int pmx_stream_sync(void*) { return 0; }
extern "C" int pmx_event_create(void** ev) { *ev = new int(0); return 0; }
extern "C" int pmx_event_destroy(void* ev) { delete (int*)ev; return 0; }
extern "C" int pmx_event_record(void*, void*) { return 0; }
extern "C" int pmx_event_elapsed_ms(void*, void*, float* ms) { *ms = 0.f; return 0; }
extern "C" int pmx_graph_begin(void*) { return fail("emulator: graph capture is not emulated"); }
extern "C" int pmx_graph_end(void*, void**) { return fail("emulator: graph capture is not emulated"); }
extern "C" int pmx_graph_launch(void*, void*) { return fail("emulator: graph capture is not emulated"); }
extern "C" int pmx_graph_destroy(void*) { return 0; }

// ---- memory ----------------------------------------------------------------------------
extern "C" int pmx_malloc(void** p, size_t bytes) {
  *p = malloc(bytes ? bytes : 1);
  if (!*p) return fail("emulator: malloc");
  memset(*p, 0xA5, bytes);
  return 0;
}
extern "C" int pmx_free(void* p) { free(p); return 0; }
extern "C" int pmx_memcpy_h2d(void* d, const void* s, size_t n, void*) { memcpy(d, s, n); return 0; }
extern "C" int pmx_memcpy_d2h(void* d, const void* s, size_t n, void*) { memcpy(d, s, n); return 0; }
extern "C" int pmx_memset_d8(void* d, int v, size_t n, void*) { memset(d, v, n); return 0; }
extern "C" void pool_init(void) {}
extern "C" void pool_destroy(void) {}

extern "C" int pmx_mlp_param_count(const int*, int, int*) { return fail("emulator"); }
extern "C" int pmx_mlp_forward(const float*, const int*, int, const float*, int, float*, void*) { return fail("emulator"); }
extern "C" int pmx_mlp_backward(const float*, const int*, int, const float*, const float*, int, float*, float*, void*) { return fail("emulator"); }
extern "C" int pmx_mlp_adam(float*, const float*, float*, float*, int, float, float, float, float, int, void*) { return fail("emulator"); }
extern "C" int pmx_mlp_train(float*, const int*, int, const float*, const float*, int, int, float, float, float, float, float*, float*, float*, float*, float*, int, void*) { return fail("emulator"); }

// static scratch of the allocation-free entry points: one zero-initialised
// host array standing in for the per-device array
namespace {
alignas(256) unsigned char g_emu_scratch[PMX_SCRATCH_BYTES];
bool g_emu_scratch_used = false;
}  // namespace
extern "C" int pmx_scratch(void** p, size_t* bytes) {
  *p = g_emu_scratch;
  if (bytes) *bytes = sizeof g_emu_scratch;
  g_emu_scratch_used = true;
  return 0;
}
extern "C" int pmx_scratch_peek(void** p) {
  *p = g_emu_scratch_used ? (void*)g_emu_scratch : nullptr;
  return 0;
}
