/* Definitions that let generated HIP kernel source compile as host C++
 * for the CPU emulator (test infrastructure only). */
#ifndef PMX_EMU_PRELUDE_H
#define PMX_EMU_PRELUDE_H
#include <cmath>
#include <cstdint>
#include <math.h>
struct pmx_dim3 { unsigned x, y, z; };
extern thread_local pmx_dim3 threadIdx;
extern thread_local pmx_dim3 blockIdx;
extern pmx_dim3 blockDim;
extern pmx_dim3 gridDim;
#define __device__
#define __host__
#define __forceinline__ inline
#define __global__
#define __launch_bounds__(n)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))
void __syncthreads();
void pmx_emu_wave_sync();
extern unsigned long long* pmx_emu_xchg;
#include <cstring>
// HIP's __shfl on a wavefront of 64 threads (every lane takes part)
template <typename T> static inline T __shfl(T v, int src, int width = 64) {
  static_assert(sizeof(T) <= 8, "__shfl of a wider type");
  (void)width;
  unsigned long long w = 0;
  memcpy(&w, &v, sizeof(T));
  pmx_emu_xchg[threadIdx.x] = w;
  pmx_emu_wave_sync();
  const unsigned long long r = pmx_emu_xchg[(threadIdx.x & ~63u) + ((unsigned)src & 63u)];
  pmx_emu_wave_sync();
  T out;
  memcpy(&out, &r, sizeof(T));
  return out;
}
// HIP's 4-wide vector types as 16-byte aligned structs
struct __attribute__((aligned(16))) float4 { float x, y, z, w; };
struct __attribute__((aligned(16))) int4 { int x, y, z, w; };
struct __attribute__((aligned(16))) uint4 { unsigned x, y, z, w; };
// 128-bit loads must be 16-byte aligned on the device: a misaligned one is
// counted (the harness fails the run) and served byte-wise
extern "C" int pmx_emu_misaligned;
#define PMX_LD4 1
template <typename V, typename T> static inline V pmx_ld4(const T* p) {
  V v;
  if (((uintptr_t)p & 15u) != 0) __atomic_fetch_add(&pmx_emu_misaligned, 1, __ATOMIC_SEQ_CST);
  memcpy(&v, p, sizeof(V));
  return v;
}
#define __builtin_nontemporal_store(v, p) (*(p) = (v))
#define __builtin_amdgcn_readfirstlane(v) (v)
#define PMX_GLOBAL
#define PMX_CONST
#define PMX_OPAQUE(x)
static inline int __mul24(int a, int b) { return a * b; }
static inline unsigned __umul24(unsigned a, unsigned b) { return a * b; }
static inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
static inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
static inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
  return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
static inline unsigned atomicAdd(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
template <typename T> static inline T pmx_emu_cas(T* p, T cmp, T val) {
  __atomic_compare_exchange_n(p, &cmp, val, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
  return cmp;
}
static inline int atomicCAS(int* p, int cmp, int val) { return pmx_emu_cas(p, cmp, val); }
static inline unsigned atomicCAS(unsigned* p, unsigned cmp, unsigned val) { return pmx_emu_cas(p, cmp, val); }
template <typename T> static inline T pmx_emu_minmax(T* p, T v, bool mn) {
  T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while ((mn ? v < old : v > old) &&
         !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return old;
}
static inline int atomicMin(int* p, int v) { return pmx_emu_minmax(p, v, true); }
static inline unsigned atomicMin(unsigned* p, unsigned v) { return pmx_emu_minmax(p, v, true); }
static inline int atomicMax(int* p, int v) { return pmx_emu_minmax(p, v, false); }
static inline unsigned atomicMax(unsigned* p, unsigned v) { return pmx_emu_minmax(p, v, false); }
#endif
