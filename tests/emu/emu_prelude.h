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
#endif
