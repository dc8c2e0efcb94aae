/* Host-only stand-in for the two HIP runtime names a generated host stub
 * uses directly, so the stub can be compiled against the CPU emulation of
 * libpmx_rt (tests/emu/emu_rt.cpp).  Test infrastructure only. */
#ifndef PMX_EMU_HIP_RUNTIME_API_H
#define PMX_EMU_HIP_RUNTIME_API_H
typedef int hipError_t;
enum { hipSuccess = 0 };
static inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
#endif
