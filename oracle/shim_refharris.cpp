// C-ABI adapter for the reference's generated refHarris, whose output
// parameter is a C++ reference (ipp_bench/harris/harris_function.cpp:9):
// it mallocs and zeroes the (R+2)x(C+2) output itself.  This shim copies
// that buffer into a caller-owned array and frees it.  ORACLE ONLY.
#include <cstdlib>
#include <cstring>
void refHarris(int C, int R, float* img, float*& harris);
extern "C" void ref_harris(int C, int R, float* img, float* out) {
  float* h = nullptr;
  refHarris(C, R, img, h);
  std::memcpy(out, h, sizeof(float) * (size_t)(R + 2) * (size_t)(C + 2));
  std::free(h);
}
