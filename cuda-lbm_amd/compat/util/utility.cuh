// utility.cuh — error-check helper names of src/util/utility.cuh. Errors are reported,
// never turned into exit().
#ifndef UTILITY_H
#define UTILITY_H
#include <hip/hip_runtime.h>
#include <stdio.h>
#define checkCudaErrors(val) lbm_compat_check((val), #val, __FILE__, __LINE__)
inline bool lbm_compat_check(hipError_t e, const char* what, const char* file, int line) {
    if (e == hipSuccess) return true;
    fprintf(stderr, "HIP error %s at %s:%d (%s)\n", hipGetErrorString(e), file, line, what);
    return false;
}
#endif
