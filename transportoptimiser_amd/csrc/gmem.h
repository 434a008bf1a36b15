// Global-memory pointers in device code.
//
// A pointer read from a structure in device memory (job tables, parameter
// blocks) is a generic pointer to the compiler, which then accesses it with
// flat instructions.  Flat accesses count against the LDS/scalar wait counter
// as well as the vector-memory one, so every wait for an LDS result also
// waits for all outstanding global loads and stores.  Casting such pointers to
// the global address space (gmem) makes the accesses global_* instructions.
#ifndef ROADMC_GMEM_H
#define ROADMC_GMEM_H

#include <hip/hip_runtime.h>

#define RMC_GLOBAL __attribute__((address_space(1)))

template <class T>
__device__ __forceinline__ RMC_GLOBAL T* gmem(T* p) {
  return (RMC_GLOBAL T*)p;
}

#endif
