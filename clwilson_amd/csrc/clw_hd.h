// clw_hd.h -- host/device qualifiers. The arithmetic headers compile both as
// gfx950 device code (hipcc) and as plain host C++ (g++), so the same source
// can be unit-tested on the CPU.
#pragma once
#include <stdint.h>
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CLW_HD __host__ __device__ __forceinline__
#else
#define CLW_HD static inline
#endif
