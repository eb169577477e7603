// putil.h -- command-line number parsing (see putil.cpp).
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
int clw_parse_uint64(uint64_t* result, const char* str, uint64_t lo, uint64_t hi);
#ifdef __cplusplus
}
#endif
