// Device view of the fixed-point fields of a family (prep.h, quantize): the
// integer a 32-bit word stands for, in units of 2^-e.  Plain words are the
// integer; block words hold a 27-bit signed mantissa over a 5-bit shift
// (integer = mantissa * 2^shift).  Bit 0 of a family's mode flags marks v in
// block mode, bit 1 marks w.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbls {

__device__ __forceinline__ double fixed_dvalue(int32_t word, bool block) {
  return block ? ldexp((double)(word >> 5), word & 31) : (double)word;
}

__device__ __forceinline__ long long fixed_ivalue(int32_t word, bool block) {
  return block ? (long long)(word >> 5) << (word & 31) : (long long)word;
}

}  // namespace fbls
