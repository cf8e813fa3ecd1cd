// Shared helpers for the cdft HIP library: error recording, launch checks, XCD-aware block mapping.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>

namespace cdft {

// Error codes returned by every extern "C" entry point (0 = success, >0 = hipError_t).
enum : int {
  CDFT_OK = 0,
  CDFT_E_ARG = -1,       // invalid argument (null pointer, bad size, unsupported order)
  CDFT_E_LAUNCH = -2,    // kernel launch failed (message holds the HIP error string)
  CDFT_E_LIBRARY = -3,   // rocFFT / hipFFT call failed
  CDFT_E_CAPACITY = -4,  // a fixed-size workspace is too small for the request
};

// Last error message of the calling thread.
const char* last_error();
int fail(int code, const char* fmt, ...);

// Records a launch failure (if any) and returns the code.
int check_launch(const char* what);

// Remaps a linear block id so that consecutive logical tiles land on the same XCD (8 XCDs; the
// hardware dispatches block ids round-robin across them). Logical neighbours then share an L2.
__device__ __forceinline__ unsigned xcd_tile(unsigned block, unsigned nblocks) {
  const unsigned nxcd = 8u;
  const unsigned per = (nblocks + nxcd - 1u) / nxcd;
  const unsigned xcd = block % nxcd;
  const unsigned local = block / nxcd;
  const unsigned full = nblocks % nxcd;  // XCDs that receive `per` blocks (others per-1); 0 = all
  unsigned base;
  if (full == 0u || xcd < full) {
    base = xcd * per;
  } else {
    base = full * per + (xcd - full) * (per - 1u);
  }
  return base + local;
}

// Float64 products, sums and differences rounded one operation at a time, never contracted into a
// fused multiply-add. (The HIP headers define __dmul_rn / __dadd_rn as the plain operators, which
// hipcc contracts like any other.) The reference's CPU expressions round every operation, and the
// geometry compares lattice coordinates and distances against thresholds such as r <= 2h: one ulp
// moves a point that sits on a threshold across it.
__device__ __forceinline__ double rn_mul(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ double rn_add(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ double rn_sub(double a, double b) {
#pragma clang fp contract(off)
  return a - b;
}

}  // namespace cdft
