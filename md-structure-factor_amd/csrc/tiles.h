// Tile geometry shared by the deposition and the adjoint reduction.
#pragma once
#include <hip/hip_runtime.h>

namespace mdsf {

// Does the window [ir - A, ir + A) of a periodic axis of N cells (images at
// multiples of N) reach the cells [t0, t1)?  The candidate test of the tiled
// deposition; the adjoint reduction uses the same test to know which (atom,
// tile) result slots were written.
__device__ __forceinline__ bool window_reaches(int ir, int A, int t0, int t1, int N) {
  const int lo = ir - A, hi = ir + A;
#pragma unroll
  for (int j = -1; j <= 1; ++j) {
    const int a = max(lo, t0 + j * N), b = min(hi, t1 + j * N);
    if (a < b) return true;
  }
  return false;
}

}  // namespace mdsf
