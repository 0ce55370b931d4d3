// Device helpers shared by the normalisation kernels (cvm_norm.hip) and the
// fused direct correlation + scoring kernel (cvm_direct.hip): the
// first-occurrence extrema order, the channel masks of packed pixel words,
// the DPP wave scan of the horizontal prefix sums and the FP64 window score.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cvm {

// Extrema are (value, raster index) pairs compared lexicographically: strict
// value compares, the smaller index on equal values -- the sequential
// first-occurrence scan of impl.go:1023-1047.
__device__ __forceinline__ bool better_min(float v, int i, float bv, int bi) { return v < bv || (v == bv && i < bi); }
__device__ __forceinline__ bool better_max(float v, int i, float bv, int bi) { return v > bv || (v == bv && i < bi); }

// Channel-byte mask of a packed pixel word: bytes of channels >= cn cleared.
__device__ __forceinline__ uint32_t chan_mask(int cn) { return cn >= 4 ? 0xFFFFFFFFu : ((1u << (8 * cn)) - 1u); }

// Inclusive add-scan of a uint32 over the 64 lanes of a wave with DPP moves
// (row_shr 1/2/4/8 inside each 16-lane row, then row_bcast 15/31 across
// rows): VALU-latency steps instead of LDS permutes.  Each move writes 0 to
// the lanes that take nothing in its step (bound_ctrl for a source outside
// the row, the row mask for the broadcasts), so every step is one add that
// the compiler folds with its move into a single DPP add.
template <int CTRL, int ROW_MASK, bool BOUND_CTRL>
__device__ __forceinline__ uint32_t dpp_or_zero(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, BOUND_CTRL);
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x, int lane) {
  (void)lane;
  x += dpp_or_zero<0x111, 0xf, true>(x);  // row_shr:1
  x += dpp_or_zero<0x112, 0xf, true>(x);  // row_shr:2
  x += dpp_or_zero<0x114, 0xf, true>(x);  // row_shr:4
  x += dpp_or_zero<0x118, 0xf, true>(x);  // row_shr:8
  x += dpp_or_zero<0x142, 0xa, false>(x);  // row_bcast:15 into rows 1 and 3
  x += dpp_or_zero<0x143, 0xc, false>(x);  // row_bcast:31 into rows 2 and 3
  return x;
}

// The FP64 score of one window from exact integers (the FP64 oracle's
// formula): num = area * corr - sum_k S_k T_k, idiff = area * s2 - sum_k S_k^2,
// every input an exact integer below 2^53.
__device__ __forceinline__ double score_fp64(double numI, double idiff, double s2, double aread, double varSum) {
  // The guard is fl(idiff / area) > lim with lim = min(0.5, 10 * 2^-23 * s2)
  // >= 0.  idiff and area are exact non-negative integers, so idiff > area
  // (a quotient above 1) always passes it and idiff == 0 (a quotient of +0)
  // always fails it: only 0 < idiff <= area needs the division.
  if (!(idiff > aread)) {
    if (idiff == 0.0) return 0.0;
    const double diff2 = idiff / aread;
    double lim = 1.1920928955078125e-06 * s2;  // 10 * 2^-23 * s2
    if (lim > 0.5) lim = 0.5;
    if (!(diff2 > lim)) return 0.0;
  }
  const double denI = sqrt(idiff) * sqrt(varSum);
  const double an = fabs(numI);
  if (an < denI) return numI / denI;
  if (an < denI * 1.125) return numI > 0.0 ? 1.0 : -1.0;
  return 0.0;
}

// The reference's binary32 sqrt, float32(math.Sqrt(float64(x))), and its
// binary32 quotient: both are the correctly rounded binary32 results (for
// sqrt the binary64 root rounded again to binary32 is the correctly rounded
// one, since 53 >= 2 * 24 + 2).  sqrtf and '/' compile to correctly rounded
// binary32 sequences here (no fast-math), a third of the instructions of the
// binary64 root; sqrtf was checked against the binary64 root rounded to
// binary32 on every finite non-negative float on the device, and
// cvm_selftest_fp runs these very functions.
__device__ __forceinline__ float sqrt_rn_f32(float x) { return sqrtf(x); }
__device__ __forceinline__ float div_rn_f32(float a, float b) { return a / b; }

// Orders the LDS accesses of one wave around a cross-lane exchange: the
// LDS unit executes one wave's accesses in issue order, so only the compiler
// has to be kept from moving them across this point.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace cvm
