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
// rows): VALU-latency steps instead of LDS permutes.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xf, 0xf, false);
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x, int lane) {
  const int rl = lane & 15;
  uint32_t t;
  t = dpp_mov<0x111>(x);
  if (rl >= 1) x += t;
  t = dpp_mov<0x112>(x);
  if (rl >= 2) x += t;
  t = dpp_mov<0x114>(x);
  if (rl >= 4) x += t;
  t = dpp_mov<0x118>(x);
  if (rl >= 8) x += t;
  t = dpp_mov<0x142>(x);
  if ((lane & 31) >= 16) x += t;
  t = dpp_mov<0x143>(x);
  if (lane >= 32) x += t;
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

// Orders the LDS accesses of one wave around a cross-lane exchange: the
// LDS unit executes one wave's accesses in issue order, so only the compiler
// has to be kept from moving them across this point.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace cvm
