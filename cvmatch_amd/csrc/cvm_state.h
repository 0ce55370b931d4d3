// Device bodies of the call-state and column-sum kernels, shared by their own
// launches (cvm_norm.hip) and by the launches that run them beside other
// work (cvm_fft.hip): the template statistics, the alpha scans, the state's
// reset, and the chunk column sums of the wave-form normalisation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cvm_internal.h"
#include "cvm_score.h"

namespace cvm {

// Rows of one chunk of the column sums (and of one normalisation band).
constexpr int NB_ROWS = 16;

// The state's accumulators and results cleared (thread 0).
__device__ __forceinline__ void state_clear(CallState* st) {
  if (threadIdx.x == 0) {
    for (int k = 0; k < 4; k++) {
      st->tsum[k] = 0;
      st->tsq[k] = 0;
    }
    st->varSum = 0;
    st->cn = 0;
    st->flat = 0;
    st->alpha_bad[0] = st->alpha_bad[1] = 0;
    st->templNorm = st->numScale = st->varScale = 0.f;
  }
}

// cn choice (Match cvmatch.go:82-85), varSum, flat test (impl.go:1137),
// templNorm (impl.go:1087), varScale / numScale (impl.go:862-870), from the
// state's accumulated statistics and alpha flags into its result fields
// (one thread).
__device__ __forceinline__ void state_finalize_fields(CallState* st, int tw, int th, int cn_mode) {
  int cn = cn_mode;
  if (cn_mode == 0) cn = (st->alpha_bad[0] == 0 && st->alpha_bad[1] == 0) ? 3 : 4;
  long long area = (long long)tw * th;
  long long var = 0;
  for (int k = 0; k < cn; k++) {
    long long s = (long long)st->tsum[k], q = (long long)st->tsq[k];
    var += area * q - s * s;
  }
  st->cn = cn;
  st->varSum = var;
  st->flat = var == 0;
  double dtw = (double)tw, dth = (double)th;
  st->templNorm = (float)sqrt((double)var / (dtw * dth));
  st->varScale = (float)(1.0 / (dtw * dth));
  st->numScale = cn == 1 ? (float)((double)(long long)st->tsum[0] / (dtw * dth)) : st->varScale;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// templStats impl.go:1063-1079 (sums and square sums per channel, exact).
// grid.y walks template rows, threads walk columns; no divisions.
// One workgroup (bx, by) of a gx x gy grid over the template.
__device__ __forceinline__ void tpl_stats_block(const uint8_t* __restrict__ tpl, int64_t tstride, int tw, int th,
                                                int cnMax, int step, CallState* st, int bx, int by, int gx, int gy) {
  __shared__ unsigned long long red[4][16];  // [k][w]: sums, [k][8 + w]: squares (<= 8 waves)
  unsigned long long s[4] = {0, 0, 0, 0}, q[4] = {0, 0, 0, 0};
  for (int y = by; y < th; y += gy) {
    const uint8_t* row = tpl + (int64_t)y * tstride;
    for (int x = bx * blockDim.x + threadIdx.x; x < tw; x += gx * blockDim.x) {
      const uint8_t* p = row + (int64_t)x * step;
      for (int k = 0; k < 4; k++)
        if (k < cnMax) {
          unsigned long long v = p[k];
          s[k] += v;
          q[k] += v * v;
        }
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < 4; k++) {
    unsigned long long a = wave_sum_u64(s[k]), b = wave_sum_u64(q[k]);
    if (lane == 0) {
      red[k][w] = a;
      red[k][8 + w] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x < 4 && (int)threadIdx.x < cnMax) {
    int k = threadIdx.x;
    unsigned long long a = 0, b = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
      a += red[k][i];
      b += red[k][8 + i];
    }
    atomicAdd(&st->tsum[k], a);
    atomicAdd(&st->tsq[k], b);
  }
}

// alphaConst cvmatch.go:146-168: every alpha byte equals the first one.
// Like the reference's scan it stops at the first mismatch: workgroups poll
// the flag between rows and leave once any workgroup has published one (the
// answer is then "not constant" whatever the rest of the image holds).
// grid.y walks rows, grid.x x threads cover a row in 16-byte groups of four
// pixels when rows are 16-byte aligned, else in single pixels.
// One workgroup (bx, by) of a gx x gy grid over rows [y_begin, h).
__device__ __forceinline__ void alpha_scan_block(const uint8_t* __restrict__ pix, int64_t stride, int w, int y_begin,
                                                 int h, CallState* st, int which, int vec, int bx, int by, int gx,
                                                 int gy) {
  const uint32_t a0 = pix[3];
  volatile unsigned int* flag = &st->alpha_bad[which];
  int bad = 0;
  const int w4 = vec ? (w >> 2) : 0;
  const uint32_t want = a0 * 0x01010101u;
  for (int y = y_begin + by; y < h; y += gy) {
    if (*flag) break;
    const uint8_t* row = pix + (int64_t)y * stride;
    for (int q = bx * blockDim.x + threadIdx.x; q < w4; q += gx * blockDim.x) {
      uint4 v = *(const uint4*)(row + (int64_t)q * 16);
      uint32_t al = (v.x >> 24) | ((v.y >> 24) << 8) | ((v.z >> 24) << 16) | ((v.w >> 24) << 24);
      if (al != want) bad = 1;
    }
    for (int x = 4 * w4 + bx * blockDim.x + threadIdx.x; x < w; x += gx * blockDim.x)
      if (row[(int64_t)x * 4 + 3] != a0) bad = 1;
    if (bad) break;
  }
  if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(&st->alpha_bad[which], 1u);
}

// The same statistics over a flat pixel index: workgroup b of nb takes
// pixels b * blockDim + tid + k * nb * blockDim, four of them in flight at
// a time, so a small template needs a workgroup or two and few atomics.
__device__ __forceinline__ void tpl_stats_flat(const uint8_t* __restrict__ tpl, int64_t tstride, int tw, int th,
                                               int cnMax, int step, CallState* st, int b, int nb) {
  __shared__ unsigned long long red[4][16];  // [k][w]: sums, [k][8 + w]: squares (<= 8 waves)
  unsigned long long s[4] = {0, 0, 0, 0}, q[4] = {0, 0, 0, 0};
  const int n = tw * th, stride = nb * (int)blockDim.x;
  // four-byte pixels on four-byte boundaries: one dword a pixel
  const bool words = step == 4 && ((((uintptr_t)tpl) | (uintptr_t)tstride) & 3) == 0;
  const uint32_t cm = chan_mask(cnMax);
  for (int i0 = b * (int)blockDim.x + (int)threadIdx.x; i0 < n; i0 += 4 * stride) {
    uint32_t w[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i = i0 + u * stride;
      w[u] = 0;
      if (i < n) {
        const int y = i / tw, x = i - y * tw;
        const uint8_t* p = tpl + (int64_t)y * tstride + (int64_t)x * step;
        if (words) {
          w[u] = *(const uint32_t*)p & cm;
        } else {
          for (int k = 0; k < 4; k++)
            if (k < cnMax) w[u] |= (uint32_t)p[k] << (8 * k);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      for (int k = 0; k < 4; k++) {
        const unsigned long long v = (w[u] >> (8 * k)) & 255u;
        s[k] += v;
        q[k] += v * v;
      }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < 4; k++) {
    unsigned long long a = wave_sum_u64(s[k]), c = wave_sum_u64(q[k]);
    if (lane == 0) {
      red[k][wv] = a;
      red[k][8 + wv] = c;
    }
  }
  __syncthreads();
  if (threadIdx.x < 4 && (int)threadIdx.x < cnMax) {
    int k = threadIdx.x;
    unsigned long long a = 0, c = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
      a += red[k][i];
      c += red[k][8 + i];
    }
    atomicAdd(&st->tsum[k], a);
    atomicAdd(&st->tsq[k], c);
  }
}

// The alpha scan without polling: workgroup (bx, by) checks rows
// [8 by, 8 by + 8) (rows past h re-read the last one) of its columns, the
// eight rows' loads in flight together.
__device__ __forceinline__ void alpha_scan_rows8(const uint8_t* __restrict__ pix, int64_t stride, int w, int h,
                                                 CallState* st, int which, int vec, int bx, int by, int gx) {
  const uint32_t a0 = pix[3];
  const uint32_t want = a0 * 0x01010101u;
  const int w4 = vec ? (w >> 2) : 0;
  const int y0 = by * 8;
  int bad = 0;
  for (int q = bx * blockDim.x + threadIdx.x; q < w4; q += gx * blockDim.x) {
    uint4 v[8];
#pragma unroll
    for (int r = 0; r < 8; r++) v[r] = *(const uint4*)(pix + (int64_t)min(y0 + r, h - 1) * stride + (int64_t)q * 16);
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const uint32_t al = (v[r].x >> 24) | ((v[r].y >> 24) << 8) | ((v[r].z >> 24) << 16) | ((v[r].w >> 24) << 24);
      if (al != want) bad = 1;
    }
  }
  for (int x = 4 * w4 + bx * blockDim.x + threadIdx.x; x < w; x += gx * blockDim.x)
    for (int r = 0; r < 8; r++)
      if (y0 + r < h && pix[(int64_t)(y0 + r) * stride + (int64_t)x * 4 + 3] != a0) bad = 1;
  if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(&st->alpha_bad[which], 1u);
}

// The independent parts of a call state in one launch: workgroups [0, nS)
// take the template statistics (an sgx x sgy grid), the next pgx x pgy the
// first ph0 rows of the parent's alpha scan (a varying alpha plane shows up
// there, and the rest of the scan then leaves at its first flag check; with
// ph0 = ih the whole scan), the last tgx x tgy the template's alpha scan;
// rows8: the compact layout (tpl_stats_flat, alpha_scan_rows8).  b is the
// workgroup's index among them (256 threads).
__device__ __forceinline__ void call_state_block(int b, const CallStateArgs& a) {
  const int nS = a.sgx * a.sgy, nP = a.pgx * a.pgy;
  if (a.rows8) {
    // the compact layout: flat statistics (sgx workgroups), scans of 8 rows
    if (b < nS) {
      tpl_stats_flat(a.tpl, a.tstride, a.tw, a.th, a.cnMax, a.step, a.st, b, nS);
    } else if (b < nS + nP) {
      b -= nS;
      alpha_scan_rows8(a.img, a.istride, a.iw, a.ph0, a.st, 0, a.ivec, b % a.pgx, b / a.pgx, a.pgx);
    } else {
      b -= nS + nP;
      alpha_scan_rows8(a.tpl, a.tstride, a.tw, a.th, a.st, 1, a.tvec, b % a.tgx, b / a.tgx, a.tgx);
    }
    return;
  }
  if (b < nS) {
    tpl_stats_block(a.tpl, a.tstride, a.tw, a.th, a.cnMax, a.step, a.st, b % a.sgx, b / a.sgx, a.sgx, a.sgy);
  } else if (b < nS + nP) {
    b -= nS;
    alpha_scan_block(a.img, a.istride, a.iw, 0, a.ph0, a.st, 0, a.ivec, b % a.pgx, b / a.pgx, a.pgx, a.pgy);
  } else {
    b -= nS + nP;
    alpha_scan_block(a.tpl, a.tstride, a.tw, 0, a.th, a.st, 1, a.tvec, b % a.tgx, b / a.tgx, a.tgx, a.tgy);
  }
}

// Partial column sums of chunk c (rows [c*NB_ROWS, c*NB_ROWS + NB_ROWS)),
// stored at row c + 1 of the scratch (row 0 is the empty prefix),
// three words per image column: ch0 | ch2 << 16, ch1 | ch3 << 16 (each at
// most NB_ROWS * 255) and the sum of squares over the first cn channels.
// Column x of chunk c.
__device__ __forceinline__ void colsum_chunks_at(int x, int c, const uint8_t* __restrict__ img, int64_t istride,
                                                 int iw, int ih, int step, int words,
                                                 const CallState* __restrict__ st, uint32_t* __restrict__ ps) {
  if (x >= iw || st->flat) return;
  const int cn = st->cn;
  const uint32_t m = chan_mask(cn);
  const int y0 = c * NB_ROWS, y1 = min(ih, y0 + NB_ROWS);
  uint32_t s02 = 0, s13 = 0, q = 0;
  for (int y = y0; y < y1; y++) {
    const uint8_t* p = img + (int64_t)y * istride + (int64_t)x * step;
    uint32_t w;
    if (words) {
      w = *(const uint32_t*)p & m;
    } else {
      w = 0;
      for (int k = 0; k < 4; k++)
        if (k < cn) w |= (uint32_t)p[k] << (8 * k);
    }
    s02 += w & 0x00FF00FFu;
    s13 += (w >> 8) & 0x00FF00FFu;
    q = __builtin_amdgcn_udot4(w, w, q, false);
  }
  uint32_t* o = ps + (int64_t)(c + 1) * 3 * iw + x;
  o[0] = s02;
  o[iw] = s13;
  o[2 * iw] = q;
}

// Running sums of the chunk partials down each column, modulo 2^32: row c
// becomes the sum of chunks [0, c).  A window of whole chunks is then the
// difference of two rows; the difference is exact because every window sum
// fits its field (each packed 16-bit channel half <= 255 * 257, the square
// sum < 2^32), so the wrap of the running sums cancels.
// Word i of the scratch rows.
__device__ __forceinline__ void colsum_prefix_at(int i, int nwords, int nch, const CallState* __restrict__ st,
                                                 uint32_t* __restrict__ ps) {
  if (i >= nwords || st->flat) return;
  uint32_t* p = ps + i;
  p[0] = 0u;
  uint32_t acc = 0u;
  int c = 1;
  for (; c + 8 <= nch + 1; c += 8) {
    uint32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = p[(int64_t)(c + k) * nwords];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      acc += v[k];
      p[(int64_t)(c + k) * nwords] = acc;
    }
  }
  for (; c <= nch; c++) {
    acc += p[(int64_t)c * nwords];
    p[(int64_t)c * nwords] = acc;
  }
}

// The same running sums with the rows of each column split over PSEG
// threads: block b takes columns [b * cols, b * cols + cols), cols =
// blockDim / PSEG, thread t column t % cols and segment t / cols of the
// rows 1..nch.  Each thread sums its segment (up to eight rows' loads in
// flight), the segment totals meet in LDS, and each thread writes its rows'
// running sums from the total of the segments above it: a chain of two
// segment walks instead of one walk over every row.  Modular uint32 sums,
// so any grouping gives the same words.  Every thread reaches the barrier.
constexpr int PSEG = 8;
__device__ __forceinline__ void colsum_prefix_block(int b, int nwords, int nch, const CallState* __restrict__ st,
                                                    uint32_t* __restrict__ ps) {
  __shared__ uint32_t tot[PSEG][64];
  const int cols = (int)blockDim.x / PSEG;
  const int t = threadIdx.x, col = t % cols, seg = t / cols;
  const int i = b * cols + col;
  const bool on = i < nwords && !st->flat;
  const int per = (nch + PSEG - 1) / PSEG;
  const int r0 = 1 + seg * per, r1 = min(nch + 1, r0 + per);
  uint32_t* p = ps + (on ? i : 0);
  uint32_t sum = 0;
  uint32_t v[8];
  if (on)
    for (int r = r0; r < r1; r += 8) {
#pragma unroll
      for (int k = 0; k < 8; k++) v[k] = r + k < r1 ? p[(int64_t)(r + k) * nwords] : 0u;
#pragma unroll
      for (int k = 0; k < 8; k++) sum += v[k];
    }
  tot[seg][col] = sum;
  __syncthreads();
  if (!on) return;
  uint32_t acc = 0;
  for (int s = 0; s < seg; s++) acc += tot[s][col];
  if (seg == 0) p[0] = 0u;
  if (per <= 8) {
    // the segment's rows are still in registers
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (r0 + k < r1) {
        acc += v[k];
        p[(int64_t)(r0 + k) * nwords] = acc;
      }
    return;
  }
  for (int r = r0; r < r1; r += 8) {
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = r + k < r1 ? p[(int64_t)(r + k) * nwords] : 0u;
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (r + k < r1) {
        acc += v[k];
        p[(int64_t)(r + k) * nwords] = acc;
      }
  }
}

}  // namespace cvm
