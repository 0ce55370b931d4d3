// Exact-statistics normalisation, template statistics, alpha scans and the
// first-occurrence extrema reduction for gfx950.
//
// Window statistics are exact integers (normalizeBandGo impl.go:859-1054):
// per-column sums over the th window rows slide down a row band in
// registers; horizontal window sums come from a block prefix scan in LDS.
// Channel sums use wrapping uint32 arithmetic -- a window sum is below
// 255 * statsCap(1) < 2^32, so differences of wrapped prefixes are exact --
// and square sums use uint64.  The float tail then follows normOne
// impl.go:825-849 op for op.  Extrema are (value, raster index) pairs reduced
// lexicographically: strict value compares, the smaller index on equal
// values, which is exactly the sequential first-occurrence scan of
// impl.go:1023-1047 / 1104-1113 (-0 and +0 compare equal there too).
#include <float.h>
#include <limits.h>

#include "cvm_internal.h"

namespace cvm {

// ------------------------------------------------------------ call state --

__global__ void k_state_init(CallState* st) {
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

void launch_state_init(CallState* st, hipStream_t s) { hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, s, st); }

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// templStats impl.go:1063-1079 (sums and square sums per channel, exact).
__global__ __launch_bounds__(256) void k_tpl_stats(const uint8_t* __restrict__ tpl, int64_t tstride, int tw, int th,
                                                   int cnMax, int step, CallState* st) {
  __shared__ unsigned long long red[4][8];
  unsigned long long s[4] = {0, 0, 0, 0}, q[4] = {0, 0, 0, 0};
  const int64_t npx = (int64_t)tw * th;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npx; i += (int64_t)gridDim.x * blockDim.x) {
    int y = (int)(i / tw), x = (int)(i - (int64_t)y * tw);
    const uint8_t* p = tpl + (int64_t)y * tstride + (int64_t)x * step;
    for (int k = 0; k < 4; k++)
      if (k < cnMax) {
        unsigned long long v = p[k];
        s[k] += v;
        q[k] += v * v;
      }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < 4; k++) {
    unsigned long long a = wave_sum_u64(s[k]), b = wave_sum_u64(q[k]);
    if (lane == 0) {
      red[k][w] = a;
      red[k][4 + w] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x < 4 && (int)threadIdx.x < cnMax) {
    int k = threadIdx.x;
    unsigned long long a = 0, b = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
      a += red[k][i];
      b += red[k][4 + i];
    }
    atomicAdd(&st->tsum[k], a);
    atomicAdd(&st->tsq[k], b);
  }
}

void launch_tpl_stats(const uint8_t* tpl, int64_t tstride, int tw, int th, int cnMax, int step, CallState* st,
                      hipStream_t s) {
  int64_t npx = (int64_t)tw * th;
  int blocks = (int)((npx + 256 * 16 - 1) / (256 * 16));
  if (blocks < 1) blocks = 1;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_tpl_stats, dim3(blocks), dim3(256), 0, s, tpl, tstride, tw, th, cnMax, step, st);
}

// alphaConst cvmatch.go:146-168: every alpha byte equals the first one.
__global__ __launch_bounds__(256) void k_alpha_scan(const uint8_t* __restrict__ pix, int64_t stride, int w, int h,
                                                    CallState* st, int which) {
  const uint8_t a0 = pix[3];
  int bad = 0;
  const int64_t npx = (int64_t)w * h;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npx; i += (int64_t)gridDim.x * blockDim.x) {
    int y = (int)(i / w), x = (int)(i - (int64_t)y * w);
    if (pix[(int64_t)y * stride + (int64_t)x * 4 + 3] != a0) bad = 1;
  }
  if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(&st->alpha_bad[which], 1u);
}

void launch_alpha_scan(const uint8_t* pix, int64_t stride, int w, int h, CallState* st, int which, hipStream_t s) {
  int64_t npx = (int64_t)w * h;
  int blocks = (int)((npx + 256 * 8 - 1) / (256 * 8));
  if (blocks < 1) blocks = 1;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_alpha_scan, dim3(blocks), dim3(256), 0, s, pix, stride, w, h, st, which);
}

// cn choice (Match cvmatch.go:82-85), varSum, flat test (impl.go:1137),
// templNorm (impl.go:1087), varScale / numScale (impl.go:862-870).
__global__ void k_state_finalize(CallState* st, int tw, int th, int cn_mode) {
  if (threadIdx.x != 0) return;
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

void launch_state_finalize(CallState* st, int tw, int th, int cn_mode, hipStream_t s) {
  hipLaunchKernelGGL(k_state_finalize, dim3(1), dim3(64), 0, s, st, tw, th, cn_mode);
}

// rgbToGray cvmatch.go:222-240: (4899 R + 9617 G + 1868 B + 8192) >> 14.
__global__ void k_rgba_to_gray(const uint8_t* __restrict__ pix, int64_t stride, int w, int h, uint8_t* __restrict__ out) {
  const int64_t npx = (int64_t)w * h;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npx; i += (int64_t)gridDim.x * blockDim.x) {
    int y = (int)(i / w), x = (int)(i - (int64_t)y * w);
    const uint8_t* p = pix + (int64_t)y * stride + (int64_t)x * 4;
    uint32_t r = p[0], g = p[1], b = p[2];
    out[i] = (uint8_t)((4899u * r + 9617u * g + 1868u * b + 8192u) >> 14);
  }
}

void launch_rgba_to_gray(const uint8_t* pix, int64_t stride, int w, int h, uint8_t* out, hipStream_t s) {
  int64_t npx = (int64_t)w * h;
  int blocks = (int)((npx + 255) / 256);
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_rgba_to_gray, dim3(blocks), dim3(256), 0, s, pix, stride, w, h, out);
}

// --------------------------------------------------------- normalisation --

// normOne impl.go:825-849.  sqrt and the final division are evaluated in
// binary64 and rounded once to binary32: for sqrt and division that double
// rounding equals the correctly rounded binary32 result (53 >= 2*24+2).
__device__ __forceinline__ float norm_one(float num, float lane0, float idiff, float s2, float numScale,
                                          float varScale, float templNorm) {
  const float eps = 10.0f * 0x1p-23f;
  float t = lane0 * numScale;
  num = num - t;
  float diff2 = idiff * varScale;
  float lim = eps * s2;
  if (lim > 0.5f) lim = 0.5f;
  float den = 0.0f;
  if (diff2 > lim) {
    float sq = (float)__dsqrt_rn((double)diff2);
    den = sq * templNorm;
  }
  float an = fabsf(num);
  if (an < den) return (float)__ddiv_rn((double)num, (double)den);
  float den2 = den * 1.125f;
  if (an < den2) return num > 0.0f ? 1.0f : -1.0f;
  return 0.0f;
}

__device__ __forceinline__ bool better_min(float v, int i, float bv, int bi) { return v < bv || (v == bv && i < bi); }
__device__ __forceinline__ bool better_max(float v, int i, float bv, int bi) { return v > bv || (v == bv && i < bi); }

constexpr int NORM_NX = 256;  // output columns per workgroup (one per thread)

// Rows of one band: enough workgroups to fill the chip while keeping the
// th-row column-sum start-up small relative to the band.
static int norm_band_rows(int rh) {
  (void)rh;
  return 64;
}

int norm_partials_count(int rw, int rh) {
  int br = norm_band_rows(rh);
  return ((rw + NORM_NX - 1) / NORM_NX) * ((rh + br - 1) / br);
}

// One workgroup: result columns [xs, xs+256) x rows [ys, ye).  Each thread
// owns CPT consecutive image columns of the strip [xs, xs+256+tw-1) and keeps
// their exact column sums over the current th window rows in registers.
template <int CPT>
__global__ __launch_bounds__(256) void k_normalize(const uint8_t* __restrict__ img, int64_t istride, int iw, int step,
                                                   int tw, int th, int rw, int rh, int bandRows,
                                                   const CallState* __restrict__ st, const float* corr, float* map,
                                                   Extrema* __restrict__ parts) {
  constexpr int NW = NORM_NX * CPT;
  __shared__ uint32_t P[4][NW + 1];
  __shared__ unsigned long long P2[NW + 1];
  __shared__ uint32_t wtot[4][4];
  __shared__ unsigned long long wtot2[4];
  __shared__ float red_v[2][4];
  __shared__ int red_i[2][4];

  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int xs = blockIdx.x * NORM_NX, ys = blockIdx.y * bandRows;
  const int ye = min(rh, ys + bandRows);
  const int cn = st->cn;
  const bool flat = st->flat != 0;
  const int W = min(NORM_NX + tw - 1, iw - xs);
  const long long area = (long long)tw * th;
  const float numScale = st->numScale, varScale = st->varScale, templNorm = st->templNorm;
  long long tsum[4];
  for (int k = 0; k < 4; k++) tsum[k] = (long long)st->tsum[k];

  uint32_t cs[CPT][4];
  unsigned long long c2[CPT];
  for (int j = 0; j < CPT; j++) {
    for (int k = 0; k < 4; k++) cs[j][k] = 0;
    c2[j] = 0;
  }
  const int cbase = tid * CPT;
  if (!flat) {
    for (int y = ys; y < ys + th; y++) {
      const uint8_t* row = img + (int64_t)y * istride;
      for (int j = 0; j < CPT; j++) {
        int c = cbase + j;
        if (c < W) {
          const uint8_t* p = row + (int64_t)(xs + c) * step;
          for (int k = 0; k < 4; k++)
            if (k < cn) {
              uint32_t v = p[k];
              cs[j][k] += v;
              c2[j] += v * v;
            }
        }
      }
    }
  }

  float bmin = FLT_MAX, bmax = -FLT_MAX;
  int bmin_i = INT_MAX, bmax_i = INT_MAX;
  const int x = xs + tid;
  const bool active = x < rw;

  for (int y = ys; y < ye; y++) {
    float v = 1.0f;
    if (!flat) {
      // block prefix scan of the column sums
      uint32_t loc[CPT][4];
      unsigned long long loc2[CPT];
      uint32_t run[4] = {0, 0, 0, 0};
      unsigned long long run2 = 0;
      for (int j = 0; j < CPT; j++) {
        for (int k = 0; k < 4; k++) {
          run[k] += cs[j][k];
          loc[j][k] = run[k];
        }
        run2 += c2[j];
        loc2[j] = run2;
      }
      // inclusive wave scan of the thread totals
      uint32_t inc[4] = {run[0], run[1], run[2], run[3]};
      unsigned long long inc2 = run2;
      for (int o = 1; o < 64; o <<= 1) {
        uint32_t up[4];
        for (int k = 0; k < 4; k++) up[k] = __shfl_up(inc[k], o, 64);
        unsigned long long up2 = __shfl_up(inc2, o, 64);
        if (lane >= o) {
          for (int k = 0; k < 4; k++) inc[k] += up[k];
          inc2 += up2;
        }
      }
      if (lane == 63) {
        for (int k = 0; k < 4; k++) wtot[wv][k] = inc[k];
        wtot2[wv] = inc2;
      }
      __syncthreads();
      uint32_t off[4] = {0, 0, 0, 0};
      unsigned long long off2 = 0;
      for (int w = 0; w < wv; w++) {
        for (int k = 0; k < 4; k++) off[k] += wtot[w][k];
        off2 += wtot2[w];
      }
      for (int k = 0; k < 4; k++) off[k] += inc[k] - run[k];  // exclusive within the wave
      off2 += inc2 - run2;
      for (int j = 0; j < CPT; j++) {
        for (int k = 0; k < 4; k++) P[k][cbase + j + 1] = off[k] + loc[j][k];
        P2[cbase + j + 1] = off2 + loc2[j];
      }
      if (tid == 0) {
        for (int k = 0; k < 4; k++) P[k][0] = 0;
        P2[0] = 0;
      }
      __syncthreads();
      if (active) {
        long long s[4];
        long long lane0 = 0, sq = 0;
        for (int k = 0; k < 4; k++) {
          s[k] = (long long)(uint32_t)(P[k][tid + tw] - P[k][tid]);
          if (k < cn) {
            lane0 += s[k] * tsum[k];
            sq += s[k] * s[k];
          }
        }
        long long s2 = (long long)(P2[tid + tw] - P2[tid]);
        if (cn == 1) lane0 = s[0];
        float l0 = (float)(double)lane0;
        float l1 = (float)(double)(area * s2 - sq);
        float l2 = (float)(double)s2;
        v = norm_one(corr[(int64_t)y * rw + x], l0, l1, l2, numScale, varScale, templNorm);
      }
      __syncthreads();  // P is rewritten for the next row
      // slide the column sums down one row (colSlideGo impl.go:761-805)
      if (y + 1 < ye) {
        const uint8_t* rsub = img + (int64_t)y * istride;
        const uint8_t* radd = img + (int64_t)(y + th) * istride;
        for (int j = 0; j < CPT; j++) {
          int c = cbase + j;
          if (c < W) {
            const uint8_t* ps = rsub + (int64_t)(xs + c) * step;
            const uint8_t* pa = radd + (int64_t)(xs + c) * step;
            for (int k = 0; k < 4; k++)
              if (k < cn) {
                uint32_t a = pa[k], b = ps[k];
                cs[j][k] += a - b;
                c2[j] += (unsigned long long)(a * a) - (unsigned long long)(b * b);
              }
          }
        }
      }
    }
    if (active) {
      int idx = y * rw + x;
      if (map) map[(int64_t)y * rw + x] = v;
      if (better_min(v, idx, bmin, bmin_i)) {
        bmin = v;
        bmin_i = idx;
      }
      if (better_max(v, idx, bmax, bmax_i)) {
        bmax = v;
        bmax_i = idx;
      }
    }
  }

  // workgroup reduction (lexicographic)
  for (int o = 32; o > 0; o >>= 1) {
    float ov = __shfl_xor(bmin, o, 64);
    int oi = __shfl_xor(bmin_i, o, 64);
    if (better_min(ov, oi, bmin, bmin_i)) {
      bmin = ov;
      bmin_i = oi;
    }
    ov = __shfl_xor(bmax, o, 64);
    oi = __shfl_xor(bmax_i, o, 64);
    if (better_max(ov, oi, bmax, bmax_i)) {
      bmax = ov;
      bmax_i = oi;
    }
  }
  if (lane == 0) {
    red_v[0][wv] = bmin;
    red_i[0][wv] = bmin_i;
    red_v[1][wv] = bmax;
    red_i[1][wv] = bmax_i;
  }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      if (better_min(red_v[0][w], red_i[0][w], bmin, bmin_i)) {
        bmin = red_v[0][w];
        bmin_i = red_i[0][w];
      }
      if (better_max(red_v[1][w], red_i[1][w], bmax, bmax_i)) {
        bmax = red_v[1][w];
        bmax_i = red_i[1][w];
      }
    }
    Extrema e;
    e.min_v = bmin;
    e.min_i = bmin_i;
    e.max_v = bmax;
    e.max_i = bmax_i;
    parts[blockIdx.y * gridDim.x + blockIdx.x] = e;
  }
}

void launch_normalize(const uint8_t* img, int64_t istride, int iw, int step, int tw, int th, int rw, int rh,
                      const CallState* st, const float* corr, float* map, Extrema* partials, hipStream_t s) {
  int br = norm_band_rows(rh);
  dim3 grid((rw + NORM_NX - 1) / NORM_NX, (rh + br - 1) / br);
  int W = NORM_NX + tw - 1;
  int cpt = (W + NORM_NX - 1) / NORM_NX;
#define CVM_NORM(C)                                                                                              \
  hipLaunchKernelGGL(k_normalize<C>, grid, dim3(NORM_NX), 0, s, img, istride, iw, step, tw, th, rw, rh, br, st, \
                     corr, map, partials)
  if (cpt <= 1)
    CVM_NORM(1);
  else if (cpt <= 2)
    CVM_NORM(2);
  else if (cpt <= 4)
    CVM_NORM(4);
  else
    CVM_NORM(8);
#undef CVM_NORM
}

// Final reduction of the workgroup partials and conversion to the
// reference's six-tuple (min_v, min_x, min_y, max_v, max_x, max_y).
struct Res6 {
  float min_v;
  int32_t min_x, min_y;
  float max_v;
  int32_t max_x, max_y;
};

__global__ __launch_bounds__(1024) void k_reduce(const Extrema* __restrict__ parts, int n, int rw, Res6* out) {
  __shared__ float rv[2][16];
  __shared__ int ri[2][16];
  float bmin = FLT_MAX, bmax = -FLT_MAX;
  int bmin_i = INT_MAX, bmax_i = INT_MAX;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    Extrema e = parts[i];
    if (better_min(e.min_v, e.min_i, bmin, bmin_i)) {
      bmin = e.min_v;
      bmin_i = e.min_i;
    }
    if (better_max(e.max_v, e.max_i, bmax, bmax_i)) {
      bmax = e.max_v;
      bmax_i = e.max_i;
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) {
    float ov = __shfl_xor(bmin, o, 64);
    int oi = __shfl_xor(bmin_i, o, 64);
    if (better_min(ov, oi, bmin, bmin_i)) {
      bmin = ov;
      bmin_i = oi;
    }
    ov = __shfl_xor(bmax, o, 64);
    oi = __shfl_xor(bmax_i, o, 64);
    if (better_max(ov, oi, bmax, bmax_i)) {
      bmax = ov;
      bmax_i = oi;
    }
  }
  if (lane == 0) {
    rv[0][wv] = bmin;
    ri[0][wv] = bmin_i;
    rv[1][wv] = bmax;
    ri[1][wv] = bmax_i;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      if (better_min(rv[0][w], ri[0][w], bmin, bmin_i)) {
        bmin = rv[0][w];
        bmin_i = ri[0][w];
      }
      if (better_max(rv[1][w], ri[1][w], bmax, bmax_i)) {
        bmax = rv[1][w];
        bmax_i = ri[1][w];
      }
    }
    Res6 r;
    r.min_v = bmin;
    r.min_x = bmin_i % rw;
    r.min_y = bmin_i / rw;
    r.max_v = bmax;
    r.max_x = bmax_i % rw;
    r.max_y = bmax_i / rw;
    *out = r;
  }
}

void launch_reduce(const Extrema* partials, int nparts, const CallState* st, int rw, void* res6, float* map_if_flat,
                   int rh, hipStream_t s) {
  (void)st;
  (void)map_if_flat;
  (void)rh;
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, s, partials, nparts, rw, (Res6*)res6);
}

}  // namespace cvm
