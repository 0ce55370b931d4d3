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

#include <type_traits>

#include <algorithm>

#include "cvm_internal.h"
#include "cvm_score.h"
#include "cvm_state.h"

namespace cvm {

// ------------------------------------------------------------ call state --

__global__ void k_state_init(CallState* st) { state_clear(st); }

void launch_state_init(CallState* st, hipStream_t s) { hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, s, st); }

__global__ __launch_bounds__(256) void k_tpl_stats(const uint8_t* __restrict__ tpl, int64_t tstride, int tw, int th,
                                                   int cnMax, int step, CallState* st) {
  tpl_stats_block(tpl, tstride, tw, th, cnMax, step, st, blockIdx.x, blockIdx.y, gridDim.x, gridDim.y);
}

void launch_tpl_stats(const uint8_t* tpl, int64_t tstride, int tw, int th, int cnMax, int step, CallState* st,
                      hipStream_t s) {
  int gx = (tw + 255) / 256;
  if (gx > 16) gx = 16;
  int gy = th < 256 ? th : 256;
  hipLaunchKernelGGL(k_tpl_stats, dim3(gx, gy), dim3(256), 0, s, tpl, tstride, tw, th, cnMax, step, st);
}

__global__ __launch_bounds__(256) void k_alpha_scan(const uint8_t* __restrict__ pix, int64_t stride, int w, int y_begin,
                                                    int h, CallState* st, int which, int vec) {
  alpha_scan_block(pix, stride, w, y_begin, h, st, which, vec, blockIdx.x, blockIdx.y, gridDim.x, gridDim.y);
}

void launch_alpha_scan(const uint8_t* pix, int64_t stride, int w, int h, CallState* st, int which, hipStream_t s) {
  int vec = ((((uintptr_t)pix) | (uintptr_t)stride) & 15) == 0;
  int per_row = vec ? (w / 4 + 1) : w;
  int gx = (per_row + 255) / 256;
  if (gx > 16) gx = 16;
  // the first rows alone first: a varying alpha plane shows up there and
  // the full scan's workgroups then leave at their first flag check
  int h0 = h < 8 ? h : 8;
  hipLaunchKernelGGL(k_alpha_scan, dim3(gx, h0), dim3(256), 0, s, pix, stride, w, 0, h0, st, which, vec);
  if (h > h0) {
    int gy = (h - h0) < 512 ? (h - h0) : 512;
    hipLaunchKernelGGL(k_alpha_scan, dim3(gx, gy), dim3(256), 0, s, pix, stride, w, h0, h, st, which, vec);
  }
}

__global__ void k_state_finalize(CallState* st, int tw, int th, int cn_mode) {
  if (threadIdx.x == 0) state_finalize_fields(st, tw, th, cn_mode);
}

void launch_state_finalize(CallState* st, int tw, int th, int cn_mode, hipStream_t s) {
  hipLaunchKernelGGL(k_state_finalize, dim3(1), dim3(64), 0, s, st, tw, th, cn_mode);
}

// One launch for the independent parts of a call state (call_state_block).
__global__ __launch_bounds__(256) void k_call_state(CallStateArgs a) { call_state_block(blockIdx.x, a); }

void launch_call_state(const uint8_t* img, int64_t istride, int iw, int ih, const uint8_t* tpl, int64_t tstride,
                       int tw, int th, int cnMax, int step, int cn_mode, CallState* st, hipStream_t s) {
  const CallStateArgs a = call_state_args(img, istride, iw, ih, tpl, tstride, tw, th, cnMax, step, cn_mode, st, false);
  hipLaunchKernelGGL(k_call_state, dim3(a.nb()), dim3(256), 0, s, a);
  if (cn_mode == 0 && ih > a.ph0) {
    const int gy = std::min(512, ih - a.ph0);
    hipLaunchKernelGGL(k_alpha_scan, dim3(a.pgx, gy), dim3(256), 0, s, img, istride, iw, a.ph0, ih, st, 0, a.ivec);
  }
  launch_state_finalize(st, tw, th, cn_mode, s);
}

CallStateArgs call_state_args(const uint8_t* img, int64_t istride, int iw, int ih, const uint8_t* tpl,
                              int64_t tstride, int tw, int th, int cnMax, int step, int cn_mode, CallState* st,
                              bool whole_parent) {
  CallStateArgs a{};
  a.img = img;
  a.istride = istride;
  a.iw = iw;
  a.tpl = tpl;
  a.tstride = tstride;
  a.tw = tw;
  a.th = th;
  a.cnMax = cnMax;
  a.step = step;
  a.st = st;
  a.rows8 = whole_parent;
  if (whole_parent) {
    a.sgx = (int)std::min<long long>(256, ((long long)tw * th + 2047) / 2048);
    a.sgy = 1;
  } else {
    a.sgx = std::min(16, (tw + 255) / 256);
    a.sgy = std::min(256, th);
  }
  if (cn_mode == 0) {
    a.ivec = ((((uintptr_t)img) | (uintptr_t)istride) & 15) == 0;
    a.tvec = ((((uintptr_t)tpl) | (uintptr_t)tstride) & 15) == 0;
    a.ph0 = whole_parent ? ih : std::min(ih, 8);
    a.pgx = std::min(16, ((a.ivec ? iw / 4 + 1 : iw) + 255) / 256);
    a.pgy = whole_parent ? (ih + 7) / 8 : std::min(512, a.ph0);
    a.tgx = std::min(16, ((a.tvec ? tw / 4 + 1 : tw) + 255) / 256);
    a.tgy = whole_parent ? (th + 7) / 8 : std::min(512, th);
  }
  return a;
}

// rgbToGray cvmatch.go:222-240: (4899 R + 9617 G + 1868 B + 8192) >> 14.
__global__ void k_rgba_to_gray(const uint8_t* __restrict__ pix, int64_t stride, int w, int h, uint8_t* __restrict__ out) {
  for (int y = blockIdx.y; y < h; y += gridDim.y) {
    const uint8_t* row = pix + (int64_t)y * stride;
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < w; x += gridDim.x * blockDim.x) {
      uint32_t px = row[(int64_t)x * 4], g = row[(int64_t)x * 4 + 1], b = row[(int64_t)x * 4 + 2];
      out[(int64_t)y * w + x] = (uint8_t)((4899u * px + 9617u * g + 1868u * b + 8192u) >> 14);
    }
  }
}

void launch_rgba_to_gray(const uint8_t* pix, int64_t stride, int w, int h, uint8_t* out, hipStream_t s) {
  int gx = (w + 255) / 256;
  if (gx > 16) gx = 16;
  int gy = h < 1024 ? h : 1024;
  hipLaunchKernelGGL(k_rgba_to_gray, dim3(gx, gy), dim3(256), 0, s, pix, stride, w, h, out);
}

// --------------------------------------------------------- normalisation --

// normOne impl.go:825-849.  sqrt and the final division are the correctly
// rounded binary32 operations (sqrt_rn_f32, div_rn_f32), which is what the
// reference's float32(math.Sqrt(float64(x))) and float32 division give.
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
    float sq = sqrt_rn_f32(diff2);
    den = sq * templNorm;
  }
  float an = fabsf(num);
  if (an < den) return div_rn_f32(num, den);
  float den2 = den * 1.125f;
  if (an < den2) return num > 0.0f ? 1.0f : -1.0f;
  return 0.0f;
}


constexpr int NORM_NX = 256;  // output columns per workgroup (one per thread)

// Rows of one band: enough workgroups to fill the chip while keeping the
// th-row column-sum start-up small relative to the band.
static int norm_band_rows(int rh) {
  (void)rh;
  return 64;
}

// Wave form (k_norm_wave below): one 64-lane workgroup per 64*Q result
// columns x NB_ROWS rows.  It is used when every exact statistic of the
// window fits the 53-bit significand (cnMax * 65025 * area^2 < 2^53, any
// template up to 431x431 at four channels): the cross term, the sum of
// squared window sums and area * s2 are then evaluated in binary64 with no
// rounding at all, exactly the integers the int64 form computes.  Larger
// templates take the 256-thread int64 form (k_normalize).

struct NormPlan {
  int wave;  // 0: int64 form
  int q;     // result columns per lane (the strip is 64*q columns)
  int sq32;  // window square sums fit uint32
  int cpt;   // image columns per lane: the fewest of 2, 4, 8 that cover the strip
  int band;  // result rows per wave: NB_ROWS, or NB_ROWS / 2 (a half band)
};

// Each lane of the wave form owns CPT consecutive image columns of the strip
// (64*q + tw - 1 <= 64*CPT columns), loaded with vector loads; the widest
// strip has CPT = 8.  A narrow strip with a narrow template needs fewer
// columns per lane, and every column costs its sums' slide and prefix on
// every row.
constexpr int NORM_CPT = 8;
// Prefix row length in LDS: prefix index c lives at c + c/32 (lane l writes
// columns CPT*l..CPT*l+CPT-1, and the pad word spreads the 64 lanes of one
// store over all banks).
__host__ __device__ constexpr int norm_nwp(int cpt) { return 64 * cpt + 1 + (64 * cpt + 1) / 32 + 1; }
constexpr int NORM_NWP = norm_nwp(NORM_CPT);

// qpref: 4, 2 or 1 result columns per lane when that strip holds the
// template (64 q + tw - 1 <= 512 columns), else (or 0) the widest that does;
// plus NORM_HALF for half bands (twice the waves, each walking half the
// rows: a small map's waves are latency-bound row chains).
static NormPlan norm_plan(int tw, int th, int cnMax, int qpref = 0) {
  NormPlan p{0, 0, 0, 0, NB_ROWS};
  if (qpref & NORM_HALF) p.band = NB_ROWS / 2;
  qpref &= NORM_HALF - 1;
  double area = (double)tw * th;
  if ((double)cnMax * 65025.0 * area * area >= 9007199254740992.0) return p;
  if (th > 257) return p;  // packed 16-bit channel column sums: 255 * th < 2^16
  for (int q : {qpref, 4, 2}) {
    if (q != 1 && q != 2 && q != 4) continue;
    if (64 * q + tw - 1 <= 64 * NORM_CPT) {
      p.wave = 1;
      p.q = q;
      p.sq32 = (double)cnMax * 65025.0 * area < 4294967296.0;
      p.cpt = NORM_CPT;
      for (int c : {2, 4})  // built: (4, 8), (2, 4), (2, 8), (1, 2), (1, 4), (1, 8)
        if (c > q && 64 * q + tw - 1 <= 64 * c) {
          p.cpt = c;
          break;
        }
      return p;
    }
  }
  return p;
}

bool norm_wave_ok(int tw, int th, int cnMax) { return norm_plan(tw, th, cnMax).wave != 0; }

bool norm_q_ok(int tw, int th, int cnMax, int q) {
  const NormPlan np = norm_plan(tw, th, cnMax, q);
  return np.wave && np.q == (q & (NORM_HALF - 1));
}

int norm_partials_count(int rw, int rh, int tw, int th, int cnMax, int qpref) {
  NormPlan np = norm_plan(tw, th, cnMax, qpref);
  if (np.wave) return ((rw + 64 * np.q - 1) / (64 * np.q)) * ((rh + np.band - 1) / np.band);
  int br = norm_band_rows(rh);
  return ((rw + NORM_NX - 1) / NORM_NX) * ((rh + br - 1) / br);
}

size_t norm_scratch_bytes(int iw, int ih, int tw, int th, int cnMax) {
  if (!norm_plan(tw, th, cnMax).wave) return 0;
  size_t nch = (size_t)(ih + NB_ROWS - 1) / NB_ROWS;
  return (nch + 1) * (size_t)iw * 3 * sizeof(uint32_t);
}

// One workgroup: result columns [xs, xs+256) x rows [ys, ye).  Each thread
// owns CPT consecutive image columns of the strip [xs, xs+256+tw-1) and keeps
// their exact column sums over the current th window rows in registers.
// Result column xs + t takes its window sums from the strip's prefix at
// columns t and t + tw, so of the whole prefix only the two 256-entry
// windows [0, 256) (PL) and [tw, tw + 256) (PH) are kept in LDS, whatever
// the template width.
template <int CPT>
__global__ __launch_bounds__(256) void k_normalize(const uint8_t* __restrict__ img, int64_t istride, int iw, int step,
                                                   int tw, int th, int rw, int rh, int bandRows,
                                                   const CallState* __restrict__ st, const float* corr, float* map,
                                                   Extrema* __restrict__ parts) {
  __shared__ uint32_t PL[4][NORM_NX], PH[4][NORM_NX];
  __shared__ unsigned long long PL2[NORM_NX], PH2[NORM_NX];
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
      // prefix entry c (the sums of strip columns [0, c)) is kept when a
      // result reads it: c < 256 (PL) or tw <= c < tw + 256 (PH)
      for (int j = 0; j < CPT; j++) {
        const int c = cbase + j + 1;
        if (c < NORM_NX) {
          for (int k = 0; k < 4; k++) PL[k][c] = off[k] + loc[j][k];
          PL2[c] = off2 + loc2[j];
        }
        if (c >= tw && c < tw + NORM_NX) {
          for (int k = 0; k < 4; k++) PH[k][c - tw] = off[k] + loc[j][k];
          PH2[c - tw] = off2 + loc2[j];
        }
      }
      if (tid == 0) {
        for (int k = 0; k < 4; k++) PL[k][0] = 0;
        PL2[0] = 0;
      }
      __syncthreads();
      if (active) {
        long long s[4];
        long long lane0 = 0, sq = 0;
        for (int k = 0; k < 4; k++) {
          s[k] = (long long)(uint32_t)(PH[k][tid] - PL[k][tid]);
          if (k < cn) {
            lane0 += s[k] * tsum[k];
            sq += s[k] * s[k];
          }
        }
        long long s2 = (long long)(PH2[tid] - PL2[tid]);
        if (cn == 1) lane0 = s[0];
        float l0 = (float)(double)lane0;
        float l1 = (float)(double)(area * s2 - sq);
        float l2 = (float)(double)s2;
        v = norm_one(corr[(int64_t)y * rw + x], l0, l1, l2, numScale, varScale, templNorm);
      }
      __syncthreads();  // PL / PH are rewritten for the next row
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

// ------------------------------------------------- normalisation, wave form --
//
// One wave (a 64-thread workgroup) owns result columns [xs, xs + 64*Q) and
// rows [ys, ys + NB_ROWS).  Lane l keeps the exact column sums of CPT
// consecutive image columns of the strip over the current th window rows.
// The sums at the band start are the difference of two column prefix sums
// over NB_ROWS-row chunks (k_colsum_chunks reads every pixel once,
// k_colsum_prefix accumulates the chunks) plus th % NB_ROWS rows read
// directly; the window then slides down one row per step.
// Horizontal window sums are differences of a wave-level prefix scan in LDS.


__global__ __launch_bounds__(256) void k_colsum_chunks(const uint8_t* __restrict__ img, int64_t istride, int iw,
                                                       int ih, int step, int words, const CallState* __restrict__ st,
                                                       uint32_t* __restrict__ ps) {
  colsum_chunks_at(blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y, img, istride, iw, ih, step, words, st, ps);
}

__global__ __launch_bounds__(256) void k_colsum_prefix(int nwords, int nch, const CallState* __restrict__ st,
                                                       uint32_t* __restrict__ ps) {
  colsum_prefix_at(blockIdx.x * blockDim.x + threadIdx.x, nwords, nch, st, ps);
}

// CPT (2, 4 or 8) consecutive pixels x0.. of one row as packed channel
// words (channel k in bits 8k..).  VMODE 1: RGBA rows 16-byte aligned, one
// 8-byte or 16-byte loads; VMODE 2: gray rows 8-byte aligned, one 2-, 4- or
// 8-byte load (both with all CPT columns inside the row, x0 a multiple of
// CPT); VMODE 0: byte loads of the nvalid valid columns, zeros past them.
template <int VMODE, int CPT>
__device__ __forceinline__ void load_pxn(const uint8_t* row, int x0, int step, int cn, int nvalid,
                                         uint32_t out[CPT]) {
  if constexpr (VMODE == 1) {
    if constexpr (CPT == 2) {
      const uint2 a = *(const uint2*)(row + (int64_t)x0 * 4);
      out[0] = a.x;
      out[1] = a.y;
    } else {
      const uint4* p = (const uint4*)(row + (int64_t)x0 * 4);
#pragma unroll
      for (int h = 0; h < CPT / 4; h++) {
        const uint4 a = p[h];
        out[4 * h] = a.x;
        out[4 * h + 1] = a.y;
        out[4 * h + 2] = a.z;
        out[4 * h + 3] = a.w;
      }
    }
  } else if constexpr (VMODE == 2) {
    if constexpr (CPT == 8) {
      const uint2 v = *(const uint2*)(row + x0);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        out[j] = (v.x >> (8 * j)) & 255u;
        out[4 + j] = (v.y >> (8 * j)) & 255u;
      }
    } else if constexpr (CPT == 4) {
      const uint32_t v = *(const uint32_t*)(row + x0);
#pragma unroll
      for (int j = 0; j < 4; j++) out[j] = (v >> (8 * j)) & 255u;
    } else {
      const uint32_t v = *(const uint16_t*)(row + x0);
      out[0] = v & 255u;
      out[1] = v >> 8;
    }
  } else {
    for (int j = 0; j < CPT; j++) {
      uint32_t w = 0;
      if (j < nvalid) {
        const uint8_t* p = row + (int64_t)(x0 + j) * step;
        for (int k = 0; k < 4; k++)
          if (k < cn) w |= (uint32_t)p[k] << (8 * k);
      }
      out[j] = w;
    }
  }
}

// CPT consecutive uint32 partials: vector loads when VEC (aligned to their
// size, all CPT inside the row), else the nvalid valid ones.
template <bool VEC, int CPT>
__device__ __forceinline__ void load_u32xn(const uint32_t* p, int nvalid, uint32_t out[CPT]) {
  if constexpr (VEC) {
    if constexpr (CPT == 2) {
      const uint2 a = *(const uint2*)p;
      out[0] = a.x;
      out[1] = a.y;
    } else {
#pragma unroll
      for (int h = 0; h < CPT / 4; h++) {
        const uint4 a = ((const uint4*)p)[h];
        out[4 * h] = a.x;
        out[4 * h + 1] = a.y;
        out[4 * h + 2] = a.z;
        out[4 * h + 3] = a.w;
      }
    }
  } else {
    for (int j = 0; j < CPT; j++) out[j] = j < nvalid ? p[j] : 0u;
  }
}

// Slide inputs of one result row: the entering and leaving pixel rows and
// the correlation values of the row.
template <int Q, int CPT>
struct NormRow {
  uint32_t a[CPT], b[CPT];
  float c[Q];
};

// VMODE 1 / 2 (RGBA / gray rows with iw % 8 == 0, aligned for vector loads):
// every lane loads eight whole columns, lanes past the image clamped to its
// last eight, with no branch around any load; VMODE 0: byte loads.
// EXACT: corr holds the int32 C' = sum (P - 128)(T - 128) of the direct
// correlation (bit pattern read through the float pointer) and every score
// is the FP64 one rounded once to float32.
// NCH: channel sums kept (1 gray, 3 RGB, 4 any cn <= 4); channels at or past
// cn are zero, so dropping them changes no sum.  P / P2: the workgroup's
// prefix rows in LDS.
template <int Q, bool SQ32, int VMODE, bool EXACT, int NCH, typename sq_t, int CPT>
__device__ __forceinline__ void norm_wave_block(const uint8_t* __restrict__ img, int64_t istride, int iw, int ih,
                                                int step, int tw, int th, int rw, int rh,
                                                const CallState* __restrict__ st, const uint32_t* __restrict__ ps,
                                                const float* corr, float* map, Extrema* __restrict__ parts, int bx,
                                                int by, int gx, uint32_t (*P)[norm_nwp(CPT)], sq_t* P2,
                                                int band) {
  constexpr int SX = 64 * Q;
  const int lane = threadIdx.x;
  const int xs = bx * SX, ys = by * band;
  const int ye = min(rh, ys + band);
  // the window sums start at the chunk boundary ys0 <= ys and slide to ys
  const int ys0 = ys - ys % NB_ROWS, by0 = ys0 / NB_ROWS;
  float bmin = FLT_MAX, bmax = -FLT_MAX;
  int bmin_i = INT_MAX, bmax_i = INT_MAX;

  if (st->flat) {
    // templNorm below the flat threshold: the map is 1 everywhere
    for (int y = ys; y < ye; y++)
      for (int q = 0; q < Q; q++) {
        int x = xs + lane + 64 * q;
        if (x < rw) {
          int idx = y * rw + x;
          if (map) map[(int64_t)y * rw + x] = 1.0f;
          if (better_min(1.0f, idx, bmin, bmin_i)) {
            bmin = 1.0f;
            bmin_i = idx;
          }
          if (better_max(1.0f, idx, bmax, bmax_i)) {
            bmax = 1.0f;
            bmax_i = idx;
          }
        }
      }
  } else {
    const int cn = st->cn;
    const bool one = cn == 1;
    const double aread = (double)tw * (double)th;
    const float numScale = st->numScale, varScale = st->varScale, templNorm = st->templNorm;
    double tsd[4];
    for (int k = 0; k < 4; k++) tsd[k] = k < cn ? (double)st->tsum[k] : 0.0;
    // EXACT: sum P*T = C' + 128 (S + sum_k T_k) - 2^14 area cn
    const double varSumD = (double)st->varSum;
    double tsumAll = tsd[0];
#pragma unroll
    for (int k = 1; k < NCH; k++) tsumAll += tsd[k];
    const double cpOff = 16384.0 * aread * (double)cn;
    const int cbase = lane * CPT;
    // every in-image column of the lane is summed; columns past the strip
    // (or, clamped, repeats of the image's last eight) only reach prefix
    // entries that no result of the strip reads
    const int x0 = VMODE ? min(xs + cbase, iw - CPT) : xs + cbase;
    const int nload = VMODE ? CPT : min(CPT, iw - x0);  // may be <= 0
    const uint32_t cm = chan_mask(cn);

    // column sums over the th window rows, channels packed in 16-bit halves:
    // cs02 = ch0 | ch2 << 16, cs13 = ch1 | ch3 << 16 (each <= 255 * th < 2^16)
    uint32_t cs02[CPT], cs13[CPT];
    sq_t c2[CPT];
#pragma unroll
    for (int j = 0; j < CPT; j++) {
      cs02[j] = cs13[j] = 0;
      c2[j] = 0;
    }
    if (nload > 0) {
      // whole chunks [by0, by0 + nfull): prefix row by0 + nfull minus row
      // by0 (ys0 + nfull * NB_ROWS <= ys0 + th <= ih, so both rows exist)
      const int nfull = th / NB_ROWS;
      const uint32_t* p0 = ps + (int64_t)by0 * 3 * iw + x0;
      const uint32_t* p1 = ps + (int64_t)(by0 + nfull) * 3 * iw + x0;
      uint32_t a0[CPT], a1[CPT], a2[CPT], b0[CPT], b1[CPT], b2[CPT];
      load_u32xn<VMODE != 0, CPT>(p0, nload, a0);
      load_u32xn<VMODE != 0, CPT>(p0 + iw, nload, a1);
      load_u32xn<VMODE != 0, CPT>(p0 + 2 * iw, nload, a2);
      load_u32xn<VMODE != 0, CPT>(p1, nload, b0);
      load_u32xn<VMODE != 0, CPT>(p1 + iw, nload, b1);
      load_u32xn<VMODE != 0, CPT>(p1 + 2 * iw, nload, b2);
#pragma unroll
      for (int j = 0; j < CPT; j++) {
        cs02[j] = b0[j] - a0[j];
        cs13[j] = b1[j] - a1[j];
        c2[j] = (sq_t)(uint32_t)(b2[j] - a2[j]);
      }
      for (int y = ys0 + nfull * NB_ROWS; y < ys0 + th; y++) {
        uint32_t w[CPT];
        load_pxn<VMODE, CPT>(img + (int64_t)y * istride, x0, step, cn, nload, w);
#pragma unroll
        for (int j = 0; j < CPT; j++) {
          uint32_t v = w[j] & cm;
          cs02[j] += v & 0x00FF00FFu;
          cs13[j] += (v >> 8) & 0x00FF00FFu;
          c2[j] += (sq_t)__builtin_amdgcn_udot4(v, v, 0u, false);
        }
      }
      // a half band's start: the window slides from rows [ys0, ys0 + th) to
      // [ys, ys + th) (y + th < ys + th <= ih: every row exists).  For a
      // half band ys - ys0 is 0 or NB_ROWS / 2: a fixed run of rows whose
      // loads go out SB rows (2 SB row loads) at a time instead of one row
      // after another, where the lane's columns are few enough to keep the
      // register count (two image columns, or four gray ones; wider RGBA
      // lanes measured 112 -> 203 registers that way and slide row by row)
      constexpr bool BATCH = CPT <= 2 || (VMODE == 2 && CPT <= 4);
      if (BATCH && ys - ys0 == NB_ROWS / 2) {
        constexpr int SB = 4;
#pragma unroll
        for (int y0 = 0; y0 < NB_ROWS / 2; y0 += SB) {
          uint32_t a[SB][CPT], b[SB][CPT];
#pragma unroll
          for (int u = 0; u < SB; u++) {
            const int y = ys0 + y0 + u;
            load_pxn<VMODE, CPT>(img + (int64_t)(y + th) * istride, x0, step, cn, nload, a[u]);
            load_pxn<VMODE, CPT>(img + (int64_t)y * istride, x0, step, cn, nload, b[u]);
          }
#pragma unroll
          for (int u = 0; u < SB; u++)
#pragma unroll
            for (int j = 0; j < CPT; j++) {
              const uint32_t va = a[u][j] & cm, vb = b[u][j] & cm;
              cs02[j] = cs02[j] - (vb & 0x00FF00FFu) + (va & 0x00FF00FFu);
              cs13[j] = cs13[j] - ((vb >> 8) & 0x00FF00FFu) + ((va >> 8) & 0x00FF00FFu);
              c2[j] += (sq_t)__builtin_amdgcn_udot4(va, va, 0u, false) -
                       (sq_t)__builtin_amdgcn_udot4(vb, vb, 0u, false);
            }
        }
      } else {
        for (int y = ys0; y < ys; y++) {
          uint32_t a[CPT], b[CPT];
          load_pxn<VMODE, CPT>(img + (int64_t)(y + th) * istride, x0, step, cn, nload, a);
          load_pxn<VMODE, CPT>(img + (int64_t)y * istride, x0, step, cn, nload, b);
#pragma unroll
          for (int j = 0; j < CPT; j++) {
            const uint32_t va = a[j] & cm, vb = b[j] & cm;
            cs02[j] = cs02[j] - (vb & 0x00FF00FFu) + (va & 0x00FF00FFu);
            cs13[j] = cs13[j] - ((vb >> 8) & 0x00FF00FFu) + ((va >> 8) & 0x00FF00FFu);
            c2[j] += (sq_t)__builtin_amdgcn_udot4(va, va, 0u, false) -
                     (sq_t)__builtin_amdgcn_udot4(vb, vb, 0u, false);
          }
        }
      }
    }

    // Row yy's inputs; rows past the band are clamped into the image (loaded
    // and never used) so the loads need no branch.
    auto fetch = [&](NormRow<Q, CPT>& r, int yy) {
      const int yl = min(yy, ye - 1);
      const int ya = min(yl + th, ih - 1);
      if (VMODE) {
        load_pxn<VMODE, CPT>(img + (int64_t)ya * istride, x0, step, cn, nload, r.a);
        load_pxn<VMODE, CPT>(img + (int64_t)yl * istride, x0, step, cn, nload, r.b);
      } else {
#pragma unroll
        for (int j = 0; j < CPT; j++) r.a[j] = r.b[j] = 0;
        if (nload > 0) {
          load_pxn<0, CPT>(img + (int64_t)ya * istride, x0, step, cn, nload, r.a);
          load_pxn<0, CPT>(img + (int64_t)yl * istride, x0, step, cn, nload, r.b);
        }
      }
#pragma unroll
      for (int q = 0; q < Q; q++) {
        int x = min(xs + lane + 64 * q, rw - 1);
        r.c[q] = corr[(int64_t)yl * rw + x];
      }
    };

    auto row = [&](const NormRow<Q, CPT>& r, int y) {
      uint32_t run[NCH];
#pragma unroll
      for (int k = 0; k < NCH; k++) run[k] = 0;
      sq_t run2 = 0;
      uint32_t loc[CPT][NCH];
      sq_t loc2[CPT];
#pragma unroll
      for (int j = 0; j < CPT; j++) {
        const uint32_t cv[4] = {cs02[j] & 0xFFFFu, cs13[j] & 0xFFFFu, cs02[j] >> 16, cs13[j] >> 16};
#pragma unroll
        for (int k = 0; k < NCH; k++) {
          run[k] += cv[k];
          loc[j][k] = run[k];
        }
        run2 += c2[j];
        loc2[j] = run2;
      }
      uint32_t off[NCH];
#pragma unroll
      for (int k = 0; k < NCH; k++) off[k] = wave_incl_scan_u32(run[k], lane) - run[k];
      sq_t off2;
      if (SQ32) {
        off2 = (sq_t)wave_incl_scan_u32((uint32_t)run2, lane) - run2;
      } else {
        sq_t inc2 = run2;
        for (int o = 1; o < 64; o <<= 1) {
          sq_t up2 = __shfl_up(inc2, o, 64);
          if (lane >= o) inc2 += up2;
        }
        off2 = inc2 - run2;
      }
#pragma unroll
      for (int j = 0; j < CPT; j++) {
        const int pc = cbase + j + 1;
        const int pi = pc + (pc >> 5);
#pragma unroll
        for (int k = 0; k < NCH; k++) P[k][pi] = off[k] + loc[j][k];
        P2[pi] = off2 + loc2[j];
      }
      if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NCH; k++) P[k][0] = 0;
        P2[0] = 0;
      }
      __syncthreads();
      float vq[Q];
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const int xl = lane + 64 * q, xb = xl + tw;
        const int ia = xl + (xl >> 5), ib = xb + (xb >> 5);
        // exact integers carried in binary64 (all below 2^53, so every sum
        // below is exact in any order)
        double sd[NCH];
#pragma unroll
        for (int k = 0; k < NCH; k++) sd[k] = (double)(uint32_t)(P[k][ib] - P[k][ia]);
        const double s2 = (double)(sq_t)(P2[ib] - P2[ia]);
        // (every product and sum here is an exact integer, so a fused
        // multiply-add rounds nothing either: the same bits in fewer steps)
        double sq = sd[0] * sd[0], lane0 = sd[0] * tsd[0];
#pragma unroll
        for (int k = 1; k < NCH; k++) {
          sq = __builtin_fma(sd[k], sd[k], sq);
          lane0 = __builtin_fma(sd[k], tsd[k], lane0);
        }
        const double idiff = __builtin_fma(aread, s2, -sq);
        if constexpr (EXACT) {
          double S = sd[0];
#pragma unroll
          for (int k = 1; k < NCH; k++) S += sd[k];
          const double corrI = __builtin_fma(128.0, S + tsumAll, (double)__float_as_int(r.c[q]) - cpOff);
          vq[q] = (float)score_fp64(__builtin_fma(aread, corrI, -lane0), idiff, s2, aread, varSumD);
        } else {
          if (one) lane0 = sd[0];
          vq[q] = norm_one(r.c[q], (float)lane0, (float)idiff, (float)s2, numScale, varScale, templNorm);
        }
      }
      __syncthreads();
      // slide to row y+1: the leaving row is inside the current sums, so
      // each 16-bit half stays in [0, 2^16) and no borrow crosses halves
#pragma unroll
      for (int j = 0; j < CPT; j++) {
        const uint32_t a = r.a[j] & cm, b = r.b[j] & cm;
        cs02[j] = cs02[j] - (b & 0x00FF00FFu) + (a & 0x00FF00FFu);
        cs13[j] = cs13[j] - ((b >> 8) & 0x00FF00FFu) + ((a >> 8) & 0x00FF00FFu);
        c2[j] += (sq_t)__builtin_amdgcn_udot4(a, a, 0u, false) - (sq_t)__builtin_amdgcn_udot4(b, b, 0u, false);
      }
#pragma unroll
      for (int q = 0; q < Q; q++) {
        const int x = xs + lane + 64 * q;
        if (x < rw) {
          const int idx = y * rw + x;
          if (map) map[(int64_t)y * rw + x] = vq[q];
          // a lane meets its windows in raster order, so a strict
          // compare keeps the first occurrence (the index decides only
          // between lanes and workgroups)
          if (vq[q] < bmin) {
            bmin = vq[q];
            bmin_i = idx;
          }
          if (vq[q] > bmax) {
            bmax = vq[q];
            bmax_i = idx;
          }
        }
      }
    };

    // two register sets in turn: row y+1's loads are in flight while row y
    // is evaluated, and no loaded register is copied before use
    NormRow<Q, CPT> ra, rb;
    fetch(ra, ys);
    for (int y = ys; y < ye; y += 2) {
      fetch(rb, y + 1);
      row(ra, y);
      if (y + 1 >= ye) break;
      fetch(ra, y + 2);
      row(rb, y + 1);
    }
  }
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
    Extrema e;
    e.min_v = bmin;
    e.min_i = bmin_i;
    e.max_v = bmax;
    e.max_i = bmax_i;
    parts[by * gx + bx] = e;
  }
}

template <int Q, bool SQ32, int VMODE, bool EXACT, int CPT>
__global__ __launch_bounds__(64) void k_norm_wave(const uint8_t* __restrict__ img, int64_t istride, int iw, int ih,
                                                  int step, int tw, int th, int rw, int rh,
                                                  const CallState* __restrict__ st, const uint32_t* __restrict__ ps,
                                                  const float* corr, float* map, Extrema* __restrict__ parts,
                                                  int band) {
  typedef typename std::conditional<SQ32, uint32_t, unsigned long long>::type sq_t;
  __shared__ uint32_t P[4][norm_nwp(CPT)];
  __shared__ sq_t P2[norm_nwp(CPT)];
  const int cn = st->cn;
#define CVM_NWB(N)                                                                                             \
  norm_wave_block<Q, SQ32, VMODE, EXACT, N, sq_t, CPT>(img, istride, iw, ih, step, tw, th, rw, rh, st, ps, corr, \
                                                       map, parts, blockIdx.x, blockIdx.y, gridDim.x, P, P2, band)
  if (VMODE == 2 || (VMODE == 0 && cn == 1))
    CVM_NWB(1);
  else if (VMODE == 1 && cn == 3)
    CVM_NWB(3);
  else
    CVM_NWB(4);
#undef CVM_NWB
}

// Many templates against one frame in one launch (a Fleet's members, all of
// one (Q, SQ32) form): workgroup b runs block b - job.wg0 of job wg_job[b].
template <int Q, bool SQ32, int VMODE, bool EXACT>
__global__ __launch_bounds__(64) void k_norm_batch(const uint8_t* __restrict__ img, int64_t istride, int iw, int ih,
                                                   int step, const NormJob* __restrict__ jobs,
                                                   const int* __restrict__ wg_job) {
  typedef typename std::conditional<SQ32, uint32_t, unsigned long long>::type sq_t;
  __shared__ uint32_t P[4][NORM_NWP];
  __shared__ sq_t P2[NORM_NWP];
  const NormJob J = jobs[wg_job[blockIdx.x]];
  const int local = blockIdx.x - J.wg0;
  const int cn = J.st->cn;
#define CVM_NBB(N)                                                                                              \
  norm_wave_block<Q, SQ32, VMODE, EXACT, N, sq_t, NORM_CPT>(img, istride, iw, ih, step, J.tw, J.th, J.rw, J.rh, J.st, J.ps, \
                                                  J.corr, nullptr, J.parts, local % J.gx, local / J.gx, J.gx, P, P2, \
                                                  NB_ROWS)
  if (VMODE == 2 || (VMODE == 0 && cn == 1))
    CVM_NBB(1);
  else if (VMODE == 1 && cn == 3)
    CVM_NBB(3);
  else
    CVM_NBB(4);
#undef CVM_NBB
}

void launch_norm_prep(const uint8_t* img, int64_t istride, int iw, int ih, int step, int tw, int th, int cnMax,
                      const CallState* st, void* scratch, hipStream_t s) {
  if (!norm_plan(tw, th, cnMax).wave || !scratch) return;
  int nch = (ih + NB_ROWS - 1) / NB_ROWS;
  int words = (step == 4 && ((((uintptr_t)img) | (uintptr_t)istride) & 3) == 0) ? 1 : 0;
  hipLaunchKernelGGL(k_colsum_chunks, dim3((iw + 255) / 256, nch), dim3(256), 0, s, img, istride, iw, ih, step, words,
                     st, (uint32_t*)scratch);
  const int nwords = 3 * iw;
  hipLaunchKernelGGL(k_colsum_prefix, dim3((nwords + 255) / 256), dim3(256), 0, s, nwords, nch, st,
                     (uint32_t*)scratch);
}

void launch_normalize(const uint8_t* img, int64_t istride, int iw, int ih, int step, int tw, int th, int rw, int rh,
                      int cnMax, const CallState* st, const float* corr, float* map, Extrema* partials,
                      void* scratch, hipStream_t s, bool exact, int qpref) {
  NormPlan np = norm_plan(tw, th, cnMax, qpref);
  if (np.wave && scratch) {
    uint32_t* ps = (uint32_t*)scratch;
    const bool a16 = ((((uintptr_t)img) | (uintptr_t)istride) & 15) == 0;
    const bool a8 = ((((uintptr_t)img) | (uintptr_t)istride) & 7) == 0;
    int vmode = 0;
    if (iw % np.cpt == 0 && step == 4 && a16) vmode = 1;
    if (iw % np.cpt == 0 && step == 1 && a8) vmode = 2;
    dim3 grid((rw + 64 * np.q - 1) / (64 * np.q), (rh + np.band - 1) / np.band);
#define CVM_NW(Q, C, S, V, X)                                                                                       \
  hipLaunchKernelGGL((k_norm_wave<Q, S, V, X, C>), grid, dim3(64), 0, s, img, istride, iw, ih, step, tw, th, rw, rh, \
                     st, ps, corr, map, partials, np.band)
#define CVM_NWX(Q, C, S, V)  \
  if (exact)                 \
    CVM_NW(Q, C, S, V, true); \
  else                       \
    CVM_NW(Q, C, S, V, false)
#define CVM_NWV(Q, C, S)      \
  if (vmode == 1) {           \
    CVM_NWX(Q, C, S, 1);      \
  } else if (vmode == 2) {    \
    CVM_NWX(Q, C, S, 2);      \
  } else {                    \
    CVM_NWX(Q, C, S, 0);      \
  }
#define CVM_NWS(Q, C)         \
  if (np.sq32) {              \
    CVM_NWV(Q, C, true);      \
  } else {                    \
    CVM_NWV(Q, C, false);     \
  }
    switch (np.q * 10 + np.cpt) {
      case 48:
        CVM_NWS(4, 8);
        break;
      case 24:
        CVM_NWS(2, 4);
        break;
      case 28:
        CVM_NWS(2, 8);
        break;
      case 12:
        CVM_NWS(1, 2);
        break;
      case 14:
        CVM_NWS(1, 4);
        break;
      default:
        CVM_NWS(1, 8);
        break;
    }
#undef CVM_NWS
#undef CVM_NWV
#undef CVM_NWX
#undef CVM_NW
    return;
  }
  int br = norm_band_rows(rh);
  dim3 grid((rw + NORM_NX - 1) / NORM_NX, (rh + br - 1) / br);
  int W = NORM_NX + tw - 1;
  int cpt = (W + NORM_NX - 1) / NORM_NX;
#define CVM_NORM(C)                                                                                              \
  hipLaunchKernelGGL(k_normalize<C>, grid, dim3(NORM_NX), 0, s, img, istride, iw, step, tw, th, rw, rh, br, st, \
                     corr, map, partials)
  // templates up to the widest row transform (4096): cpt <= 17
  if (cpt <= 1)
    CVM_NORM(1);
  else if (cpt <= 2)
    CVM_NORM(2);
  else if (cpt <= 4)
    CVM_NORM(4);
  else if (cpt <= 8)
    CVM_NORM(8);
  else if (cpt <= 12)
    CVM_NORM(12);
  else
    CVM_NORM(17);
#undef CVM_NORM
}

// Wave-form variant of a shape: (q, sq32), or q = 0 when the int64 form applies.
void norm_form(int tw, int th, int cnMax, int* q, int* sq32) {
  NormPlan np = norm_plan(tw, th, cnMax);
  *q = np.wave ? np.q : 0;
  *sq32 = np.sq32;
}

int norm_grid_x(int rw, int q) { return (rw + 64 * q - 1) / (64 * q); }
int norm_grid_y(int rh) { return (rh + NB_ROWS - 1) / NB_ROWS; }

int launch_norm_batch(const uint8_t* img, int64_t istride, int iw, int ih, int step, int q, int sq32, bool exact,
                      const NormJob* jobs, const int* wg_job, int nwg, hipStream_t s) {
  if (nwg <= 0) return -1;
  const bool a16 = ((((uintptr_t)img) | (uintptr_t)istride) & 15) == 0;
  const bool a8 = ((((uintptr_t)img) | (uintptr_t)istride) & 7) == 0;
  int vmode = 0;
  if (iw % NORM_CPT == 0 && step == 4 && a16) vmode = 1;
  if (iw % NORM_CPT == 0 && step == 1 && a8) vmode = 2;
#define CVM_NB(Q, S, V, X) \
  hipLaunchKernelGGL((k_norm_batch<Q, S, V, X>), dim3(nwg), dim3(64), 0, s, img, istride, iw, ih, step, jobs, wg_job)
#define CVM_NBX(Q, S, V)  \
  if (exact)              \
    CVM_NB(Q, S, V, true); \
  else                    \
    CVM_NB(Q, S, V, false)
#define CVM_NBV(Q, S)      \
  if (vmode == 1) {        \
    CVM_NBX(Q, S, 1);      \
  } else if (vmode == 2) { \
    CVM_NBX(Q, S, 2);      \
  } else {                 \
    CVM_NBX(Q, S, 0);      \
  }
  if (q == 4) {
    if (sq32) {
      CVM_NBV(4, true);
    } else {
      CVM_NBV(4, false);
    }
  } else if (q == 2) {
    if (sq32) {
      CVM_NBV(2, true);
    } else {
      CVM_NBV(2, false);
    }
  } else {
    return -1;
  }
#undef CVM_NBV
#undef CVM_NBX
#undef CVM_NB
  return 0;
}

// Final reduction of the workgroup partials and conversion to the
// reference's six-tuple (min_v, min_x, min_y, max_v, max_x, max_y).
struct Res6 {
  float min_v;
  int32_t min_x, min_y;
  float max_v;
  int32_t max_x, max_y;
};

__device__ void reduce_block(const Extrema* __restrict__ parts, int n, int rw, Res6* out);

__global__ __launch_bounds__(1024) void k_reduce(const Extrema* __restrict__ parts, int n, int rw, Res6* out) {
  reduce_block(parts, n, rw, out);
}

__global__ __launch_bounds__(1024) void k_reduce_batch(const NormJob* __restrict__ jobs) {
  const NormJob J = jobs[blockIdx.x];
  reduce_block(J.parts, J.gx * J.gy, J.rw, (Res6*)J.res);
}

// The compact call's reduction, which also clears its state's accumulators
// for the next compact call (nothing reads them after the normalisation).
__global__ __launch_bounds__(1024) void k_reduce_clear(const Extrema* __restrict__ parts, int n, int rw, Res6* out,
                                                       CallState* st) {
  reduce_block(parts, n, rw, out);
  if (threadIdx.x == 0) state_clear(st);
}

void launch_reduce_clear(const Extrema* partials, int nparts, int rw, void* res6, CallState* st, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_clear, dim3(1), dim3(1024), 0, s, partials, nparts, rw, (Res6*)res6, st);
}

void launch_reduce_batch(const NormJob* jobs, int njobs, hipStream_t s) {
  if (njobs > 0) hipLaunchKernelGGL(k_reduce_batch, dim3(njobs), dim3(1024), 0, s, jobs);
}

__device__ void reduce_block(const Extrema* __restrict__ parts, int n, int rw, Res6* out) {
  __shared__ float rv[2][16];
  __shared__ int ri[2][16];
  float bmin = FLT_MAX, bmax = -FLT_MAX;
  int bmin_i = INT_MAX, bmax_i = INT_MAX;
  // four partials' loads in flight at a time
  for (int i0 = threadIdx.x; i0 < n; i0 += 4 * blockDim.x) {
    Extrema e[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i = i0 + u * (int)blockDim.x;
      if (i < n) {
        e[u] = parts[i];
      } else {
        e[u].min_v = FLT_MAX;
        e[u].max_v = -FLT_MAX;
        e[u].min_i = e[u].max_i = INT_MAX;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (better_min(e[u].min_v, e[u].min_i, bmin, bmin_i)) {
        bmin = e[u].min_v;
        bmin_i = e[u].min_i;
      }
      if (better_max(e[u].max_v, e[u].max_i, bmax, bmax_i)) {
        bmax = e[u].max_v;
        bmax_i = e[u].max_i;
      }
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
  if (wv != 0) return;
  // the waves' results across the first wave's lanes, reduced by shuffles
  // (the order is total, so any grouping gives the same winner)
  const int nw = (int)(blockDim.x >> 6);
  bmin = lane < nw ? rv[0][lane] : FLT_MAX;
  bmin_i = lane < nw ? ri[0][lane] : INT_MAX;
  bmax = lane < nw ? rv[1][lane] : -FLT_MAX;
  bmax_i = lane < nw ? ri[1][lane] : INT_MAX;
  for (int o = 8; o > 0; o >>= 1) {
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
