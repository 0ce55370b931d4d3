// Block-FFT cross-correlation kernels for gfx950.
//
// Each kernel performs, per output element, exactly the reference's scalar
// op sequence (radix4.go:75-149, radix4cols.go:33-125, impl.go:301-464), so
// the bits do not depend on the launch shape.  What differs from the
// reference is only the schedule and the storage:
//   * row transforms run one packed row pair per workgroup, in LDS, with the
//     bit-reversal permutation folded into the LDS store (radix4.go:93-96);
//   * the forward column transform, the conjugate product with the template
//     spectrum and the inverse column transform are ONE kernel over an LDS
//     slab of CG columns, so the spectrum makes one HBM round trip instead of
//     three; zero pad rows are never read (they are synthesised in LDS);
//   * the inverse row kernel accumulates the channels in LDS in channel
//     order and stores each result row once.
#include "cvm_internal.h"

namespace cvm {

// ------------------------------------------------------ complex helpers --

// mulPlain radix4.go:80-85: each product and the sum/difference rounds once.
__device__ __forceinline__ float2 mul_plain(float2 v, float2 w) {
  float a = v.x * w.x;
  float b = v.y * w.y;
  float c = v.x * w.y;
  float d = v.y * w.x;
  return make_float2(a - b, c + d);
}

// One radix-4 DIT butterfly (fftR4Scalar radix4.go:121-145).  A = a[j],
// C = a[j+h], B = a[j+2h], D = a[j+3h]; twiddles already direction-adjusted.
__device__ __forceinline__ void bfly4(float2& A, float2& Cq, float2& B, float2& D, float2 w1, float2 w2, float2 w3,
                                      bool inverse) {
  float2 av = A;
  float2 tb = mul_plain(B, w1);
  float2 tc = mul_plain(Cq, w2);
  float2 td = mul_plain(D, w3);
  float2 s0 = make_float2(av.x + tc.x, av.y + tc.y);
  float2 s1 = make_float2(av.x - tc.x, av.y - tc.y);
  float2 s2 = make_float2(tb.x + td.x, tb.y + td.y);
  float2 s3 = make_float2(tb.x - td.x, tb.y - td.y);
  float rr, ri;
  if (inverse) {
    rr = -s3.y;
    ri = s3.x;
  } else {
    rr = s3.y;
    ri = -s3.x;
  }
  A = make_float2(s0.x + s2.x, s0.y + s2.y);
  Cq = make_float2(s1.x + rr, s1.y + ri);
  B = make_float2(s0.x - s2.x, s0.y - s2.y);
  D = make_float2(s1.x - rr, s1.y - ri);
}

__device__ __forceinline__ float2 conj_if(float2 w, bool inverse) { return inverse ? make_float2(w.x, -w.y) : w; }

// Full cascade over an LDS vector a[n] that already holds bit-reversed input
// (fftR4Scalar radix4.go:106-149).  Caller synchronises before the call.
__device__ void cascade_row(float2* a, int n, int log2n, const float2* __restrict__ tri, bool inverse) {
  const int tid = threadIdx.x, nt = blockDim.x;
  int h = 1;
  if (log2n & 1) {
    for (int i = tid; i < (n >> 1); i += nt) {
      float2 u = a[2 * i], v = a[2 * i + 1];
      a[2 * i] = make_float2(u.x + v.x, u.y + v.y);
      a[2 * i + 1] = make_float2(u.x - v.x, u.y - v.y);
    }
    __syncthreads();
    h = 2;
  }
  for (; 4 * h <= n; h *= 4) {
    for (int b = tid; b < (n >> 2); b += nt) {
      int j = b & (h - 1);
      int base = (b - j) * 4 + j;
      float2 w1 = conj_if(tri[j], inverse), w2 = conj_if(tri[h + j], inverse), w3 = conj_if(tri[2 * h + j], inverse);
      float2 A = a[base], Cq = a[base + h], B = a[base + 2 * h], D = a[base + 3 * h];
      bfly4(A, Cq, B, D, w1, w2, w3, inverse);
      a[base] = A;
      a[base + h] = Cq;
      a[base + 2 * h] = B;
      a[base + 3 * h] = D;
    }
    __syncthreads();
    tri += 3 * h;
  }
}

// The same cascade on every column of an LDS slab a[n][CG] (colsR4Go
// radix4cols.go:260-274 with logical row L at LDS row L).
template <int CG>
__device__ void cascade_cols(float2* a, int n, int log2n, const float2* __restrict__ tri, bool inverse) {
  const int tid = threadIdx.x, nt = blockDim.x;
  int h = 1;
  if (log2n & 1) {
    for (int idx = tid; idx < (n >> 1) * CG; idx += nt) {
      int i = idx / CG, c = idx % CG;
      float2 u = a[(2 * i) * CG + c], v = a[(2 * i + 1) * CG + c];
      a[(2 * i) * CG + c] = make_float2(u.x + v.x, u.y + v.y);
      a[(2 * i + 1) * CG + c] = make_float2(u.x - v.x, u.y - v.y);
    }
    __syncthreads();
    h = 2;
  }
  for (; 4 * h <= n; h *= 4) {
    for (int idx = tid; idx < (n >> 2) * CG; idx += nt) {
      int b = idx / CG, c = idx % CG;
      int j = b & (h - 1);
      int base = (b - j) * 4 + j;
      float2 w1 = conj_if(tri[j], inverse), w2 = conj_if(tri[h + j], inverse), w3 = conj_if(tri[2 * h + j], inverse);
      float2 A = a[base * CG + c], Cq = a[(base + h) * CG + c];
      float2 B = a[(base + 2 * h) * CG + c], D = a[(base + 3 * h) * CG + c];
      bfly4(A, Cq, B, D, w1, w2, w3, inverse);
      a[base * CG + c] = A;
      a[(base + h) * CG + c] = Cq;
      a[(base + 2 * h) * CG + c] = B;
      a[(base + 3 * h) * CG + c] = D;
    }
    __syncthreads();
    tri += 3 * h;
  }
}

// -------------------------------------------------------------- twiddles --

// sincospiFrac impl.go:98-133 in FP64 on the device: exact dyadic range
// reduction and the fdlibm-style kernels, each double op rounded once.
__device__ void sincospi_frac(int j, int half, float& co, float& si) {
  int m = 2 * j;
  int k = m / half;
  int rem = m - k * half;
  double u = (double)rem / (double)half;
  bool swap = false;
  if (u > 0.5) {
    u = 1.0 - u;
    swap = true;
  }
  double y = u * 1.57079632679489661923;
  double z = y * y;
  double r = z * (4.16666666666666019037e-02 +
                  z * (-1.38888888888741095749e-03 +
                       z * (2.48015872894767294178e-05 +
                            z * (-2.75573143513906633035e-07 +
                                 z * (2.08757232129817482790e-09 + z * -1.13596475577881948265e-11)))));
  double hz = 0.5 * z;
  double w = 1.0 - hz;
  double c = w + (((1.0 - w) - hz) + z * r);
  double r2 = 8.33333333332248946124e-03 +
              z * (-1.98412698298579493134e-04 +
                   z * (2.75573137070700676789e-06 + z * (-2.50507602534068634195e-08 + z * 1.58969099521155010221e-10)));
  double v = z * y;
  double s = y + v * (-1.66666666666666324348e-01 + z * r2);
  if (swap) {
    double t = c;
    c = s;
    s = t;
  }
  if (k != 0) {
    double t = c;
    c = -s;
    s = t;
  }
  co = (float)c;
  si = (float)s;
}

int tri_len(int n) {
  int len = 0;
  int log2n = 0;
  while ((1 << log2n) < n) log2n++;
  int h = (log2n & 1) ? 2 : 1;
  for (; 4 * h <= n; h *= 4) len += 3 * h;
  return len;
}

// makeTriTwiddles radix4.go:45-73 (entry i of the flattened table) and the
// bit-reversal map of makeBitrev impl.go:148-165.
__global__ void k_twiddles(float2* tri, int32_t* brev, int n, int log2n, int len) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) brev[i] = log2n == 0 ? 0 : (int32_t)(__brev((unsigned)i) >> (32 - log2n));
  if (i >= len) return;
  int h = (log2n & 1) ? 2 : 1;
  int off = 0;
  while (i >= off + 3 * h) {
    off += 3 * h;
    h *= 4;
  }
  int q = i - off;
  int blk = q / h, j = q - blk * h;
  float c, s;
  if (blk == 0) {
    sincospi_frac(j, 2 * h, c, s);
  } else if (blk == 1) {
    sincospi_frac(j, h, c, s);
  } else {
    int m = 3 * j;
    bool neg = false;
    if (m >= 2 * h) {
      m -= 2 * h;
      neg = true;
    }
    sincospi_frac(m, 2 * h, c, s);
    if (neg) {
      c = -c;
      s = -s;
    }
  }
  tri[i] = make_float2(c, -s);
}

void launch_twiddles(float2* tri, int32_t* brev, int n, hipStream_t s) {
  int log2n = 0;
  while ((1 << log2n) < n) log2n++;
  int len = tri_len(n);
  int total = len > n ? len : n;
  hipLaunchKernelGGL(k_twiddles, dim3((total + 255) / 256), dim3(256), 0, s, tri, brev, n, log2n, len);
}

// ------------------------------------------------------ forward row FFTs --

// forwardRowPair impl.go:301-344: rows r, r+1 of channel k of tile t packed
// as one complex row (zero padded to dftW), transformed, untangled into the
// two half spectra, stored at spectrum rows r and r+1 (natural order; the
// column kernel applies the bit reversal when it loads).
__global__ __launch_bounds__(256) void k_row_fwd(const uint8_t* __restrict__ img, int64_t istride, int step,
                                                 const CallState* __restrict__ st, TileGrid g,
                                                 const float2* __restrict__ triW, const int32_t* __restrict__ brevW,
                                                 float2* __restrict__ spec, int64_t tileStride, int64_t chanStride) {
  extern __shared__ float2 z[];
  const int k = blockIdx.z, t = blockIdx.y, r = 2 * blockIdx.x;
  if (st->flat || k >= st->cn) return;
  int x0, y0, bw, bh;
  tile_rect(g, t, x0, y0, bw, bh);
  const int loadW = bw + g.padW, loadH = bh + g.padH;
  const int dftH = tile_shrunk(g, y0) ? g.dftH2 : g.dftH;
  if (r >= loadH || r >= dftH) return;
  const int n = g.dftW, tid = threadIdx.x, nt = blockDim.x;
  const uint8_t* ra = img + (int64_t)(y0 + r) * istride + (int64_t)x0 * step + k;
  const bool two = r + 1 < loadH;
  for (int x = tid; x < n; x += nt) {
    float2 v = make_float2(0.f, 0.f);
    if (x < loadW) {
      v.x = (float)ra[(int64_t)x * step];
      if (two) v.y = (float)ra[istride + (int64_t)x * step];
    }
    z[brevW[x]] = v;
  }
  __syncthreads();
  cascade_row(z, n, g.log2W, triW, false);
  float2* sa = spec + (int64_t)t * tileStride + (int64_t)k * chanStride + (int64_t)r * g.hwPad;
  float2* sb = sa + g.hwPad;
  const int mask = n - 1;
  for (int q = tid; q < g.hw; q += nt) {
    float2 zk = z[q], zn = z[(n - q) & mask];
    sa[q] = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
    sb[q] = make_float2(0.5f * (zk.y + zn.y), 0.5f * (zn.x - zk.x));
  }
}

void launch_row_fwd(const uint8_t* img, int64_t istride, int step, int cnMax, const CallState* st,
                    const TileGrid& g, const PlanTables& pt, float2* spec, int64_t tileStride, int64_t chanStride,
                    int ntiles, hipStream_t s) {
  int nt = g.dftW >= 1024 ? 256 : (g.dftW >= 256 ? 64 : 64);
  size_t lds = sizeof(float2) * (size_t)g.dftW;
  hipLaunchKernelGGL(k_row_fwd, dim3(g.dftH / 2, ntiles, cnMax), dim3(nt), lds, s, img, istride, step, st, g, pt.triW,
                     pt.brevW, spec, tileStride, chanStride);
}

// ----------------------------------------------------------- column pass --

// One workgroup owns CG spectrum columns of one tile/channel for the whole
// column height.  Modes:
//   COL_TSPEC : load rows (pad rows = 0), forward, scale, store all rows
//   COL_FWD   : load rows (pad rows = 0), forward, store all rows
//   COL_CORR  : load rows, forward, *conj(tspec), inverse, store needed rows
//   COL_MULINV: load all rows (already forward), *conj(tspec), inverse, store
// The forward input of row r goes to LDS row brev[r] (rows at bit-reversed
// slots, blockForwardGo impl.go:350-374); after the forward cascade LDS row
// k holds frequency k.  The product is written to LDS row brev[k] so the
// inverse cascade again sees bit-reversed input (the rmap = brev relabelling
// of colsR4Go), and leaves spatial row y at LDS row y.
template <int CG>
__global__ __launch_bounds__(256) void k_cols(int mode, const float2* __restrict__ src, float2* __restrict__ dst,
                                              int64_t srcTileStride, int64_t dstTileStride, int64_t chanStride,
                                              const CallState* __restrict__ st, TileGrid g,
                                              const float2* __restrict__ triH, const int32_t* __restrict__ brevH,
                                              const float2* __restrict__ triH2, const int32_t* __restrict__ brevH2,
                                              const float2* __restrict__ tspec, const float2* __restrict__ tspec2,
                                              int64_t tchanStride, int64_t tchanStride2, float scale) {
  extern __shared__ float2 a[];
  const int k = blockIdx.z, t = blockIdx.y, c0 = blockIdx.x * CG;
  if (st->flat || k >= st->cn || c0 >= g.hw) return;
  int x0, y0, bw, bh;
  tile_rect(g, t, x0, y0, bw, bh);
  const bool shr = tile_shrunk(g, y0);
  const int n = shr ? g.dftH2 : g.dftH;
  const int log2n = shr ? g.log2H2 : g.log2H;
  const float2* tri = shr ? triH2 : triH;
  const int32_t* brev = shr ? brevH2 : brevH;
  const float2* ts = shr ? tspec2 + (int64_t)k * tchanStride2 : tspec + (int64_t)k * tchanStride;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int hwPad = g.hwPad;
  const float2* sp = src + (int64_t)t * srcTileStride + (int64_t)k * chanStride + c0;
  float2* dp = dst + (int64_t)t * dstTileStride + (int64_t)k * chanStride + c0;

  int loaded = n;
  if (mode != COL_MULINV) {
    int loadH = bh + g.padH;
    loaded = (loadH + 1) & ~1;
    if (loaded > n) loaded = n;
  }
  const int ncol = (g.hw - c0) < CG ? (g.hw - c0) : CG;
  for (int idx = tid; idx < n * CG; idx += nt) {
    int r = idx / CG, c = idx % CG;
    float2 v = make_float2(0.f, 0.f);
    if (r < loaded && c < ncol) v = sp[(int64_t)r * hwPad + c];
    int pos = (mode == COL_MULINV) ? r : brev[r];
    a[pos * CG + c] = v;
  }
  __syncthreads();
  if (mode != COL_MULINV) cascade_cols<CG>(a, n, log2n, tri, false);

  if (mode == COL_TSPEC || mode == COL_FWD) {
    for (int idx = tid; idx < n * CG; idx += nt) {
      int r = idx / CG, c = idx % CG;
      if (c < ncol) {
        float2 v = a[r * CG + c];
        if (mode == COL_TSPEC) v = make_float2(v.x * scale, v.y * scale);
        dp[(int64_t)r * hwPad + c] = v;
      }
    }
    return;
  }

  // conjugate product (mulConjGo impl.go:376-388) fused with the bit
  // reversal: rows k and brev[k] are handled by the same thread.
  for (int idx = tid; idx < n * CG; idx += nt) {
    int r = idx / CG, c = idx % CG;
    int br = brev[r];
    if (r > br) continue;
    float2 x = a[r * CG + c], y = a[br * CG + c];
    float2 tx = make_float2(0.f, 0.f), ty = make_float2(0.f, 0.f);
    if (c < ncol) {
      tx = ts[(int64_t)r * hwPad + c0 + c];
      ty = ts[(int64_t)br * hwPad + c0 + c];
    }
    float p0 = x.x * tx.x, p1 = x.y * tx.y, p2 = x.y * tx.x, p3 = x.x * tx.y;
    float q0 = y.x * ty.x, q1 = y.y * ty.y, q2 = y.y * ty.x, q3 = y.x * ty.y;
    a[br * CG + c] = make_float2(p0 + p1, p2 - p3);
    a[r * CG + c] = make_float2(q0 + q1, q2 - q3);
  }
  __syncthreads();
  cascade_cols<CG>(a, n, log2n, tri, true);

  int need = (bh + 1) & ~1;
  if (need > n) need = n;
  for (int idx = tid; idx < need * CG; idx += nt) {
    int r = idx / CG, c = idx % CG;
    if (c < ncol) dp[(int64_t)r * hwPad + c] = a[r * CG + c];
  }
}

static int col_group(int dftH) {
  // LDS slab of dftH x CG complex kept at or below 64 KiB
  int cg = 16;
  while (cg > 1 && (size_t)dftH * cg * sizeof(float2) > 65536) cg >>= 1;
  return cg;
}

void launch_cols(ColMode mode, const float2* src, float2* dst, int64_t srcTileStride, int64_t dstTileStride,
                 int64_t chanStride, int cnMax, const CallState* st, const TileGrid& g, const PlanTables& pt,
                 const float2* tspec, const float2* tspec2, int64_t tchanStride, int64_t tchanStride2, float scale,
                 int ntiles, hipStream_t s) {
  int cg = col_group(g.dftH);
  size_t lds = sizeof(float2) * (size_t)g.dftH * cg;
  dim3 grid((g.hw + cg - 1) / cg, ntiles, cnMax);
  dim3 block(256);
#define CVM_COLS(CGV)                                                                                          \
  hipLaunchKernelGGL(k_cols<CGV>, grid, block, lds, s, (int)mode, src, dst, srcTileStride, dstTileStride,      \
                     chanStride, st, g, pt.triH, pt.brevH, pt.triH2, pt.brevH2, tspec, tspec2, tchanStride,    \
                     tchanStride2, scale)
  switch (cg) {
    case 16: CVM_COLS(16); break;
    case 8: CVM_COLS(8); break;
    case 4: CVM_COLS(4); break;
    case 2: CVM_COLS(2); break;
    default: CVM_COLS(1); break;
  }
#undef CVM_COLS
}

// shrinkInto impl.go:591-610: the shrunk band's template spectrum is the
// stride-2^shift row gather of the full one, times 2^shift (exact).
__global__ void k_shrink(const float2* __restrict__ tspec, float2* __restrict__ tspec2, int rows2, int shift, int hwPad,
                         int64_t chanStride, int64_t chanStride2) {
  int k = blockIdx.z, r = blockIdx.y;
  float up = (float)(1 << shift);
  const float2* srow = tspec + (int64_t)k * chanStride + (int64_t)(r << shift) * hwPad;
  float2* drow = tspec2 + (int64_t)k * chanStride2 + (int64_t)r * hwPad;
  for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < hwPad; x += gridDim.x * blockDim.x) {
    float2 v = srow[x];
    drow[x] = make_float2(v.x * up, v.y * up);
  }
}

void launch_shrink(const float2* tspec, float2* tspec2, int cnMax, int rows2, int shift, int hwPad,
                   int64_t chanStride, int64_t chanStride2, hipStream_t s) {
  hipLaunchKernelGGL(k_shrink, dim3((hwPad + 255) / 256, rows2, cnMax), dim3(256), 0, s, tspec, tspec2, rows2, shift,
                     hwPad, chanStride, chanStride2);
}

// ------------------------------------------------------ inverse row FFTs --

// inverseRowPair impl.go:394-444 for every channel of tile t, rows r, r+1:
// combine the two half spectra into one complex row, inverse transform,
// accumulate the real/imaginary parts into result rows y0+r / y0+r+1 with
// channel 0 stored and channels 1.. added in order.
__global__ __launch_bounds__(256) void k_row_inv(const float2* __restrict__ spec, int64_t tileStride,
                                                 int64_t chanStride, const CallState* __restrict__ st, TileGrid g,
                                                 const float2* __restrict__ triW, const int32_t* __restrict__ brevW,
                                                 float* __restrict__ corr, int rw) {
  extern __shared__ float2 sm[];
  const int n = g.dftW;
  float2* z = sm;
  float2* acc = sm + n;
  const int t = blockIdx.y, r = 2 * blockIdx.x;
  if (st->flat) return;
  int x0, y0, bw, bh;
  tile_rect(g, t, x0, y0, bw, bh);
  if (r >= bh) return;
  const int cn = st->cn, tid = threadIdx.x, nt = blockDim.x, hw = g.hw;
  for (int k = 0; k < cn; k++) {
    const float2* sa = spec + (int64_t)t * tileStride + (int64_t)k * chanStride + (int64_t)r * g.hwPad;
    const float2* sb = sa + g.hwPad;
    for (int q = tid; q < n; q += nt) {
      float2 v;
      if (q < hw) {
        float2 A = sa[q], B = sb[q];
        v = make_float2(A.x - B.y, A.y + B.x);
      } else {
        int m = n - q;
        float2 A = sa[m], B = sb[m];
        v = make_float2(A.x + B.y, B.x - A.y);
      }
      z[brevW[q]] = v;
    }
    __syncthreads();
    cascade_row(z, n, g.log2W, triW, true);
    for (int x = tid; x < bw; x += nt) {
      float2 v = z[x];
      if (k == 0) {
        acc[x] = v;
      } else {
        float2 o = acc[x];
        acc[x] = make_float2(o.x + v.x, o.y + v.y);
      }
    }
    __syncthreads();
  }
  float* o1 = corr + (int64_t)(y0 + r) * rw + x0;
  for (int x = tid; x < bw; x += nt) o1[x] = acc[x].x;
  if (r + 1 < bh) {
    float* o2 = o1 + rw;
    for (int x = tid; x < bw; x += nt) o2[x] = acc[x].y;
  }
}

void launch_row_inv(const float2* spec, int64_t tileStride, int64_t chanStride, const CallState* st, int cnMax,
                    const TileGrid& g, const PlanTables& pt, float* corr, int rw, int ntiles, hipStream_t s) {
  (void)cnMax;
  int nt = g.dftW >= 1024 ? 256 : 64;
  size_t lds = 2 * sizeof(float2) * (size_t)g.dftW;
  int maxPairs = (g.blockH > g.extH ? g.blockH : g.extH) / 2 + 1;
  // last-band tiles may be taller than blockH (fleet members); cover the
  // largest bh any tile can have
  int bhMax = g.blockH;
  {
    int y0last = (g.nty - 1) * g.blockH;
    int bhl = g.extH - y0last;
    if (bhl > bhMax) bhMax = bhl;
  }
  maxPairs = (bhMax + 1) / 2;
  hipLaunchKernelGGL(k_row_inv, dim3(maxPairs, ntiles), dim3(nt), lds, s, spec, tileStride, chanStride, st, g, pt.triW,
                     pt.brevW, corr, rw);
}

}  // namespace cvm
