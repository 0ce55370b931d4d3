// Block-FFT cross-correlation kernels for gfx950.
//
// Each kernel performs, per output element, exactly the reference's scalar
// op sequence (radix4.go:75-149, radix4cols.go:33-125, impl.go:301-464), so
// the bits do not depend on the launch shape.  What differs from the
// reference is only the schedule and the storage:
//   * transforms run in LDS as register passes: a thread loads the 8, 16 or
//     32 elements one or two radix-4 stages (and the radix-2 head) combine,
//     runs those butterflies in registers and stores them back, so an
//     n-point transform makes ceil(stages/2) LDS round trips;
//   * the bit-reversal permutation (radix4.go:93-96) is folded into the LDS
//     stores that feed a transform;
//   * the forward column transform, the conjugate product with the template
//     spectrum and the inverse column transform are ONE kernel over an LDS
//     slab of CG columns, so the spectrum makes one HBM round trip; zero pad
//     rows are synthesised in LDS, never read;
//   * the inverse row kernel accumulates the channels in registers in
//     channel order and stores each result row once;
//   * workgroup numbers are remapped so that workgroups which share input
//     (the channels of one row pair, the tiles of one template-spectrum slab)
//     run on the same XCD and share its L2.
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

// Row vectors in LDS get one pad element per 32, so threads that read 8-32
// consecutive elements each do not pile onto one bank.
__device__ __forceinline__ int padi(int i) { return i + (i >> 5); }
__host__ __device__ inline int padded_len(int n) { return n + (n >> 5); }

// makeBitrev impl.go:148-165 as one bit-reverse instruction (log2n >= 1).
__device__ __forceinline__ int bitrev_idx(int i, int log2n) { return (int)(__brev((unsigned)i) >> (32 - log2n)); }

// Workgroups are dispatched to the 8 XCDs round robin.  Renumber them so
// that consecutive logical ids run on the same XCD.
__device__ __forceinline__ int xcd_remap(int p, int g) {
  int per = g >> 3, rem = g & 7;
  int x = p & 7, slot = p >> 3;
  return x * per + (x < rem ? x : rem) + slot;
}

// ---------------------------------------------------------- LDS transform --
//
// nvec independent n-point vectors live in LDS: element i of vector v at
// a[v * vstride + (PAD ? padi(i) : i) * es].  fft_lds runs the whole radix-4
// cascade (fftR4Scalar radix4.go:106-149) on all of them, input already in
// bit-reversed order.  A pass covers index bits [lo, lo+NB): the radix-2
// head (bit 0, HEAD = 1) and NST radix-4 stages whose digit sits at bits
// [lo+HEAD+2s, lo+HEAD+2s+2); stage s has quarter size h = 2^(lo+HEAD+2s)
// and its twiddle index j = index mod h.
template <int HEAD, int NST, bool PAD, bool VFAST>
__device__ __forceinline__ void fft_pass(float2* a, int nvec, int lvec, int vstride, int es, int log2n, int lo,
                                         const float2* __restrict__ tri0, const float2* __restrict__ tri1,
                                         bool inverse) {
  constexpr int NB = HEAD + 2 * NST;
  constexpr int E = 1 << NB;
  const int lgrp = log2n - NB;  // log2 of the groups per vector
  const int total = 1 << (lgrp + lvec);
  const int h0 = 1 << (lo + HEAD);
  const int h1 = h0 << 2;
  for (int gi = threadIdx.x; gi < total; gi += blockDim.x) {
    int v, gg;
    if (VFAST) {
      v = gi & ((1 << lvec) - 1);
      gg = gi >> lvec;
    } else {
      gg = gi & ((1 << lgrp) - 1);
      v = gi >> lgrp;
    }
    if (v >= nvec) continue;  // lvec rounds nvec up to a power of two
    const int low = gg & ((1 << lo) - 1);
    const int base = low | ((gg >> lo) << (lo + NB));
    float2* av = a + v * vstride;
    float2 r[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
      int idx = base | (e << lo);
      r[e] = av[(PAD ? padi(idx) : idx) * es];
    }
    if (HEAD) {
#pragma unroll
      for (int e = 0; e < E; e += 2) {
        float2 u = r[e], w = r[e + 1];
        r[e] = make_float2(u.x + w.x, u.y + w.y);
        r[e + 1] = make_float2(u.x - w.x, u.y - w.y);
      }
    }
    if (NST >= 1) {
      constexpr int d = HEAD;
#pragma unroll
      for (int e = 0; e < E; e++) {
        if ((e >> d) & 3) continue;
        int j = low | ((e & ((1 << d) - 1)) << lo);
        float2 w1 = conj_if(tri0[j], inverse), w2 = conj_if(tri0[h0 + j], inverse),
               w3 = conj_if(tri0[2 * h0 + j], inverse);
        bfly4(r[e], r[e | (1 << d)], r[e | (2 << d)], r[e | (3 << d)], w1, w2, w3, inverse);
      }
    }
    if (NST >= 2) {
      constexpr int d = HEAD + 2;
#pragma unroll
      for (int e = 0; e < E; e++) {
        if ((e >> d) & 3) continue;
        int j = low | ((e & ((1 << d) - 1)) << lo);
        float2 w1 = conj_if(tri1[j], inverse), w2 = conj_if(tri1[h1 + j], inverse),
               w3 = conj_if(tri1[2 * h1 + j], inverse);
        bfly4(r[e], r[e | (1 << d)], r[e | (2 << d)], r[e | (3 << d)], w1, w2, w3, inverse);
      }
    }
#pragma unroll
    for (int e = 0; e < E; e++) {
      int idx = base | (e << lo);
      av[(PAD ? padi(idx) : idx) * es] = r[e];
    }
  }
}

// The full cascade: the caller has stored the input and synchronised; on
// return every vector holds its transform and the block is synchronised.
// Passes hold at most 16 elements per thread: [head + 1 stage] when log2 n
// is odd, then two radix-4 stages per pass.
template <bool PAD, bool VFAST>
__device__ void fft_lds(float2* a, int nvec, int vstride, int es, int log2n, const float2* __restrict__ tri,
                        bool inverse) {
  int lvec = 0;
  while ((1 << lvec) < nvec) lvec++;
  const int S = log2n >> 1;  // radix-4 stages
  int s = 0, lo = 0;
  const float2* t = tri;
  if (log2n & 1) {
    if (S >= 1) {
      fft_pass<1, 1, PAD, VFAST>(a, nvec, lvec, vstride, es, log2n, 0, t, t, inverse);
      t += 3 * 2;
      s = 1;
      lo = 3;
    } else {
      fft_pass<1, 0, PAD, VFAST>(a, nvec, lvec, vstride, es, log2n, 0, t, t, inverse);
      lo = 1;
    }
    __syncthreads();
  }
  while (s < S) {
    int nst = (S - s) < 2 ? (S - s) : 2;
    int h0 = 1 << lo;
    const float2* t1 = t + 3 * h0;
    if (nst == 2)
      fft_pass<0, 2, PAD, VFAST>(a, nvec, lvec, vstride, es, log2n, lo, t, t1, inverse);
    else
      fft_pass<0, 1, PAD, VFAST>(a, nvec, lvec, vstride, es, log2n, lo, t, t1, inverse);
    __syncthreads();
    for (int q = 0; q < nst; q++) t += 3 * (h0 << (2 * q));
    s += nst;
    lo += 2 * nst;
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
// ------------------------------------------------------ forward row FFTs --

// Channels per workgroup for row transforms of length n: as many as fit a
// 144 KiB LDS budget (gfx950 gives a workgroup up to 160 KiB), so one
// workgroup transforms every channel of its row pair.
static int row_cpg(int n, int cnMax) {
  int cpg = (int)(147456 / ((size_t)padded_len(n) * sizeof(float2)));
  if (cpg > cnMax) cpg = cnMax;
  return cpg < 1 ? 1 : cpg;
}

// forwardRowPair impl.go:301-344: rows r, r+1 of channel k of tile t packed
// as one complex row (zero padded to dftW), transformed, untangled into the
// two half spectra, stored at spectrum rows r and r+1 (natural order; the
// column kernel applies the bit reversal when it loads).  One workgroup
// handles up to cpg channels of one row pair.
__global__ __launch_bounds__(256) void k_row_fwd(const uint8_t* __restrict__ img, int64_t istride, int step,
                                                 const CallState* __restrict__ st, TileGrid g,
                                                 const float2* __restrict__ triW, const int32_t* __restrict__ brevW,
                                                 float2* __restrict__ spec, int64_t tileStride, int64_t chanStride,
                                                 int cpg, int ngroups, int npairs) {
  extern __shared__ float2 z[];
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int cgi = L % ngroups, rest = L / ngroups;
  const int rp = rest % npairs, t = rest / npairs;
  if (st->flat) return;
  const int k0 = cgi * cpg;
  const int nv = min(cpg, st->cn - k0);
  if (nv <= 0) return;
  int x0, y0, bw, bh;
  tile_rect(g, t, x0, y0, bw, bh);
  const int loadW = bw + g.padW, loadH = bh + g.padH;
  const int dftH = tile_shrunk(g, y0) ? g.dftH2 : g.dftH;
  const int r = 2 * rp;
  if (r >= loadH || r >= dftH) return;
  const int n = g.dftW, np = padded_len(n), tid = threadIdx.x, nt = blockDim.x;
  const uint8_t* ra = img + (int64_t)(y0 + r) * istride + (int64_t)x0 * step + k0;
  const bool two = r + 1 < loadH;
  // all of this thread's pixels are loaded before any is stored
  constexpr int MAXI = 16;  // n / threads: 4096 / 256, 1024 / 64
  uint32_t wa[MAXI], wb[MAXI];
#pragma unroll
  for (int i = 0; i < MAXI; i++) {
    int x = tid + i * nt;
    wa[i] = wb[i] = 0;
    if (x < loadW) {
      const uint8_t* pa = ra + (int64_t)x * step;
      for (int v = 0; v < nv; v++) {
        wa[i] |= (uint32_t)pa[v] << (8 * v);
        if (two) wb[i] |= (uint32_t)pa[istride + v] << (8 * v);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < MAXI; i++) {
    int x = tid + i * nt;
    if (x < n) {
      const int pos = padi(bitrev_idx(x, g.log2W));
      for (int v = 0; v < nv; v++)
        z[v * np + pos] = make_float2((float)((wa[i] >> (8 * v)) & 255u), (float)((wb[i] >> (8 * v)) & 255u));
    }
  }
  __syncthreads();
  fft_lds<true, false>(z, nv, np, 1, g.log2W, triW, false);
  const int mask = n - 1;
  for (int v = 0; v < nv; v++) {
    float2* sa = spec + (int64_t)t * tileStride + (int64_t)(k0 + v) * chanStride + (int64_t)r * g.hwPad;
    float2* sb = sa + g.hwPad;
    const float2* zv = z + v * np;
    for (int q = tid; q < g.hw; q += nt) {
      float2 zk = zv[padi(q)], zn = zv[padi((n - q) & mask)];
      sa[q] = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
      sb[q] = make_float2(0.5f * (zk.y + zn.y), 0.5f * (zn.x - zk.x));
    }
  }
}

void launch_row_fwd(const uint8_t* img, int64_t istride, int step, int cnMax, const CallState* st,
                    const TileGrid& g, const PlanTables& pt, float2* spec, int64_t tileStride, int64_t chanStride,
                    int ntiles, hipStream_t s) {
  int nt = g.dftW >= 1024 ? 256 : 64;
  int cpg = row_cpg(g.dftW, cnMax);
  int ngroups = (cnMax + cpg - 1) / cpg;
  int npairs = g.dftH / 2;
  size_t lds = sizeof(float2) * (size_t)padded_len(g.dftW) * cpg;
  hipLaunchKernelGGL(k_row_fwd, dim3(ngroups * npairs * ntiles), dim3(nt), lds, s, img, istride, step, st, g, pt.triW,
                     pt.brevW, spec, tileStride, chanStride, cpg, ngroups, npairs);
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
// of colsR4Go), and leaves spatial row y at LDS row y.  Logical workgroup
// order is tile-fastest, so the tiles that read one template-spectrum slab
// run together on one XCD.
template <int CG>
__global__ __launch_bounds__(256) void k_cols(int mode, const float2* __restrict__ src, float2* __restrict__ dst,
                                              int64_t srcTileStride, int64_t dstTileStride, int64_t chanStride,
                                              const CallState* __restrict__ st, TileGrid g,
                                              const float2* __restrict__ triH, const int32_t* __restrict__ brevH,
                                              const float2* __restrict__ triH2, const int32_t* __restrict__ brevH2,
                                              const float2* __restrict__ tspec, const float2* __restrict__ tspec2,
                                              int64_t tchanStride, int64_t tchanStride2, float scale, int ntiles,
                                              int nslabs) {
  extern __shared__ float2 a[];
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int t = L % ntiles, rest = L / ntiles;
  const int slab = rest % nslabs, k = rest / nslabs;
  const int c0 = slab * CG;
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
  const int tot = n * CG;
  for (int b0 = tid; b0 < tot; b0 += nt * 8) {
    float2 v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      int idx = b0 + u * nt;
      int r = idx / CG, c = idx % CG;
      v[u] = make_float2(0.f, 0.f);
      if (idx < tot && r < loaded && c < ncol) v[u] = sp[(int64_t)r * hwPad + c];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      int idx = b0 + u * nt;
      if (idx < tot) {
        int r = idx / CG, c = idx % CG;
        int pos = (mode == COL_MULINV) ? r : bitrev_idx(r, log2n);
        a[pos * CG + c] = v[u];
      }
    }
  }
  __syncthreads();
  if (mode != COL_MULINV) fft_lds<false, true>(a, CG, 1, CG, log2n, tri, false);

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
  for (int b0 = tid; b0 < tot; b0 += nt * 4) {
    float2 tx[4], ty[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      int idx = b0 + u * nt;
      int r = idx / CG, c = idx % CG;
      int br = bitrev_idx(r, log2n);
      tx[u] = ty[u] = make_float2(0.f, 0.f);
      if (idx < tot && r <= br && c < ncol) {
        tx[u] = ts[(int64_t)r * hwPad + c0 + c];
        ty[u] = ts[(int64_t)br * hwPad + c0 + c];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      int idx = b0 + u * nt;
      int r = idx / CG, c = idx % CG;
      int br = bitrev_idx(r, log2n);
      if (idx < tot && r <= br) {
        float2 x = a[r * CG + c], y = a[br * CG + c];
        float p0 = x.x * tx[u].x, p1 = x.y * tx[u].y, p2 = x.y * tx[u].x, p3 = x.x * tx[u].y;
        float q0 = y.x * ty[u].x, q1 = y.y * ty[u].y, q2 = y.y * ty[u].x, q3 = y.x * ty[u].y;
        a[br * CG + c] = make_float2(p0 + p1, p2 - p3);
        a[r * CG + c] = make_float2(q0 + q1, q2 - q3);
      }
    }
  }
  __syncthreads();
  fft_lds<false, true>(a, CG, 1, CG, log2n, tri, true);

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
  int nslabs = (g.hw + cg - 1) / cg;
  dim3 grid(nslabs * ntiles * cnMax);
  dim3 block(256);
#define CVM_COLS(CGV)                                                                                          \
  hipLaunchKernelGGL(k_cols<CGV>, grid, block, lds, s, (int)mode, src, dst, srcTileStride, dstTileStride,      \
                     chanStride, st, g, pt.triH, pt.brevH, pt.triH2, pt.brevH2, tspec, tspec2, tchanStride,    \
                     tchanStride2, scale, ntiles, nslabs)
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

constexpr int ROWINV_MAXI = 16;  // dftW / threads: 4096/256 and 1024/64

// inverseRowPair impl.go:394-444 for every channel of tile t, rows r, r+1:
// combine the two half spectra into one complex row, inverse transform,
// accumulate the real/imaginary parts into result rows y0+r / y0+r+1 with
// channel 0 stored and channels 1.. added in order (in registers; each
// thread owns the same columns for every channel).
__global__ __launch_bounds__(256) void k_row_inv(const float2* __restrict__ spec, int64_t tileStride,
                                                 int64_t chanStride, const CallState* __restrict__ st, TileGrid g,
                                                 const float2* __restrict__ triW, const int32_t* __restrict__ brevW,
                                                 float* __restrict__ corr, int rw, int cpg, int npairs) {
  extern __shared__ float2 z[];
  const int n = g.dftW, np = padded_len(n);
  const int L = blockIdx.x;
  const int rp = L % npairs, t = L / npairs;
  const int r = 2 * rp;
  if (st->flat) return;
  int x0, y0, bw, bh;
  tile_rect(g, t, x0, y0, bw, bh);
  if (r >= bh) return;
  const int cn = st->cn, tid = threadIdx.x, nt = blockDim.x, hw = g.hw;
  float2 acc[ROWINV_MAXI];
#pragma unroll
  for (int i = 0; i < ROWINV_MAXI; i++) acc[i] = make_float2(0.f, 0.f);
  for (int k0 = 0; k0 < cn; k0 += cpg) {
    const int nv = min(cpg, cn - k0);
    for (int v = 0; v < nv; v++) {
      const float2* sa = spec + (int64_t)t * tileStride + (int64_t)(k0 + v) * chanStride + (int64_t)r * g.hwPad;
      const float2* sb = sa + g.hwPad;
      float2* zv = z + v * np;
      for (int q0 = tid; q0 < n; q0 += nt * 8) {
        float2 A[8], B[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          int q = q0 + u * nt;
          int m = q < hw ? q : n - q;
          A[u] = B[u] = make_float2(0.f, 0.f);
          if (q < n) {
            A[u] = sa[m];
            B[u] = sb[m];
          }
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          int q = q0 + u * nt;
          if (q < n) {
            float2 w = q < hw ? make_float2(A[u].x - B[u].y, A[u].y + B[u].x)
                              : make_float2(A[u].x + B[u].y, B[u].x - A[u].y);
            zv[padi(bitrev_idx(q, g.log2W))] = w;
          }
        }
      }
    }
    __syncthreads();
    fft_lds<true, false>(z, nv, np, 1, g.log2W, triW, true);
    for (int v = 0; v < nv; v++) {
      const float2* zv = z + v * np;
#pragma unroll
      for (int i = 0; i < ROWINV_MAXI; i++) {
        int x = tid + i * nt;
        if (x < bw) {
          float2 w = zv[padi(x)];
          if (k0 + v == 0)
            acc[i] = w;
          else
            acc[i] = make_float2(acc[i].x + w.x, acc[i].y + w.y);
        }
      }
    }
    __syncthreads();
  }
  float* o1 = corr + (int64_t)(y0 + r) * rw + x0;
  float* o2 = o1 + rw;
  const bool second = r + 1 < bh;
#pragma unroll
  for (int i = 0; i < ROWINV_MAXI; i++) {
    int x = tid + i * nt;
    if (x < bw) {
      o1[x] = acc[i].x;
      if (second) o2[x] = acc[i].y;
    }
  }
}

void launch_row_inv(const float2* spec, int64_t tileStride, int64_t chanStride, const CallState* st, int cnMax,
                    const TileGrid& g, const PlanTables& pt, float* corr, int rw, int ntiles, hipStream_t s) {
  int nt = g.dftW >= 1024 ? 256 : 64;
  int cpg = row_cpg(g.dftW, cnMax);
  size_t lds = sizeof(float2) * (size_t)padded_len(g.dftW) * cpg;
  // last-band tiles may be taller than blockH (fleet members): cover the
  // largest bh any tile can have
  int bhMax = g.blockH;
  int bhl = g.extH - (g.nty - 1) * g.blockH;
  if (bhl > bhMax) bhMax = bhl;
  int npairs = (bhMax + 1) / 2;
  hipLaunchKernelGGL(k_row_inv, dim3(npairs * ntiles), dim3(nt), lds, s, spec, tileStride, chanStride, st, g, pt.triW,
                     pt.brevW, corr, rw, cpg, npairs);
}

}  // namespace cvm
