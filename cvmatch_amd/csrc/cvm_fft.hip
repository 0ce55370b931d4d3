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
#include "cvm_state.h"

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
// NVEC independent 2^LOG2N-point vectors live in LDS: element i of vector v
// at a[v * vstride + (PAD ? padi(i) : i) * ES].  fft_lds runs the whole
// radix-4 cascade (fftR4Scalar radix4.go:106-149) on them, input already in
// bit-reversed order.  All sizes are compile-time constants, so the pass
// schedule, the group indices and the twiddle offsets fold away.
//
// A pass covers index bits [LO, LO+NB): the radix-2 head (bit 0, HEAD = 1)
// and NST radix-4 stages whose digit sits at bits [LO+HEAD+2s, +2); stage s
// has quarter size h = 2^(LO+HEAD+2s), twiddle index j = index mod h, and its
// w1/w2/w3 blocks start at tri + (h - base) (base = 2 for odd log2 n, else 1:
// the stages before it hold 3*(base + 4 base + ...) = h - base entries).

__host__ __device__ constexpr int ceil_log2c(int n) { return n <= 1 ? 0 : 1 + ceil_log2c((n + 1) >> 1); }

// The butterflies of one pass on the E = 2^(HEAD + 2*NST) elements of one
// group held in registers: r[e] is index base | (e << LO), low = base mod
// 2^LO.  The first radix-4 stage uses twiddles at tri + (h0 - base), the
// second at tri + (4*h0 - base).
template <int LOG2N, int LO, int HEAD, int NST>
__device__ __forceinline__ void bfly_regs(float2* r, int low, const float2* __restrict__ tri, bool inverse) {
  constexpr int E = 1 << (HEAD + 2 * NST);
  constexpr int tbase = (LOG2N & 1) ? 2 : 1;
  constexpr int h0 = 1 << (LO + HEAD);
  constexpr int h1 = h0 << 2;
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
    const float2* t0 = tri + (h0 - tbase);
#pragma unroll
    for (int e = 0; e < E; e++) {
      if ((e >> d) & 3) continue;
      int j = low | ((e & ((1 << d) - 1)) << LO);
      float2 w1 = conj_if(t0[j], inverse), w2 = conj_if(t0[h0 + j], inverse), w3 = conj_if(t0[2 * h0 + j], inverse);
      bfly4(r[e], r[e | (1 << d)], r[e | (2 << d)], r[e | (3 << d)], w1, w2, w3, inverse);
    }
  }
  if (NST >= 2) {
    constexpr int d = HEAD + 2;
    const float2* t1 = tri + (h1 - tbase);
#pragma unroll
    for (int e = 0; e < E; e++) {
      if ((e >> d) & 3) continue;
      int j = low | ((e & ((1 << d) - 1)) << LO);
      float2 w1 = conj_if(t1[j], inverse), w2 = conj_if(t1[h1 + j], inverse), w3 = conj_if(t1[2 * h1 + j], inverse);
      bfly4(r[e], r[e | (1 << d)], r[e | (2 << d)], r[e | (3 << d)], w1, w2, w3, inverse);
    }
  }
}

// Group gg of a pass over bits [LO, LO+NB): its lowest index and low bits.
template <int LO, int NB>
__device__ __forceinline__ int grp_base(int gg) {
  return (gg & ((1 << LO) - 1)) | ((gg >> LO) << (LO + NB));
}

// One pass in place in LDS over every group of every vector (threads take
// group-vector pairs, vector index fastest when VFAST).
template <int LOG2N, int LO, int HEAD, int NST, int NVEC, int NT, int ES, bool PAD, bool VFAST>
__device__ __forceinline__ void fft_pass(float2* a, int nvec, int vstride, const float2* __restrict__ tri,
                                         bool inverse) {
  constexpr int NB = HEAD + 2 * NST;
  constexpr int E = 1 << NB;
  constexpr int LGRP = LOG2N - NB;
  constexpr int LVEC = ceil_log2c(NVEC);
  constexpr int TOTAL = 1 << (LGRP + LVEC);
  constexpr int ITERS = (TOTAL + NT - 1) / NT;
#pragma unroll
  for (int it = 0; it < ITERS; it++) {
    const int gi = threadIdx.x + it * NT;
    if ((TOTAL % NT) != 0 && gi >= TOTAL) break;
    int v, gg;
    if (VFAST) {
      v = gi & ((1 << LVEC) - 1);
      gg = gi >> LVEC;
    } else {
      gg = gi & ((1 << LGRP) - 1);
      v = gi >> LGRP;
    }
    if (NVEC != (1 << LVEC) || !VFAST) {
      if (v >= nvec) continue;
    }
    const int base = grp_base<LO, NB>(gg);
    float2* av = a + v * vstride;
    float2 r[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
      int idx = base | (e << LO);
      r[e] = av[(PAD ? padi(idx) : idx) * ES];
    }
    bfly_regs<LOG2N, LO, HEAD, NST>(r, gg & ((1 << LO) - 1), tri, inverse);
#pragma unroll
    for (int e = 0; e < E; e++) {
      int idx = base | (e << LO);
      av[(PAD ? padi(idx) : idx) * ES] = r[e];
    }
  }
}

// Radix-4 passes from bit LO on, two stages per pass (one for the last odd
// stage out), each followed by a barrier.
template <int LOG2N, int LO, int NVEC, int NT, int ES, bool PAD, bool VFAST>
__device__ __forceinline__ void fft_rest(float2* a, int nvec, int vstride, const float2* __restrict__ tri,
                                         bool inverse) {
  constexpr int SREM = (LOG2N - LO) / 2;
  if constexpr (SREM >= 1) {
    constexpr int NST = SREM >= 2 ? 2 : 1;
    fft_pass<LOG2N, LO, 0, NST, NVEC, NT, ES, PAD, VFAST>(a, nvec, vstride, tri, inverse);
    __syncthreads();
    fft_rest<LOG2N, LO + 2 * NST, NVEC, NT, ES, PAD, VFAST>(a, nvec, vstride, tri, inverse);
  }
}

// The full cascade: the caller has stored the input and synchronised; on
// return every vector holds its transform and the block is synchronised.
template <int LOG2N, int NVEC, int NT, int ES, bool PAD, bool VFAST>
__device__ __forceinline__ void fft_lds(float2* a, int nvec, int vstride, const float2* __restrict__ tri,
                                        bool inverse) {
  if constexpr (LOG2N & 1) {
    // head + the h = 2 stage (a lone head for n = 2)
    constexpr int NST = LOG2N >= 3 ? 1 : 0;
    fft_pass<LOG2N, 0, 1, NST, NVEC, NT, ES, PAD, VFAST>(a, nvec, vstride, tri, inverse);
    __syncthreads();
    fft_rest<LOG2N, 1 + 2 * NST, NVEC, NT, ES, PAD, VFAST>(a, nvec, vstride, tri, inverse);
  } else {
    fft_rest<LOG2N, 0, NVEC, NT, ES, PAD, VFAST>(a, nvec, vstride, tri, inverse);
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

// Threads for a row transform of 2^LOG2N points (the forward kernel, with
// no per-thread accumulators, takes 512 from 2048 points on: more waves per
// CU behind the same LDS).
template <int LOG2N>
struct RowNT {
  static constexpr int value = LOG2N >= 10 ? 256 : 64;
};
template <int LOG2N>
struct RowNTF {
  static constexpr int value = LOG2N >= 11 ? 512 : RowNT<LOG2N>::value;
};
#ifndef CVM_ROWINV_WIDE
#define CVM_ROWINV_WIDE 1
#endif
template <int LOG2N>
struct RowNTI {
  static constexpr int value = (CVM_ROWINV_WIDE && LOG2N >= 11) ? 512 : RowNT<LOG2N>::value;
};
// Waves a SIMD the inverse row kernel is compiled for: two 512-thread
// workgroups of 2048 points fit a CU's LDS, so their registers are held to
// 128 (124 used, no spills); 4096 points fit one workgroup a CU.
template <int LOG2N>
struct RowInvWaves {
  static constexpr int value = (RowNTI<LOG2N>::value == 512 && LOG2N == 11) ? 4 : 1;
};
static int row_inv_nt(int log2n) { return (CVM_ROWINV_WIDE && log2n >= 11) ? 512 : log2n >= 10 ? 256 : 64; }

// Channels per workgroup for row transforms of length n: as many as fit a
// 144 KiB LDS budget (gfx950 gives a workgroup up to 160 KiB), so one
// workgroup usually transforms every channel of its row pair.
static int row_cpg(int n, int cnMax) {
  int cpg = (int)(147456 / ((size_t)padded_len(n) * sizeof(float2)));
  if (cpg > cnMax) cpg = cnMax;
  return cpg < 1 ? 1 : cpg;
}

// forwardRowPair impl.go:301-344: rows r, r+1 of channel k of tile t packed
// as one complex row (zero padded to dftW), transformed, untangled into the
// two half spectra, stored at spectrum rows r and r+1 (natural order; the
// column kernel applies the bit reversal when it loads).  One workgroup
// handles up to NVEC channels of one row pair.
// One forward row launch's arguments (a buffer of tiles of one image).
struct RowFwdArgs {
  const uint8_t* img;
  int64_t istride;
  int step;
  const CallState* st;
  TileGrid g;
  float2* spec;
  int64_t tileStride, chanStride;
  int cpg, ngroups, npairs, words, cnAll;
  int t0;   // first tile
  int nwg;  // workgroups: ngroups * npairs * ntiles
};

// Workgroup L of a forward row launch; z: the LDS vectors.
template <int LOG2N, int NVEC>
__device__ __forceinline__ void row_fwd_body(float2* z, int L, const RowFwdArgs& A, const float2* __restrict__ triW) {
  constexpr int n = 1 << LOG2N, NT = RowNTF<LOG2N>::value, np = n + (n >> 5);
  const uint8_t* __restrict__ img = A.img;
  const int64_t istride = A.istride, tileStride = A.tileStride, chanStride = A.chanStride;
  const int step = A.step, cpg = A.cpg, ngroups = A.ngroups, npairs = A.npairs, words = A.words, cnAll = A.cnAll;
  const CallState* __restrict__ st = A.st;
  const TileGrid& g = A.g;
  float2* __restrict__ spec = A.spec;
  const int cgi = L % ngroups, rest = L / ngroups;
  const int rp = rest % npairs, t = A.t0 + rest / npairs;
  // without a call state every channel of the buffer is transformed
  if (st && st->flat) return;
  const int k0 = cgi * cpg;
  const int nv = min(cpg, (st ? st->cn : cnAll) - k0);
  if (nv <= 0) return;
  int x0, y0, bw, bh;
  tile_rect(g, t, x0, y0, bw, bh);
  const int loadW = bw + g.padW, loadH = bh + g.padH;
  const int dftH = tile_shrunk(g, y0) ? g.dftH2 : g.dftH;
  const int r = 2 * rp;
  if (r >= loadH || r >= dftH) return;
  const int tid = threadIdx.x;
  const uint8_t* rpx = img + (int64_t)(y0 + r) * istride + (int64_t)x0 * step;
  const uint8_t* ra = rpx + k0;
  const bool two = r + 1 < loadH;
  // all of this thread's pixels are loaded before any is stored
  constexpr int MAXI = (n + NT - 1) / NT;
  uint32_t wa[MAXI], wb[MAXI];
  if (words) {
    // four-byte pixels, one dword load each, shifted so the group's first
    // channel sits in the low byte (bytes of channels past the group are
    // never used)
    const int64_t rb = two ? istride : 0;
    const int sh = 8 * k0;
#pragma unroll
    for (int i = 0; i < MAXI; i++) {
      int x = tid + i * NT;
      wa[i] = wb[i] = 0;
      if (x < loadW) {
        const uint8_t* pa = rpx + (int64_t)x * 4;
        wa[i] = *(const uint32_t*)pa >> sh;
        wb[i] = *(const uint32_t*)(pa + rb) >> sh;
      }
    }
    if (!two)
#pragma unroll
      for (int i = 0; i < MAXI; i++) wb[i] = 0;
  } else {
#pragma unroll
    for (int i = 0; i < MAXI; i++) {
      int x = tid + i * NT;
      wa[i] = wb[i] = 0;
      if (x < loadW) {
        const uint8_t* pa = ra + (int64_t)x * step;
        for (int v = 0; v < nv; v++) {
          wa[i] |= (uint32_t)pa[v] << (8 * v);
          if (two) wb[i] |= (uint32_t)pa[istride + v] << (8 * v);
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < MAXI; i++) {
    int x = tid + i * NT;
    if (x < n) {
      const int pos = padi(bitrev_idx(x, LOG2N));
#pragma unroll
      for (int v = 0; v < NVEC; v++)
        if (v < nv)
          z[v * np + pos] = make_float2((float)((wa[i] >> (8 * v)) & 255u), (float)((wb[i] >> (8 * v)) & 255u));
    }
  }
  __syncthreads();
  fft_lds<LOG2N, NVEC, NT, 1, true, false>(z, nv, np, triW, false);
  constexpr int mask = n - 1;
  const int hw = g.hw;
#pragma unroll
  for (int v = 0; v < NVEC; v++) {
    if (v >= nv) break;
    float2* sa = spec + (int64_t)t * tileStride + (int64_t)(k0 + v) * chanStride + (int64_t)r * g.hwPad;
    float2* sb = sa + g.hwPad;
    const float2* zv = z + v * np;
    for (int q = tid; q < hw; q += NT) {
      float2 zk = zv[padi(q)], zn = zv[padi((n - q) & mask)];
      sa[q] = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
      sb[q] = make_float2(0.5f * (zk.y + zn.y), 0.5f * (zn.x - zk.x));
    }
  }
}

template <int LOG2N, int NVEC>
__global__ __launch_bounds__(RowNTF<LOG2N>::value) void k_row_fwd(RowFwdArgs A, const float2* __restrict__ triW) {
  extern __shared__ float2 z[];
  row_fwd_body<LOG2N, NVEC>(z, xcd_remap(blockIdx.x, gridDim.x), A, triW);
}

// The forward row launch of one image on plan width 2^LOG2N.
template <int LOG2N>
static RowFwdArgs row_fwd_args(const uint8_t* img, int64_t istride, int step, int cnMax, const CallState* st,
                               const TileGrid& g, float2* spec, int64_t tileStride, int64_t chanStride, int ntiles) {
  RowFwdArgs a;
  a.img = img;
  a.istride = istride;
  a.step = step;
  a.st = st;
  a.g = g;
  a.spec = spec;
  a.tileStride = tileStride;
  a.chanStride = chanStride;
  a.cpg = row_cpg(1 << LOG2N, cnMax);
  a.ngroups = (cnMax + a.cpg - 1) / a.cpg;
  a.npairs = g.dftH / 2;
  a.words = (step == 4 && ((((uintptr_t)img) | (uintptr_t)istride) & 3) == 0) ? 1 : 0;
  a.cnAll = cnMax;
  a.t0 = 0;
  a.nwg = a.ngroups * a.npairs * ntiles;
  return a;
}

template <int LOG2N>
static void row_fwd_n(const uint8_t* img, int64_t istride, int step, int cnMax, const CallState* st,
                      const TileGrid& g, const PlanTables& pt, float2* spec, int64_t tileStride, int64_t chanStride,
                      int ntiles, hipStream_t s, int t0 = 0) {
  RowFwdArgs a = row_fwd_args<LOG2N>(img, istride, step, cnMax, st, g, spec, tileStride, chanStride, ntiles);
  a.t0 = t0;
  size_t lds = sizeof(float2) * (size_t)padded_len(1 << LOG2N) * a.cpg;
  dim3 grid(a.nwg), block(RowNTF<LOG2N>::value);
  if (a.cpg <= 1)
    hipLaunchKernelGGL((k_row_fwd<LOG2N, 1>), grid, block, lds, s, a, pt.triW);
  else if (a.cpg == 2)
    hipLaunchKernelGGL((k_row_fwd<LOG2N, 2>), grid, block, lds, s, a, pt.triW);
  else
    hipLaunchKernelGGL((k_row_fwd<LOG2N, 4>), grid, block, lds, s, a, pt.triW);
}

void launch_row_fwd(const uint8_t* img, int64_t istride, int step, int cnMax, const CallState* st,
                    const TileGrid& g, const PlanTables& pt, float2* spec, int64_t tileStride, int64_t chanStride,
                    int ntiles, hipStream_t s) {
#define CVM_RF(L) \
  case L:         \
    row_fwd_n<L>(img, istride, step, cnMax, st, g, pt, spec, tileStride, chanStride, ntiles, s); \
    break
  switch (g.log2W) {
    CVM_RF(1); CVM_RF(2); CVM_RF(3); CVM_RF(4); CVM_RF(5); CVM_RF(6);
    CVM_RF(7); CVM_RF(8); CVM_RF(9); CVM_RF(10); CVM_RF(11); CVM_RF(12);
    default: break;
  }
#undef CVM_RF
}

// Tiles [t0, t0 + ntiles) only.
void launch_row_fwd_range(const uint8_t* img, int64_t istride, int step, int cnMax, const CallState* st,
                          const TileGrid& g, const PlanTables& pt, float2* spec, int64_t tileStride,
                          int64_t chanStride, int t0, int ntiles, hipStream_t s) {
#define CVM_RF(L) \
  case L:         \
    row_fwd_n<L>(img, istride, step, cnMax, st, g, pt, spec, tileStride, chanStride, ntiles, s, t0); \
    break
  switch (g.log2W) {
    CVM_RF(1); CVM_RF(2); CVM_RF(3); CVM_RF(4); CVM_RF(5); CVM_RF(6);
    CVM_RF(7); CVM_RF(8); CVM_RF(9); CVM_RF(10); CVM_RF(11); CVM_RF(12);
    default: break;
  }
#undef CVM_RF
}

// ----------------------------------------------------------- column pass --

// Columns per workgroup for a column transform of 2^LOG2N points: the LDS
// slab n x CG complex stays at 64 KiB.
template <int LOG2N>
struct ColCG {
  static constexpr int value = LOG2N <= 9 ? 16 : (1 << (13 - LOG2N));
};

// Pass schedule of a column transform.  The first pass takes the radix-2
// head (odd sizes) and up to two radix-4 stages (32 or 16 points per group;
// NST1 >= 0 caps them, for a workgroup with more threads than groups of
// that size); every later pass two stages, the last one one or two.
template <int LOG2N, int NST1 = -1>
struct ColFirst {
  static constexpr int HEAD = LOG2N & 1;
  static constexpr int NSTD = (LOG2N - HEAD) / 2 >= 2 ? 2 : (LOG2N - HEAD) / 2;
  static constexpr int NST = (NST1 >= 0 && NST1 < NSTD) ? NST1 : NSTD;
  static constexpr int NB = HEAD + 2 * NST;
};

// Lowest index bit of the last pass when the passes after the first start
// at bit lo (remaining bit counts are even).
__host__ __device__ constexpr int col_last_lo(int log2n, int lo) {
  return log2n - lo <= 4 ? lo : col_last_lo(log2n, lo + 4);
}

// The two-stage passes from bit LO up to (not including) the last pass, in
// place in the LDS slab, each followed by a barrier.
template <int LOG2N, int LO, int LOL, int CG, int NT = 256>
__device__ __forceinline__ void col_mid(float2* a, const float2* __restrict__ tri, bool inverse) {
  if constexpr (LO < LOL) {
    fft_pass<LOG2N, LO, 0, 2, CG, NT, CG, false, true>(a, CG, 1, tri, inverse);
    __syncthreads();
    col_mid<LOG2N, LO + 4, LOL, CG, NT>(a, tri, inverse);
  }
}

// The inverse column transform of the LDS slab (input at the bit-reversed
// rows, as the products are stored) and the store of its first `need`
// spatial rows: the passes in LDS, the last one in registers straight to the
// spectrum rows.  The caller has synchronised after filling the slab.
__device__ __forceinline__ void store_f2(float2* p, float2 v) { *p = v; }
__device__ __forceinline__ void store_f2(float __attribute__((ext_vector_type(2))) __attribute__((address_space(1)))* p,
                                         float2 v) {
  float __attribute__((ext_vector_type(2))) u;
  u.x = v.x;
  u.y = v.y;
  *p = u;
}

template <int LOG2N, int NT, int NST1, typename DstPtr, int CGO = 0>
__device__ __forceinline__ void col_inverse_store(float2* a, const float2* __restrict__ tri, DstPtr dp, int hwPad,
                                                  int ncol, int need) {
  constexpr int CG = CGO ? CGO : ColCG<LOG2N>::value, LCG = ceil_log2c(CG);
  using F = ColFirst<LOG2N, NST1>;
  constexpr int NB1 = F::NB;
  constexpr bool SINGLE = NB1 == LOG2N;
  constexpr int LOL = SINGLE ? 0 : col_last_lo(LOG2N, NB1);
  constexpr int HEADL = SINGLE ? F::HEAD : 0;
  constexpr int NSTL = SINGLE ? F::NST : (LOG2N - LOL) / 2;
  constexpr int NBL = HEADL + 2 * NSTL, EL = 1 << NBL;
  constexpr int GL = CG << (LOG2N - NBL), ITL = (GL + NT - 1) / NT;
  const int tid = threadIdx.x;
  if constexpr (!SINGLE) {
    fft_pass<LOG2N, 0, F::HEAD, F::NST, CG, NT, CG, false, true>(a, CG, 1, tri, true);
    __syncthreads();
    col_mid<LOG2N, NB1, LOL, CG, NT>(a, tri, true);
  }
#pragma unroll
  for (int it = 0; it < ITL; it++) {
    const int gi = tid + it * NT, v = gi & (CG - 1), gg = gi >> LCG;
    const int base = grp_base<LOL, NBL>(gg);
    if (GL % NT == 0 || gi < GL) {
      float2 r[EL];
#pragma unroll
      for (int e = 0; e < EL; e++) r[e] = a[(base | (e << LOL)) * CG + v];
      bfly_regs<LOG2N, LOL, HEADL, NSTL>(r, gg & ((1 << LOL) - 1), tri, true);
      if (v < ncol)
#pragma unroll
        for (int e = 0; e < EL; e++) {
          const int y = base | (e << LOL);
          if (y < need) store_f2(dp + (int64_t)y * hwPad + v, r[e]);
        }
    }
  }
}

// One workgroup owns CG spectrum columns of one tile/channel for the whole
// column height; threads take (group, column) pairs with the column fastest,
// so each row's CG columns are one contiguous access.  Modes:
//   COL_TSPEC : rows (pad rows = 0) -> forward -> scaled, all rows stored
//   COL_FWD   : rows (pad rows = 0) -> forward, all rows stored
//   COL_CORR  : rows -> forward -> * conj(tspec) -> inverse, needed rows
//   COL_MULINV: forward rows -> * conj(tspec) -> inverse, needed rows
//   COL_CORRT : as COL_CORR, the template spectrum's slab first computed in
//               the workgroup (COL_TSPEC's values, kept in registers) from
//               the template's row spectra trows on its grid *gtp
// The forward input in bit-reversed row order (blockForwardGo
// impl.go:350-374) is read straight from the spectrum by the first pass:
// transform index i is source row brev(i).  The product for frequency k
// (mulConjGo impl.go:376-388) goes to slab row brev(k), the bit-reversed
// input of the inverse (the rmap = brev relabelling of colsR4Go), whose last
// pass writes spatial rows straight to the spectrum.  The template spectrum
// values of the last forward pass are loaded before the first pass starts.
// Logical workgroup order is tile-fastest, so the tiles that read one
// template-spectrum slab run together on one XCD.  A body covers tiles
// [t0, t0 + ntiles), all of column height 2^LOG2N; L is its logical
// workgroup index.
// NT threads per workgroup; NST1: the first pass's radix-4 stages (ColFirst).
template <int LOG2N, int MODE, int NT = 256, int NST1 = -1, int CGO = 0>
__device__ __forceinline__ void cols_body(float2* a, int L, const float2* __restrict__ src,
                                          float2* __restrict__ dst, int64_t srcTileStride, int64_t dstTileStride,
                                          int64_t chanStride, const CallState* __restrict__ st, const TileGrid& g,
                                          const float2* __restrict__ tri, const float2* __restrict__ tspec,
                                          int64_t tchanStride, float scale, int t0, int ntiles, int nslabs,
                                          const TileGrid* gtp = nullptr, const float2* __restrict__ trows = nullptr) {
  constexpr int n = 1 << LOG2N, CG = CGO ? CGO : ColCG<LOG2N>::value, LCG = ceil_log2c(CG);
  constexpr bool CORR = MODE == COL_CORR || MODE == COL_CORRT;
  using F = ColFirst<LOG2N, NST1>;
  constexpr int NB1 = F::NB, E1 = 1 << NB1;
  constexpr int G1 = CG << (LOG2N - NB1), IT1 = (G1 + NT - 1) / NT;
  constexpr bool SINGLE = NB1 == LOG2N;
  constexpr int LOL = SINGLE ? 0 : col_last_lo(LOG2N, NB1);
  constexpr int HEADL = SINGLE ? F::HEAD : 0;
  constexpr int NSTL = SINGLE ? F::NST : (LOG2N - LOL) / 2;
  constexpr int NBL = HEADL + 2 * NSTL, EL = 1 << NBL;
  constexpr int GL = CG << (LOG2N - NBL), ITL = (GL + NT - 1) / NT;
  const int t = t0 + L % ntiles, rest = L / ntiles;
  const int slab = rest % nslabs, k = rest / nslabs;
  const int c0 = slab * CG;
  // st null: every channel (a template spectrum built before its call state)
  if ((st && (st->flat || k >= st->cn)) || c0 >= g.hw) return;
  int x0, y0, bw, bh;
  tile_rect(g, t, x0, y0, bw, bh);
  const float2* ts = tspec + (int64_t)k * tchanStride + c0;
  const int tid = threadIdx.x;
  const int hwPad = g.hwPad;
  const float2* sp = src + (int64_t)t * srcTileStride + (int64_t)k * chanStride + c0;
  float2* dp = dst + (int64_t)t * dstTileStride + (int64_t)k * chanStride + c0;
  const int ncol = (g.hw - c0) < CG ? (g.hw - c0) : CG;
  int need = (bh + 1) & ~1;
  if (need > n) need = n;

  // The forward transform of the slab whose first row is at spp (rows from
  // `loaded` on read as zero), left in registers: out[it][e] is index
  // grp_base(gg) | (e << LOL) of this thread's last-pass group.
  auto forward = [&](const float2* __restrict__ spp, int loaded, float2 (&out)[ITL][EL]) {
    // first forward pass straight from the spectrum rows
    float2 r1[IT1][E1];
#pragma unroll
    for (int it = 0; it < IT1; it++) {
      const int gi = tid + it * NT, v = gi & (CG - 1), gg = gi >> LCG;
      const bool ok = (G1 % NT == 0 || gi < G1) && v < ncol;
#pragma unroll
      for (int e = 0; e < E1; e++) {
        const int row = bitrev_idx((gg << NB1) | e, LOG2N);
        r1[it][e] = make_float2(0.f, 0.f);
        if (ok && row < loaded) r1[it][e] = spp[(int64_t)row * hwPad + v];
      }
    }
#pragma unroll
    for (int it = 0; it < IT1; it++) bfly_regs<LOG2N, 0, F::HEAD, F::NST>(r1[it], 0, tri, false);
    if constexpr (SINGLE) {
#pragma unroll
      for (int it = 0; it < ITL; it++)
#pragma unroll
        for (int e = 0; e < EL; e++) out[it][e] = r1[it][e];
    } else {
#pragma unroll
      for (int it = 0; it < IT1; it++) {
        const int gi = tid + it * NT, v = gi & (CG - 1), gg = gi >> LCG;
        if (G1 % NT == 0 || gi < G1)
#pragma unroll
          for (int e = 0; e < E1; e++) a[((gg << NB1) | e) * CG + v] = r1[it][e];
      }
      __syncthreads();
      col_mid<LOG2N, NB1, LOL, CG, NT>(a, tri, false);
#pragma unroll
      for (int it = 0; it < ITL; it++) {
        const int gi = tid + it * NT, v = gi & (CG - 1), gg = gi >> LCG;
        const int base = grp_base<LOL, NBL>(gg);
        if (GL % NT == 0 || gi < GL) {
#pragma unroll
          for (int e = 0; e < EL; e++) out[it][e] = a[(base | (e << LOL)) * CG + v];
          bfly_regs<LOG2N, LOL, 0, NSTL>(out[it], gg & ((1 << LOL) - 1), tri, false);
        }
      }
    }
  };

  // template spectrum for the elements this thread owns in the last
  // forward pass
  float2 tsv[CORR ? ITL : 1][CORR ? EL : 1];
  if constexpr (MODE == COL_CORR) {
#pragma unroll
    for (int it = 0; it < ITL; it++) {
      const int gi = tid + it * NT, v = gi & (CG - 1), gg = gi >> LCG;
      const int base = grp_base<LOL, NBL>(gg);
#pragma unroll
      for (int e = 0; e < EL; e++) {
        const int idx = base | (e << LOL);
        tsv[it][e] = make_float2(0.f, 0.f);
        if ((GL % NT == 0 || gi < GL) && v < ncol) tsv[it][e] = ts[(int64_t)idx * hwPad + v];
      }
    }
  }
  if constexpr (MODE == COL_CORRT) {
    // the template's slab of channel k: its spectrum values scaled as
    // COL_TSPEC stores them; the slab in LDS is then free again
    int tx0, ty0, tbw, tbh;
    tile_rect(*gtp, 0, tx0, ty0, tbw, tbh);
    int tl = (tbh + gtp->padH + 1) & ~1;
    if (tl > n) tl = n;
#pragma unroll
    for (int it = 0; it < ITL; it++)
#pragma unroll
      for (int e = 0; e < EL; e++) tsv[it][e] = make_float2(0.f, 0.f);
    forward(trows + (int64_t)k * tchanStride + c0, tl, tsv);
#pragma unroll
    for (int it = 0; it < ITL; it++)
#pragma unroll
      for (int e = 0; e < EL; e++) tsv[it][e] = make_float2(tsv[it][e].x * scale, tsv[it][e].y * scale);
    __syncthreads();
  }

  float2 rl[ITL][EL];
  if constexpr (MODE != COL_MULINV) {
    int loaded = (bh + g.padH + 1) & ~1;
    if (loaded > n) loaded = n;
    forward(sp, loaded, rl);
    if constexpr (MODE == COL_TSPEC || MODE == COL_FWD) {
#pragma unroll
      for (int it = 0; it < ITL; it++) {
        const int gi = tid + it * NT, v = gi & (CG - 1), gg = gi >> LCG;
        const int base = grp_base<LOL, NBL>(gg);
        if ((GL % NT == 0 || gi < GL) && v < ncol)
#pragma unroll
          for (int e = 0; e < EL; e++) {
            float2 val = rl[it][e];
            if (MODE == COL_TSPEC) val = make_float2(val.x * scale, val.y * scale);
            dp[(int64_t)(base | (e << LOL)) * hwPad + v] = val;
          }
      }
      return;
    }
    // conjugate product at frequency idx, to slab row brev(idx) once every
    // thread has read its last-pass inputs
#pragma unroll
    for (int it = 0; it < ITL; it++)
#pragma unroll
      for (int e = 0; e < EL; e++) {
        const float2 x = rl[it][e], w = tsv[it][e];
        const float p0 = x.x * w.x, p1 = x.y * w.y, p2 = x.y * w.x, p3 = x.x * w.y;
        rl[it][e] = make_float2(p0 + p1, p2 - p3);
      }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITL; it++) {
      const int gi = tid + it * NT, v = gi & (CG - 1), gg = gi >> LCG;
      const int base = grp_base<LOL, NBL>(gg);
      if (GL % NT == 0 || gi < GL)
#pragma unroll
        for (int e = 0; e < EL; e++) a[bitrev_idx(base | (e << LOL), LOG2N) * CG + v] = rl[it][e];
    }
  } else {
    // forward rows of a fleet frame: product straight from global memory
    constexpr int TOT = n * CG, PER = (TOT + NT - 1) / NT, LB = PER < 16 ? PER : 16;
#pragma unroll
    for (int b0 = 0; b0 < PER; b0 += LB) {
      float2 x[LB], w[LB];
#pragma unroll
      for (int u = 0; u < LB; u++) {
        const int idx = (b0 + u) * NT + tid, r = idx >> LCG, c = idx & (CG - 1);
        x[u] = w[u] = make_float2(0.f, 0.f);
        if ((TOT % NT == 0 || idx < TOT) && c < ncol) {
          x[u] = sp[(int64_t)r * hwPad + c];
          w[u] = ts[(int64_t)r * hwPad + c];
        }
      }
#pragma unroll
      for (int u = 0; u < LB; u++) {
        const int idx = (b0 + u) * NT + tid, r = idx >> LCG, c = idx & (CG - 1);
        if (TOT % NT == 0 || idx < TOT) {
          const float p0 = x[u].x * w[u].x, p1 = x[u].y * w[u].y, p2 = x[u].y * w[u].x, p3 = x[u].x * w[u].y;
          a[bitrev_idx(r, LOG2N) * CG + c] = make_float2(p0 + p1, p2 - p3);
        }
      }
    }
  }
  __syncthreads();
  col_inverse_store<LOG2N, NT, NST1, float2*, CGO>(a, tri, dp, hwPad, ncol, need);
}

// Two members of one Fleet bucket against the same frame spectrum in one
// launch (COL_MULINV for both): each workgroup reads its frame slab once,
// keeps it in registers, and for member A then member B forms the product
// with that member's template spectrum in LDS, runs the inverse passes and
// stores that member's needed rows.  Members share the plan, tiles and slabs;
// each has its own call state (channel count), template spectrum, product
// spectra and tile extents.  Full-height tiles only (no shrunk band).
// Global-address-space views of the members' pointers: they arrive inside a
// by-value struct, and as generic (flat) pointers their loads and stores
// could alias the LDS slab and would be kept in order with its accesses.
typedef float gfv2 __attribute__((ext_vector_type(2)));
typedef const gfv2 __attribute__((address_space(1)))* gcf2p;
typedef gfv2 __attribute__((address_space(1)))* gf2p;

template <int LOG2N>
__global__ __launch_bounds__(256) void k_cols_pair(const float2* __restrict__ src, int64_t tileStride,
                                                   int64_t chanStride, TileGrid g, const float2* __restrict__ tri,
                                                   int64_t tchanStride, int ntiles, int nslabs, ColPairMember A,
                                                   ColPairMember B) {
  constexpr int n = 1 << LOG2N, CG = ColCG<LOG2N>::value, LCG = ceil_log2c(CG), NT = 256;
  constexpr int TOT = n * CG, PER = (TOT + NT - 1) / NT;
  extern __shared__ float2 a[];
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int t = L % ntiles, rest = L / ntiles;
  const int slab = rest % nslabs, k = rest / nslabs;
  const int c0 = slab * CG;
  const bool onA = !A.st->flat && k < A.st->cn, onB = !B.st->flat && k < B.st->cn;
  if (c0 >= g.hw || !(onA || onB)) return;
  const int tid = threadIdx.x, hwPad = g.hwPad;
  const int ncol = (g.hw - c0) < CG ? (g.hw - c0) : CG;
  const float2* sp = src + (int64_t)t * tileStride + (int64_t)k * chanStride + c0;
  // the frame slab, once: element u of this thread is row (u NT + tid) >> LCG
  float2 x[PER];
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int idx = u * NT + tid, r = idx >> LCG, c = idx & (CG - 1);
    x[u] = make_float2(0.f, 0.f);
    if ((TOT % NT == 0 || idx < TOT) && c < ncol) x[u] = sp[(int64_t)r * hwPad + c];
  }
  auto member = [&](const ColPairMember& M) {
    const gcf2p ts = (gcf2p)M.tspec + (int64_t)k * tchanStride + c0;
    // template values in batches of LB, all loads of a batch before its
    // products are stored (so they are in flight together)
    constexpr int LB = PER < 16 ? PER : 16;
#pragma unroll
    for (int b0 = 0; b0 < PER; b0 += LB) {
      float2 w[LB];
#pragma unroll
      for (int u = 0; u < LB; u++) {
        const int idx = (b0 + u) * NT + tid, r = idx >> LCG, c = idx & (CG - 1);
        w[u] = make_float2(0.f, 0.f);
        if ((TOT % NT == 0 || idx < TOT) && c < ncol) {
          const gfv2 tv = ts[(int64_t)r * hwPad + c];
          w[u] = make_float2(tv.x, tv.y);
        }
      }
#pragma unroll
      for (int u = 0; u < LB; u++) {
        const int idx = (b0 + u) * NT + tid, r = idx >> LCG, c = idx & (CG - 1);
        if (TOT % NT == 0 || idx < TOT) {
          const float2 xv = x[b0 + u];
          const float p0 = xv.x * w[u].x, p1 = xv.y * w[u].y, p2 = xv.y * w[u].x, p3 = xv.x * w[u].y;
          a[bitrev_idx(r, LOG2N) * CG + c] = make_float2(p0 + p1, p2 - p3);
        }
      }
    }
    __syncthreads();
    TileGrid gm = g;
    gm.extW = M.extW;
    gm.extH = M.extH;
    int x0, y0, bw, bh;
    tile_rect(gm, t, x0, y0, bw, bh);
    int need = (bh + 1) & ~1;
    if (need > n) need = n;
    const gf2p dp = (gf2p)M.dst + (int64_t)t * tileStride + (int64_t)k * chanStride + c0;
    col_inverse_store<LOG2N, 256, -1>(a, tri, dp, hwPad, ncol, need);
  };
  if (onA) member(A);
  if (onA && onB) __syncthreads();  // the slab is rewritten for B
  if (onB) member(B);
}

// One band of tiles of a column launch: its twiddles, template spectrum and
// tile range, and the workgroups given to it (a multiple of 8 when another
// band follows it in the launch, so the next band's first workgroup starts
// on XCD 0; workgroups past the band's real count find no channel and
// return).
struct ColBand {
  const float2* tri;
  const float2* tspec;
  int64_t tchanStride;
  int t0, ntiles, nwg;
};

template <int LOG2N, int MODE>
__global__ __launch_bounds__(256) void k_cols(const float2* __restrict__ src, float2* __restrict__ dst,
                                              int64_t srcTileStride, int64_t dstTileStride, int64_t chanStride,
                                              const CallState* __restrict__ st, TileGrid g, float scale, int nslabs,
                                              ColBand b) {
  extern __shared__ float2 a[];
  cols_body<LOG2N, MODE>(a, xcd_remap(blockIdx.x, gridDim.x), src, dst, srcTileStride, dstTileStride, chanStride,
                         st, g, b.tri, b.tspec, b.tchanStride, scale, b.t0, b.ntiles, nslabs);
}

// Full-height tiles and the shrunk last band (column height 2^LOG2N2) in
// one launch: the band's workgroups come after every full-height one, spread
// over all XCDs, and fill the tail of the full-height work.
template <int LOG2N, int LOG2N2, int MODE>
__global__ __launch_bounds__(256) void k_cols2(const float2* __restrict__ src, float2* __restrict__ dst,
                                               int64_t srcTileStride, int64_t dstTileStride, int64_t chanStride,
                                               const CallState* __restrict__ st, TileGrid g, float scale, int nslabs,
                                               ColBand b1, ColBand b2) {
  extern __shared__ float2 a[];
  const int p = blockIdx.x;
  if (p < b1.nwg)
    cols_body<LOG2N, MODE>(a, xcd_remap(p, b1.nwg), src, dst, srcTileStride, dstTileStride, chanStride, st, g,
                           b1.tri, b1.tspec, b1.tchanStride, scale, b1.t0, b1.ntiles, nslabs);
  else
    cols_body<LOG2N2, MODE>(a, xcd_remap(p - b1.nwg, b2.nwg), src, dst, srcTileStride, dstTileStride, chanStride,
                            st, g, b2.tri, b2.tspec, b2.tchanStride, scale, b2.t0, b2.ntiles, nslabs);
}

// The same with 512 threads a workgroup and one radix-4 stage in the first
// pass (8 or 4 points a group, so every thread has groups): half the
// registers a thread and twice the waves a workgroup.
template <int LOG2N, int LOG2N2, int MODE>
__global__ __launch_bounds__(512) void k_cols2w(const float2* __restrict__ src, float2* __restrict__ dst,
                                                int64_t srcTileStride, int64_t dstTileStride, int64_t chanStride,
                                                const CallState* __restrict__ st, TileGrid g, float scale, int nslabs,
                                                ColBand b1, ColBand b2) {
  extern __shared__ float2 a[];
  const int p = blockIdx.x;
  if (p < b1.nwg)
    cols_body<LOG2N, MODE, 512, 1>(a, xcd_remap(p, b1.nwg), src, dst, srcTileStride, dstTileStride, chanStride, st,
                                   g, b1.tri, b1.tspec, b1.tchanStride, scale, b1.t0, b1.ntiles, nslabs);
  else
    cols_body<LOG2N2, MODE, 512, 1>(a, xcd_remap(p - b1.nwg, b2.nwg), src, dst, srcTileStride, dstTileStride,
                                    chanStride, st, g, b2.tri, b2.tspec, b2.tchanStride, scale, b2.t0, b2.ntiles,
                                    nslabs);
}

template <int LOG2N>
static void cols_n(ColMode mode, const float2* src, float2* dst, int64_t srcTileStride, int64_t dstTileStride,
                   int64_t chanStride, int cnMax, const CallState* st, const TileGrid& g, const float2* tri,
                   const float2* tspec, int64_t tchanStride, float scale, int t0, int ntiles, hipStream_t s) {
  constexpr int CG = ColCG<LOG2N>::value;
  size_t lds = sizeof(float2) * (size_t)(1 << LOG2N) * CG;
  int nslabs = (g.hw + CG - 1) / CG;
  dim3 grid(nslabs * ntiles * cnMax), block(256);
  ColBand b{tri, tspec, tchanStride, t0, ntiles, (int)grid.x};
#define CVM_CM(M)                                                                                                \
  hipLaunchKernelGGL((k_cols<LOG2N, M>), grid, block, lds, s, src, dst, srcTileStride, dstTileStride, chanStride, \
                     st, g, scale, nslabs, b)
  switch (mode) {
    case COL_TSPEC: CVM_CM(COL_TSPEC); break;
    case COL_FWD: CVM_CM(COL_FWD); break;
    case COL_CORR: CVM_CM(COL_CORR); break;
    case COL_MULINV: CVM_CM(COL_MULINV); break;
  }
#undef CVM_CM
}

// Both bands in one launch (the hot modes; shrunk band one or two sizes
// below the full height).  Returns false when no such kernel exists.
template <int LOG2N, int LOG2N2>
static bool cols2_n(ColMode mode, const float2* src, float2* dst, int64_t srcTileStride, int64_t dstTileStride,
                    int64_t chanStride, int cnMax, const CallState* st, const TileGrid& g, ColBand b1, ColBand b2,
                    float scale, hipStream_t s) {
  constexpr int CG = ColCG<LOG2N>::value;
  static_assert(ColCG<LOG2N2>::value == CG, "both bands share the column slabs");
  size_t lds = sizeof(float2) * (size_t)(1 << LOG2N) * CG;
  int nslabs = (g.hw + CG - 1) / CG;
  b1.nwg = (nslabs * b1.ntiles * cnMax + 7) & ~7;
  b2.nwg = nslabs * b2.ntiles * cnMax;
  dim3 grid(b1.nwg + b2.nwg), block(256);
  if constexpr (LOG2N == 9) {
    // 512-row columns: the 512-thread form (4 waves a SIMD instead of 2;
    // 4K RGBA column pass 111.5 -> 104.8 us); CVMATCH_COLS_WIDE=0 keeps
    // the 256-thread one
    const char* e = getenv("CVMATCH_COLS_WIDE");
    if (!(e && e[0] == '0') && (mode == COL_CORR || mode == COL_MULINV)) {
      if (mode == COL_CORR)
        hipLaunchKernelGGL((k_cols2w<LOG2N, LOG2N2, COL_CORR>), grid, dim3(512), lds, s, src, dst, srcTileStride,
                           dstTileStride, chanStride, st, g, scale, nslabs, b1, b2);
      else
        hipLaunchKernelGGL((k_cols2w<LOG2N, LOG2N2, COL_MULINV>), grid, dim3(512), lds, s, src, dst,
                           srcTileStride, dstTileStride, chanStride, st, g, scale, nslabs, b1, b2);
      return true;
    }
  }
  if (mode == COL_CORR)
    hipLaunchKernelGGL((k_cols2<LOG2N, LOG2N2, COL_CORR>), grid, block, lds, s, src, dst, srcTileStride,
                       dstTileStride, chanStride, st, g, scale, nslabs, b1, b2);
  else if (mode == COL_MULINV)
    hipLaunchKernelGGL((k_cols2<LOG2N, LOG2N2, COL_MULINV>), grid, block, lds, s, src, dst, srcTileStride,
                       dstTileStride, chanStride, st, g, scale, nslabs, b1, b2);
  else
    return false;
  return true;
}

static void cols_dispatch(int log2n, ColMode mode, const float2* src, float2* dst, int64_t srcTileStride,
                          int64_t dstTileStride, int64_t chanStride, int cnMax, const CallState* st,
                          const TileGrid& g, const float2* tri, const float2* tspec, int64_t tchanStride, float scale,
                          int t0, int ntiles, hipStream_t s) {
#define CVM_CN(L)                                                                                         \
  case L:                                                                                                 \
    cols_n<L>(mode, src, dst, srcTileStride, dstTileStride, chanStride, cnMax, st, g, tri, tspec, tchanStride, \
              scale, t0, ntiles, s);                                                                      \
    break
  switch (log2n) {
    CVM_CN(1); CVM_CN(2); CVM_CN(3); CVM_CN(4); CVM_CN(5); CVM_CN(6); CVM_CN(7);
    CVM_CN(8); CVM_CN(9); CVM_CN(10); CVM_CN(11); CVM_CN(12); CVM_CN(13);
    default: break;
  }
#undef CVM_CN
}

void launch_cols(ColMode mode, const float2* src, float2* dst, int64_t srcTileStride, int64_t dstTileStride,
                 int64_t chanStride, int cnMax, const CallState* st, const TileGrid& g, const PlanTables& pt,
                 const float2* tspec, const float2* tspec2, int64_t tchanStride, int64_t tchanStride2, float scale,
                 int ntiles, hipStream_t s) {
  // tiles of the shrunk last band (the last ntx tiles) run at height dftH2
  int nshr = g.dftH2 ? g.ntx : 0;
  if (nshr > ntiles) nshr = 0;
  int nfull = ntiles - nshr;
  if (nfull > 0 && nshr > 0 && (mode == COL_CORR || mode == COL_MULINV)) {
    ColBand b1{pt.triH, tspec, tchanStride, 0, nfull, 0}, b2{pt.triH2, tspec2, tchanStride2, nfull, nshr, 0};
    bool done = false;
#define CVM_C2(L, L2)                                                                                             \
  if (g.log2H == L && g.log2H2 == L2)                                                                             \
  done = cols2_n<L, L2>(mode, src, dst, srcTileStride, dstTileStride, chanStride, cnMax, st, g, b1, b2, scale, s)
    CVM_C2(9, 8); CVM_C2(9, 7); CVM_C2(8, 7); CVM_C2(8, 6); CVM_C2(7, 6); CVM_C2(7, 5); CVM_C2(6, 5);
    CVM_C2(6, 4);
#undef CVM_C2
    if (done) return;
  }
  if (nfull > 0)
    cols_dispatch(g.log2H, mode, src, dst, srcTileStride, dstTileStride, chanStride, cnMax, st, g, pt.triH, tspec,
                  tchanStride, scale, 0, nfull, s);
  if (nshr > 0)
    cols_dispatch(g.log2H2, mode, src, dst, srcTileStride, dstTileStride, chanStride, cnMax, st, g, pt.triH2, tspec2,
                  tchanStride2, scale, nfull, nshr, s);
}

// Tiles [t0, t0 + ntiles) of one tile row only: the full-height transform,
// or the shrunk band's for the last row when there is one.
void launch_cols_range(ColMode mode, const float2* src, float2* dst, int64_t tileStride, int64_t chanStride,
                       int cnMax, const CallState* st, const TileGrid& g, const PlanTables& pt, const float2* tspec,
                       const float2* tspec2, int64_t tchanStride, int64_t tchanStride2, int t0, int ntiles,
                       hipStream_t s) {
  const bool shrunk = g.dftH2 != 0 && (t0 / g.ntx) == g.nty - 1;
  if (shrunk)
    cols_dispatch(g.log2H2, mode, src, dst, tileStride, tileStride, chanStride, cnMax, st, g, pt.triH2, tspec2,
                  tchanStride2, 1.0f, t0, ntiles, s);
  else
    cols_dispatch(g.log2H, mode, src, dst, tileStride, tileStride, chanStride, cnMax, st, g, pt.triH, tspec,
                  tchanStride, 1.0f, t0, ntiles, s);
}

template <int LOG2N>
static void cols_pair_n(const float2* src, int64_t tileStride, int64_t chanStride, int cnMax, const TileGrid& g,
                        const float2* tri, int64_t tchanStride, int ntiles, const ColPairMember& A,
                        const ColPairMember& B, hipStream_t s) {
  constexpr int CG = ColCG<LOG2N>::value;
  const size_t lds = sizeof(float2) * (size_t)(1 << LOG2N) * CG;
  const int nslabs = (g.hw + CG - 1) / CG;
  hipLaunchKernelGGL(k_cols_pair<LOG2N>, dim3(nslabs * ntiles * cnMax), dim3(256), lds, s, src, tileStride,
                     chanStride, g, tri, tchanStride, ntiles, nslabs, A, B);
}

bool launch_cols_pair(const float2* src, int64_t tileStride, int64_t chanStride, int cnMax, const TileGrid& g,
                      const PlanTables& pt, int64_t tchanStride, int ntiles, const ColPairMember& A,
                      const ColPairMember& B, hipStream_t s) {
  if (g.dftH2 != 0) return false;  // full-height tiles only
#define CVM_CP(L)                                                                                   \
  case L:                                                                                           \
    cols_pair_n<L>(src, tileStride, chanStride, cnMax, g, pt.triH, tchanStride, ntiles, A, B, s); \
    return true
  switch (g.log2H) {
    CVM_CP(6); CVM_CP(7); CVM_CP(8); CVM_CP(9); CVM_CP(10); CVM_CP(11);
    default: return false;
  }
#undef CVM_CP
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
// channel 0 stored and channels 1.. added in order (in registers; each
// thread owns the same columns for every channel).
// One inverse row launch's arguments.
struct RowInvArgs {
  const float2* spec;
  int64_t tileStride, chanStride;
  const CallState* st;
  TileGrid g;
  float* corr;
  int rw, cpg, npairs;
  int nwg;  // workgroups: npairs * ntiles
};

// Workgroup L of an inverse row launch; z: the LDS vectors.
template <int LOG2N, int NVEC>
__device__ __forceinline__ void row_inv_body(float2* z, int L, const RowInvArgs& A, const float2* __restrict__ triW) {
  constexpr int n = 1 << LOG2N, NT = RowNTI<LOG2N>::value, np = n + (n >> 5);
  constexpr int MAXI = (n + NT - 1) / NT, H = n / 2, MAXH = (H + NT - 1) / NT;
  const float2* __restrict__ spec = A.spec;
  const int64_t tileStride = A.tileStride, chanStride = A.chanStride;
  const CallState* __restrict__ st = A.st;
  const TileGrid& g = A.g;
  float* __restrict__ corr = A.corr;
  const int rw = A.rw, cpg = A.cpg, npairs = A.npairs;
  const int rp = L % npairs, t = L / npairs;
  const int r = 2 * rp;
  if (st->flat) return;
  int x0, y0, bw, bh;
  tile_rect(g, t, x0, y0, bw, bh);
  if (r >= bh) return;
  const int cn = st->cn, tid = threadIdx.x;
  float2 acc[MAXI];
#pragma unroll
  for (int i = 0; i < MAXI; i++) acc[i] = make_float2(0.f, 0.f);
  for (int k0 = 0; k0 < cn; k0 += cpg) {
    const int nv = min(cpg, cn - k0);
#pragma unroll
    for (int v = 0; v < NVEC; v++) {
      if (v >= nv) break;
      const float2* sa = spec + (int64_t)t * tileStride + (int64_t)(k0 + v) * chanStride + (int64_t)r * g.hwPad;
      const float2* sb = sa + g.hwPad;
      float2* zv = z + v * np;
      // spectrum column m < n/2 feeds output bins m and n - m (the
      // conjugate-symmetric half), column n/2 bin n/2 alone: each column
      // is loaded once
      float2 A[MAXH], B[MAXH], Ah, Bh;
#pragma unroll
      for (int i = 0; i < MAXH; i++) {
        const int m = tid + i * NT;
        A[i] = B[i] = make_float2(0.f, 0.f);
        if (H % NT == 0 || m < H) {
          A[i] = sa[m];
          B[i] = sb[m];
        }
      }
      if (tid == 0) {
        Ah = sa[H];
        Bh = sb[H];
      }
#pragma unroll
      for (int i = 0; i < MAXH; i++) {
        const int m = tid + i * NT;
        if (H % NT == 0 || m < H) {
          zv[padi(bitrev_idx(m, LOG2N))] = make_float2(A[i].x - B[i].y, A[i].y + B[i].x);
          if (m > 0) zv[padi(bitrev_idx(n - m, LOG2N))] = make_float2(A[i].x + B[i].y, B[i].x - A[i].y);
        }
      }
      if (tid == 0) zv[padi(bitrev_idx(H, LOG2N))] = make_float2(Ah.x - Bh.y, Ah.y + Bh.x);
    }
    __syncthreads();
    fft_lds<LOG2N, NVEC, NT, 1, true, false>(z, nv, np, triW, true);
#pragma unroll
    for (int v = 0; v < NVEC; v++) {
      if (v >= nv) break;
      const float2* zv = z + v * np;
#pragma unroll
      for (int i = 0; i < MAXI; i++) {
        int x = tid + i * NT;
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
  for (int i = 0; i < MAXI; i++) {
    int x = tid + i * NT;
    if (x < bw) {
      o1[x] = acc[i].x;
      if (second) o2[x] = acc[i].y;
    }
  }
}

template <int LOG2N, int NVEC>
__global__ __launch_bounds__(RowNTI<LOG2N>::value) __attribute__((amdgpu_waves_per_eu(RowInvWaves<LOG2N>::value))) void k_row_inv(RowInvArgs A, const float2* __restrict__ triW) {
  extern __shared__ float2 z[];
  row_inv_body<LOG2N, NVEC>(z, blockIdx.x, A, triW);
}

template <int LOG2N>
static RowInvArgs row_inv_args(const float2* spec, int64_t tileStride, int64_t chanStride, const CallState* st,
                               int cnMax, const TileGrid& g, float* corr, int rw, int ntiles) {
  RowInvArgs a;
  a.spec = spec;
  a.tileStride = tileStride;
  a.chanStride = chanStride;
  a.st = st;
  a.g = g;
  a.corr = corr;
  a.rw = rw;
  a.cpg = row_cpg(1 << LOG2N, cnMax);
  // last-band tiles may be taller than blockH (fleet members): cover the
  // largest bh any tile can have
  int bhMax = g.blockH;
  int bhl = g.extH - (g.nty - 1) * g.blockH;
  if (bhl > bhMax) bhMax = bhl;
  a.npairs = (bhMax + 1) / 2;
  a.nwg = a.npairs * ntiles;
  return a;
}

template <int LOG2N>
static void row_inv_n(const float2* spec, int64_t tileStride, int64_t chanStride, const CallState* st, int cnMax,
                      const TileGrid& g, const PlanTables& pt, float* corr, int rw, int ntiles, hipStream_t s) {
  const RowInvArgs a = row_inv_args<LOG2N>(spec, tileStride, chanStride, st, cnMax, g, corr, rw, ntiles);
  size_t lds = sizeof(float2) * (size_t)padded_len(1 << LOG2N) * a.cpg;
  dim3 grid(a.nwg), block(RowNTI<LOG2N>::value);
  if (a.cpg <= 1)
    hipLaunchKernelGGL((k_row_inv<LOG2N, 1>), grid, block, lds, s, a, pt.triW);
  else if (a.cpg == 2)
    hipLaunchKernelGGL((k_row_inv<LOG2N, 2>), grid, block, lds, s, a, pt.triW);
  else
    hipLaunchKernelGGL((k_row_inv<LOG2N, 4>), grid, block, lds, s, a, pt.triW);
}

void launch_row_inv(const float2* spec, int64_t tileStride, int64_t chanStride, const CallState* st, int cnMax,
                    const TileGrid& g, const PlanTables& pt, float* corr, int rw, int ntiles, hipStream_t s) {
#define CVM_RI(L) \
  case L:         \
    row_inv_n<L>(spec, tileStride, chanStride, st, cnMax, g, pt, corr, rw, ntiles, s); \
    break
  switch (g.log2W) {
    CVM_RI(1); CVM_RI(2); CVM_RI(3); CVM_RI(4); CVM_RI(5); CVM_RI(6);
    CVM_RI(7); CVM_RI(8); CVM_RI(9); CVM_RI(10); CVM_RI(11); CVM_RI(12);
    default: break;
  }
#undef CVM_RI
}


// ------------------------------------------------- compact one-shot call --
//
// A small one-shot call runs in one stream, and its independent parts share
// launches instead of streams: a graph with a second stream costs the host
// two to three times as long to launch (most of a small call's time), while
// workgroups of different jobs in one grid run side by side just the same.
// Each launch gives its first workgroups to one job and the rest to the
// other; every job runs exactly the code of its own launch.

// One column launch's arguments (a band of tiles of one height).
struct ColArgs {
  const float2* src;
  float2* dst;
  int64_t srcTileStride, dstTileStride, chanStride;
  const CallState* st;
  TileGrid g;
  const float2* tri;
  const float2* tspec;
  int64_t tchanStride;
  float scale;
  int t0, ntiles, nslabs;
  int nwg;
  TileGrid gt;           // COL_CORRT: the template's grid
  const float2* trows;   // COL_CORRT: the template's row spectra
};

template <int LOG2N, int MODE, int NT = 256, int NST1 = -1, int CGO = 0>
__device__ __forceinline__ void cols_job(float2* a, int p, const ColArgs& C) {
  cols_body<LOG2N, MODE, NT, NST1, CGO>(a, xcd_remap(p, C.nwg), C.src, C.dst, C.srcTileStride, C.dstTileStride,
                                   C.chanStride, C.st, C.g, C.tri, C.tspec, C.tchanStride, C.scale, C.t0, C.ntiles,
                                   C.nslabs, &C.gt, C.trows);
}

// The chunk column sums (a nbx x nch grid of 256-thread workgroups).
struct ChunkArgs {
  const uint8_t* img;
  int64_t istride;
  int iw, ih, step, words;
  const CallState* st;
  uint32_t* ps;
  int nbx, nch;
};

// The chunk prefix sums (nwords words, one a thread).
struct PrefixArgs {
  int nwords, nch;
  const CallState* st;
  uint32_t* ps;
  int nwg;
};

// The parent's and the template's forward row transforms (every channel)
// and the call state's independent parts (template statistics, alpha
// scans), accumulated into a state whose accumulators the previous compact
// call's reduction left zero.
template <int LOG2N, int NVEC>
__global__ __launch_bounds__(RowNTF<LOG2N>::value) void k_front(RowFwdArgs P, RowFwdArgs T,
                                                                const float2* __restrict__ triW, CallStateArgs S) {
  extern __shared__ float2 z[];
  const int b = blockIdx.x;
  if (b < P.nwg)
    row_fwd_body<LOG2N, NVEC>(z, xcd_remap(b, P.nwg), P, triW);
  else if (b < P.nwg + T.nwg)
    row_fwd_body<LOG2N, NVEC>(z, b - P.nwg, T, triW);
  else
    call_state_block(b - P.nwg - T.nwg, S);
}

// The state's finalisation as one launch's prologue: each workgroup
// finalises a copy of the accumulated state in LDS and works from it, and
// workgroup 0 publishes the result fields for the later launches.
struct FinalizeArgs {
  CallState* st;
  int tw, th, cn_mode;
};

// The correlation's column pass (each workgroup first transforms the
// template's slab it multiplies by) and the chunk column sums, after the
// state's finalisation.
template <int LOG2N, int NT = 256, int NST1 = -1, int CGO = 0>
__global__ __launch_bounds__(NT) void k_corr_chunks(ColArgs C, ChunkArgs K, FinalizeArgs F) {
  extern __shared__ float2 a[];
  __shared__ CallState sst;
  if (threadIdx.x == 0) {
    sst = *F.st;
    state_finalize_fields(&sst, F.tw, F.th, F.cn_mode);
    if (blockIdx.x == 0) {
      F.st->cn = sst.cn;
      F.st->varSum = sst.varSum;
      F.st->flat = sst.flat;
      F.st->templNorm = sst.templNorm;
      F.st->varScale = sst.varScale;
      F.st->numScale = sst.numScale;
    }
  }
  __syncthreads();
  if ((int)blockIdx.x < C.nwg) {
    C.st = &sst;
    cols_job<LOG2N, COL_CORRT, NT, NST1, CGO>(a, blockIdx.x, C);
  } else {
    const int b = blockIdx.x - C.nwg;
    colsum_chunks_at((b % K.nbx) * NT + threadIdx.x, b / K.nbx, K.img, K.istride, K.iw, K.ih, K.step, K.words, &sst,
                     K.ps);
  }
}

// The inverse row pass and the chunk prefix sums.
template <int LOG2N, int NVEC>
__global__ __launch_bounds__(RowNTI<LOG2N>::value) __attribute__((amdgpu_waves_per_eu(RowInvWaves<LOG2N>::value))) void k_rowinv_prefix(RowInvArgs R, PrefixArgs Q,
                                                                        const float2* __restrict__ triW) {
  extern __shared__ float2 z[];
  if ((int)blockIdx.x < R.nwg)
    row_inv_body<LOG2N, NVEC>(z, blockIdx.x, R, triW);
  else
    colsum_prefix_block(blockIdx.x - R.nwg, Q.nwords, Q.nch, Q.st, Q.ps);
}

template <int LOG2W, int NVEC>
static void compact_rows(const CompactCall& c, const RowFwdArgs& P, const RowFwdArgs& T, const RowInvArgs& R,
                         const PrefixArgs& Q, bool front, hipStream_t s) {
  const size_t lds = sizeof(float2) * (size_t)padded_len(1 << LOG2W) * P.cpg;
  if (front) {
    // the template statistics in workgroups of this launch's size: about
    // eight pixels a thread, as with the 256-thread layout of the state
    CallStateArgs S = c.csa;
    constexpr int NT = RowNTF<LOG2W>::value;
    S.sgx = (int)std::min<long long>(256, ((long long)c.tw * c.th + 8 * NT - 1) / (8 * NT));
    hipLaunchKernelGGL((k_front<LOG2W, NVEC>), dim3(P.nwg + T.nwg + S.nb()), dim3(NT), lds, s, P, T, c.pt.triW,
                       S);
  } else {
    hipLaunchKernelGGL((k_rowinv_prefix<LOG2W, NVEC>), dim3(R.nwg + Q.nwg), dim3(RowNTI<LOG2W>::value), lds, s, R, Q,
                       c.pt.triW);
  }
}

template <int LOG2W>
static void compact_rows_w(const CompactCall& c, bool front, const PrefixArgs& Q, hipStream_t s) {
  const RowFwdArgs P = row_fwd_args<LOG2W>(c.img, c.istride, c.step, c.cnMax, nullptr, c.g, c.spec, c.tileStride,
                                           c.chanStride, c.ntiles);
  const RowFwdArgs T = row_fwd_args<LOG2W>(c.tpl, c.tstride, c.step, c.cnMax, nullptr, c.gt, c.tspec, 0, c.tchan, 1);
  const RowInvArgs R = row_inv_args<LOG2W>(c.spec, c.tileStride, c.chanStride, c.st, c.cnMax, c.g, c.corr, c.rw,
                                           c.ntiles);
  if (P.cpg <= 1)
    compact_rows<LOG2W, 1>(c, P, T, R, Q, front, s);
  else if (P.cpg == 2)
    compact_rows<LOG2W, 2>(c, P, T, R, Q, front, s);
  else
    compact_rows<LOG2W, 4>(c, P, T, R, Q, front, s);
}

static void compact_rows_dispatch(const CompactCall& c, bool front, const PrefixArgs& Q, hipStream_t s) {
#define CVM_CR(L)                          \
  case L:                                  \
    compact_rows_w<L>(c, front, Q, s);     \
    break
  switch (c.g.log2W) {
    CVM_CR(5); CVM_CR(6); CVM_CR(7); CVM_CR(8); CVM_CR(9); CVM_CR(10); CVM_CR(11);
    default: break;
  }
#undef CVM_CR
}

template <int LOG2H, int NT = 256, int NST1 = -1, int CGO = 0>
static void compact_cols(const CompactCall& c, ChunkArgs K, hipStream_t s) {
  const FinalizeArgs F{c.st, c.tw, c.th, c.cn_mode};
  K.nbx = (c.iw + NT - 1) / NT;
  constexpr int CG = CGO ? CGO : ColCG<LOG2H>::value;
  const size_t lds = sizeof(float2) * (size_t)(1 << LOG2H) * CG;
  const int nslabs = (c.g.hw + CG - 1) / CG;
  ColArgs C;
  C.tri = c.pt.triH;
  C.t0 = 0;
  C.nslabs = nslabs;
  C.chanStride = c.chanStride;
  C.src = c.spec;
  C.dst = c.spec;
  C.srcTileStride = C.dstTileStride = c.tileStride;
  C.st = c.st;
  C.g = c.g;
  // the template's row spectra; each workgroup transforms its slab, pad rows
  // zero, scaled
  C.gt = c.gt;
  C.trows = c.tspec;
  C.tspec = c.tspec;
  C.tchanStride = c.tchan;
  C.scale = c.scale;
  C.ntiles = c.ntiles;
  C.nwg = nslabs * c.ntiles * c.cnMax;
  hipLaunchKernelGGL((k_corr_chunks<LOG2H, NT, NST1, CGO>), dim3(C.nwg + K.nbx * K.nch), dim3(NT), lds, s, C, K, F);
}

static void compact_cols_dispatch(const CompactCall& c, const ChunkArgs& K, hipStream_t s) {
#define CVM_CC(L)              \
  case L:                      \
    compact_cols<L>(c, K, s);  \
    break
  // 512-row columns in the 512-thread form (one radix-4 stage in the first
  // pass: twice the waves a workgroup, each with half the points; 640x480
  // RGBA Match 0.075 -> 0.068 ms; 256-row columns measured slower so),
  // in slabs of 8 columns rather than 16: twice the workgroups, each with
  // half the points, and every thread still busy (the first pass has 8 x 64
  // groups): 640x480 Match 0.0656 -> 0.0615 ms.  (256-row columns in
  // 8-column slabs leave half of a 256-thread workgroup idle: 1080p
  // MatchGray 0.0746 -> 0.080 ms.)  CVMATCH_COMPACT_WIDE=0 keeps the
  // 256-thread form, CVMATCH_COMPACT_CG=16 the 16-column slabs.
  const char* e = getenv("CVMATCH_COMPACT_WIDE");
  const bool wide = !(e && e[0] == '0');
  const char* cge = getenv("CVMATCH_COMPACT_CG");
  const bool cg8 = !(cge && atoi(cge) == 16);
  switch (c.g.log2H) {
    CVM_CC(5); CVM_CC(6); CVM_CC(7); CVM_CC(8); CVM_CC(10); CVM_CC(11);
    case 9:
      if (wide && cg8)
        compact_cols<9, 512, 1, 8>(c, K, s);
      else if (wide)
        compact_cols<9, 512, 1>(c, K, s);
      else
        compact_cols<9>(c, K, s);
      break;
    default: break;
  }
#undef CVM_CC
}

bool compact_ok(const TileGrid& g) { return g.dftH2 == 0 && g.log2W >= 5 && g.log2W <= 11 && g.log2H >= 5 && g.log2H <= 11; }

void launch_compact_fft(const CompactCall& c, hipStream_t s) {
  const int nch = (c.ih + NB_ROWS - 1) / NB_ROWS;
  ChunkArgs K;
  K.img = c.img;
  K.istride = c.istride;
  K.iw = c.iw;
  K.ih = c.ih;
  K.step = c.step;
  K.words = (c.step == 4 && ((((uintptr_t)c.img) | (uintptr_t)c.istride) & 3) == 0) ? 1 : 0;
  K.st = c.st;
  K.ps = (uint32_t*)c.nscratch;
  K.nbx = (c.iw + 255) / 256;
  K.nch = nch;
  PrefixArgs Q;
  Q.nwords = 3 * c.iw;
  Q.nch = nch;
  Q.st = c.st;
  Q.ps = (uint32_t*)c.nscratch;
  const int nt = row_inv_nt(c.g.log2W);  // the inverse row launch's workgroup size
  Q.nwg = (Q.nwords + nt / PSEG - 1) / (nt / PSEG);  // nt / PSEG columns a workgroup
  compact_rows_dispatch(c, true, Q, s);   // parent + template rows, state statistics
  compact_cols_dispatch(c, K, s);         // state finalised; template + correlation columns, chunk sums
  compact_rows_dispatch(c, false, Q, s);  // inverse rows + chunk prefix sums
}

}  // namespace cvm
