// Every occurrence on a response map, on the device: threshold + local-
// maximum suppression, the recipe of TestFindAllOccurrences
// (cvmatch_test.go:346-365) and README.md:82-90.  (x, y) is a peak unless
// resp < thresh or an in-bounds element within |dx| <= rx, |dy| <= ry is
// strictly greater.
//
// "No neighbour strictly greater" is "not (window max > v)" when the window
// max is formed by strict comparisons (it is then one of the window's
// elements), so the window test is separable:
//   k_peak_rowmax  row pass: max over [x-rx, x+rx] of each row, staged in LDS
//                  up to 2048 outputs a workgroup (-inf outside the row stands
//                  for "no element there"), by doubling: maxima over 1, 4,
//                  16 ... then 2x as needed, P <= 2rx+1 elements (two
//                  doublings per LDS round trip), then the window is two
//                  overlapping length-P runs (log2 steps per point instead of
//                  2rx+1); very wide windows fall back to a linear scan of the
//                  staged row;
//   k_peak_mark    column pass over the row maxima, only for points that pass
//                  the threshold, stopping at the first greater one; four
//                  passes of 64 consecutive points per wave, four 64-bit
//                  ballots per wave and a count per 1024-point segment;
//   k_peak_tilescan  exclusive scan of the segment counts within tiles of
//                  1024 segments, and each tile's sum;
//   k_peak_emit    each hit's rank = the sums of the tiles before its own +
//                  its segment's offset in the tile + the hits before it in
//                  the segment, so peaks come out in raster order, as the
//                  reference's loop appends them.
// The map is read about twice (once per pass) plus the window re-reads that
// the caches absorb; the column pass touches only thresholded points.
#include <math.h>
#include <stdlib.h>

#include <algorithm>

#include "cvm_internal.h"

namespace cvm {

constexpr int PK_T = 256;               // threads per workgroup (4 waves)
constexpr int PK_PTS = 4;               // points per lane, 64 apart
constexpr int PK_SEG = PK_T * PK_PTS;   // raster points per segment
constexpr int PK_TILE = 1024;           // segments per scan tile
constexpr int PK_MASKS = (PK_T / 64) * PK_PTS;  // ballots per segment

// both ping-pong buffers fit 64 KiB
constexpr int PK_DBL_MAX_N = 65536 / (2 * (int)sizeof(float));
// largest row segment one row-pass workgroup takes
constexpr int PK_SEG_MAX = 2048;

__device__ __forceinline__ float pk_max(float u, float v) { return v > u ? v : u; }

// One workgroup per `seg` consecutive outputs of a row (a multiple of PK_T, up
// to the whole row), staged with their 2rx halo: a wide segment reads its
// halo once per row instead of once per 256 outputs, and spreads each
// pass's barrier over more points.
template <bool DOUBLING>
__global__ __launch_bounds__(PK_T) void k_peak_rowmax(const float* __restrict__ map, int rw, int rh, int rx, int seg,
                                                      float* __restrict__ rmax) {
  extern __shared__ float lds[];  // DOUBLING: 2 x n, else n
  const int x0 = blockIdx.x * seg;
  const int n = seg + 2 * rx;
  const int k = 2 * rx + 1;
  for (int y = blockIdx.y; y < rh; y += gridDim.y) {
    const float* row = map + (size_t)y * rw;
    float* a = lds;
    for (int i = threadIdx.x; i < n; i += PK_T) {
      const int x = x0 - rx + i;
      a[i] = (x >= 0 && x < rw) ? row[x] : -INFINITY;
    }
    __syncthreads();
    float* out = rmax + (size_t)y * rw + x0;
    const int cnt = min(seg, rw - x0);
    if (DOUBLING) {
      // a holds the maxima over runs of p elements starting at each index
      // (valid while index + p <= n); a pass writes those over 4p (or 2p)
      // into b, two doublings per LDS round trip while 4p <= k
      float* b = lds + n;
      int p = 1;
      for (; 4 * p <= k; p *= 4) {
        for (int i = threadIdx.x; i + 4 * p <= n; i += PK_T)
          b[i] = pk_max(pk_max(a[i], a[i + p]), pk_max(a[i + 2 * p], a[i + 3 * p]));
        __syncthreads();
        float* t = a;
        a = b;
        b = t;
      }
      if (2 * p <= k) {
        for (int i = threadIdx.x; i + 2 * p <= n; i += PK_T) b[i] = pk_max(a[i], a[i + p]);
        __syncthreads();
        a = b;
        p *= 2;
      }
      // p <= k < 2p: window [j, j + k) = runs [j, j + p) and [j + k - p, j + k)
      for (int j = threadIdx.x; j < cnt; j += PK_T) out[j] = pk_max(a[j], a[j + k - p]);
    } else {
      for (int j = threadIdx.x; j < cnt; j += PK_T) {
        float m = a[j];
        for (int i = 1; i < k; i++) m = pk_max(m, a[j + i]);
        out[j] = m;
      }
    }
    __syncthreads();
  }
}

// Point j of lane l in wave w of segment g is g*PK_SEG + w*256 + j*64 + l:
// each of the four passes reads 64 consecutive points, and a horizontal run
// of candidates spreads over lanes.  A candidate's column walk issues its row
// maxima eight at a time, so a long window costs few memory round trips.
__global__ __launch_bounds__(PK_T) void k_peak_mark(const float* __restrict__ map, const float* __restrict__ rmax,
                                                    int rw, int rh, int ry, float thresh,
                                                    unsigned long long* __restrict__ masks, int* __restrict__ counts,
                                                    int* __restrict__ total) {
  constexpr int kBatch = 8;
  __shared__ int wc[PK_T / 64];
  const unsigned npts = (unsigned)rw * (unsigned)rh;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const unsigned base = (unsigned)blockIdx.x * PK_SEG + (unsigned)w * 64 * PK_PTS + l;
  int c = 0;
#pragma unroll
  for (int j = 0; j < PK_PTS; j++) {
    const unsigned idx = base + j * 64;
    bool hit = false;
    if (idx < npts) {
      const float v = map[idx];
      if (!(v < thresh)) {
        const int y = (int)(idx / (unsigned)rw), x = (int)(idx - (unsigned)y * (unsigned)rw);
        const int ya = max(0, y - ry), yb = min(rh - 1, y + ry);
        hit = true;
        for (int y0 = ya; y0 <= yb && hit; y0 += kBatch) {
          float r[kBatch];
#pragma unroll
          for (int k = 0; k < kBatch; k++) r[k] = rmax[(size_t)min(y0 + k, yb) * rw + x];
#pragma unroll
          for (int k = 0; k < kBatch; k++) hit = hit && !(r[k] > v);
        }
      }
    }
    const unsigned long long b = __ballot(hit);
    if (l == 0) masks[(size_t)blockIdx.x * PK_MASKS + w * PK_PTS + j] = b;
    c += __popcll(b);
  }
  if (l == 0) wc[w] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int i = 0; i < PK_T / 64; i++) s += wc[i];
    counts[blockIdx.x] = s;
    if (blockIdx.x == 0) *total = 0;  // k_peak_tilescan adds the tile sums
  }
}

// One workgroup per tile of PK_TILE segments.
__global__ __launch_bounds__(PK_TILE) void k_peak_tilescan(const int* __restrict__ counts, int nseg,
                                                           int* __restrict__ offsets, int* __restrict__ tilesum,
                                                           int* __restrict__ total) {
  __shared__ int sh[PK_TILE];
  const int i = blockIdx.x * PK_TILE + threadIdx.x;
  const int c = i < nseg ? counts[i] : 0;
  sh[threadIdx.x] = c;
  __syncthreads();
  for (int off = 1; off < PK_TILE; off <<= 1) {
    const int t = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < nseg) offsets[i] = sh[threadIdx.x] - c;
  if (threadIdx.x == PK_TILE - 1) {
    tilesum[blockIdx.x] = sh[PK_TILE - 1];
    atomicAdd(total, sh[PK_TILE - 1]);
  }
}

__global__ __launch_bounds__(PK_T) void k_peak_emit(const float* __restrict__ map, int rw,
                                                    const unsigned long long* __restrict__ masks,
                                                    const int* __restrict__ counts, const int* __restrict__ offsets,
                                                    const int* __restrict__ tilesum, int cap, int32_t* __restrict__ xy,
                                                    float* __restrict__ vals) {
  __shared__ int part[PK_T / 64];
  const int sgi = blockIdx.x;
  if (counts[sgi] == 0) return;  // uniform over the workgroup
  // sum of the tiles before this segment's own
  const int ntile = sgi / PK_TILE;
  int acc = 0;
  for (int i = threadIdx.x; i < ntile; i += PK_T) acc += tilesum[i];
  for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) part[w] = acc;
  __syncthreads();
  int rank = offsets[sgi];
  for (int i = 0; i < PK_T / 64; i++) rank += part[i];
  const unsigned long long* m = masks + (size_t)sgi * PK_MASKS;
  for (int i = 0; i < w * PK_PTS; i++) rank += __popcll(m[i]);
  const unsigned long long lt = (1ull << l) - 1ull;
#pragma unroll
  for (int j = 0; j < PK_PTS; j++) {
    const unsigned long long b = m[w * PK_PTS + j];
    if ((b >> l) & 1ull) {
      const int r = rank + __popcll(b & lt);
      if (r < cap) {
        const unsigned idx = (unsigned)sgi * PK_SEG + (unsigned)w * 64 * PK_PTS + (unsigned)j * 64 + l;
        const int y = (int)(idx / (unsigned)rw), x = (int)(idx - (unsigned)y * (unsigned)rw);
        xy[2 * r] = x;
        xy[2 * r + 1] = y;
        vals[r] = map[idx];
      }
    }
    rank += __popcll(b);
  }
}

static int peaks_segments(int rw, int rh) { return (int)(((long long)rw * rh + PK_SEG - 1) / PK_SEG); }
static int peaks_tiles(int nseg) { return (nseg + PK_TILE - 1) / PK_TILE; }

// scratch: row maxima [rh][rw] | (256-aligned) ballots [nseg][PK_MASKS] |
// counts [nseg] | offsets [nseg] | tile sums [ntile] | total
struct PeakScratch {
  float* rmax;
  unsigned long long* masks;
  int *counts, *offsets, *tilesum, *total;
};

static PeakScratch peaks_layout(void* scratch, int rw, int rh) {
  const int nseg = peaks_segments(rw, rh);
  PeakScratch p;
  p.rmax = (float*)scratch;
  uintptr_t a = (uintptr_t)scratch + sizeof(float) * (size_t)rw * rh;
  a = (a + 255) & ~(uintptr_t)255;
  p.masks = (unsigned long long*)a;
  p.counts = (int*)(a + sizeof(unsigned long long) * (size_t)nseg * PK_MASKS);
  p.offsets = p.counts + nseg;
  p.tilesum = p.offsets + nseg;
  p.total = p.tilesum + peaks_tiles(nseg);
  return p;
}

size_t peaks_scratch_bytes(int rw, int rh) {
  const size_t nseg = (size_t)peaks_segments(rw, rh);
  return sizeof(float) * (size_t)rw * rh + 256 + sizeof(unsigned long long) * nseg * PK_MASKS +
         sizeof(int) * (2 * nseg + (size_t)peaks_tiles((int)nseg) + 1);
}

// Outputs per row-pass workgroup: the row split into the fewest segments of
// at most PK_SEG_MAX outputs whose doubling buffers fit 64 KiB, balanced and
// rounded up to PK_T; one PK_T segment for the linear scan of very wide
// windows.  CVMATCH_PEAK_SEG overrides the cap (a multiple of PK_T).
static int peaks_row_segment(int rw, int rx) {
  static const int cap = [] {
    const char* e = getenv("CVMATCH_PEAK_SEG");
    const int v = e ? atoi(e) : 0;
    return v >= PK_T && v % PK_T == 0 ? v : PK_SEG_MAX;
  }();
  const int fit = (PK_DBL_MAX_N - 2 * rx) / PK_T * PK_T;
  if (fit < PK_T) return PK_T;
  const int most = std::min(cap, fit);
  const int nseg = (rw + most - 1) / most;
  const int per = (rw + nseg - 1) / nseg;
  return (per + PK_T - 1) / PK_T * PK_T;
}

int peaks_max_rx() { return (65536 / (int)sizeof(float) - PK_T) / 2; }

int* peaks_total(void* scratch, int rw, int rh) { return peaks_layout(scratch, rw, rh).total; }

void launch_peaks(const float* map, int rw, int rh, float thresh, int rx, int ry, void* scratch, int cap,
                  int32_t* xy, float* vals, hipStream_t s) {
  const int nseg = peaks_segments(rw, rh);
  const PeakScratch p = peaks_layout(scratch, rw, rh);
  const int seg = peaks_row_segment(rw, rx);
  const int n = seg + 2 * rx;
  const dim3 grow((rw + seg - 1) / seg, (unsigned)std::min(rh, 65535));
  if (n <= PK_DBL_MAX_N)
    hipLaunchKernelGGL(k_peak_rowmax<true>, grow, dim3(PK_T), 2 * sizeof(float) * n, s, map, rw, rh, rx, seg,
                       p.rmax);
  else
    hipLaunchKernelGGL(k_peak_rowmax<false>, grow, dim3(PK_T), sizeof(float) * n, s, map, rw, rh, rx, seg, p.rmax);
  hipLaunchKernelGGL(k_peak_mark, dim3(nseg), dim3(PK_T), 0, s, map, p.rmax, rw, rh, ry, thresh, p.masks, p.counts,
                     p.total);
  hipLaunchKernelGGL(k_peak_tilescan, dim3(peaks_tiles(nseg)), dim3(PK_TILE), 0, s, p.counts, nseg, p.offsets,
                     p.tilesum, p.total);
  hipLaunchKernelGGL(k_peak_emit, dim3(nseg), dim3(PK_T), 0, s, map, rw, p.masks, p.counts, p.offsets, p.tilesum,
                     cap, xy, vals);
}

}  // namespace cvm
