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
//                  (-inf outside the row stands for "no element there");
//   k_peak_mark    column pass over the row maxima, only for points that pass
//                  the threshold, with early exit at the first greater one;
//                  one 64-bit ballot per wave and a count per 256-point raster
//                  segment;
//   k_peak_scan    exclusive scan of the segment counts (one workgroup);
//   k_peak_emit    each hit's rank = its segment's offset + the hits before it
//                  in the segment, so peaks come out in raster order, as the
//                  reference's loop appends them.
// The map is read about twice (once per pass) plus the window re-reads that
// the caches absorb; the column pass touches only thresholded points.
#include <math.h>

#include <algorithm>

#include "cvm_internal.h"

namespace cvm {

constexpr int PK_SEG = 256;  // raster points per workgroup (4 waves)

__global__ __launch_bounds__(PK_SEG) void k_peak_rowmax(const float* __restrict__ map, int rw, int rh, int rx,
                                                        float* __restrict__ rmax) {
  extern __shared__ float seg[];  // PK_SEG + 2 rx
  const int x0 = blockIdx.x * PK_SEG;
  const int n = PK_SEG + 2 * rx;
  for (int y = blockIdx.y; y < rh; y += gridDim.y) {
    const float* row = map + (size_t)y * rw;
    for (int i = threadIdx.x; i < n; i += PK_SEG) {
      const int x = x0 - rx + i;
      seg[i] = (x >= 0 && x < rw) ? row[x] : -INFINITY;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x;
    if (x < rw) {
      float m = seg[threadIdx.x];
      for (int i = 1; i <= 2 * rx; i++) {
        const float v = seg[threadIdx.x + i];
        m = v > m ? v : m;
      }
      rmax[(size_t)y * rw + x] = m;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(PK_SEG) void k_peak_mark(const float* __restrict__ map, const float* __restrict__ rmax,
                                                      int rw, int rh, int ry, float thresh,
                                                      unsigned long long* __restrict__ masks, int* __restrict__ counts) {
  __shared__ int wc[PK_SEG / 64];
  const long long npts = (long long)rw * rh;
  const long long idx = (long long)blockIdx.x * PK_SEG + threadIdx.x;
  bool hit = false;
  if (idx < npts) {
    const float v = map[idx];
    if (!(v < thresh)) {
      const int y = (int)(idx / rw), x = (int)(idx - (long long)y * rw);
      const int ya = max(0, y - ry), yb = min(rh - 1, y + ry);
      hit = true;
      for (int yy = ya; yy <= yb; yy++) {
        if (rmax[(size_t)yy * rw + x] > v) {
          hit = false;
          break;
        }
      }
    }
  }
  const unsigned long long b = __ballot(hit);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    masks[(size_t)blockIdx.x * (PK_SEG / 64) + w] = b;
    wc[w] = __popcll(b);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int c = 0;
    for (int k = 0; k < PK_SEG / 64; k++) c += wc[k];
    counts[blockIdx.x] = c;
  }
}

__global__ __launch_bounds__(1024) void k_peak_scan(const int* __restrict__ counts, int nseg, int* __restrict__ offsets,
                                                    int* __restrict__ total) {
  __shared__ int sh[1024];
  const int per = (nseg + 1023) / 1024;
  const int a = min(nseg, (int)threadIdx.x * per), b = min(nseg, a + per);
  int s = 0;
  for (int i = a; i < b; i++) s += counts[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int t = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int run = sh[threadIdx.x] - s;
  for (int i = a; i < b; i++) {
    offsets[i] = run;
    run += counts[i];
  }
  if (threadIdx.x == 1023) *total = sh[1023];
}

__global__ __launch_bounds__(PK_SEG) void k_peak_emit(const float* __restrict__ map, int rw,
                                                      const unsigned long long* __restrict__ masks,
                                                      const int* __restrict__ counts, const int* __restrict__ offsets,
                                                      int cap, int32_t* __restrict__ xy, float* __restrict__ vals) {
  const int sgi = blockIdx.x;
  if (counts[sgi] == 0) return;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const unsigned long long* m = masks + (size_t)sgi * (PK_SEG / 64);
  const unsigned long long b = m[w];
  if (!((b >> l) & 1ull)) return;
  int rank = offsets[sgi];
  for (int k = 0; k < w; k++) rank += __popcll(m[k]);
  rank += __popcll(b & ((1ull << l) - 1ull));
  if (rank >= cap) return;
  const long long idx = (long long)sgi * PK_SEG + threadIdx.x;
  const int y = (int)(idx / rw), x = (int)(idx - (long long)y * rw);
  xy[2 * rank] = x;
  xy[2 * rank + 1] = y;
  vals[rank] = map[idx];
}

int peaks_segments(int rw, int rh) { return (int)(((long long)rw * rh + PK_SEG - 1) / PK_SEG); }

size_t peaks_scratch_bytes(int rw, int rh) {
  const size_t nseg = (size_t)peaks_segments(rw, rh);
  return sizeof(float) * (size_t)rw * rh + sizeof(unsigned long long) * nseg * (PK_SEG / 64) +
         sizeof(int) * (2 * nseg + 1) + 256;
}

int peaks_max_rx() { return (65536 / (int)sizeof(float) - PK_SEG) / 2; }

// scratch: row maxima [rh][rw] | (256-aligned) ballots [nseg][4] | counts
// [nseg] | offsets [nseg] | total
static int* peaks_counts(void* scratch, int rw, int rh) {
  const int nseg = peaks_segments(rw, rh);
  uintptr_t p = (uintptr_t)scratch + sizeof(float) * (size_t)rw * rh;
  p = (p + 255) & ~(uintptr_t)255;
  return (int*)(p + sizeof(unsigned long long) * (size_t)nseg * (PK_SEG / 64));
}

int* peaks_total(void* scratch, int rw, int rh) { return peaks_counts(scratch, rw, rh) + 2 * peaks_segments(rw, rh); }

void launch_peaks(const float* map, int rw, int rh, float thresh, int rx, int ry, void* scratch, int cap,
                  int32_t* xy, float* vals, hipStream_t s) {
  const int nseg = peaks_segments(rw, rh);
  float* rmax = (float*)scratch;
  int* counts = peaks_counts(scratch, rw, rh);
  unsigned long long* masks = (unsigned long long*)((uint8_t*)counts - sizeof(unsigned long long) * (size_t)nseg * (PK_SEG / 64));
  int* offsets = counts + nseg;
  int* total = offsets + nseg;
  const dim3 grow((rw + PK_SEG - 1) / PK_SEG, (unsigned)std::min(rh, 65535));
  hipLaunchKernelGGL(k_peak_rowmax, grow, dim3(PK_SEG), sizeof(float) * (PK_SEG + 2 * rx), s, map, rw, rh, rx, rmax);
  hipLaunchKernelGGL(k_peak_mark, dim3(nseg), dim3(PK_SEG), 0, s, map, rmax, rw, rh, ry, thresh, masks, counts);
  hipLaunchKernelGGL(k_peak_scan, dim3(1), dim3(1024), 0, s, counts, nseg, offsets, total);
  hipLaunchKernelGGL(k_peak_emit, dim3(nseg), dim3(PK_SEG), 0, s, map, rw, masks, counts, offsets, cap, xy, vals);
}

}  // namespace cvm
