// Direct cross-correlation on int8 MFMA (gfx950): the window x template sum
// of TM_CCOEFF_NORMED as an exact integer.
//
// Bytes are shifted to int8 (v - 128 == v ^ 0x80), so every product is exact
// and the int32 accumulation is exact while area * cn * 2^14 < 2^31.  The
// kernel writes C'(x, y) = sum (P - 128)(T - 128) over the window; the
// normalisation recovers sum P*T = C' + 128 (sum S_P + sum S_T) - 2^14 area cn
// from the exact window sums it already has.
//
// GEMM shape.  One MFMA tile is 32 output rows (M) x 32 output columns (N).
// K runs over the staged bytes of one image row segment, channels
// interleaved: KSTEP bytes per pixel, 1 for gray, 4 for RGBA, 3 for RGBA
// staged without alpha (cn = 3).  For template row i,
//   A_i[m][u] = P'[y0 + m + i][KSTEP*x0 + u] and
//   B_i[u][n] = T'[i][u - KSTEP*n]  when 0 <= u - KSTEP*n < KSTEP*tw and the
//                                   byte's channel is below cn, else 0,
// a banded Toeplitz matrix that is the same for every tile.  Summing the
// th products A_i B_i into one accumulator gives the 32 x 32 block.  B_i is
// built once per template in MFMA fragment order (k_bfrag).  The A fragment
// of column tile t and K chunk kc is chunk kc + KSTEP*t of the row segment, so
// a wave walking kc in steps of KSTEP (one residue class at a time) slides a
// window of T A fragments by one load per T MFMAs.  Image rows stream through
// an LDS ring, one new row per template row.
//
// v_mfma_i32_32x32x32_i8 lane layout: A byte b of lane l is A[l%32][16(l/32)+b],
// B byte b of lane l is B[16(l/32)+b][l%32], and accumulator register j of
// lane l is D[8(j/4) + 4(l/32) + j%4][l%32].
#include <float.h>
#include <limits.h>

#include <type_traits>

#include "cvm_internal.h"
#include "cvm_score.h"

namespace cvm {

typedef int dv4i __attribute__((ext_vector_type(4)));
typedef int dv16i __attribute__((ext_vector_type(16)));
// B fragments are read through a global-address-space pointer: a job table
// hands the kernel generic pointers, and generic (flat) loads are counted
// with the LDS reads, so every wait for an A fragment would drain them too.
typedef const dv4i __attribute__((address_space(1)))* gdv4p;

// The direct kernels run two waves per SIMD (their LDS ring and accumulators
// allow no more): the register budget of two waves lets the compiler keep
// loads in flight in registers of their own instead of reusing registers
// that a load may still be writing (each reuse is a wait for that load).
// (Six or more chunks per class need more than the two-wave budget.)
#define CVM_DIRECT_OCC __attribute__((amdgpu_waves_per_eu(CMAX >= 6 ? 1 : 2, 2)))

constexpr int DIR_T = 4;      // column tiles per wave (128 output columns)
constexpr int DIR_WAVES = 1;  // waves per workgroup, 32 output rows each
constexpr int DIR_ROWS = 32 * DIR_WAVES;
// Ring rows: the DIR_ROWS a step reads plus the one written at its start (a
// single wave's LDS accesses run in program order; queued rows wait in
// registers, not in the ring).
constexpr int DIR_RING = DIR_ROWS + 1;
static_assert(DIR_WAVES == 1, "a ring of DIR_ROWS + 1 rows relies on one wave's program order");
// Ring rows of a workgroup of `waves` waves (the multi-wave form below).
__host__ __device__ constexpr int direct_ring_rows(int waves) { return waves == 1 ? DIR_RING : 3 * waves + 31; }
constexpr int DIR_CMAX = 9;   // K chunks per residue class: nk <= kstep * DIR_CMAX
// LDS behind the ring that the unused A fragments of the last ring row may
// read: at most chunk KSTEP - 1 + KSTEP (CMAX + DIR_T - 2) of a row, i.e.
// 32 (4 - 1 + 4 * 11) + 32 = 1536 bytes from the row's start.
constexpr int DIR_SLACK = 2048;

// Staged bytes per pixel: gray 1, RGBA 4, RGBA with cn = 3 staged as RGB.
int direct_kstep(int step, int cn) { return step == 1 ? 1 : (cn == 3 ? 3 : 4); }

int direct_nk(int tw, int kstep) { return (kstep * (32 + tw - 1) + 31) / 32; }

// B fragments plus DIR_CMAX of padding: a class load may run past the last
// fragment (those registers are never used).
size_t direct_bfrag_bytes(int tw, int th, int kstep) {
  return ((size_t)th * direct_nk(tw, kstep) + DIR_CMAX) * 64 * 16;
}

static int stage_unit(int kstep) { return kstep == 3 ? 12 : 16; }  // staged bytes per 16 source bytes

static int direct_units(int nk, int kstep) {
  const int bytes = 32 * (nk + kstep * (DIR_T - 1));
  const int spu = stage_unit(kstep);
  return (bytes + spu - 1) / spu;
}

static int direct_pitch(int nk, int kstep) {
  // staged row bytes, padded so consecutive rows start 16 bytes (4 banks)
  // apart modulo 128: the 32 rows of one A fragment spread over all banks
  int bytes = direct_units(nk, kstep) * stage_unit(kstep);
  return (bytes + 127) / 128 * 128 + 16;
}


// Visit order of the K chunks within one template row: residue classes of
// kc modulo kstep, each in increasing kc (the A window slides inside a class).
__host__ __device__ inline void direct_visit(int s, int nk, int kstep, int& rc, int& kc) {
  for (rc = 0; rc < kstep; rc++) {
    const int c = (nk - rc + kstep - 1) / kstep;
    if (s < c) {
      kc = rc + kstep * s;
      return;
    }
    s -= c;
  }
  kc = 0;
}

// B_i fragments in visit order: bf[((i * nk + s) * 64 + lane) * 16 + b] =
// B_i[32 kc + 16 (lane/32) + b][lane % 32] with kc the s-th visited chunk.
// The template has sstep bytes per pixel (4 when kstep is 3 or 4).
__global__ void k_bfrag(const uint8_t* __restrict__ tpl, int64_t tstride, int tw, int th, int kstep, int sstep,
                        int cn, const CallState* __restrict__ st, int nk, int8_t* __restrict__ bf) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // one 16-byte fragment slice
  const int64_t total = (int64_t)th * nk * 64;
  if (id >= total) return;
  if (st) cn = st->cn;
  const int lane = (int)(id & 63);
  const int64_t is = id >> 6;
  const int sv = (int)(is % nk), i = (int)(is / nk);
  int rc, kc;
  direct_visit(sv, nk, kstep, rc, kc);
  const int n = lane & 31, h = lane >> 5;
  uint32_t w[4] = {0, 0, 0, 0};
  for (int b = 0; b < 16; b++) {
    const int u = 32 * kc + 16 * h + b;
    const int d = u - kstep * n;
    uint32_t v = 0;
    if (d >= 0 && d < kstep * tw) {
      const int j = d / kstep, k = d - j * kstep;
      if (k < cn) v = (uint32_t)(tpl[(int64_t)i * tstride + (int64_t)j * sstep + k] ^ 0x80u) & 255u;
    }
    w[b >> 2] |= v << (8 * (b & 3));
  }
  uint4* o = (uint4*)(bf + id * 16);
  *o = make_uint4(w[0], w[1], w[2], w[3]);
}

void launch_bfrag(const uint8_t* tpl, int64_t tstride, int tw, int th, int kstep, int cn, const CallState* st,
                  int8_t* bf, hipStream_t s) {
  int nk = direct_nk(tw, kstep);
  int64_t total = (int64_t)th * nk * 64;
  hipLaunchKernelGGL(k_bfrag, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, tpl, tstride, tw, th, kstep,
                     kstep == 1 ? 1 : 4, cn, st, nk, bf);
}

// Staging of image rows for one output block (columns from X0, rows from
// Y0): the 16 source bytes of staging unit c of staged row rho (image row
// Y0 + rho), and their write into a staged LDS row.  VEC: a unit is wholly
// inside or wholly outside the image; the load is clamped into the image and
// `ok` says whether to keep it.  The int8 shift, the zeroing and the RGB
// repacking are applied when the unit is written (put), so the load itself
// is never waited on early.
template <int KSTEP, bool VEC>
struct RowStager {
  static constexpr int SSTEP = KSTEP == 1 ? 1 : 4;  // source bytes per pixel
  const uint8_t* __restrict__ img;
  int64_t istride;
  int ih, Y0;
  int64_t xb0, imgBytes;
  __device__ RowStager(const uint8_t* im, int64_t is, int iw, int ih_, int X0, int Y0_)
      : img(im), istride(is), ih(ih_), Y0(Y0_), xb0((int64_t)SSTEP * X0), imgBytes((int64_t)SSTEP * iw) {}

  __device__ __forceinline__ uint4 fetch(int rho, int c, bool& ok) const {
    const int r = Y0 + rho;
    const int64_t b0 = xb0 + 16 * (int64_t)c;
    if (VEC) {
      ok = r < ih && b0 < imgBytes;
      const int rr = r < ih ? r : ih - 1;
      const int64_t bb = b0 < imgBytes ? b0 : imgBytes - 16;
      return *(const uint4*)(img + (int64_t)rr * istride + bb);
    }
    ok = true;
    uint32_t w[4] = {0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u};
    if (r < ih) {
      const uint8_t* p = img + (int64_t)r * istride;
      for (int b = 0; b < 16; b++)
        if (b0 + b < imgBytes)
          w[b >> 2] = (w[b >> 2] & ~(255u << (8 * (b & 3)))) | ((uint32_t)p[b0 + b] << (8 * (b & 3)));
    }
    return make_uint4(w[0], w[1], w[2], w[3]);
  }
  // write unit c of a staged row whose first byte is dst (outside bytes: 0)
  __device__ static __forceinline__ void put(uint8_t* dst, int c, uint4 v, bool ok) {
    if (!ok) v = make_uint4(0x80808080u, 0x80808080u, 0x80808080u, 0x80808080u);
    v.x ^= 0x80808080u;
    v.y ^= 0x80808080u;
    v.z ^= 0x80808080u;
    v.w ^= 0x80808080u;
    if (KSTEP == 3) {
      // four RGBA pixels -> twelve RGB bytes
      uint32_t* o = (uint32_t*)(dst + 12 * c);
      o[0] = (v.x & 0x00FFFFFFu) | (v.y << 24);
      o[1] = ((v.y >> 8) & 0x0000FFFFu) | (v.z << 16);
      o[2] = ((v.z >> 16) & 0x000000FFu) | (v.w << 8);
    } else {
      *(uint4*)(dst + 16 * c) = v;
    }
  }
};

// One workgroup: output rows [Y0, Y0 + DIR_ROWS) x columns [X0, X0 + 32 T).
// VEC: image rows 16-byte aligned with sstep*iw % 16 == 0 (16-byte loads).
// CMAX: chunks per residue class, ceil(nk / KSTEP) <= CMAX.
template <int KSTEP, bool VEC, int CMAX>
__device__ __forceinline__ void direct_block(uint8_t* ring, const uint8_t* __restrict__ img, int64_t istride, int iw,
                                             int ih, const int8_t* __restrict__ bf, int th, int nk, int pitch, int bx,
                                             int by, dv16i (&acc)[DIR_T]) {
  constexpr int T = DIR_T;
  constexpr int SPU = KSTEP == 3 ? 12 : 16;  // staged bytes per 16 source bytes
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int X0 = bx * 32 * T, Y0 = by * DIR_ROWS;
  const int units = (32 * (nk + KSTEP * (T - 1)) + SPU - 1) / SPU;  // staging units per row
  const RowStager<KSTEP, VEC> stg(img, istride, iw, ih, X0, Y0);
  auto fetch = [&](int rho, int c, bool& ok) { return stg.fetch(rho, c, ok); };
  auto put = [&](uint8_t* dst, int c, uint4 v, bool ok) { RowStager<KSTEP, VEC>::put(dst, c, v, ok); };

  // initial rows [0, DIR_ROWS)
  for (int e = tid; e < DIR_ROWS * units; e += 64 * DIR_WAVES) {
    const int rho = e / units, c = e - rho * units;
    bool ok;
    const uint4 v = fetch(rho, c, ok);
    put(ring + rho * pitch, c, v, ok);
  }
  __syncthreads();

#pragma unroll
  for (int t = 0; t < T; t++)
#pragma unroll
    for (int j = 0; j < 16; j++) acc[t][j] = 0;

  const int half = lane >> 5, m = lane & 31;
  // Per template row, K chunk residue class rc has c_rc = ceil((nk - rc) / KSTEP)
  // <= CMAX chunks; its c_rc + T - 1 A fragments and c_rc B fragments sit in
  // statically indexed registers.  A fragments (LDS) are loaded one class
  // ahead.  B fragments (global, L2) have one buffer per class when KSTEP > 1:
  // each class loads the B of the class KSTEP - 1 ahead (the next row's when
  // past this row) into the buffer the previous class has just used, so a B
  // load has KSTEP - 1 classes of MFMAs to arrive; gray rows (one class) keep
  // the next row's B in a second buffer loaded at the row start.  Loads are
  // unconditional (A chunk indices clamp to the staged row, B rows clamp to
  // the template and run into the padding), so the compiler's wait counts
  // are exact.
  constexpr int NA = CMAX + T - 1;
  int cls[KSTEP], base[KSTEP];
  {
    int acc_s = 0;
#pragma unroll
    for (int rc = 0; rc < KSTEP; rc++) {
      cls[rc] = rc < nk ? (nk - rc + KSTEP - 1) / KSTEP : 0;
      base[rc] = acc_s;
      acc_s += cls[rc];
    }
  }
  constexpr int NBUF = KSTEP == 1 ? 2 : KSTEP;
  dv4i A0[NA], A1[NA], Bk[NBUF][CMAX];
  // A fragment j of class rc: chunk rc + KSTEP j of the row, a compile-time
  // offset.  The fragments past the class's chunk count are loaded but feed
  // no MFMA; they may read past the row (into the next ring row, or past
  // the ring's last row into the DIR_SLACK bytes allocated behind it).
  auto loadA = [&](dv4i* buf, const uint8_t* arow, int rc) {
#pragma unroll
    for (int j = 0; j < NA; j++) buf[j] = *(const dv4i*)(arow + 32 * (rc + KSTEP * j));
  };
  // template rows past the last clamp to it (loaded, not used)
  // (the fragment's address: a wave-uniform row base plus this lane's 16
  // bytes, so the loads take a scalar base and one vector offset)
  auto loadB = [&](dv4i* buf, int i, int rc) {
    const int f0 = __builtin_amdgcn_readfirstlane(min(i, th - 1) * nk + base[rc]);
    const gdv4p p = (gdv4p)(const dv4i*)(bf + (int64_t)f0 * 1024) + lane;
#pragma unroll
    for (int j = 0; j < CMAX; j++) buf[j] = p[j * 64];
  };
  // (the class's chunk count is wave-uniform: scalar branches)
  auto mma = [&](const dv4i* a, const dv4i* bb, int c) {
    const int cu = __builtin_amdgcn_readfirstlane(c);
#pragma unroll
    for (int j = 0; j < CMAX; j++)
      if (j < cu)
#pragma unroll
        for (int t = 0; t < T; t++) acc[t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[j + t], bb[j], acc[t], 0, 0, 0);
  };

  // Staged row DIR_ROWS + i (first read in step i + 1) is fetched at the
  // start of step i - 2 (before the loop for i < 2) into register slot i & 1
  // and written to its ring slot at the start of step i, before step i
  // prefetches step i + 1's first class (the workgroup is one wave, so its
  // LDS accesses take effect in program order).  Its slot held row
  // i + DIR_ROWS - DIR_RING = i - 1, last read in step i - 1.  The rows are
  // walked in pairs, so the two register slots are static: no copies of
  // registers a load may still be writing.  Unit nc[p] of each staged row is
  // this thread's (units past the row reload the last one, not written).
  int nc[2];
#pragma unroll
  for (int p = 0; p < 2; p++) nc[p] = tid + p * 64 * DIR_WAVES;
  const int ncl[2] = {min(nc[0], units - 1), min(nc[1], units - 1)};
  uint4 rq[2][2];
  bool rok[2][2];
#pragma unroll
  for (int p = 0; p < 2; p++) rq[0][p] = fetch(DIR_ROWS, ncl[p], rok[0][p]);

  // Odd KSTEP: class c of the whole walk (row i, class rc: c = i * KSTEP + rc)
  // uses A buffer c & 1 and loads class c + 1's A into the other one, so a
  // row's first class takes buffer i & 1 and the rows are walked in pairs,
  // every buffer index static (gray rows keep their B the same way: row i
  // in Bk[i & 1]).  Even KSTEP: every row starts in A0, and the next row's
  // first class is loaded into a third buffer AN (skipped after the last
  // row), copied at the row start.  That keeps AN live beside A0 and the
  // RGBA kernels at one wave per SIMD, which measured faster than two (4K
  // RGBA 128x128: 0.66 against 0.81 ms; the 33-row ring of these rows lets
  // four single-wave workgroups share a CU).
  constexpr bool PAIRED = (KSTEP & 1) != 0;
  dv4i AN[PAIRED ? 1 : NA];
  // this lane's A rows start at rbase + roff, roff = ((m + i) % DIR_RING)
  // pitch for step i, advanced by one row a step (no division per step)
  const uint8_t* const rbase = ring + 16 * half;
  const int ringBytes = DIR_RING * pitch;
  int roff = ((32 * wave + m) % DIR_RING) * pitch;
  {
    loadA(PAIRED ? A0 : AN, rbase + roff, 0);
    if (KSTEP == 1) {
      loadB(Bk[0], 0, 0);
    } else {
#pragma unroll
      for (int rc = 0; rc + 1 < KSTEP; rc++) loadB(Bk[rc], 0, rc);
    }
  }
#pragma unroll
  for (int p = 0; p < 2; p++) rq[1][p] = fetch(DIR_ROWS + 1, ncl[p], rok[1][p]);
  // one template row; P0 = parity of its first class, QS = i & 1 (its
  // register slot).  The rows run in pairs; with odd th the pair's second
  // step is a row past the template that issues its loads and stores (into
  // ring slots no real step reads again) but no MFMA, so every step has the
  // same loads and the compiler counts the loads in flight exactly.  For the
  // same reason the row stores are unconditional: lanes past the row's
  // units rewrite the last unit with the same bytes.
  auto step = [&](int i, auto P0, auto QS) {
    constexpr int p0 = decltype(P0)::value;
    constexpr int qs = decltype(QS)::value;
    const bool live = i < th;
    {
      const int ns = (DIR_ROWS + i) % DIR_RING;
#pragma unroll
      for (int p = 0; p < 2; p++) put(ring + ns * pitch, ncl[p], rq[qs][p], rok[qs][p]);
    }
    // The register slot is refilled right after its store (for step i + 2),
    // so it has a whole step of MFMAs to arrive; the scheduling barrier keeps
    // the new loads behind the store, whose wait would otherwise cover them.
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int p = 0; p < 2; p++) rq[qs][p] = fetch(DIR_ROWS + i + 2, ncl[p], rok[qs][p]);
    // this wave's A rows: staged rows 32 wave + m + i
    int roffN = roff + pitch;
    if (roffN >= ringBytes) roffN -= ringBytes;
    const uint8_t* arow = rbase + roff;
    const uint8_t* arowN = rbase + roffN;
    roff = roffN;
    if (!PAIRED) {
#pragma unroll
      for (int j = 0; j < NA; j++) A0[j] = AN[j];
    }
    if (KSTEP == 1) loadB(Bk[(p0 + 1) & 1], i + 1, 0);
#pragma unroll
    for (int rc = 0; rc < KSTEP; rc++) {
      dv4i* ca = ((p0 + rc) & 1) ? A1 : A0;
      dv4i* na = ((p0 + rc) & 1) ? A0 : A1;
      if (KSTEP > 1) {
        const int lc = rc + KSTEP - 1;
        if (lc < KSTEP)
          loadB(Bk[lc], i, lc);
        else
          loadB(Bk[lc - KSTEP], i + 1, lc - KSTEP);
      }
      if (rc + 1 < KSTEP) loadA(na, arow, rc + 1);
      if (rc == KSTEP - 1) {
        // the next step's first class reads the row written at this step's
        // start (program order within the wave; the barrier is a compiler
        // fence here); past the last row it reads an unused ring row
        if (KSTEP == 1) __syncthreads();
        loadA(PAIRED ? na : AN, arowN, 0);
      }
      mma(ca, Bk[KSTEP == 1 ? (p0 & 1) : rc], live ? cls[rc] : 0);
      if (rc == 0 && KSTEP > 1) __syncthreads();
    }
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using IP = std::integral_constant<int, PAIRED ? 1 : 0>;
  for (int i = 0; i < th; i += 2) {
    step(i, I0(), I0());
    step(i + 1, IP(), I1());
  }
}

// The same output block (32 rows x 32 T columns) computed by W waves that
// share one ring of staged rows: wave w takes the template rows i = w mod W,
// so the W waves of step s read the overlapping row windows [W s + w,
// W s + w + 32).  The ring holds 3 W + 31 rows: those of step s, the W rows
// step s + 1 adds (written during step s - 1) and the W being written during
// step s for step s + 2 -- about (3 W + 31) / W rows per wave instead of 33,
// which lets more waves share a CU.  Each step: the waves write their row for
// step s + 2 (fetched a step earlier), run their template row (residue
// classes with double-buffered A and B fragments) and meet at one barrier;
// the last class already loads the next step's first A fragments (rows that
// the barrier of step s - 1 published) and B fragments, so no step starts on
// a load.  The slot a step writes held a row only step s - 1 read.  Each
// wave ends with its partial sum of the block (the caller adds them).
template <int W, int KSTEP, bool VEC, int CMAX>
__device__ __forceinline__ void direct_block_mw(uint8_t* ring, const uint8_t* __restrict__ img, int64_t istride,
                                                int iw, int ih, const int8_t* __restrict__ bf, int th, int nk,
                                                int pitch, int bx, int by, dv16i (&acc)[DIR_T]) {
  constexpr int T = DIR_T;
  constexpr int R = direct_ring_rows(W);
  constexpr int SPU = KSTEP == 3 ? 12 : 16;
  constexpr int NA = CMAX + T - 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int X0 = bx * 32 * T, Y0 = by * 32;
  const int units = (32 * (nk + KSTEP * (T - 1)) + SPU - 1) / SPU;
  const int nrows = th + 31;  // staged rows the block reads
  const RowStager<KSTEP, VEC> stg(img, istride, iw, ih, X0, Y0);

  // steps 0 and 1 read staged rows [0, 2 W + 31)
  for (int e = tid; e < (2 * W + 31) * units; e += 64 * W) {
    const int rho = e / units, c = e - rho * units;
    bool ok;
    const uint4 v = stg.fetch(rho, c, ok);
    RowStager<KSTEP, VEC>::put(ring + rho * pitch, c, v, ok);
  }
  // this wave's row written in step s is staged row W s + 2 W + 31 + wave
  // (read from step s + 2 on), fetched during step s - 1 (units past the
  // row reload the last one and are not written)
  const int nc0 = lane, nc1 = lane + 64;
  const int ncl0 = min(nc0, units - 1), ncl1 = min(nc1, units - 1);
  bool k0, k1;
  uint4 q0 = stg.fetch(2 * W + 31 + wave, ncl0, k0), q1 = stg.fetch(2 * W + 31 + wave, ncl1, k1);

#pragma unroll
  for (int t = 0; t < T; t++)
#pragma unroll
    for (int j = 0; j < 16; j++) acc[t][j] = 0;
  const int half = lane >> 5, m = lane & 31;
  int cls[KSTEP], base[KSTEP];
  {
    int acc_s = 0;
#pragma unroll
    for (int rc = 0; rc < KSTEP; rc++) {
      cls[rc] = rc < nk ? (nk - rc + KSTEP - 1) / KSTEP : 0;
      base[rc] = acc_s;
      acc_s += cls[rc];
    }
  }
  dv4i A0[NA], A1[NA], B0[CMAX], B1[CMAX];
  // A fragment j of class rc: chunk rc + KSTEP j of the row, a compile-time
  // offset.  The fragments past the class's chunk count are loaded but feed
  // no MFMA; they may read past the row (into the next ring row, or past
  // the ring's last row into the DIR_SLACK bytes allocated behind it).
  auto loadA = [&](dv4i* buf, const uint8_t* arow, int rc) {
#pragma unroll
    for (int j = 0; j < NA; j++) buf[j] = *(const dv4i*)(arow + 32 * (rc + KSTEP * j));
  };
  // template rows past the last clamp to it (loaded, not used)
  // (the fragment's address: a wave-uniform row base plus this lane's 16
  // bytes, so the loads take a scalar base and one vector offset)
  auto loadB = [&](dv4i* buf, int i, int rc) {
    const int f0 = __builtin_amdgcn_readfirstlane(min(i, th - 1) * nk + base[rc]);
    const gdv4p p = (gdv4p)(const dv4i*)(bf + (int64_t)f0 * 1024) + lane;
#pragma unroll
    for (int j = 0; j < CMAX; j++) buf[j] = p[j * 64];
  };
  // (the class's chunk count is wave-uniform: scalar branches)
  auto mma = [&](const dv4i* a, const dv4i* bb, int c) {
    const int cu = __builtin_amdgcn_readfirstlane(c);
#pragma unroll
    for (int j = 0; j < CMAX; j++)
      if (j < cu)
#pragma unroll
        for (int t = 0; t < T; t++) acc[t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[j + t], bb[j], acc[t], 0, 0, 0);
  };
  __syncthreads();

  // class c of the whole walk (step s, class rc: c = s * KSTEP + rc) uses
  // buffers c & 1 and loads class c + 1 into the others, so with odd KSTEP a
  // step's first class alternates between the buffer pairs and the steps are
  // walked in pairs, every buffer index static
  // this lane's A rows start at rbase + roff, roff = ((m + i) % R) pitch
  // for its step's template row i, advanced by W rows a step
  const uint8_t* const rbase = ring + 16 * half;
  const int ringBytes = R * pitch, stepBytes = W * pitch;
  int roff = ((m + wave) % R) * pitch;
  loadA(A0, rbase + roff, 0);
  loadB(B0, wave, 0);
  const int steps = (th + W - 1) / W;
  auto step = [&](int st, auto P0) {
    constexpr int p0 = decltype(P0)::value;
    const int wr = W * st + 2 * W + 31 + wave;
    if (wr < nrows) {
      uint8_t* dst = ring + (wr % R) * pitch;
      if (nc0 < units) RowStager<KSTEP, VEC>::put(dst, nc0, q0, k0);
      if (nc1 < units) RowStager<KSTEP, VEC>::put(dst, nc1, q1, k1);
    }
    q0 = stg.fetch(wr + W, ncl0, k0);
    q1 = stg.fetch(wr + W, ncl1, k1);
    const int i = W * st + wave;
    int roffN = roff + stepBytes;
    if (roffN >= ringBytes) roffN -= ringBytes;
    const uint8_t* arow = rbase + roff;
    const uint8_t* arowN = rbase + roffN;
    roff = roffN;
#pragma unroll
    for (int rc = 0; rc < KSTEP; rc++) {
      dv4i* ca = ((p0 + rc) & 1) ? A1 : A0;
      dv4i* na = ((p0 + rc) & 1) ? A0 : A1;
      dv4i* cb = ((p0 + rc) & 1) ? B1 : B0;
      dv4i* nb = ((p0 + rc) & 1) ? B0 : B1;
      if (rc + 1 < KSTEP) {
        loadA(na, arow, rc + 1);
        loadB(nb, i, rc + 1);
      } else {
        // the next step's first class: its rows were published by the
        // previous step's barrier (past the last step: unused ring rows)
        loadA(na, arowN, 0);
        loadB(nb, i + W, 0);
      }
      if (i < th) mma(ca, cb, cls[rc]);
    }
    __syncthreads();
  };
  if constexpr ((KSTEP & 1) != 0) {
    for (int st = 0; st < steps; st += 2) {
      step(st, std::integral_constant<int, 0>());
      if (st + 1 < steps) step(st + 1, std::integral_constant<int, 1>());
    }
  } else {
    for (int st = 0; st < steps; st++) step(st, std::integral_constant<int, 0>());
  }
}

// The block's sum over the template rows: one wave's whole sum (W = 1) or
// each of the W waves' partial sums.
template <int W, int KSTEP, bool VEC, int CMAX>
__device__ __forceinline__ void direct_accumulate(uint8_t* ring, const uint8_t* __restrict__ img, int64_t istride,
                                                  int iw, int ih, const int8_t* __restrict__ bf, int th, int nk,
                                                  int pitch, int bx, int by, dv16i (&acc)[DIR_T]) {
  if constexpr (W == 1)
    direct_block<KSTEP, VEC, CMAX>(ring, img, istride, iw, ih, bf, th, nk, pitch, bx, by, acc);
  else
    direct_block_mw<W, KSTEP, VEC, CMAX>(ring, img, istride, iw, ih, bf, th, nk, pitch, bx, by, acc);
}

// One block into the C' map: the W waves' partial sums (each a sum of exact
// int8 products, bounded like the whole) are added through LDS into wave 0's,
// which stores the block.
template <int W, int KSTEP, bool VEC, int CMAX>
__device__ __forceinline__ void direct_block_w(uint8_t* ring, const uint8_t* __restrict__ img, int64_t istride,
                                               int iw, int ih, const int8_t* __restrict__ bf, int th, int nk,
                                               int pitch, int rw, int rh, int32_t* __restrict__ out, int bx, int by) {
  constexpr int T = DIR_T;
  dv16i acc[T];
  direct_accumulate<W, KSTEP, VEC, CMAX>(ring, img, istride, iw, ih, bf, th, nk, pitch, bx, by, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if constexpr (W > 1) {
    int32_t* red = (int32_t*)ring;  // (W - 1) x 16 x 64 words, lane-minor; the ring is free
#pragma unroll
    for (int t = 0; t < T; t++) {
      if (wave > 0)
#pragma unroll
        for (int j = 0; j < 16; j++) red[((wave - 1) * 16 + j) * 64 + lane] = acc[t][j];
      __syncthreads();
      if (wave == 0)
        for (int w = 1; w < W; w++)
#pragma unroll
          for (int j = 0; j < 16; j++) acc[t][j] += red[((w - 1) * 16 + j) * 64 + lane];
      __syncthreads();
    }
  }
  if (wave != 0) return;
  const int half = lane >> 5, m = lane & 31;
  const int X0 = bx * 32 * T, yb = by * 32 + 4 * half;
#pragma unroll
  for (int t = 0; t < T; t++) {
    const int x = X0 + 32 * t + m;
    if (x < rw) {
#pragma unroll
      for (int j = 0; j < 16; j++) {
        const int y = yb + 8 * (j >> 2) + (j & 3);
        if (y < rh) out[(int64_t)y * rw + x] = acc[t][j];
      }
    }
  }
}

template <int W, int KSTEP, bool VEC, int CMAX>
__global__ __launch_bounds__(64 * W) CVM_DIRECT_OCC void k_direct_i8(const uint8_t* __restrict__ img, int64_t istride, int iw,
                                                      int ih, const int8_t* __restrict__ bf, int th, int nk,
                                                      int pitch, int rw, int rh, int32_t* __restrict__ out) {
  extern __shared__ uint4 ring4[];
  direct_block_w<W, KSTEP, VEC, CMAX>((uint8_t*)ring4, img, istride, iw, ih, bf, th, nk, pitch, rw, rh, out,
                                      blockIdx.x, blockIdx.y);
}

// Many templates against one frame in one launch (a Fleet's members, all of
// one (KSTEP, CMAX) form): workgroup b runs block b - job.wg0 of job wg_job[b].
template <int W, int KSTEP, bool VEC, int CMAX>
__global__ __launch_bounds__(64 * W) CVM_DIRECT_OCC void k_direct_batch(const uint8_t* __restrict__ img, int64_t istride, int iw,
                                                         int ih, const DirectJob* __restrict__ jobs,
                                                         const int* __restrict__ wg_job) {
  extern __shared__ uint4 ring4[];
  const DirectJob J = jobs[wg_job[blockIdx.x]];
  const int local = blockIdx.x - J.wg0;
  direct_block_w<W, KSTEP, VEC, CMAX>((uint8_t*)ring4, img, istride, iw, ih, J.bf, J.th, J.nk, J.pitch, J.rw, J.rh,
                                      J.out, local % J.gx, local / J.gx);
}

int direct_pitch_of(int tw, int kstep) { return direct_pitch(direct_nk(tw, kstep), kstep); }

int direct_grid_x(int rw) { return (rw + 32 * DIR_T - 1) / (32 * DIR_T); }
int direct_grid_y(int rh) { return (rh + DIR_ROWS - 1) / DIR_ROWS; }

// The instantiated CMAX that covers ceil(nk / kstep), or 0 past DIR_CMAX.
int direct_cmax(int nk, int kstep) {
  const int need = (nk + kstep - 1) / kstep;
  for (int c : {2, 3, 4, 5, 6, DIR_CMAX})
    if (need <= c) return c;
  return 0;
}

// LDS of one workgroup of `waves` waves on rows of `pitch` bytes: the ring,
// and for several waves also the partial sums added at the end.
static size_t direct_lds_of(int pitch, int waves) {
  if (waves == 1) return (size_t)DIR_RING * pitch + DIR_SLACK;
  return std::max((size_t)direct_ring_rows(waves) * pitch + DIR_SLACK,
                  (size_t)(waves - 1) * 16 * 64 * sizeof(int32_t));
}

size_t direct_lds_bytes(int tw, int kstep, int waves) {
  return direct_lds_of(direct_pitch(direct_nk(tw, kstep), kstep), waves);
}

bool direct_waves_ok(int waves) { return waves == 1 || waves == 2 || waves == 4; }

bool direct_shape_ok(int tw, int kstep) {
  const int nk = direct_nk(tw, kstep);
  return direct_cmax(nk, kstep) != 0 && direct_units(nk, kstep) <= 2 * 64 && direct_lds_bytes(tw, kstep, 1) <= 160 * 1024;
}

static bool direct_vec(const uint8_t* img, int64_t istride, int iw, int kstep) {
  const int sstep = kstep == 1 ? 1 : 4;
  return ((((uintptr_t)img) | (uintptr_t)istride | (uintptr_t)(sstep * iw)) & 15) == 0;
}

// dispatch over (waves, kstep, cmax, vec) = key waves * 1000 + kstep * 100 + cmax * 2 + vec
#define CVM_DIRECT_KEYS(LAUNCH, W)                                                                         \
  LAUNCH(W, 1, 2) LAUNCH(W, 1, 3) LAUNCH(W, 1, 4) LAUNCH(W, 1, 5) LAUNCH(W, 1, 6) LAUNCH(W, 1, 9)          \
  LAUNCH(W, 3, 2) LAUNCH(W, 3, 3) LAUNCH(W, 3, 4) LAUNCH(W, 3, 5) LAUNCH(W, 3, 6) LAUNCH(W, 3, 9)          \
  LAUNCH(W, 4, 2) LAUNCH(W, 4, 3) LAUNCH(W, 4, 4) LAUNCH(W, 4, 5) LAUNCH(W, 4, 6) LAUNCH(W, 4, 9)
#define CVM_DIRECT_SWITCH(LAUNCH)                                                      \
  switch (waves * 1000 + kstep * 100 + cmax * 2 + (vec ? 1 : 0)) {                     \
    CVM_DIRECT_KEYS(LAUNCH, 1) CVM_DIRECT_KEYS(LAUNCH, 2) CVM_DIRECT_KEYS(LAUNCH, 4) \
    default:                                                                           \
      return -1;                                                                       \
  }

int launch_direct(const uint8_t* img, int64_t istride, int iw, int ih, int kstep, const int8_t* bf, int tw, int th,
                  int rw, int rh, int32_t* out, int waves, hipStream_t s) {
  if (!direct_shape_ok(tw, kstep) || !direct_waves_ok(waves)) return -1;
  const int nk = direct_nk(tw, kstep);
  const int pitch = direct_pitch(nk, kstep);
  const int cmax = direct_cmax(nk, kstep);
  const size_t lds = direct_lds_of(pitch, waves);
  if (lds > 160 * 1024) return -1;
  const bool vec = direct_vec(img, istride, iw, kstep);
  dim3 grid(direct_grid_x(rw), direct_grid_y(rh)), block(64 * waves);
#define CVM_DI(W, KS, CM)                                                                                       \
  case W * 1000 + KS * 100 + CM * 2 + 1:                                                                        \
    hipLaunchKernelGGL((k_direct_i8<W, KS, true, CM>), grid, block, lds, s, img, istride, iw, ih, bf, th, nk,   \
                       pitch, rw, rh, out);                                                                     \
    break;                                                                                                      \
  case W * 1000 + KS * 100 + CM * 2:                                                                            \
    hipLaunchKernelGGL((k_direct_i8<W, KS, false, CM>), grid, block, lds, s, img, istride, iw, ih, bf, th, nk,  \
                       pitch, rw, rh, out);                                                                     \
    break;
  CVM_DIRECT_SWITCH(CVM_DI)
#undef CVM_DI
  return 0;
}

int launch_direct_batch(const uint8_t* img, int64_t istride, int iw, int ih, int kstep, int cmax,
                        const DirectJob* jobs, const int* wg_job, int nwg, int max_pitch, int waves, hipStream_t s) {
  if (!direct_waves_ok(waves)) return -1;
  const size_t lds = direct_lds_of(max_pitch, waves);
  if (lds > 160 * 1024 || nwg <= 0) return -1;
  const bool vec = direct_vec(img, istride, iw, kstep);
#define CVM_DB(W, KS, CM)                                                                                           \
  case W * 1000 + KS * 100 + CM * 2 + 1:                                                                            \
    hipLaunchKernelGGL((k_direct_batch<W, KS, true, CM>), dim3(nwg), dim3(64 * W), lds, s, img, istride, iw, ih,   \
                       jobs, wg_job);                                                                               \
    break;                                                                                                          \
  case W * 1000 + KS * 100 + CM * 2:                                                                                \
    hipLaunchKernelGGL((k_direct_batch<W, KS, false, CM>), dim3(nwg), dim3(64 * W), lds, s, img, istride, iw, ih,  \
                       jobs, wg_job);                                                                               \
    break;
  CVM_DIRECT_SWITCH(CVM_DB)
#undef CVM_DB
  return 0;
}

#undef CVM_DIRECT_KEYS
#undef CVM_DIRECT_SWITCH


// ------------------------------------------- fused correlation and scoring
//
// One workgroup correlates its 32 x 128 output block on the MFMA (as above),
// parks the block's C' values in LDS and scores them there, so no C' map
// reaches HBM: a member costs one extrema partial per block.  The scoring is
// the wave-form exact normalisation (window sums from the 16-row chunk
// prefix sums of the frame, slid down the rows in registers, horizontal sums
// from a DPP prefix scan) followed by the FP64 score rounded once to float32,
// op for op the unfused k_norm_batch in its exact mode, so every score and
// extremum is bitwise the same.  Scoring waves: one wave takes all 32 rows
// x 128 columns (Q = 2 results per lane); two waves take 16 rows each; four
// waves take 16 rows x 64 columns each (Q = 1).  Each wave keeps its own
// prefix rows, so a scoring wave synchronises only with itself.

constexpr int FB_PITCH = 32 * DIR_T + 8;  // C' words per block row in LDS (rows r and r + 4 on other banks)
constexpr int FB_CBYTES = 32 * FB_PITCH * 4;
constexpr int FB_CHUNK = 16;  // rows per chunk of the frame's prefix column sums (the normalisation's band)

__host__ __device__ constexpr int fused_nwp(int cpt) { return 64 * cpt + 1 + (64 * cpt + 1) / 32 + 1; }
__host__ __device__ constexpr int fused_cpt(int cmax) { return cmax >= 6 ? 8 : 4; }
__host__ __device__ constexpr int fused_nch(int kstep) { return kstep == 1 ? 1 : (kstep == 3 ? 3 : 4); }
// LDS bytes of one scoring wave's prefix rows (NCH channel rows + the square row)
__host__ __device__ constexpr int fused_pbytes(int cpt, int nch, bool sq32) {
  return fused_nwp(cpt) * (4 * nch + (sq32 ? 4 : 8));
}

// CPT consecutive pixels x0.. of one row as packed channel words (channel k
// in bits 8k..), all inside the row: one 16-byte load per four RGBA pixels,
// one 4- or 8-byte load of gray bytes.
template <int CPT, int STEP>
__device__ __forceinline__ void fused_px(const uint8_t* row, int x0, uint32_t (&out)[CPT]) {
  if constexpr (STEP == 4) {
#pragma unroll
    for (int h = 0; h < CPT / 4; h++) {
      const uint4 a = *(const uint4*)(row + (int64_t)(x0 + 4 * h) * 4);
      out[4 * h] = a.x;
      out[4 * h + 1] = a.y;
      out[4 * h + 2] = a.z;
      out[4 * h + 3] = a.w;
    }
  } else if constexpr (CPT == 4) {
    const uint32_t v = *(const uint32_t*)(row + x0);
#pragma unroll
    for (int j = 0; j < 4; j++) out[j] = (v >> (8 * j)) & 255u;
  } else {
    const uint2 v = *(const uint2*)(row + x0);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      out[j] = (v.x >> (8 * j)) & 255u;
      out[4 + j] = (v.y >> (8 * j)) & 255u;
    }
  }
}

template <int CPT>
__device__ __forceinline__ void fused_u32(const uint32_t* p, uint32_t (&out)[CPT]) {
#pragma unroll
  for (int h = 0; h < CPT / 4; h++) {
    const uint4 a = ((const uint4*)p)[h];
    out[4 * h] = a.x;
    out[4 * h + 1] = a.y;
    out[4 * h + 2] = a.z;
    out[4 * h + 3] = a.w;
  }
}

template <int Q, int CPT>
struct FusedRow {
  uint32_t a[CPT], b[CPT];
  int32_t c[Q];
};

// Scores of result rows [ys, ye) x columns [xs, xs + 64 Q) of one member by
// one wave, folded into (bmin, bmax).  cb: the C' value of (ys, xs) in LDS,
// rows FB_PITCH words apart.  Requires iw % CPT == 0, xs % CPT == 0,
// ys % FB_CHUNK == 0, ys < ye <= rh and 64 CPT >= 64 Q + tw - 1.
template <int Q, int CPT, int NCH, int STEP, typename sq_t>
__device__ __forceinline__ void fused_score(const uint8_t* __restrict__ img, int64_t istride, int iw, int ih, int tw,
                                            int th, int rw, const CallState* __restrict__ st,
                                            const uint32_t* __restrict__ ps, const int32_t* cb, int xs, int ys,
                                            int ye, uint32_t* pl, float& bmin, int& bmin_i, float& bmax,
                                            int& bmax_i) {
  constexpr int NWP = fused_nwp(CPT);
  constexpr bool SQ32 = sizeof(sq_t) == 4;
  uint32_t(*P)[NWP] = (uint32_t(*)[NWP])pl;
  sq_t* P2 = (sq_t*)(pl + NCH * NWP);
  const int lane = threadIdx.x & 63;
  const int cn = st->cn;
  const double aread = (double)tw * (double)th;
  double tsd[NCH];
#pragma unroll
  for (int k = 0; k < NCH; k++) tsd[k] = k < cn ? (double)st->tsum[k] : 0.0;
  const double varSumD = (double)st->varSum;
  double tsumAll = tsd[0];
#pragma unroll
  for (int k = 1; k < NCH; k++) tsumAll += tsd[k];
  // sum P*T = C' + 128 (S + sum_k T_k) - 2^14 area cn
  const double cpOff = 16384.0 * aread * (double)cn;
  const int cbase = lane * CPT;
  // lanes past the image load its last CPT columns: whole lanes only (iw and
  // xs are multiples of CPT), and their prefix entries lie past every one a
  // result of the strip reads
  const int x0 = min(xs + cbase, iw - CPT);
  const uint32_t cm = chan_mask(cn);

  // column sums over the th window rows, channels packed in 16-bit halves:
  // cs02 = ch0 | ch2 << 16, cs13 = ch1 | ch3 << 16 (each <= 255 * th < 2^16)
  uint32_t cs02[CPT], cs13[CPT];
  sq_t c2[CPT];
  {
    const int c0 = ys / FB_CHUNK, nfull = th / FB_CHUNK;
    const uint32_t* p0 = ps + (int64_t)c0 * 3 * iw + x0;
    const uint32_t* p1 = ps + (int64_t)(c0 + nfull) * 3 * iw + x0;
    uint32_t a0[CPT], a1[CPT], a2[CPT], b0[CPT], b1[CPT], b2[CPT];
    fused_u32<CPT>(p0, a0);
    fused_u32<CPT>(p0 + iw, a1);
    fused_u32<CPT>(p0 + 2 * iw, a2);
    fused_u32<CPT>(p1, b0);
    fused_u32<CPT>(p1 + iw, b1);
    fused_u32<CPT>(p1 + 2 * iw, b2);
#pragma unroll
    for (int j = 0; j < CPT; j++) {
      cs02[j] = b0[j] - a0[j];
      cs13[j] = b1[j] - a1[j];
      c2[j] = (sq_t)(uint32_t)(b2[j] - a2[j]);
    }
    for (int y = ys + nfull * FB_CHUNK; y < ys + th; y++) {
      uint32_t w[CPT];
      fused_px<CPT, STEP>(img + (int64_t)y * istride, x0, w);
#pragma unroll
      for (int j = 0; j < CPT; j++) {
        const uint32_t v = w[j] & cm;
        cs02[j] += v & 0x00FF00FFu;
        cs13[j] += (v >> 8) & 0x00FF00FFu;
        c2[j] += (sq_t)__builtin_amdgcn_udot4(v, v, 0u, false);
      }
    }
  }

  // row yy's inputs (rows past the band clamped into the image: loaded, not used)
  auto fetch = [&](FusedRow<Q, CPT>& r, int yy) {
    const int yl = min(yy, ye - 1);
    const int ya = min(yl + th, ih - 1);
    fused_px<CPT, STEP>(img + (int64_t)ya * istride, x0, r.a);
    fused_px<CPT, STEP>(img + (int64_t)yl * istride, x0, r.b);
#pragma unroll
    for (int q = 0; q < Q; q++) r.c[q] = cb[(yl - ys) * FB_PITCH + lane + 64 * q];
  };

  auto row = [&](const FusedRow<Q, CPT>& r, int y) {
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
    if constexpr (SQ32) {
      off2 = (sq_t)wave_incl_scan_u32((uint32_t)run2, lane) - run2;
    } else {
      sq_t inc2 = run2;
      for (int o = 1; o < 64; o <<= 1) {
        const sq_t up2 = __shfl_up(inc2, o, 64);
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
    wave_lds_sync();
    float vq[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int xl = lane + 64 * q, xb = xl + tw;
      const int ia = xl + (xl >> 5), ib = xb + (xb >> 5);
      // exact integers carried in binary64 (all below 2^53)
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
      double S = sd[0];
#pragma unroll
      for (int k = 1; k < NCH; k++) S += sd[k];
      const double corrI = __builtin_fma(128.0, S + tsumAll, (double)r.c[q] - cpOff);
      vq[q] = (float)score_fp64(__builtin_fma(aread, corrI, -lane0), idiff, s2, aread, varSumD);
    }
    wave_lds_sync();
    // slide to row y+1: the leaving row is inside the current sums, so each
    // 16-bit half stays in [0, 2^16) and no borrow crosses halves
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

  // two register sets in turn: row y+1's loads are in flight while row y is scored
  FusedRow<Q, CPT> ra, rb;
  fetch(ra, ys);
  for (int y = ys; y < ye; y += 2) {
    fetch(rb, y + 1);
    row(ra, y);
    if (y + 1 >= ye) break;
    fetch(ra, y + 2);
    row(rb, y + 1);
  }
}

// Many members against one frame, correlated and scored in one launch (all
// of one (KSTEP, CMAX, SQ32) form): workgroup b runs block b - job.wg0 of job
// wg_job[b] and writes that block's extrema partial.  The frame's rows are
// 16-byte aligned (direct_vec) and iw % CPT == 0.
template <int W, int KSTEP, int CMAX, bool SQ32>
__global__ __launch_bounds__(64 * W) CVM_DIRECT_OCC void k_direct_fused(const uint8_t* __restrict__ img, int64_t istride, int iw,
                                                         int ih, const uint32_t* __restrict__ ps,
                                                         const FusedJob* __restrict__ jobs,
                                                         const int* __restrict__ wg_job) {
  typedef typename std::conditional<SQ32, uint32_t, unsigned long long>::type sq_t;
  constexpr int T = DIR_T;
  constexpr int CPT = fused_cpt(CMAX), NCH = fused_nch(KSTEP), STEP = KSTEP == 1 ? 1 : 4;
  constexpr int Q = W == 4 ? 1 : 2;
  constexpr int PB = fused_pbytes(CPT, NCH, SQ32);
  extern __shared__ uint4 lds4[];
  __shared__ Extrema wred[W];
  uint8_t* lds = (uint8_t*)lds4;
  const FusedJob J = jobs[wg_job[blockIdx.x]];
  const int local = blockIdx.x - J.wg0;
  const int bx = local % J.gx, by = local / J.gx;
  const int X0 = bx * 32 * T, Y0 = by * 32;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const CallState* st = J.st;
  if (st->flat) {
    // templNorm below the flat threshold: every score is 1
    if (tid == 0) {
      Extrema e;
      e.min_v = e.max_v = 1.0f;
      e.min_i = e.max_i = Y0 * J.rw + X0;
      J.parts[local] = e;
    }
    return;
  }

  dv16i acc[T];
  direct_accumulate<W, KSTEP, true, CMAX>(lds, img, istride, iw, ih, J.bf, J.th, J.nk, J.pitch, bx, by, acc);
  __syncthreads();  // the ring is no longer read
  int32_t* cbuf = (int32_t*)lds;
  {
    const int half = lane >> 5, m = lane & 31;
    // wave 0 stores its sum, the other waves then add their partial sums
    for (int w = 0; w < W; w++) {
      if (wave == w)
#pragma unroll
        for (int t = 0; t < T; t++)
#pragma unroll
          for (int j = 0; j < 16; j++) {
            int32_t* c = cbuf + (8 * (j >> 2) + 4 * half + (j & 3)) * FB_PITCH + 32 * t + m;
            if (w == 0)
              *c = acc[t][j];
            else
              atomicAdd(c, acc[t][j]);
          }
      if (w == 0 || w == W - 1) __syncthreads();
    }
  }

  float bmin = FLT_MAX, bmax = -FLT_MAX;
  int bmin_i = INT_MAX, bmax_i = INT_MAX;
  {
    const int ys = Y0 + (W == 1 ? 0 : 16 * (wave & 1));
    const int xs = X0 + (W == 4 ? 64 * (wave >> 1) : 0);
    const int ye = min(J.rh, ys + (W == 1 ? 32 : 16));
    uint32_t* pl = (uint32_t*)(lds + FB_CBYTES + wave * PB);
    if (ys < ye && xs < J.rw)
      fused_score<Q, CPT, NCH, STEP, sq_t>(img, istride, iw, ih, J.tw, J.th, J.rw, st, ps,
                                           cbuf + (ys - Y0) * FB_PITCH + (xs - X0), xs, ys, ye, pl, bmin, bmin_i,
                                           bmax, bmax_i);
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
  if (W > 1) {
    if (lane == 0) {
      wred[wave].min_v = bmin;
      wred[wave].min_i = bmin_i;
      wred[wave].max_v = bmax;
      wred[wave].max_i = bmax_i;
    }
    __syncthreads();
    if (tid == 0)
      for (int w = 1; w < W; w++) {
        if (better_min(wred[w].min_v, wred[w].min_i, bmin, bmin_i)) {
          bmin = wred[w].min_v;
          bmin_i = wred[w].min_i;
        }
        if (better_max(wred[w].max_v, wred[w].max_i, bmax, bmax_i)) {
          bmax = wred[w].max_v;
          bmax_i = wred[w].max_i;
        }
      }
  }
  if (tid == 0) {
    Extrema e;
    e.min_v = bmin;
    e.min_i = bmin_i;
    e.max_v = bmax;
    e.max_i = bmax_i;
    J.parts[local] = e;
  }
}

// CPT of the fused scoring for a chunk count, and whether a (kstep, cmax,
// sq32) form is built: the 64-bit square sums only where a window's square
// sum can pass 2^32 (cn * 65025 * area >= 2^32 needs tw >= 86 at cn 3,
// tw >= 65 at cn 4; never at cn 1 within th <= 257).
int fused_cpt_of(int cmax) { return fused_cpt(cmax); }

static bool fused_form_built(int kstep, int cmax, bool sq32) {
  if (sq32) return true;
  if (kstep == 3) return cmax >= 4;
  if (kstep == 4) return cmax >= 3;
  return false;
}

size_t fused_lds_bytes(int max_pitch, int kstep, int cmax, bool sq32, int waves) {
  const size_t ring = (size_t)direct_ring_rows(waves) * max_pitch + DIR_SLACK;
  const size_t score = (size_t)FB_CBYTES + (size_t)waves * fused_pbytes(fused_cpt(cmax), fused_nch(kstep), sq32);
  return std::max(ring, score);
}

bool fused_ok(const uint8_t* img, int64_t istride, int iw, int kstep, int cmax, bool sq32) {
  return fused_form_built(kstep, cmax, sq32) && iw % fused_cpt(cmax) == 0 && direct_vec(img, istride, iw, kstep);
}

int launch_direct_fused(const uint8_t* img, int64_t istride, int iw, int ih, const uint32_t* ps, int kstep, int cmax,
                        bool sq32, const FusedJob* jobs, const int* wg_job, int nwg, int max_pitch, int waves,
                        hipStream_t s) {
  if (!direct_waves_ok(waves) || nwg <= 0 || !fused_ok(img, istride, iw, kstep, cmax, sq32)) return -1;
  const size_t lds = fused_lds_bytes(max_pitch, kstep, cmax, sq32, waves);
  if (lds + 64 > 160 * 1024) return -1;
#define CVM_DF(W, KS, CM, SQ)                                                                                    \
  case W * 10000 + KS * 1000 + CM * 10 + SQ:                                                                    \
    hipLaunchKernelGGL((k_direct_fused<W, KS, CM, SQ == 1>), dim3(nwg), dim3(64 * W), lds, s, img, istride, iw, \
                       ih, ps, jobs, wg_job);                                                                   \
    break;
#define CVM_DF_W(W)                                                                                         \
  CVM_DF(W, 1, 2, 1) CVM_DF(W, 1, 3, 1) CVM_DF(W, 1, 4, 1) CVM_DF(W, 1, 5, 1) CVM_DF(W, 1, 6, 1)              \
  CVM_DF(W, 1, 9, 1) CVM_DF(W, 3, 2, 1) CVM_DF(W, 3, 3, 1) CVM_DF(W, 3, 4, 1) CVM_DF(W, 3, 4, 0)              \
  CVM_DF(W, 3, 5, 1) CVM_DF(W, 3, 5, 0) CVM_DF(W, 3, 6, 1) CVM_DF(W, 3, 6, 0) CVM_DF(W, 3, 9, 1)              \
  CVM_DF(W, 3, 9, 0) CVM_DF(W, 4, 2, 1) CVM_DF(W, 4, 3, 1) CVM_DF(W, 4, 3, 0) CVM_DF(W, 4, 4, 1)              \
  CVM_DF(W, 4, 4, 0) CVM_DF(W, 4, 5, 1) CVM_DF(W, 4, 5, 0) CVM_DF(W, 4, 6, 1) CVM_DF(W, 4, 6, 0)              \
  CVM_DF(W, 4, 9, 1) CVM_DF(W, 4, 9, 0)
  switch (waves * 10000 + kstep * 1000 + cmax * 10 + (sq32 ? 1 : 0)) {
    CVM_DF_W(1)
    CVM_DF_W(2)
    CVM_DF_W(4)
    default:
      return -1;
  }
#undef CVM_DF_W
#undef CVM_DF
  return 0;
}

}  // namespace cvm
