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
// K runs over the raw bytes of one image row segment, channels interleaved:
// for template row i, A_i[m][u] = P'[y0 + m + i][step*x0 + u] and
//   B_i[u][n] = T'[i][u - step*n]   when 0 <= u - step*n < step*tw and the
//                                   byte's channel is below cn, else 0,
// a banded Toeplitz matrix that is the same for every tile.  Summing the
// th products A_i B_i into one accumulator gives the 32 x 32 block.  B_i is
// built once per template in MFMA fragment order (k_bfrag).  The A fragment
// of column tile t and K chunk kc is chunk kc + step*t of the row segment, so
// a wave walking kc in steps of `step` slides a window of T A fragments by one
// load per T MFMAs.  Image rows stream through an LDS ring, one new row per
// template row.
//
// v_mfma_i32_32x32x32_i8 lane layout: A byte b of lane l is A[l%32][16(l/32)+b],
// B byte b of lane l is B[16(l/32)+b][l%32], and accumulator register j of
// lane l is D[8(j/4) + 4(l/32) + j%4][l%32].
#include "cvm_internal.h"

namespace cvm {

typedef int dv4i __attribute__((ext_vector_type(4)));
typedef int dv16i __attribute__((ext_vector_type(16)));

constexpr int DIR_T = 4;      // column tiles per wave (128 output columns)
constexpr int DIR_WAVES = 2;  // waves per workgroup, 32 output rows each
constexpr int DIR_ROWS = 32 * DIR_WAVES;
constexpr int DIR_RING = DIR_ROWS + 8;
constexpr int DIR_CMAX = 9;   // K chunks per residue class: nk <= step * DIR_CMAX
constexpr int DIR_PF = 4;     // ring rows in flight (<= DIR_RING - DIR_ROWS)

int direct_nk(int tw, int step) { return (step * (32 + tw - 1) + 31) / 32; }

// B fragments plus DIR_CMAX of padding: a class load may run past the last
// fragment (those registers are never used).
size_t direct_bfrag_bytes(int tw, int th, int step) {
  return ((size_t)th * direct_nk(tw, step) + DIR_CMAX) * 64 * 16;
}

static int direct_pitch(int nk, int step) {
  // staged row bytes, padded so consecutive rows start 16 bytes (4 banks)
  // apart modulo 128: the 32 rows of one A fragment spread over all banks
  int bytes = 32 * (nk + step * (DIR_T - 1));
  return (bytes + 127) / 128 * 128 + 16;
}

size_t direct_lds_bytes(int tw, int step) { return (size_t)DIR_RING * direct_pitch(direct_nk(tw, step), step); }

// Visit order of the K chunks within one template row: residue classes of
// kc modulo step, each in increasing kc (the A window slides inside a class).
__host__ __device__ inline void direct_visit(int s, int nk, int step, int& rc, int& kc) {
  for (rc = 0; rc < step; rc++) {
    const int c = (nk - rc + step - 1) / step;
    if (s < c) {
      kc = rc + step * s;
      return;
    }
    s -= c;
  }
  kc = 0;
}

// B_i fragments in visit order: bf[((i * nk + s) * 64 + lane) * 16 + b] =
// B_i[32 kc + 16 (lane/32) + b][lane % 32] with kc the s-th visited chunk.
__global__ void k_bfrag(const uint8_t* __restrict__ tpl, int64_t tstride, int tw, int th, int step, int cn,
                        const CallState* __restrict__ st, int nk, int8_t* __restrict__ bf) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // one 16-byte fragment slice
  const int64_t total = (int64_t)th * nk * 64;
  if (id >= total) return;
  if (st) cn = st->cn;
  const int lane = (int)(id & 63);
  const int64_t is = id >> 6;
  const int sv = (int)(is % nk), i = (int)(is / nk);
  int rc, kc;
  direct_visit(sv, nk, step, rc, kc);
  const int n = lane & 31, h = lane >> 5;
  uint32_t w[4] = {0, 0, 0, 0};
  for (int b = 0; b < 16; b++) {
    const int u = 32 * kc + 16 * h + b;
    const int d = u - step * n;
    uint32_t v = 0;
    if (d >= 0 && d < step * tw && (d % step) < cn) v = (uint32_t)(tpl[(int64_t)i * tstride + d] ^ 0x80u) & 255u;
    w[b >> 2] |= v << (8 * (b & 3));
  }
  uint4* o = (uint4*)(bf + id * 16);
  *o = make_uint4(w[0], w[1], w[2], w[3]);
}

void launch_bfrag(const uint8_t* tpl, int64_t tstride, int tw, int th, int step, int cn, const CallState* st,
                  int8_t* bf, hipStream_t s) {
  int nk = direct_nk(tw, step);
  int64_t total = (int64_t)th * nk * 64;
  hipLaunchKernelGGL(k_bfrag, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, tpl, tstride, tw, th, step, cn,
                     st, nk, bf);
}

// One workgroup: output rows [Y0, Y0 + DIR_ROWS) x columns [X0, X0 + 32 T).
// VEC: image rows 16-byte aligned (16-byte staging loads).
template <int STEP, bool VEC>
__global__ __launch_bounds__(64 * DIR_WAVES) void k_direct_i8(const uint8_t* __restrict__ img, int64_t istride,
                                                              int iw, int ih, const int8_t* __restrict__ bf, int th,
                                                              int nk, int pitch, int rw, int rh,
                                                              int32_t* __restrict__ out) {
  constexpr int T = DIR_T;
  extern __shared__ uint4 ring4[];
  uint8_t* ring = (uint8_t*)ring4;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int X0 = blockIdx.x * 32 * T, Y0 = blockIdx.y * DIR_ROWS;
  const int rowChunks = 2 * (nk + STEP * (T - 1));  // 16-byte chunks per staged row
  const int64_t xb0 = (int64_t)STEP * X0;
  const int64_t imgBytes = (int64_t)STEP * iw;

  // 16 bytes of staged row rho (image row Y0 + rho), chunk c.  VEC: the
  // caller guarantees 16-byte alignment and STEP*iw % 16 == 0, so a chunk is
  // wholly inside or wholly outside the image; the load is clamped into the
  // image and `ok` says whether to keep it.  The int8 shift (v ^ 0x80) and
  // the zeroing are applied when the chunk is written (cook), so the load
  // itself is never waited on early.
  auto fetch = [&](int rho, int c, bool& ok) -> uint4 {
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
        if (b0 + b < imgBytes) w[b >> 2] = (w[b >> 2] & ~(255u << (8 * (b & 3)))) | ((uint32_t)p[b0 + b] << (8 * (b & 3)));
    }
    return make_uint4(w[0], w[1], w[2], w[3]);
  };
  auto cook = [](uint4 v, bool ok) -> uint4 {
    if (!ok) return make_uint4(0, 0, 0, 0);
    return make_uint4(v.x ^ 0x80808080u, v.y ^ 0x80808080u, v.z ^ 0x80808080u, v.w ^ 0x80808080u);
  };

  // initial rows [0, DIR_ROWS)
  for (int e = tid; e < DIR_ROWS * rowChunks; e += 64 * DIR_WAVES) {
    const int rho = e / rowChunks, c = e - rho * rowChunks;
    bool ok;
    const uint4 v = fetch(rho, c, ok);
    *(uint4*)(ring + rho * pitch + 16 * c) = cook(v, ok);
  }
  __syncthreads();

  dv16i acc[T];
#pragma unroll
  for (int t = 0; t < T; t++)
#pragma unroll
    for (int j = 0; j < 16; j++) acc[t][j] = 0;

  const int half = lane >> 5, m = lane & 31;
  // Per template row, K chunk residue class rc has c_rc = ceil((nk - rc) / STEP)
  // chunks; its c_rc + T - 1 A fragments and c_rc B fragments sit in
  // statically indexed registers, the next class's loads issued before this
  // class's MFMAs, the next row's first B class before the barrier.
  constexpr int NA = DIR_CMAX + T - 1;
  const dv4i* bq = (const dv4i*)bf + lane;
  int cls[STEP], base[STEP];
  {
    int acc_s = 0;
#pragma unroll
    for (int rc = 0; rc < STEP; rc++) {
      cls[rc] = rc < nk ? (nk - rc + STEP - 1) / STEP : 0;
      base[rc] = acc_s;
      acc_s += cls[rc];
    }
  }
  dv4i A0[NA], A1[NA], B0[DIR_CMAX], B1[DIR_CMAX], BN[DIR_CMAX];
  // unconditional loads (a fixed count the compiler can wait on exactly):
  // A chunk indices clamp to the staged row, B runs into the padding
  const int lastChunk = nk + STEP * (T - 1) - 1;
  auto loadA = [&](dv4i* buf, const uint8_t* arow, int rc) {
#pragma unroll
    for (int j = 0; j < NA; j++) buf[j] = *(const dv4i*)(arow + 32 * min(rc + STEP * j, lastChunk));
  };
  auto loadB = [&](dv4i* buf, int i, int rc) {
    const dv4i* p = bq + ((int64_t)i * nk + base[rc]) * 64;
#pragma unroll
    for (int j = 0; j < DIR_CMAX; j++) buf[j] = p[(int64_t)j * 64];
  };
  auto mma = [&](const dv4i* a, const dv4i* bb, int c) {
#pragma unroll
    for (int j = 0; j < DIR_CMAX; j++)
      if (j < c)
#pragma unroll
        for (int t = 0; t < T; t++) acc[t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[j + t], bb[j], acc[t], 0, 0, 0);
  };

  // Staged row DIR_ROWS + i (first read in step i + 1) is fetched in step
  // i - DIR_PF + 1 into a register queue and written to its ring slot at the
  // end of step i, so DIR_PF steps of MFMAs cover its load latency.
  // chunk nc[p] of each staged row is this thread's (chunks past the row
  // reload the last one and are not written)
  int nc[2];
#pragma unroll
  for (int p = 0; p < 2; p++) nc[p] = tid + p * 64 * DIR_WAVES;
  const int ncl[2] = {min(nc[0], rowChunks - 1), min(nc[1], rowChunks - 1)};
  uint4 rq[DIR_PF][2];
  bool rok[DIR_PF][2];
#pragma unroll
  for (int d = 0; d < DIR_PF - 1; d++)
#pragma unroll
    for (int p = 0; p < 2; p++) rq[d][p] = fetch(DIR_ROWS + d, ncl[p], rok[d][p]);

  loadB(BN, 0, 0);
  for (int i = 0; i < th; i++) {
    const bool more = i + 1 < th;
#pragma unroll
    for (int p = 0; p < 2; p++) rq[DIR_PF - 1][p] = fetch(DIR_ROWS + i + DIR_PF - 1, ncl[p], rok[DIR_PF - 1][p]);
    // this wave's A rows: staged rows 32 wave + m + i
    const int slot = (32 * wave + m + i) % DIR_RING;
    const uint8_t* arow = ring + slot * pitch + 16 * half;
#pragma unroll
    for (int j = 0; j < DIR_CMAX; j++) B0[j] = BN[j];
    loadA(A0, arow, 0);
#pragma unroll
    for (int rc = 0; rc < STEP; rc++) {
      dv4i* ca = (rc & 1) ? A1 : A0;
      dv4i* cb = (rc & 1) ? B1 : B0;
      dv4i* na = (rc & 1) ? A0 : A1;
      dv4i* nb = (rc & 1) ? B0 : B1;
      if (rc + 1 < STEP) {
        loadA(na, arow, rc + 1);
        loadB(nb, i, rc + 1);
      } else if (more) {
        loadB(BN, i + 1, 0);
      }
      mma(ca, cb, cls[rc]);
    }
    if (more) {
      // slot of staged row DIR_ROWS + i: it held row i + DIR_ROWS - DIR_RING,
      // last read in that step, which every wave has left (barrier below)
      const int ns = (DIR_ROWS + i) % DIR_RING;
#pragma unroll
      for (int p = 0; p < 2; p++)
        if (nc[p] < rowChunks) *(uint4*)(ring + ns * pitch + 16 * nc[p]) = cook(rq[0][p], rok[0][p]);
    }
#pragma unroll
    for (int d = 0; d < DIR_PF - 1; d++)
#pragma unroll
      for (int p = 0; p < 2; p++) {
        rq[d][p] = rq[d + 1][p];
        rok[d][p] = rok[d + 1][p];
      }
    __syncthreads();
  }

  const int yb = Y0 + 32 * wave + 4 * half;
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

int launch_direct(const uint8_t* img, int64_t istride, int iw, int ih, int step, const int8_t* bf, int tw, int th,
                  int rw, int rh, int32_t* out, hipStream_t s) {
  const int nk = direct_nk(tw, step);
  const int pitch = direct_pitch(nk, step);
  if (32 * (nk + step * (DIR_T - 1)) > pitch || 2 * 64 * DIR_WAVES < 2 * (nk + step * (DIR_T - 1)) ||
      nk > step * DIR_CMAX)
    return -1;
  const size_t lds = (size_t)DIR_RING * pitch;
  if (lds > 160 * 1024) return -1;
  const bool vec = ((((uintptr_t)img) | (uintptr_t)istride | (uintptr_t)(step * iw)) & 15) == 0;
  dim3 grid((rw + 32 * DIR_T - 1) / (32 * DIR_T), (rh + DIR_ROWS - 1) / DIR_ROWS), block(64 * DIR_WAVES);
#define CVM_DI(S, V) \
  hipLaunchKernelGGL((k_direct_i8<S, V>), grid, block, lds, s, img, istride, iw, ih, bf, th, nk, pitch, rw, rh, out)
  if (step == 4) {
    if (vec)
      CVM_DI(4, true);
    else
      CVM_DI(4, false);
  } else if (step == 1) {
    if (vec)
      CVM_DI(1, true);
    else
      CVM_DI(1, false);
  } else {
    return -1;
  }
#undef CVM_DI
  return 0;
}

}  // namespace cvm
