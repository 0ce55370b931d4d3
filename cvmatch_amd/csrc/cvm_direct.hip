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

int direct_nk(int tw, int step) { return (step * (32 + tw - 1) + 31) / 32; }

size_t direct_bfrag_bytes(int tw, int th, int step) { return (size_t)th * direct_nk(tw, step) * 64 * 16; }

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

  // 16 bytes of staged row rho (image row Y0 + rho), chunk c, as int8 (v ^ 0x80);
  // bytes outside the image are 0 (they only reach discarded outputs)
  auto fetch = [&](int rho, int c) -> uint4 {
    const int r = Y0 + rho;
    const int64_t b0 = xb0 + 16 * (int64_t)c;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (r < ih) {
      const uint8_t* p = img + (int64_t)r * istride;
      if (VEC && b0 + 16 <= imgBytes) {
        v = *(const uint4*)(p + b0);
        v.x ^= 0x80808080u;
        v.y ^= 0x80808080u;
        v.z ^= 0x80808080u;
        v.w ^= 0x80808080u;
      } else {
        uint32_t w[4] = {0, 0, 0, 0};
        for (int b = 0; b < 16; b++)
          if (b0 + b < imgBytes) w[b >> 2] |= (uint32_t)(p[b0 + b] ^ 0x80u) << (8 * (b & 3));
        v = make_uint4(w[0], w[1], w[2], w[3]);
      }
    }
    return v;
  };

  // initial rows [0, DIR_ROWS)
  for (int e = tid; e < DIR_ROWS * rowChunks; e += 64 * DIR_WAVES) {
    const int rho = e / rowChunks, c = e - rho * rowChunks;
    *(uint4*)(ring + rho * pitch + 16 * c) = fetch(rho, c);
  }
  __syncthreads();

  dv16i acc[T];
#pragma unroll
  for (int t = 0; t < T; t++)
#pragma unroll
    for (int j = 0; j < 16; j++) acc[t][j] = 0;

  const int half = lane >> 5, m = lane & 31;
  // B fragments stream in visit order through a 4-deep register queue that
  // runs across template rows
  const dv4i* bq = (const dv4i*)bf + lane;
  const int64_t gtot = (int64_t)th * nk;
  int64_t g = 0;
  dv4i q[4];
#pragma unroll
  for (int p = 0; p < 4; p++) q[p] = p < gtot ? bq[(int64_t)p * 64] : dv4i{0, 0, 0, 0};
  for (int i = 0; i < th; i++) {
    // the row step i + 1 adds (staged row DIR_ROWS + i), loaded now and
    // written after this step's MFMAs
    const bool more = i + 1 < th;
    uint4 nxt[2];
    int nc[2];
#pragma unroll
    for (int p = 0; p < 2; p++) {
      nc[p] = tid + p * 64 * DIR_WAVES;
      nxt[p] = make_uint4(0, 0, 0, 0);
      if (more && nc[p] < rowChunks) nxt[p] = fetch(DIR_ROWS + i, nc[p]);
    }
    // this wave's A rows: staged rows 32 wave + m + i
    const int slot = (32 * wave + m + i) % DIR_RING;
    const uint8_t* arow = ring + slot * pitch + 16 * half;
#pragma unroll
    for (int rc = 0; rc < STEP; rc++) {
      if (rc >= nk) break;
      dv4i aw[T];
#pragma unroll
      for (int t = 0; t < T; t++) aw[t] = *(const dv4i*)(arow + 32 * (rc + STEP * t));
      for (int kc = rc; kc < nk; kc += STEP) {
        const dv4i bcur = q[0];
#pragma unroll
        for (int p = 0; p < 3; p++) q[p] = q[p + 1];
        q[3] = (g + 4 < gtot) ? bq[(g + 4) * 64] : dv4i{0, 0, 0, 0};
        g++;
        dv4i anew = aw[T - 1];
        if (kc + STEP < nk) anew = *(const dv4i*)(arow + 32 * (kc + STEP * T));
#pragma unroll
        for (int t = 0; t < T; t++) acc[t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(aw[t], bcur, acc[t], 0, 0, 0);
#pragma unroll
        for (int t = 0; t < T - 1; t++) aw[t] = aw[t + 1];
        aw[T - 1] = anew;
      }
    }
    if (more) {
      // slot of staged row DIR_ROWS + i: it held row i - 8, last read in
      // step i - 8, which every wave has left (barrier below)
      const int ns = (DIR_ROWS + i) % DIR_RING;
#pragma unroll
      for (int p = 0; p < 2; p++)
        if (nc[p] < rowChunks) *(uint4*)(ring + ns * pitch + 16 * nc[p]) = nxt[p];
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
  if (32 * (nk + step * (DIR_T - 1)) > pitch || 2 * 64 * DIR_WAVES < 2 * (nk + step * (DIR_T - 1))) return -1;
  const size_t lds = (size_t)DIR_RING * pitch;
  if (lds > 160 * 1024) return -1;
  const bool vec = ((((uintptr_t)img) | (uintptr_t)istride) & 15) == 0;
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
