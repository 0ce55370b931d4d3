/*
 * cvm_oracle.c -- CPU ORACLE for the TM_CCOEFF_NORMED path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (cvmatch_amd/, the HIP
 * library) links, loads or calls this file; only tests/, the smoke check and
 * the bench's cpu_baseline leg do, and only as the checker.
 *
 * Two oracles live here:
 *
 *  1. A restatement of the reference's scalar ("purego") op sequence, every
 *     floating-point operation a single IEEE-754 binary32/binary64 rounding,
 *     no FMA (build with -ffp-contract=off -fno-fast-math).  It follows:
 *       sincospiFrac ............ impl.go:98-133
 *       makeBitrev .............. impl.go:148-165
 *       newGoPlan ............... impl.go:246-295
 *       forwardRowPair/blockForwardGo impl.go:301-374
 *       mulConjGo ............... impl.go:376-388
 *       inverseRowPair/blockInverseEmitGo impl.go:394-464
 *       buildTSpecSet/shrinkDH/shrinkInto impl.go:520-610
 *       crossCorrGo ............. impl.go:616-693
 *       colBuildGo/colSlideGo ... impl.go:708-805
 *       normOne ................. impl.go:825-849
 *       normalizeBandGo ......... impl.go:859-1054
 *       templStats .............. impl.go:1063-1079
 *       normalizeParallelGo ..... impl.go:1081-1114
 *       statsCap/matchU8 ........ impl.go:1122-1155
 *       makeTriTwiddles/mulPlain/fftR4/fftR4Scalar radix4.go:45-149
 *       colsR4Head/colsR4Pass/colsR4Go radix4cols.go:33-125
 *       alphaConst/rgbToGray .... cvmatch.go:146-168, 222-240
 *       Matcher (newMatcher/Find) matcher.go:66-130
 *       Fleet (bucketize/FindAll/buildGeom) fleet.go:111-390
 *     Worker threads (tile queue, row bands) follow the reference's own
 *     decomposition; the output does not depend on the thread count.
 *
 *  2. An FP64 oracle: exact int64 direct correlation and window statistics,
 *     then float64 normalisation with OpenCV's guard ladder, written from the
 *     definition the reference's brute-force test uses (cvmatch_test.go:13-81).
 *
 * Pinning (what checks each part) is listed in tests/test_oracle_*.py.
 */
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float re, im; } cpx;

typedef struct {
  float min_v;
  int32_t min_x, min_y;
  float max_v;
  int32_t max_x, max_y;
} orc_result;

enum { ORC_OK = 0, ORC_BAD_ARGS = 1, ORC_STATS_CAP = 2, ORC_NOMEM = 3, ORC_EMPTY_FLEET = 4, ORC_MIXED_FLEET = 5 };

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

static int next_pow2(int v) {
  int n = 1;
  while (n < v) n <<= 1;
  return n;
}

/* bits.Len(uint(n)) */
static int bit_len(unsigned long long n) {
  int l = 0;
  while (n) { l++; n >>= 1; }
  return l;
}

/* oddLog2 radix4.go:36 -- log2(n) odd for a power of two n */
static int odd_log2(int n) { return bit_len((unsigned)n) % 2 == 0; }

static int trailing_zeros(int n) {
  int t = 0;
  while (n > 1 && !(n & 1)) { n >>= 1; t++; }
  return t;
}

/* ------------------------------------------------------------ twiddles -- */

/* sincospiFrac impl.go:98-133: cos/sin(pi*j/half) from exact dyadic range
 * reduction and fdlibm-style kernels, each double op rounded once. */
void orc_sincospi(int j, int half, float *c_out, float *s_out) {
  int m = 2 * j;
  int k = m / half;
  int rem = m - k * half;
  double u = (double)rem / (double)half;
  int swap = 0;
  if (u > 0.5) {
    u = 1.0 - u;
    swap = 1;
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
  *c_out = (float)c;
  *s_out = (float)s;
}

/* number of complex entries in the stage-triplet table of an n-point FFT */
int orc_tri_len(int n) {
  int len = 0;
  int h = odd_log2(n) ? 2 : 1;
  for (; 4 * h <= n; h *= 4) len += 3 * h;
  return len;
}

/* makeTriTwiddles radix4.go:45-73 */
int orc_tri_twiddles(int n, cpx *tri) {
  int at = 0;
  int h = odd_log2(n) ? 2 : 1;
  for (; 4 * h <= n; h *= 4) {
    for (int j = 0; j < h; j++) {
      float c, s;
      orc_sincospi(j, 2 * h, &c, &s);
      tri[at].re = c;
      tri[at].im = -s;
      at++;
    }
    for (int j = 0; j < h; j++) {
      float c, s;
      orc_sincospi(j, h, &c, &s);
      tri[at].re = c;
      tri[at].im = -s;
      at++;
    }
    for (int j = 0; j < h; j++) {
      int m = 3 * j, neg = 0;
      if (m >= 2 * h) {
        m -= 2 * h;
        neg = 1;
      }
      float c, s;
      orc_sincospi(m, 2 * h, &c, &s);
      if (neg) {
        c = -c;
        s = -s;
      }
      tri[at].re = c;
      tri[at].im = -s;
      at++;
    }
  }
  return at;
}

/* makeBitrev impl.go:148-165 (the table; the swap pairs are i < brev[i]) */
void orc_bitrev(int n, int32_t *brev) {
  int br = 0;
  for (int i = 0; i < n; i++) {
    brev[i] = br;
    int bit = n >> 1;
    while (br & bit) {
      br ^= bit;
      bit >>= 1;
    }
    br |= bit;
  }
}

/* ------------------------------------------------------------- FFT core -- */

/* mulPlain radix4.go:80-85 */
static inline cpx mul_plain(cpx v, cpx w) {
  float a = v.re * w.re;
  float b = v.im * w.im;
  float c = v.re * w.im;
  float d = v.im * w.re;
  cpx r;
  r.re = a - b;
  r.im = c + d;
  return r;
}

/* the radix-4 butterfly of fftR4Scalar radix4.go:121-145 on four operands:
 * A = a[j], C = a[j+h], B = a[j+2h], D = a[j+3h] */
static inline void bfly4(cpx *pA, cpx *pC, cpx *pB, cpx *pD, cpx w1, cpx w2, cpx w3, int inverse) {
  if (inverse) {
    w1.im = -w1.im;
    w2.im = -w2.im;
    w3.im = -w3.im;
  }
  cpx av = *pA;
  cpx tb = mul_plain(*pB, w1);
  cpx tc = mul_plain(*pC, w2);
  cpx td = mul_plain(*pD, w3);
  cpx s0 = {av.re + tc.re, av.im + tc.im};
  cpx s1 = {av.re - tc.re, av.im - tc.im};
  cpx s2 = {tb.re + td.re, tb.im + td.im};
  cpx s3 = {tb.re - td.re, tb.im - td.im};
  float rr, ri;
  if (inverse) {
    rr = -s3.im;
    ri = s3.re;
  } else {
    rr = s3.im;
    ri = -s3.re;
  }
  pA->re = s0.re + s2.re;
  pA->im = s0.im + s2.im;
  pC->re = s1.re + rr;
  pC->im = s1.im + ri;
  pB->re = s0.re - s2.re;
  pB->im = s0.im - s2.im;
  pD->re = s1.re - rr;
  pD->im = s1.im - ri;
}

/* fftR4Scalar radix4.go:106-149 (post-swap cascade) */
static void fft_r4_scalar(cpx *a, int n, const cpx *tri, int inverse) {
  int h = 1;
  if (odd_log2(n)) {
    for (int i = 0; i < n; i += 2) {
      cpx u = a[i], v = a[i + 1];
      a[i].re = u.re + v.re;
      a[i].im = u.im + v.im;
      a[i + 1].re = u.re - v.re;
      a[i + 1].im = u.im - v.im;
    }
    h = 2;
  }
  for (; 4 * h <= n; h *= 4) {
    const cpx *w1 = tri, *w2 = tri + h, *w3 = tri + 2 * h;
    tri += 3 * h;
    for (int base = 0; base < n; base += 4 * h)
      for (int j = 0; j < h; j++)
        bfly4(&a[base + j], &a[base + j + h], &a[base + j + 2 * h], &a[base + j + 3 * h], w1[j], w2[j], w3[j],
              inverse);
  }
}

/* fftR4 radix4.go:91-102: swap pairs, then the cascade */
static void fft_r4(cpx *a, int n, const cpx *tri, const int32_t *brev, int inverse) {
  for (int i = 0; i < n; i++) {
    int j = brev[i];
    if (j > i) {
      cpx t = a[i];
      a[i] = a[j];
      a[j] = t;
    }
  }
  fft_r4_scalar(a, n, tri, inverse);
}

/* colsR4Go radix4cols.go:260-274 with colsR4Head/colsR4Pass folded in:
 * rmap == NULL indexes rows directly, else every row access goes through rmap */
static void cols_r4(cpx *d, int n, int width, const cpx *tri, int inverse, const int32_t *rmap) {
  int h = 1;
  if (odd_log2(n)) {
    for (int r = 0; r < n; r += 2) {
      int i0 = r, i1 = r + 1;
      if (rmap) {
        i0 = rmap[i0];
        i1 = rmap[i1];
      }
      cpx *p = d + (size_t)i0 * width, *q = d + (size_t)i1 * width;
      for (int c = 0; c < width; c++) {
        cpx u = p[c], v = q[c];
        p[c].re = u.re + v.re;
        p[c].im = u.im + v.im;
        q[c].re = u.re - v.re;
        q[c].im = u.im - v.im;
      }
    }
    h = 2;
  }
  for (; 4 * h <= n; h *= 4) {
    const cpx *w1 = tri, *w2 = tri + h, *w3 = tri + 2 * h;
    tri += 3 * h;
    for (int base = 0; base < n; base += 4 * h)
      for (int j = 0; j < h; j++) {
        int r = base + j;
        int i0 = r, i1 = r + h, i2 = r + 2 * h, i3 = r + 3 * h;
        if (rmap) {
          i0 = rmap[i0];
          i1 = rmap[i1];
          i2 = rmap[i2];
          i3 = rmap[i3];
        }
        cpx *pA = d + (size_t)i0 * width, *pC = d + (size_t)i1 * width;
        cpx *pB = d + (size_t)i2 * width, *pD = d + (size_t)i3 * width;
        for (int c = 0; c < width; c++) bfly4(&pA[c], &pC[c], &pB[c], &pD[c], w1[j], w2[j], w3[j], inverse);
      }
  }
}

/* exported single-vector/slab engines for the FFT tests */
void orc_fft_r4(float *a, int n, int inverse) {
  cpx *tri = malloc(sizeof(cpx) * (orc_tri_len(n) + 1));
  int32_t *brev = malloc(sizeof(int32_t) * n);
  orc_tri_twiddles(n, tri);
  orc_bitrev(n, brev);
  fft_r4((cpx *)a, n, tri, brev, inverse);
  free(tri);
  free(brev);
}

void orc_cols_r4(float *d, int n, int width, int inverse, int use_rmap) {
  cpx *tri = malloc(sizeof(cpx) * (orc_tri_len(n) + 1));
  int32_t *brev = malloc(sizeof(int32_t) * n);
  orc_tri_twiddles(n, tri);
  orc_bitrev(n, brev);
  cols_r4((cpx *)d, n, width, tri, inverse, use_rmap ? brev : NULL);
  free(tri);
  free(brev);
}

/* ---------------------------------------------------------------- plan -- */

typedef struct {
  int dftW, dftH, hw, blockW, blockH;
  int64_t cost;
  cpx *triW, *triH;
  int32_t *brevW, *brevH;
} plan_t;

/* newGoPlan impl.go:246-295: integer cost-model argmin over pow2 sizes */
static void plan_geometry(int tw, int th, int rw, int rh, plan_t *p) {
  const int64_t areaCap = (int64_t)1 << 22;
  int minW = imax(2, next_pow2(tw)), minH = imax(2, next_pow2(th));
  int maxW = imax(minW, next_pow2(rw + tw - 1)), maxH = imax(minH, next_pow2(rh + th - 1));
  int bestW = minW, bestH = minH;
  int64_t bestCost = -1, bestArea = 0;
  for (int dftW = minW; dftW <= maxW; dftW <<= 1) {
    for (int dftH = minH; dftH <= maxH; dftH <<= 1) {
      int64_t area = (int64_t)dftW * dftH;
      if (area > areaCap) continue;
      int blockW = imin(dftW - tw + 1, rw);
      int blockH = imin(dftH - th + 1, rh);
      int64_t ntx = (rw + blockW - 1) / blockW;
      int64_t hw = dftW / 2 + 1;
      int64_t rowFFT = 5 * (int64_t)dftW * (bit_len(dftW) - 1);
      int64_t colPass = 5 * (int64_t)dftH * (bit_len(dftH) - 1) * hw;
      int64_t specN = (int64_t)dftH * hw;
      int64_t cost = 0;
      for (int y0 = 0; y0 < rh; y0 += blockH) {
        int bh = imin(blockH, rh - y0);
        int loadH = imin(bh + th - 1, dftH);
        cost += ntx * ((int64_t)((loadH + 1) / 2 + (bh + 1) / 2) * rowFFT + 2 * colPass + 16 * specN);
      }
      if (bestCost < 0 || cost < bestCost || (cost == bestCost && area < bestArea)) {
        bestCost = cost;
        bestArea = area;
        bestW = dftW;
        bestH = dftH;
      }
    }
  }
  p->dftW = bestW;
  p->dftH = bestH;
  p->hw = bestW / 2 + 1;
  p->cost = bestCost;
  p->blockW = imin(bestW - tw + 1, rw);
  p->blockH = imin(bestH - th + 1, rh);
}

static int plan_tables(plan_t *p) {
  p->triW = malloc(sizeof(cpx) * (orc_tri_len(p->dftW) + 1));
  p->triH = malloc(sizeof(cpx) * (orc_tri_len(p->dftH) + 1));
  p->brevW = malloc(sizeof(int32_t) * p->dftW);
  p->brevH = malloc(sizeof(int32_t) * p->dftH);
  if (!p->triW || !p->triH || !p->brevW || !p->brevH) return ORC_NOMEM;
  orc_tri_twiddles(p->dftW, p->triW);
  orc_tri_twiddles(p->dftH, p->triH);
  orc_bitrev(p->dftW, p->brevW);
  orc_bitrev(p->dftH, p->brevH);
  return ORC_OK;
}

static void plan_free(plan_t *p) {
  free(p->triW);
  free(p->triH);
  free(p->brevW);
  free(p->brevH);
  p->triW = p->triH = NULL;
  p->brevW = p->brevH = NULL;
}

/* exported for the planner pins: out = {dftW, dftH, blockW, blockH} */
int64_t orc_plan(int tw, int th, int rw, int rh, int32_t *out) {
  plan_t p;
  plan_geometry(tw, th, rw, rh, &p);
  out[0] = p.dftW;
  out[1] = p.dftH;
  out[2] = p.blockW;
  out[3] = p.blockH;
  return p.cost;
}

/* shrinkDH impl.go:573-581 */
static int shrink_dh(const plan_t *p, int rh, int th) {
  int lastY0 = ((rh - 1) / p->blockH) * p->blockH;
  int lastBh = rh - lastY0;
  if (lastY0 > 0) {
    int dh2 = imax(2, next_pow2(lastBh + th - 1));
    if (dh2 < p->dftH) return dh2;
  }
  return 0;
}

int orc_shrink_dh(int tw, int th, int rw, int rh) {
  plan_t p;
  plan_geometry(tw, th, rw, rh, &p);
  return shrink_dh(&p, rh, th);
}

/* ------------------------------------------------- forward / inverse -- */

/* forwardRowPair impl.go:301-344 */
static void forward_row_pair(const uint8_t *chan, int64_t stride, int step, int x0, int y0, int loadW, int loadH,
                             const plan_t *p, cpx *spec, cpx *z, int r) {
  int n = p->dftW, hw = p->hw, mask = p->dftW - 1;
  cpx *sa = spec + (size_t)p->brevH[r] * hw;
  cpx *sb = spec + (size_t)p->brevH[r + 1] * hw;
  const uint8_t *ra = chan + (int64_t)(y0 + r) * stride + (int64_t)x0 * step;
  if (r + 1 < loadH) {
    const uint8_t *rb = chan + (int64_t)(y0 + r + 1) * stride + (int64_t)x0 * step;
    for (int x = 0; x < loadW; x++) {
      z[x].re = (float)ra[x * step];
      z[x].im = (float)rb[x * step];
    }
  } else {
    for (int x = 0; x < loadW; x++) {
      z[x].re = (float)ra[x * step];
      z[x].im = 0.0f;
    }
  }
  for (int x = loadW; x < n; x++) z[x].re = z[x].im = 0.0f;
  fft_r4(z, n, p->triW, p->brevW, 0);
  for (int k = 0; k < hw; k++) {
    cpx zk = z[k], zn = z[(n - k) & mask];
    sa[k].re = 0.5f * (zk.re + zn.re);
    sa[k].im = 0.5f * (zk.im - zn.im);
    sb[k].re = 0.5f * (zk.im + zn.im);
    sb[k].im = 0.5f * (zn.re - zk.re);
  }
}

/* blockForwardGo impl.go:350-374 (spec rows at brev slots, then columns) */
static void block_forward(const uint8_t *chan, int64_t stride, int step, int x0, int y0, int loadW, int loadH,
                          const plan_t *p, cpx *spec, cpx *z) {
  int hw = p->hw;
  int loaded = imin(p->dftH, (loadH + 1) & ~1);
  for (int r = loaded; r < p->dftH; r++) memset(spec + (size_t)p->brevH[r] * hw, 0, sizeof(cpx) * hw);
  for (int r = 0; r < loadH && r < p->dftH; r += 2) forward_row_pair(chan, stride, step, x0, y0, loadW, loadH, p, spec, z, r);
  cols_r4(spec, p->dftH, hw, p->triH, 0, NULL);
}

/* mulConjGo impl.go:376-388 */
static void mul_conj(cpx *spec, const cpx *tspec, size_t n) {
  for (size_t i = 0; i < n; i++) {
    cpx a = spec[i], b = tspec[i];
    float p0 = a.re * b.re, p1 = a.im * b.im;
    float p2 = a.im * b.re, p3 = a.re * b.im;
    spec[i].re = p0 + p1;
    spec[i].im = p2 - p3;
  }
}

/* inverseRowPair impl.go:394-444 */
static void inverse_row_pair(const plan_t *p, const cpx *spec, cpx *z, float *res, int rw, int x0, int y0, int bw,
                             int bh, int add, int r) {
  int n = p->dftW, hw = p->hw;
  const cpx *sa = spec + (size_t)p->brevH[r] * hw;
  const cpx *sb = spec + (size_t)p->brevH[r + 1] * hw;
  for (int k = 0; k < hw; k++) {
    z[k].re = sa[k].re - sb[k].im;
    z[k].im = sa[k].im + sb[k].re;
  }
  for (int k = hw; k < n; k++) {
    int m = n - k;
    z[k].re = sa[m].re + sb[m].im;
    z[k].im = sb[m].re - sa[m].im;
  }
  fft_r4(z, n, p->triW, p->brevW, 1);
  float *o = res + (size_t)(y0 + r) * rw + x0;
  float *o2 = (r + 1 < bh) ? res + (size_t)(y0 + r + 1) * rw + x0 : NULL;
  if (add) {
    for (int x = 0; x < bw; x++) o[x] += z[x].re;
    if (o2)
      for (int x = 0; x < bw; x++) o2[x] += z[x].im;
  } else {
    for (int x = 0; x < bw; x++) o[x] = z[x].re;
    if (o2)
      for (int x = 0; x < bw; x++) o2[x] = z[x].im;
  }
}

/* blockInverseEmitGo impl.go:446-464 */
static void block_inverse_emit(const plan_t *p, cpx *spec, cpx *z, float *res, int rw, int x0, int y0, int bw, int bh,
                               int add) {
  cols_r4(spec, p->dftH, p->hw, p->triH, 1, p->brevH);
  for (int r = 0; r < bh; r += 2) inverse_row_pair(p, spec, z, res, rw, x0, y0, bw, bh, add, r);
}

/* --------------------------------------------------- template spectra -- */

typedef struct {
  cpx *spec;   /* cn * dftH * hw */
  int has_p2;  /* shrunk last band */
  plan_t p2;   /* same dftW, dftH = dh2 */
  cpx *spec2;  /* cn * dh2 * hw */
} tspec_set;

/* buildTSpecSet impl.go:520-561 + shrinkInto impl.go:591-610 */
static int build_tspec_set(const uint8_t *tpl, int64_t tstride, int step, int cn, int tw, int th, int rh, int shrinkTH,
                           const plan_t *p, tspec_set *set) {
  size_t specN = (size_t)p->dftH * p->hw;
  memset(set, 0, sizeof(*set));
  set->spec = malloc(sizeof(cpx) * cn * specN);
  cpx *z = malloc(sizeof(cpx) * p->dftW);
  if (!set->spec || !z) {
    free(z);
    return ORC_NOMEM;
  }
  float scale = 1.0f / ((float)p->dftW * (float)p->dftH);
  for (int k = 0; k < cn; k++) {
    cpx *ts = set->spec + k * specN;
    block_forward(tpl + k, tstride, step, 0, 0, tw, th, p, ts, z);
    for (size_t i = 0; i < specN; i++) {
      ts[i].re = ts[i].re * scale;
      ts[i].im = ts[i].im * scale;
    }
  }
  free(z);
  int dh2 = shrink_dh(p, rh, shrinkTH);
  if (dh2 > 0) {
    set->has_p2 = 1;
    set->p2 = *p;
    set->p2.dftH = dh2;
    set->p2.triW = set->p2.triH = NULL;
    set->p2.brevW = set->p2.brevH = NULL;
    /* the shrunk plan shares dftW; give it its own tables */
    if (plan_tables(&set->p2) != ORC_OK) return ORC_NOMEM;
    set->spec2 = malloc(sizeof(cpx) * cn * (size_t)dh2 * p->hw);
    if (!set->spec2) return ORC_NOMEM;
    int shift = trailing_zeros(p->dftH) - trailing_zeros(dh2);
    float up = (float)(1 << shift);
    for (int k = 0; k < cn; k++)
      for (int r = 0; r < dh2; r++) {
        const cpx *src = set->spec + k * specN + (size_t)(r << shift) * p->hw;
        cpx *dst = set->spec2 + (size_t)k * dh2 * p->hw + (size_t)r * p->hw;
        for (int x = 0; x < p->hw; x++) {
          dst[x].re = src[x].re * up;
          dst[x].im = src[x].im * up;
        }
      }
  }
  return ORC_OK;
}

static void tspec_free(tspec_set *s) {
  free(s->spec);
  free(s->spec2);
  if (s->has_p2) plan_free(&s->p2);
  memset(s, 0, sizeof(*s));
}

/* --------------------------------------------------- cross-correlation -- */

typedef struct {
  const uint8_t *img;
  int64_t istride;
  int step, cn, tw, th, rw, rh;
  const plan_t *p;
  const tspec_set *set;
  float *result;
  int ntx, ntiles, lastY0;
  atomic_int next;
  int err;
} corr_job;

static void *corr_worker(void *arg) {
  corr_job *J = arg;
  const plan_t *p = J->p;
  size_t specN = (size_t)p->dftH * p->hw;
  cpx *spec = malloc(sizeof(cpx) * specN);
  cpx *z = malloc(sizeof(cpx) * p->dftW);
  if (!spec || !z) {
    J->err = ORC_NOMEM;
    free(spec);
    free(z);
    return NULL;
  }
  for (;;) {
    int t = atomic_fetch_add(&J->next, 1);
    if (t >= J->ntiles) break;
    int x0 = (t % J->ntx) * p->blockW, y0 = (t / J->ntx) * p->blockH;
    int bw = imin(p->blockW, J->rw - x0), bh = imin(p->blockH, J->rh - y0);
    const plan_t *bp = p;
    const cpx *bt = J->set->spec;
    if (y0 == J->lastY0 && J->set->has_p2) {
      bp = &J->set->p2;
      bt = J->set->spec2;
    }
    size_t sn = (size_t)bp->dftH * bp->hw;
    for (int k = 0; k < J->cn; k++) {
      block_forward(J->img + k, J->istride, J->step, x0, y0, bw + J->tw - 1, bh + J->th - 1, bp, spec, z);
      mul_conj(spec, bt + k * sn, sn);
      block_inverse_emit(bp, spec, z, J->result, J->rw, x0, y0, bw, bh, k > 0);
    }
  }
  free(spec);
  free(z);
  return NULL;
}

/* crossCorrGo impl.go:616-693 (tile queue; tiles own disjoint output) */
static int cross_corr(const uint8_t *img, int64_t istride, int step, int cn, int tw, int th, int rw, int rh,
                      int threads, const plan_t *p, const tspec_set *set, float *result) {
  corr_job J;
  J.img = img;
  J.istride = istride;
  J.step = step;
  J.cn = cn;
  J.tw = tw;
  J.th = th;
  J.rw = rw;
  J.rh = rh;
  J.p = p;
  J.set = set;
  J.result = result;
  J.ntx = (rw + p->blockW - 1) / p->blockW;
  int nty = (rh + p->blockH - 1) / p->blockH;
  J.ntiles = J.ntx * nty;
  J.lastY0 = ((rh - 1) / p->blockH) * p->blockH;
  atomic_init(&J.next, 0);
  J.err = 0;
  int nw = imin(threads, J.ntiles);
  if (nw <= 1) {
    corr_worker(&J);
    return J.err;
  }
  pthread_t tid[64];
  for (int i = 1; i < nw; i++) pthread_create(&tid[i], NULL, corr_worker, &J);
  corr_worker(&J);
  for (int i = 1; i < nw; i++) pthread_join(tid[i], NULL);
  return J.err;
}

/* ------------------------------------------------------- normalisation -- */

/* normOne impl.go:825-849 */
static inline float norm_one(float num, float lane0, float idiff, float s2, float numScale, float varScale, float eps,
                             float templNorm) {
  float t = lane0 * numScale;
  num = num - t;
  float diff2 = idiff * varScale;
  float lim = eps * s2;
  if (lim > 0.5f) lim = 0.5f;
  float den = 0.0f;
  if (diff2 > lim) {
    float sq = (float)sqrt((double)diff2);
    den = sq * templNorm;
  }
  float an = fabsf(num);
  if (an < den) {
    num = num / den;
  } else {
    float den2 = den * 1.125f;
    if (an < den2)
      num = num > 0 ? 1.0f : -1.0f;
    else
      num = 0.0f;
  }
  return num;
}

typedef struct {
  float minV, maxV;
  int minX, minY, maxX, maxY;
} extrema;

/* colBuildGo impl.go:708-759 */
static void col_build(int32_t *colSum, int64_t *colSum2, const uint8_t *img, int64_t istride, int iw, int cn, int cs,
                      int step, int y0, int th) {
  memset(colSum, 0, sizeof(int32_t) * (size_t)iw * cs);
  memset(colSum2, 0, sizeof(int64_t) * (size_t)iw);
  for (int y = y0; y < y0 + th; y++) {
    const uint8_t *row = img + (int64_t)y * istride;
    for (int x = 0; x < iw; x++)
      for (int k = 0; k < cn; k++) {
        int32_t v = row[(int64_t)x * step + k];
        colSum[x * cs + k] += v;
        colSum2[x] += (int64_t)(v * v);
      }
  }
}

/* colSlideGo impl.go:761-805 */
static void col_slide(int32_t *colSum, int64_t *colSum2, const uint8_t *rsub, const uint8_t *radd, int iw, int cn, int cs,
                      int step) {
  for (int x = 0; x < iw; x++)
    for (int k = 0; k < cn; k++) {
      int32_t a = radd[(int64_t)x * step + k], b = rsub[(int64_t)x * step + k];
      colSum[x * cs + k] += a - b;
      colSum2[x] += (int64_t)(a * a - b * b);
    }
}

/* normalizeBandGo impl.go:859-1054 (rows [y0, y1)); corr may alias result */
static extrema normalize_band(const uint8_t *img, int64_t istride, int iw, int cn, int step, int tw, int th, int rw,
                              int y0, int y1, const int64_t *tsum, float templNorm, const float *corr, float *result,
                              int *err) {
  int64_t area = (int64_t)tw * th;
  float varScale = (float)(1.0 / ((double)tw * (double)th));
  float numScale = varScale;
  if (cn == 1) numScale = (float)((double)tsum[0] / ((double)tw * (double)th));
  int cs = cn;
  if (step == 4 && (cn == 3 || cn == 4)) cs = 4;
  int32_t *colSum = malloc(sizeof(int32_t) * (size_t)iw * cs);
  int64_t *colSum2 = malloc(sizeof(int64_t) * (size_t)iw);
  extrema ext = {3.40282346638528859811704183484516925440e+38f, -3.40282346638528859811704183484516925440e+38f, 0, y0,
                 0, y0};
  if (!colSum || !colSum2) {
    *err = ORC_NOMEM;
    free(colSum);
    free(colSum2);
    return ext;
  }
  col_build(colSum, colSum2, img, istride, iw, cn, cs, step, y0, th);
  const float eps = 10.0f * 0x1p-23f;
  for (int y = y0;; y++) {
    int64_t s[4] = {0, 0, 0, 0}, s2 = 0;
    for (int x = 0; x < tw; x++) {
      for (int k = 0; k < cn; k++) s[k] += colSum[x * cs + k];
      s2 += colSum2[x];
    }
    float *rrow = result + (size_t)y * rw;
    const float *crow = corr + (size_t)y * rw;
    for (int x = 0; x < rw; x++) {
      int64_t lane0 = 0, sq = 0;
      for (int k = 0; k < cn; k++) {
        lane0 += s[k] * tsum[k];
        sq += s[k] * s[k];
      }
      if (cn == 1) lane0 = s[0];
      float l0 = (float)(double)lane0;
      float l1 = (float)(double)(area * s2 - sq);
      float l2 = (float)(double)s2;
      rrow[x] = norm_one(crow[x], l0, l1, l2, numScale, varScale, eps, templNorm);
      if (x + 1 < rw) {
        for (int k = 0; k < cn; k++) s[k] += (int64_t)(colSum[(x + tw) * cs + k] - colSum[x * cs + k]);
        s2 += colSum2[x + tw] - colSum2[x];
      }
    }
    for (int x = 0; x < rw; x++) {
      float v = rrow[x];
      if (v < ext.minV) {
        ext.minV = v;
        ext.minX = x;
        ext.minY = y;
      }
      if (v > ext.maxV) {
        ext.maxV = v;
        ext.maxX = x;
        ext.maxY = y;
      }
    }
    if (y + 1 >= y1) break;
    col_slide(colSum, colSum2, img + (int64_t)y * istride, img + (int64_t)(y + th) * istride, iw, cn, cs, step);
  }
  free(colSum);
  free(colSum2);
  return ext;
}

typedef struct {
  const uint8_t *img;
  int64_t istride;
  int iw, cn, step, tw, th, rw;
  const int64_t *tsum;
  float templNorm;
  const float *corr;
  float *result;
  int y0, y1;
  extrema ext;
  int err;
} band_job;

static void *band_worker(void *arg) {
  band_job *B = arg;
  B->ext = normalize_band(B->img, B->istride, B->iw, B->cn, B->step, B->tw, B->th, B->rw, B->y0, B->y1, B->tsum,
                          B->templNorm, B->corr, B->result, &B->err);
  return NULL;
}

/* normalizeParallelGo impl.go:1081-1114 */
static int normalize_parallel(const uint8_t *img, int64_t istride, int iw, int tw, int th, int cn, int step, int rw,
                              int rh, int threads, const int64_t *tsum, int64_t varSum, const float *corr,
                              float *result, orc_result *out) {
  float templNorm = (float)sqrt((double)varSum / ((double)tw * (double)th));
  int nb = threads;
  int maxb = rh / imax(th, 32);
  if (maxb >= 1 && nb > maxb) nb = maxb;
  nb = imax(imin(nb, rh), 1);
  if (nb > 64) nb = 64;
  band_job jobs[64];
  pthread_t tid[64];
  for (int b = 0; b < nb; b++) {
    band_job *B = &jobs[b];
    B->img = img;
    B->istride = istride;
    B->iw = iw;
    B->cn = cn;
    B->step = step;
    B->tw = tw;
    B->th = th;
    B->rw = rw;
    B->tsum = tsum;
    B->templNorm = templNorm;
    B->corr = corr;
    B->result = result;
    B->y0 = (int)((int64_t)rh * b / nb);
    B->y1 = (int)((int64_t)rh * (b + 1) / nb);
    B->err = 0;
  }
  for (int b = 1; b < nb; b++) pthread_create(&tid[b], NULL, band_worker, &jobs[b]);
  band_worker(&jobs[0]);
  for (int b = 1; b < nb; b++) pthread_join(tid[b], NULL);
  extrema r = jobs[0].ext;
  int err = jobs[0].err;
  for (int b = 1; b < nb; b++) {
    if (jobs[b].err) err = jobs[b].err;
    if (jobs[b].ext.minV < r.minV) {
      r.minV = jobs[b].ext.minV;
      r.minX = jobs[b].ext.minX;
      r.minY = jobs[b].ext.minY;
    }
    if (jobs[b].ext.maxV > r.maxV) {
      r.maxV = jobs[b].ext.maxV;
      r.maxX = jobs[b].ext.maxX;
      r.maxY = jobs[b].ext.maxY;
    }
  }
  out->min_v = r.minV;
  out->min_x = r.minX;
  out->min_y = r.minY;
  out->max_v = r.maxV;
  out->max_x = r.maxX;
  out->max_y = r.maxY;
  return err;
}

/* ----------------------------------------------------------- entrypoint -- */

/* statsCap impl.go:1122-1124 */
int64_t orc_stats_cap(int cn) {
  static const int64_t caps[5] = {0, 11909805, 8421504, 6876129, 5954902};
  return (cn >= 1 && cn <= 4) ? caps[cn] : 0;
}

/* templStats impl.go:1063-1079 */
void orc_templ_stats(const uint8_t *tpl, int64_t tstride, int tw, int th, int cn, int step, int64_t *tsum,
                     int64_t *varSum) {
  int64_t area = (int64_t)tw * th;
  *varSum = 0;
  for (int k = 0; k < 4; k++) tsum[k] = 0;
  for (int k = 0; k < cn; k++) {
    int64_t s = 0, s2 = 0;
    for (int y = 0; y < th; y++) {
      const uint8_t *row = tpl + (int64_t)y * tstride + k;
      for (int x = 0; x < tw; x++) {
        int64_t v = row[(int64_t)x * step];
        s += v;
        s2 += v * v;
      }
    }
    tsum[k] = s;
    *varSum += area * s2 - s * s;
  }
}

static int check_args(int64_t istride, int iw, int ih, int64_t tstride, int tw, int th, int cn, int step) {
  if (cn < 1 || cn > 4 || step < cn || tw < 1 || th < 1 || tw > iw || th > ih || istride < (int64_t)iw * step ||
      tstride < (int64_t)tw * step)
    return ORC_BAD_ARGS;
  if ((int64_t)tw * th > orc_stats_cap(cn)) return ORC_STATS_CAP;
  return ORC_OK;
}

/* the correlate+normalise core shared by matchU8 and Matcher.Find once the
 * template statistics and the cn choice are known */
static int match_core(const uint8_t *img, int64_t istride, int iw, int ih, const uint8_t *tpl, int64_t tstride, int tw,
                      int th, int cn, int step, int threads, const int64_t *tsum, int64_t varSum, float *result,
                      orc_result *out) {
  int rw = iw - tw + 1, rh = ih - th + 1;
  float *res = result;
  if (!res) {
    res = malloc(sizeof(float) * (size_t)rw * rh);
    if (!res) return ORC_NOMEM;
  }
  plan_t p;
  plan_geometry(tw, th, rw, rh, &p);
  int err = plan_tables(&p);
  tspec_set set;
  memset(&set, 0, sizeof(set));
  if (!err) err = build_tspec_set(tpl, tstride, step, cn, tw, th, rh, th, &p, &set);
  if (!err) err = cross_corr(img, istride, step, cn, tw, th, rw, rh, threads, &p, &set, res);
  tspec_free(&set);
  plan_free(&p);
  if (!err) err = normalize_parallel(img, istride, iw, tw, th, cn, step, rw, rh, threads, tsum, varSum, res, res, out);
  if (!result) free(res);
  return err;
}

/* matchU8 impl.go:1126-1155 */
int orc_match_u8(const uint8_t *img, int64_t istride, int iw, int ih, const uint8_t *tpl, int64_t tstride, int tw,
                 int th, int cn, int step, int threads, float *result, orc_result *out) {
  int err = check_args(istride, iw, ih, tstride, tw, th, cn, step);
  if (err) return err;
  if (threads < 1) threads = 1;
  if (threads > 64) threads = 64;
  int rw = iw - tw + 1, rh = ih - th + 1;
  int64_t tsum[4], varSum;
  orc_templ_stats(tpl, tstride, tw, th, cn, step, tsum, &varSum);
  if (varSum == 0) {
    if (result)
      for (size_t i = 0; i < (size_t)rw * rh; i++) result[i] = 1.0f;
    out->min_v = out->max_v = 1.0f;
    out->min_x = out->min_y = out->max_x = out->max_y = 0;
    return ORC_OK;
  }
  return match_core(img, istride, iw, ih, tpl, tstride, tw, th, cn, step, threads, tsum, varSum, result, out);
}

/* ------------------------------------------------------- image plumbing -- */

/* alphaConst cvmatch.go:146-168 */
int orc_alpha_const(const uint8_t *pix, int64_t stride, int w, int h) {
  uint8_t a = pix[3];
  for (int y = 0; y < h; y++) {
    const uint8_t *row = pix + (int64_t)y * stride;
    for (int x = 0; x < w; x++)
      if (row[x * 4 + 3] != a) return 0;
  }
  return 1;
}

/* rgbToGray cvmatch.go:222-240 (BT.601 fixed point, alpha ignored) */
void orc_rgba_to_gray(const uint8_t *pix, int64_t stride, int w, int h, uint8_t *out) {
  for (int y = 0; y < h; y++) {
    const uint8_t *src = pix + (int64_t)y * stride;
    for (int x = 0; x < w; x++) {
      uint32_t r = src[x * 4], g = src[x * 4 + 1], b = src[x * 4 + 2];
      out[(size_t)y * w + x] = (uint8_t)((4899 * r + 9617 * g + 1868 * b + 8192) >> 14);
    }
  }
}

/* Match cvmatch.go:75-94 / MatchMap :100-117 on interleaved RGBA */
int orc_match_rgba(const uint8_t *img, int64_t istride, int iw, int ih, const uint8_t *tpl, int64_t tstride, int tw,
                   int th, int threads, float *result, orc_result *out) {
  int cn = 4;
  if (orc_alpha_const(img, istride, iw, ih) && orc_alpha_const(tpl, tstride, tw, th)) cn = 3;
  return orc_match_u8(img, istride, iw, ih, tpl, tstride, tw, th, cn, 4, threads, result, out);
}

/* -------------------------------------------------------------- Fleet -- */

/* one prepared template (newMatcher matcher.go:66-76): compact pixels with
 * stride w*step, exact stats over cn = step channels, alpha constancy */
typedef struct {
  const uint8_t *pix;
  int w, h, step, aConst;
  int64_t tsum[4], varSum;
} member_t;

static void member_init(member_t *m, const uint8_t *pix, int w, int h, int step) {
  m->pix = pix;
  m->w = w;
  m->h = h;
  m->step = step;
  orc_templ_stats(pix, (int64_t)w * step, w, h, step, step, m->tsum, &m->varSum);
  m->aConst = step == 4 ? orc_alpha_const(pix, (int64_t)w * 4, w, h) : 0;
}

/* Matcher.Find matcher.go:83-130 on a compact template (stride w*step).
 * Bit-identical to Match / MatchGray by construction (the reference caches
 * the plan and spectrum per geometry; recomputing them gives the same bits). */
int orc_matcher_find(const uint8_t *img, int64_t istride, int iw, int ih, const uint8_t *tpl, int tw, int th, int step,
                     int threads, float *result, orc_result *out) {
  if (step != 1 && step != 4) return ORC_BAD_ARGS;
  if ((int64_t)tw * th > orc_stats_cap(step)) return ORC_STATS_CAP;
  if (tw > iw || th > ih) return ORC_BAD_ARGS;
  member_t m;
  member_init(&m, tpl, tw, th, step);
  if (m.varSum == 0) {
    out->min_v = out->max_v = 1.0f;
    out->min_x = out->min_y = out->max_x = out->max_y = 0;
    return ORC_OK;
  }
  int cn = step;
  if (step == 4 && m.aConst && orc_alpha_const(img, istride, iw, ih)) cn = 3;
  if (threads < 1) threads = 1;
  return match_core(img, istride, iw, ih, tpl, (int64_t)tw * step, tw, th, cn, step, threads, m.tsum, m.varSum, result,
                    out);
}

typedef struct {
  int pw2, ph2;
  int tw, th;
  int n;
  int *members;
} bucket_t;

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* bucketize fleet.go:111-171. Returns the number of buckets; each bucket's
 * member list is sorted ascending. keys sort: larger area, then larger w, then larger h. */
static int bucketize(const member_t *ms, int n, bucket_t *out) {
  int *kw = malloc(sizeof(int) * n), *kh = malloc(sizeof(int) * n);
  int nk = 0;
  for (int i = 0; i < n; i++) {
    if (ms[i].varSum == 0) continue;
    int w2 = next_pow2(ms[i].w), h2 = next_pow2(ms[i].h);
    int seen = 0;
    for (int j = 0; j < nk; j++)
      if (kw[j] == w2 && kh[j] == h2) seen = 1;
    if (!seen) {
      kw[nk] = w2;
      kh[nk] = h2;
      nk++;
    }
  }
  /* sort keys (distinct, so the order is total) */
  for (int a = 0; a < nk; a++)
    for (int b = a + 1; b < nk; b++) {
      long long aa = (long long)kw[a] * kh[a], ab = (long long)kw[b] * kh[b];
      int swap = 0;
      if (aa != ab)
        swap = ab > aa;
      else if (kw[a] != kw[b])
        swap = kw[b] > kw[a];
      else
        swap = kh[b] > kh[a];
      if (swap) {
        int t = kw[a];
        kw[a] = kw[b];
        kw[b] = t;
        t = kh[a];
        kh[a] = kh[b];
        kh[b] = t;
      }
    }
  int nb = 0;
  for (int c = 0; c < nk; c++) {
    int slot = -1;
    for (int s = 0; s < nb; s++)
      if (out[s].pw2 >= kw[c] && out[s].ph2 >= kh[c]) {
        slot = s;
        break;
      }
    if (slot < 0) {
      slot = nb++;
      out[slot].pw2 = kw[c];
      out[slot].ph2 = kh[c];
      out[slot].n = 0;
      out[slot].members = malloc(sizeof(int) * n);
    }
    for (int i = 0; i < n; i++)
      if (ms[i].varSum != 0 && next_pow2(ms[i].w) == kw[c] && next_pow2(ms[i].h) == kh[c])
        out[slot].members[out[slot].n++] = i;
  }
  for (int s = 0; s < nb; s++) {
    qsort(out[s].members, out[s].n, sizeof(int), cmp_int);
    out[s].tw = out[s].th = 0;
    for (int j = 0; j < out[s].n; j++) {
      out[s].tw = imax(out[s].tw, ms[out[s].members[j]].w);
      out[s].th = imax(out[s].th, ms[out[s].members[j]].h);
    }
  }
  free(kw);
  free(kh);
  return nb;
}

/* exported for the bucketize test: bucket b gets (tw, th, count) in
 * dims[3b..3b+2] and its members appended to member_out in order */
int orc_bucketize(const int32_t *ws, const int32_t *hs, const int32_t *flat, int n, int32_t *dims, int32_t *member_out) {
  member_t *ms = calloc(n, sizeof(member_t));
  bucket_t *bs = calloc(n > 0 ? n : 1, sizeof(bucket_t));
  for (int i = 0; i < n; i++) {
    ms[i].w = ws[i];
    ms[i].h = hs[i];
    ms[i].varSum = flat[i] ? 0 : 1;
  }
  int nb = bucketize(ms, n, bs);
  int at = 0;
  for (int b = 0; b < nb; b++) {
    dims[3 * b] = bs[b].tw;
    dims[3 * b + 1] = bs[b].th;
    dims[3 * b + 2] = bs[b].n;
    for (int j = 0; j < bs[b].n; j++) member_out[at++] = bs[b].members[j];
    free(bs[b].members);
  }
  free(ms);
  free(bs);
  return nb;
}

/* Fleet.FindAll fleet.go:176-332 with buildGeom fleet.go:337-390, one thread.
 * tpls[i] is member i's compact template (stride w*step).  The member-order
 * and the claim order do not change any output (outputs are disjoint). */
int orc_fleet_find_all(const uint8_t *img, int64_t istride, int iw, int ih, int step, const uint8_t *const *tpls,
                       const int32_t *tws, const int32_t *ths, int n, orc_result *out) {
  if (n <= 0) return ORC_EMPTY_FLEET;
  if (step != 1 && step != 4) return ORC_MIXED_FLEET;
  member_t *ms = calloc(n, sizeof(member_t));
  for (int i = 0; i < n; i++) {
    if ((int64_t)tws[i] * ths[i] > orc_stats_cap(step)) {
      free(ms);
      return ORC_STATS_CAP;
    }
    member_init(&ms[i], tpls[i], tws[i], ths[i], step);
  }
  for (int i = 0; i < n; i++) {
    if (ms[i].w > iw || ms[i].h > ih) {
      free(ms);
      return ORC_BAD_ARGS;
    }
    if (ms[i].varSum == 0) {
      out[i].min_v = out[i].max_v = 1.0f;
      out[i].min_x = out[i].min_y = out[i].max_x = out[i].max_y = 0;
    }
  }
  bucket_t *bs = calloc(n, sizeof(bucket_t));
  int nb = bucketize(ms, n, bs);
  int pConst = step == 4 && orc_alpha_const(img, istride, iw, ih);
  int err = ORC_OK;
  for (int bi = 0; bi < nb && !err; bi++) {
    bucket_t *b = &bs[bi];
    int rwF = iw - b->tw + 1, rhF = ih - b->th + 1;
    plan_t p;
    plan_geometry(b->tw, b->th, rwF, rhF, &p);
    plan_tables(&p);
    int cnMax = 0;
    int *cns = malloc(sizeof(int) * b->n);
    tspec_set *sets = calloc(b->n, sizeof(tspec_set));
    for (int j = 0; j < b->n; j++) {
      member_t *m = &ms[b->members[j]];
      int cn = m->step;
      if (m->step == 4) cn = (m->aConst && pConst) ? 3 : 4;
      cns[j] = cn;
      cnMax = imax(cnMax, cn);
      err |= build_tspec_set(m->pix, (int64_t)m->w * m->step, m->step, cn, m->w, m->h, rhF, b->th, &p, &sets[j]);
    }
    int dh2 = shrink_dh(&p, rhF, b->th);
    plan_t p2;
    memset(&p2, 0, sizeof(p2));
    if (dh2 > 0) {
      p2 = p;
      p2.dftH = dh2;
      plan_tables(&p2);
    }
    int ntx = (rwF + p.blockW - 1) / p.blockW;
    int nty = (rhF + p.blockH - 1) / p.blockH;
    int ntiles = ntx * nty;
    int lastY0 = ((rhF - 1) / p.blockH) * p.blockH;
    size_t *off = malloc(sizeof(size_t) * (ntiles + 1));
    off[0] = 0;
    for (int t = 0; t < ntiles; t++) {
      int y0 = (t / ntx) * p.blockH;
      const plan_t *bp = (y0 == lastY0 && dh2 > 0) ? &p2 : &p;
      off[t + 1] = off[t] + (size_t)cnMax * bp->dftH * bp->hw;
    }
    cpx *frame = malloc(sizeof(cpx) * (off[ntiles] ? off[ntiles] : 1));
    cpx *z = malloc(sizeof(cpx) * p.dftW);
    /* Phase 1 fleet.go:240-263 */
    for (int t = 0; t < ntiles; t++) {
      int x0 = (t % ntx) * p.blockW, y0 = (t / ntx) * p.blockH;
      int bwF = imin(p.blockW, rwF - x0), bhF = imin(p.blockH, rhF - y0);
      const plan_t *bp = (y0 == lastY0 && dh2 > 0) ? &p2 : &p;
      size_t sn = (size_t)bp->dftH * bp->hw;
      for (int k = 0; k < cnMax; k++)
        block_forward(img + k, istride, step, x0, y0, bwF + b->tw - 1, bhF + b->th - 1, bp, frame + off[t] + k * sn, z);
    }
    /* Phase 2 fleet.go:265-328 */
    cpx *spec = malloc(sizeof(cpx) * (size_t)p.dftH * p.hw);
    for (int j = 0; j < b->n && !err; j++) {
      int mi = b->members[j];
      member_t *m = &ms[mi];
      int cn = cns[j];
      int rw = iw - m->w + 1, rh = ih - m->h + 1;
      float *res = malloc(sizeof(float) * (size_t)rw * rh);
      for (int t = 0; t < ntiles; t++) {
        int x0 = (t % ntx) * p.blockW, y0 = (t / ntx) * p.blockH;
        const plan_t *bp = &p;
        const cpx *ts = sets[j].spec;
        if (y0 == lastY0 && sets[j].has_p2) {
          bp = &p2;
          ts = sets[j].spec2;
        }
        size_t sn = (size_t)bp->dftH * bp->hw;
        int bw = p.blockW, bh = p.blockH;
        if (t % ntx == ntx - 1) bw = rw - x0;
        if (t / ntx == nty - 1) bh = rh - y0;
        for (int k = 0; k < cn; k++) {
          memcpy(spec, frame + off[t] + k * sn, sizeof(cpx) * sn);
          mul_conj(spec, ts + k * sn, sn);
          block_inverse_emit(bp, spec, z, res, rw, x0, y0, bw, bh, k > 0);
        }
      }
      err |= normalize_parallel(img, istride, iw, m->w, m->h, cn, step, rw, rh, 1, m->tsum, m->varSum, res, res,
                                &out[mi]);
      free(res);
    }
    free(spec);
    free(z);
    free(frame);
    free(off);
    for (int j = 0; j < b->n; j++) tspec_free(&sets[j]);
    free(sets);
    free(cns);
    if (dh2 > 0) plan_free(&p2);
    plan_free(&p);
  }
  for (int bi = 0; bi < nb; bi++) free(bs[bi].members);
  free(bs);
  free(ms);
  return err;
}

/* ---------------------------------------------------------- FP64 oracle -- */

/* Exact-integer direct correlation at window (x, y): corr = sum img*tpl,
 * per-channel window sums and the window square sum, all int64. */
static void window_exact(const uint8_t *img, int64_t istride, const uint8_t *tpl, int64_t tstride, int tw, int th,
                         int cn, int step, int x, int y, int64_t *corr, int64_t *s, int64_t *s2) {
  int64_t c = 0, q = 0;
  for (int k = 0; k < 4; k++) s[k] = 0;
  for (int dy = 0; dy < th; dy++) {
    const uint8_t *ir = img + (int64_t)(y + dy) * istride + (int64_t)x * step;
    const uint8_t *tr = tpl + (int64_t)dy * tstride;
    for (int dx = 0; dx < tw; dx++)
      for (int k = 0; k < cn; k++) {
        int64_t iv = ir[dx * step + k], tv = tr[dx * step + k];
        c += iv * tv;
        s[k] += iv;
        q += iv * iv;
      }
  }
  *corr = c;
  *s2 = q;
}

/* TM_CCOEFF_NORMED in float64 from exact integers, with OpenCV's guard
 * ladder as the reference's brute force states it (cvmatch_test.go:57-76):
 *   num   = corr - sum_k s_k*mean_k        = (area*corr - sum_k s_k*tsum_k)/area
 *   diff2 = s2 - sum_k s_k^2/area          = idiff/area
 *   den   = sqrt(diff2)*sqrt(varSum/area)  if diff2 > min(0.5, 10*FLT_EPSILON*s2)
 * The ratio num/den is evaluated as numI/sqrt(idiff*varSum), one division. */
static double fp64_score(int64_t area, int64_t corr, const int64_t *s, int64_t s2, const int64_t *tsum, int64_t varSum,
                         int cn) {
  if (varSum == 0) return 1.0;
  int64_t cross = 0, sq = 0;
  for (int k = 0; k < cn; k++) {
    cross += s[k] * tsum[k];
    sq += s[k] * s[k];
  }
  int64_t numI = area * corr - cross;
  int64_t idiff = area * s2 - sq;
  double diff2 = (double)idiff / (double)area;
  double lim = 10.0 * 1.1920928955078125e-07 * (double)s2;
  if (lim > 0.5) lim = 0.5;
  if (!(diff2 > lim)) return 0.0;
  double denI = sqrt((double)idiff) * sqrt((double)varSum);
  double an = fabs((double)numI);
  if (an < denI) return (double)numI / denI;
  if (an < denI * 1.125) return numI > 0 ? 1.0 : -1.0;
  return 0.0;
}

/* full FP64 map (small inputs only: O(rw*rh*tw*th*cn)) */
int orc_fp64_map(const uint8_t *img, int64_t istride, int iw, int ih, const uint8_t *tpl, int64_t tstride, int tw, int th,
                 int cn, int step, double *out) {
  int err = check_args(istride, iw, ih, tstride, tw, th, cn, step);
  if (err) return err;
  int rw = iw - tw + 1, rh = ih - th + 1;
  int64_t tsum[4], varSum;
  orc_templ_stats(tpl, tstride, tw, th, cn, step, tsum, &varSum);
  int64_t area = (int64_t)tw * th;
  for (int y = 0; y < rh; y++)
    for (int x = 0; x < rw; x++) {
      int64_t corr, s[4], s2;
      window_exact(img, istride, tpl, tstride, tw, th, cn, step, x, y, &corr, s, &s2);
      out[(size_t)y * rw + x] = fp64_score(area, corr, s, s2, tsum, varSum, cn);
    }
  return ORC_OK;
}

/* FP64 scores at sampled windows (xs[i], ys[i]) */
int orc_fp64_points(const uint8_t *img, int64_t istride, int iw, int ih, const uint8_t *tpl, int64_t tstride, int tw,
                    int th, int cn, int step, const int32_t *xs, const int32_t *ys, int npts, double *out) {
  int err = check_args(istride, iw, ih, tstride, tw, th, cn, step);
  if (err) return err;
  int64_t tsum[4], varSum;
  orc_templ_stats(tpl, tstride, tw, th, cn, step, tsum, &varSum);
  int64_t area = (int64_t)tw * th;
  for (int i = 0; i < npts; i++) {
    int64_t corr, s[4], s2;
    window_exact(img, istride, tpl, tstride, tw, th, cn, step, xs[i], ys[i], &corr, s, &s2);
    out[i] = fp64_score(area, corr, s, s2, tsum, varSum, cn);
  }
  return ORC_OK;
}

/* FNV-1a 64 over the map's float bits and the extrema, little-endian words
 * (goldenHash impl_golden_test.go:27-44) */
uint64_t orc_golden_hash(const float *res, size_t n, const orc_result *r) {
  uint64_t h = 14695981039346656037ULL;
  const uint64_t prime = 1099511628211ULL;
#define PUT32(v)                                   \
  do {                                             \
    uint32_t w_ = (v);                             \
    for (int b_ = 0; b_ < 4; b_++) {               \
      h ^= (uint8_t)(w_ >> (8 * b_));              \
      h *= prime;                                  \
    }                                              \
  } while (0)
  for (size_t i = 0; i < n; i++) {
    uint32_t bits;
    memcpy(&bits, &res[i], 4);
    PUT32(bits);
  }
  uint32_t bits;
  memcpy(&bits, &r->min_v, 4);
  PUT32(bits);
  PUT32((uint32_t)r->min_x);
  PUT32((uint32_t)r->min_y);
  memcpy(&bits, &r->max_v, 4);
  PUT32(bits);
  PUT32((uint32_t)r->max_x);
  PUT32((uint32_t)r->max_y);
#undef PUT32
  return h;
}

/* ------------------------------------------------------------- peaks -- */

/* Every occurrence on a response map: the threshold + local-maximum
 * suppression recipe of TestFindAllOccurrences (cvmatch_test.go:346-365),
 * with the threshold and the neighbourhood half-sizes as parameters.  (x, y)
 * is a peak unless resp < thresh or some in-bounds element within |dx| <= rx,
 * |dy| <= ry is strictly greater.  Peaks are written in raster order, at most
 * cap of them; returns the total count. */
int orc_find_peaks(const float *resp, int w, int h, float thresh, int rx, int ry, int32_t *xy, float *vals, int cap) {
  int count = 0;
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      float v = resp[(size_t)y * w + x];
      if (v < thresh) continue;
      int is_max = 1;
      for (int dy = -ry; dy <= ry && is_max; dy++) {
        for (int dx = -rx; dx <= rx; dx++) {
          int nx = x + dx, ny = y + dy;
          if (nx >= 0 && ny >= 0 && nx < w && ny < h && resp[(size_t)ny * w + nx] > v) {
            is_max = 0;
            break;
          }
        }
      }
      if (is_max) {
        if (count < cap) {
          xy[2 * count] = x;
          xy[2 * count + 1] = y;
          vals[count] = v;
        }
        count++;
      }
    }
  }
  return count;
}
