// Internal declarations shared by the HIP translation units of libcvmatch_hip.
//
// Numerical contract (the reference's scalar op sequence, radix4.go:17-31,
// impl.go:376-464, 816-849): every float op is one IEEE-754 rounding, never a
// fused multiply-add.  All sources are compiled with -ffp-contract=off and
// without fast-math, so `a * b - c * d` stays two roundings of products plus
// one of the difference, exactly as the reference writes it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <string>
#include <vector>

namespace cvm {

// ------------------------------------------------------------ geometry --

// Tiling of one correlation (crossCorrGo impl.go:616-693, FindAll
// fleet.go:214-328).  Tile t covers result columns [x0, x0+bw) and rows
// [y0, y0+bh) with x0 = (t % ntx) * blockW, y0 = (t / ntx) * blockH; interior
// tiles have bw = blockW / bh = blockH, the last tile on each axis runs to
// the extent (extW - x0, extH - y0).  Forward loads are bw+padW columns by
// bh+padH rows.  Tiles whose y0 == lastY0 use the shrunk height dftH2 when
// dftH2 != 0 (shrinkDH impl.go:573-581).
struct TileGrid {
  int ntx, nty;
  int blockW, blockH;
  int extW, extH;
  int padW, padH;
  int lastY0;
  int dftW, log2W, hw, hwPad;   // row transform and spectrum row stride (complex)
  int dftH, log2H;              // full column height
  int dftH2, log2H2;            // shrunk last-band height (0: none)
};

__host__ __device__ inline void tile_rect(const TileGrid& g, int t, int& x0, int& y0, int& bw, int& bh) {
  int tx = t % g.ntx, ty = t / g.ntx;
  x0 = tx * g.blockW;
  y0 = ty * g.blockH;
  bw = (tx == g.ntx - 1) ? g.extW - x0 : g.blockW;
  bh = (ty == g.nty - 1) ? g.extH - y0 : g.blockH;
}

__host__ __device__ inline bool tile_shrunk(const TileGrid& g, int y0) { return g.dftH2 != 0 && y0 == g.lastY0; }

// Device-resident per-call state.  Kernels read the channel count and the
// flat flag from here, so no host round trip sits inside a call.
struct CallState {
  unsigned long long tsum[4];   // template channel sums (exact)
  unsigned long long tsq[4];    // template channel square sums (exact)
  long long varSum;             // sum_k area*tsq_k - tsum_k^2
  int cn;                       // effective channel count used by every kernel
  int flat;                     // varSum == 0
  unsigned int alpha_bad[2];    // parent / template alpha not constant
  float templNorm, numScale, varScale;
  int pad_;
};

// Geometry of one call-state launch (call_state_block): the template
// statistics grid, the parent's alpha scan over rows [0, ph0), the
// template's alpha scan; the scans only under the alpha rule.
struct CallStateArgs {
  const uint8_t* img;
  int64_t istride;
  int iw, ph0, ivec;
  const uint8_t* tpl;
  int64_t tstride;
  int tw, th, tvec, cnMax, step;
  int sgx, sgy, pgx, pgy, tgx, tgy;
  int rows8;  // the compact layout: sgx flat statistics workgroups, scans of 8 rows
  CallState* st;
  int nb() const { return sgx * sgy + pgx * pgy + tgx * tgy; }
};

// Exact template statistics of up to four channels (templStats impl.go:1063-1079).
struct TemplStats {
  unsigned long long tsum[4];
  unsigned long long tsq[4];
};

// One template of a batched direct-correlation launch: its B fragments,
// C' map, shape, ring pitch and grid (gx blocks along x), first workgroup.
struct DirectJob {
  const int8_t* bf;
  int32_t* out;
  int th, nk, pitch, rw, rh, gx, wg0, pad_;
};

struct Extrema;

// One template of a fused correlation + scoring launch: its B fragments,
// call state, one extrema partial per 32 x 128 block, shape, ring pitch and
// grid (gx blocks along x), first workgroup.
struct FusedJob {
  const int8_t* bf;
  const CallState* st;
  Extrema* parts;
  int tw, th, nk, pitch, rw, rh, gx, wg0;
};

// One template of a batched normalisation launch: its call state, chunk
// partials (per channel count), correlation, extrema partials and result.
struct NormJob {
  const CallState* st;
  const uint32_t* ps;
  const float* corr;
  Extrema* parts;
  void* res;  // Res6 of the batched reduction
  int tw, th, rw, rh, gx, gy, wg0, pad_;
};

// Min/max partial of one normalisation workgroup: (value, raster index).
struct Extrema {
  float min_v;
  int min_i;
  float max_v;
  int max_i;
};

// Tables for one FFT size: stage twiddle triplets and the bit-reversal map.
struct FftTables {
  int n = 0;
  float2* tri = nullptr;   // device
  int32_t* brev = nullptr; // device
};

// Tables for both axes of a plan.
struct PlanTables {
  const float2* triW;
  const int32_t* brevW;
  const float2* triH;
  const int32_t* brevH;
  const float2* triH2;
  const int32_t* brevH2;
};

// ---------------------------------------------------------- launchers --
// Implemented in cvm_fft.hip / cvm_norm.hip; all stream-ordered, no sync.

int tri_len(int n);
void launch_twiddles(float2* tri, int32_t* brev, int n, hipStream_t s);

// Spectrum column modes.
enum ColMode { COL_TSPEC = 0, COL_CORR = 1, COL_FWD = 2, COL_MULINV = 3, COL_CORRT = 4 };

// One member of a paired column launch (two Fleet members of one bucket,
// COL_MULINV, full-height tiles): its call state, template spectrum, product
// spectra (same layout as the frame's) and result extents.
struct ColPairMember {
  const CallState* st;
  const float2* tspec;
  float2* dst;
  int extW, extH;
};
// false when the shape has no paired kernel (a shrunk band, or a column
// height outside 2^6..2^11): launch the members one by one instead.
bool launch_cols_pair(const float2* src, int64_t tileStride, int64_t chanStride, int cnMax, const TileGrid& g,
                      const PlanTables& pt, int64_t tchanStride, int ntiles, const ColPairMember& A,
                      const ColPairMember& B, hipStream_t s);

// Forward packed row FFTs of every loaded row pair of every tile/channel.
// spec layout: [tile][channel][dftH][hwPad] complex (tile stride given).
// With st == nullptr all cnMax channels are transformed and the call state
// is not read (it may still be under construction on another stream).
void launch_row_fwd(const uint8_t* img, int64_t istride, int step, int cnMax, const CallState* st,
                    const TileGrid& g, const PlanTables& pt, float2* spec, int64_t tileStride, int64_t chanStride,
                    int ntiles, hipStream_t s);

// Column pass: see ColMode.  tspec/tspec2 are [channel][rows][hwPad].
void launch_cols(ColMode mode, const float2* src, float2* dst, int64_t srcTileStride, int64_t dstTileStride,
                 int64_t chanStride, int cnMax, const CallState* st, const TileGrid& g, const PlanTables& pt,
                 const float2* tspec, const float2* tspec2, int64_t tchanStride, int64_t tchanStride2, float scale,
                 int ntiles, hipStream_t s);

// Shrunk-band template spectrum: rows r<<shift of the full one times 2^shift.
// One tile row of a call (tiles [t0, t0 + ntiles)): its forward rows, and
// its column pass (the shrunk band's transform for the last row of a plan
// that has one).
void launch_row_fwd_range(const uint8_t* img, int64_t istride, int step, int cnMax, const CallState* st,
                          const TileGrid& g, const PlanTables& pt, float2* spec, int64_t tileStride,
                          int64_t chanStride, int t0, int ntiles, hipStream_t s);
void launch_cols_range(ColMode mode, const float2* src, float2* dst, int64_t tileStride, int64_t chanStride,
                       int cnMax, const CallState* st, const TileGrid& g, const PlanTables& pt, const float2* tspec,
                       const float2* tspec2, int64_t tchanStride, int64_t tchanStride2, int t0, int ntiles,
                       hipStream_t s);
void launch_shrink(const float2* tspec, float2* tspec2, int cnMax, int rows2, int shift, int hwPad,
                   int64_t chanStride, int64_t chanStride2, hipStream_t s);

// Inverse packed row FFTs + channel accumulation + emit into corr[rh][rw].
void launch_row_inv(const float2* spec, int64_t tileStride, int64_t chanStride, const CallState* st, int cnMax,
                    const TileGrid& g, const PlanTables& pt, float* corr, int rw, int ntiles, hipStream_t s);

// Call-state set-up: zero, template statistics, alpha scans, finalisation.
void launch_state_init(CallState* st, hipStream_t s);
void launch_state_set(CallState* st, const TemplStats& ts, int tpl_alpha_bad, hipStream_t s);
void launch_tpl_stats(const uint8_t* tpl, int64_t tstride, int tw, int th, int cnMax, int step, CallState* st,
                      hipStream_t s);
void launch_alpha_scan(const uint8_t* pix, int64_t stride, int w, int h, CallState* st, int which, hipStream_t s);
// cn_mode: 1..4 fixed; 0 = RGBA auto (3 iff both alpha planes constant)
void launch_state_finalize(CallState* st, int tw, int th, int cn_mode, hipStream_t s);
// A one-shot call's state after launch_state_init in three launches instead
// of six: the template statistics, the template's alpha scan and the
// parent's first rows together, then the rest of the parent's alpha scan
// (cn_mode 0), then the finalisation.
void launch_call_state(const uint8_t* img, int64_t istride, int iw, int ih, const uint8_t* tpl, int64_t tstride,
                       int tw, int th, int cnMax, int step, int cn_mode, CallState* st, hipStream_t s);
// The geometry of launch_call_state's first launch; whole_parent: the
// whole parent's alpha scan in it, in the compact layout (rows8), for a
// launch that takes the state's independent parts beside other work.
CallStateArgs call_state_args(const uint8_t* img, int64_t istride, int iw, int ih, const uint8_t* tpl,
                              int64_t tstride, int tw, int th, int cnMax, int step, int cn_mode, CallState* st,
                              bool whole_parent);

// The compact form of a small one-shot call (one stream, independent jobs
// sharing launches): the parent's and the template's row transforms with
// the call state's statistics and alpha scans (csa, the whole parent scan
// in it; st's accumulators must be zero on entry); the correlation's column
// pass, each workgroup transforming its template slab first, with the chunk
// column sums, each workgroup finalising the state first; the inverse rows
// with the chunk prefix sums.  The normalisation and the reduction (which
// clears st's accumulators again) follow.  Every job runs the code of its
// own launch.
struct CompactCall {
  const uint8_t* img;
  int64_t istride;
  int iw, ih, step, cnMax, cn_mode, tw, th, rw, ntiles;
  TileGrid g, gt;  // the parent's tiles; the template as one tile
  float2* spec;    // the parent's tile spectra
  int64_t tileStride, chanStride;
  const uint8_t* tpl;
  int64_t tstride;
  float2* tspec;  // the template's row spectra, every channel
  int64_t tchan;
  float scale;  // 1 / (dftW dftH)
  PlanTables pt;
  CallState* st;
  CallStateArgs csa;
  float* corr;
  void* nscratch;  // the chunk sums
};
// Plans the compact form takes: no shrunk band, transforms of 32 to 2048.
bool compact_ok(const TileGrid& g);
void launch_compact_fft(const CompactCall& c, hipStream_t s);

// Exact-statistics normalisation + tail + per-workgroup extrema.
// map may alias corr (in place) or be null (extrema only).
// cnMax: the largest channel count the call can take (the device decides).
// qpref (0, or 4, 2, 1 result columns per lane of the wave form's strips)
// selects the strip width; a width that cannot hold the template is ignored.
// qpref | NORM_HALF: the same strips in half bands (NB_ROWS / 2 rows a wave).
constexpr int NORM_HALF = 8;
int norm_partials_count(int rw, int rh, int tw, int th, int cnMax, int qpref = 0);
bool norm_q_ok(int tw, int th, int cnMax, int q);
// bytes of scratch launch_normalize needs (chunk partial column sums)
size_t norm_scratch_bytes(int iw, int ih, int tw, int th, int cnMax);
// Per-chunk partial column sums the wave-form normalisation starts from
// (nothing for the int64 form); scratch as for launch_normalize, which must
// run after it.
void launch_norm_prep(const uint8_t* img, int64_t istride, int iw, int ih, int step, int tw, int th, int cnMax,
                      const CallState* st, void* scratch, hipStream_t s);
// exact: corr holds the int32 C' of launch_direct, scores are FP64-grade
// (needs the wave form: norm_wave_ok).
void launch_normalize(const uint8_t* img, int64_t istride, int iw, int ih, int step, int tw, int th, int rw, int rh,
                      int cnMax, const CallState* st, const float* corr, float* map, Extrema* partials,
                      void* scratch, hipStream_t s, bool exact = false, int qpref = 0);
// The wave-form normalisation applies (it is the one with an exact mode).
bool norm_wave_ok(int tw, int th, int cnMax);

// Direct correlation on int8 MFMA (cvm_direct.hip).  kstep: staged bytes per
// pixel (1 gray, 4 RGBA, 3 RGBA with cn = 3 staged without alpha).  B
// fragments of one template (cn from st when st is set), then C' per window.
int direct_kstep(int step, int cn);
int direct_nk(int tw, int kstep);
size_t direct_bfrag_bytes(int tw, int th, int kstep);
// LDS of a workgroup of `waves` waves (1, 2 or 4: direct_waves_ok)
size_t direct_lds_bytes(int tw, int kstep, int waves);
bool direct_waves_ok(int waves);
bool direct_shape_ok(int tw, int kstep);
int direct_cmax(int nk, int kstep);
void launch_bfrag(const uint8_t* tpl, int64_t tstride, int tw, int th, int kstep, int cn, const CallState* st,
                  int8_t* bf, hipStream_t s);
int launch_direct(const uint8_t* img, int64_t istride, int iw, int ih, int kstep, const int8_t* bf, int tw, int th,
                  int rw, int rh, int32_t* out, int waves, hipStream_t s);
// Final first-occurrence reduction into res6 (device) = {min_v, min_x, min_y, max_v, max_x, max_y}.
// Batched forms (one launch for many templates against one frame).
int direct_pitch_of(int tw, int kstep);
int direct_grid_x(int rw);
int direct_grid_y(int rh);
int launch_direct_batch(const uint8_t* img, int64_t istride, int iw, int ih, int kstep, int cmax,
                        const DirectJob* jobs, const int* wg_job, int nwg, int max_pitch, int waves, hipStream_t s);
// Fused correlation + FP64 scoring of many members against one frame (no
// C' map; one extrema partial per direct block; reduce with
// launch_reduce_batch over jobs whose gx x gy is the direct grid).  ps: the
// frame's chunk column sums (launch_norm_prep) for the members' channel
// count.  fused_ok: the form is built and the frame's layout allows it
// (rows 16-byte aligned for the staging, width a multiple of fused_cpt_of).
int fused_cpt_of(int cmax);
bool fused_ok(const uint8_t* img, int64_t istride, int iw, int kstep, int cmax, bool sq32);
size_t fused_lds_bytes(int max_pitch, int kstep, int cmax, bool sq32, int waves);
int launch_direct_fused(const uint8_t* img, int64_t istride, int iw, int ih, const uint32_t* ps, int kstep, int cmax,
                        bool sq32, const FusedJob* jobs, const int* wg_job, int nwg, int max_pitch, int waves,
                        hipStream_t s);
void norm_form(int tw, int th, int cnMax, int* q, int* sq32);
int norm_grid_x(int rw, int q);
int norm_grid_y(int rh);
int launch_norm_batch(const uint8_t* img, int64_t istride, int iw, int ih, int step, int q, int sq32, bool exact,
                      const NormJob* jobs, const int* wg_job, int nwg, hipStream_t s);
// one workgroup per job: its extrema partials -> its Res6
void launch_reduce_batch(const NormJob* jobs, int njobs, hipStream_t s);
// launch_reduce for the compact form, clearing st's accumulators afterwards.
void launch_reduce_clear(const Extrema* partials, int nparts, int rw, void* res6, CallState* st, hipStream_t s);
void launch_reduce(const Extrema* partials, int nparts, const CallState* st, int rw, void* res6, float* map_if_flat,
                   int rh, hipStream_t s);

void launch_rgba_to_gray(const uint8_t* pix, int64_t stride, int w, int h, uint8_t* out, hipStream_t s);

// Host transfers (cvm_xfer.hip).  A pageable host buffer of 1 MiB or more
// moves through pinned staging chunks with host_threads() copy threads;
// buffers from host_alloc and small ones go straight to the driver.
// xfer_h2d returns with a staged (pageable, 1 MiB or more) source already
// copied out; a host_alloc or small source is still read by the copy in
// flight on s (xfer_h2d_source_busy says which), so a caller that returns
// before s completes waits for that copy first.  xfer_d2h returns with the
// destination written, unless it is a host_alloc buffer (then the copy is
// left in flight on s).
int set_host_threads(int n);
int host_threads();
hipError_t host_alloc(size_t bytes, void** out);
hipError_t host_free(void* p);
hipError_t xfer_h2d(void* dst, const void* src, size_t n, hipStream_t s);
// xfer_h2d with done(bytes) called as each chunk's copy is enqueued on s.
hipError_t xfer_h2d_chunks(void* dst, const void* src, size_t n, hipStream_t s,
                           const std::function<int(size_t)>& done);
bool xfer_h2d_source_busy(const void* src, size_t n);
hipError_t xfer_release_stream(hipStream_t s);
hipError_t xfer_d2h(void* dst, const void* src, size_t n, hipStream_t s);
// Peaks (cvm_peaks.hip): threshold + local-maximum suppression over a map
// with rx/ry already clamped to the map (rx <= peaks_max_rx()).  Writes up to
// cap (x, y) pairs and values in raster order and the total count into the
// scratch's peaks_total slot.
size_t peaks_scratch_bytes(int rw, int rh);
int peaks_max_rx();
int* peaks_total(void* scratch, int rw, int rh);
void launch_peaks(const float* map, int rw, int rh, float thresh, int rx, int ry, void* scratch, int cap,
                  int32_t* xy, float* vals, hipStream_t s);

// the copy-thread pool alone (host memory only), for its self-test
void host_copy(void* dst, const void* src, size_t n, int threads);

// Autotuner (cvm_tune.hip): measured choices among candidates that change no
// output bit, per key, cached for the process and in a file.
// tune_choice: the override environment variable's value when it names a
// candidate, else the cached choice, else each candidate timed with time_of
// (ms; negative = not applicable) and the fastest taken (the first candidate
// when nothing was measured).  tune_note records a choice made elsewhere.
bool tuning_enabled();
int tune_choice(const std::string& what, const std::vector<int>& cands, const char* override_env,
                const std::function<float(int)>& time_of);
// tune_value: a measured value rather than a choice (the override's value,
// else the cached one, else measure(times) -- dflt when tuning is off)
int tune_value(const std::string& what, const char* override_env, int dflt,
               const std::function<int(std::vector<float>&)>& measure);
void tune_note(const std::string& what, int choice, const std::vector<float>& ms, const char* source);
int tune_report(char* buf, int cap);
// average ms of one run of `launch` on s (after one untimed run); -1 on failure
float time_launches(hipStream_t s, const std::function<int()>& launch, int reps);

}  // namespace cvm
