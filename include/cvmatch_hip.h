/*
 * cvmatch_hip.h -- C ABI of libcvmatch_hip.so, the MI355X (gfx950) engine for
 * OpenCV-compatible TM_CCOEFF_NORMED template matching + minMaxLoc.
 *
 * The entry points are what the reference's Go API binds to through a thin
 * cgo shim.  Each declaration names the reference function it replaces
 * (paths relative to the hkloudou/cvmatch source tree).  Conventions:
 *   - images are row-major uint8 with a byte stride >= width*step; step is 4
 *     for interleaved RGBA and 1 for gray; channel k sits at byte offset k;
 *   - result maps are row-major float32 of rw x rh, rw = iw-tw+1, rh = ih-th+1,
 *     element (x, y) at map[y*rw + x];
 *   - every call is synchronous: on return the result (and map) are written;
 *   - functions return CVM_OK or a CVM_E_* code; cvm_last_error_detail()
 *     returns the calling thread's last message, worded as the reference's
 *     panic strings so a shim can re-panic with them verbatim;
 *   - all functions are safe to call from several host threads.
 * Output bits are identical to the reference's scalar op sequence on the same
 * inputs (same tile plan, twiddles, packing, rounding order; no FMA).
 */
#ifndef CVMATCH_HIP_H
#define CVMATCH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The six-tuple of Match/MatchGray/Find and Result (fleet.go:13-18). */
typedef struct {
  float min_v;
  int32_t min_x, min_y;
  float max_v;
  int32_t max_x, max_y;
} cvm_result;

enum {
  CVM_OK = 0,
  CVM_E_BAD_ARGS = 1,    /* "cvmatch: bad match arguments (...)"  impl.go:1129, matcher.go:98 */
  CVM_E_STATS_CAP = 2,   /* "cvmatch: template area ... exact-statistics bound" impl.go:1132 */
  CVM_E_EMPTY_FLEET = 3, /* "cvmatch: empty fleet" fleet.go:87 */
  CVM_E_MIXED_FLEET = 4, /* "cvmatch: fleet mixes color and gray matchers" fleet.go:95 */
  CVM_E_HIP = 5,         /* HIP runtime failure (no reference counterpart) */
  CVM_E_UNSUPPORTED = 6, /* geometry beyond the library's kernel limits (no reference counterpart) */
  CVM_E_NIL_IMAGE = 7,   /* "cvmatch: nil image" cvmatch.go:172 */
  CVM_E_EMPTY_IMAGE = 8  /* "cvmatch: empty image" cvmatch.go:176 */
};

typedef struct cvm_matcher cvm_matcher;
typedef struct cvm_fleet cvm_fleet;

/* Thread-local detail of the last failing call on this thread. */
const char *cvm_last_error_detail(void);

/* Library version string. */
const char *cvm_version(void);

/* ---------------------------------------------------------------- core -- */

/* matchU8 impl.go:1126-1155: the whole pipeline on host buffers.
 * cn in [1,4], step >= cn.  map: NULL (extrema only) or rw*rh floats. */
int cvm_match_u8(const uint8_t *img, int64_t istride, int32_t iw, int32_t ih, const uint8_t *tpl, int64_t tstride,
                 int32_t tw, int32_t th, int32_t cn, int32_t step, float *map, cvm_result *out);

/* Match cvmatch.go:75-94 on interleaved RGBA host buffers: cn = 3 when both
 * alpha planes are constant (cvmatch.go:82-85), else 4. */
int cvm_match(const uint8_t *img, int64_t istride, int32_t iw, int32_t ih, const uint8_t *tpl, int64_t tstride,
              int32_t tw, int32_t th, cvm_result *out);

/* MatchMap cvmatch.go:100-117: as cvm_match plus the full response map
 * (caller-owned, rw*rh floats). */
int cvm_match_map(const uint8_t *img, int64_t istride, int32_t iw, int32_t ih, const uint8_t *tpl, int64_t tstride,
                  int32_t tw, int32_t th, float *map, cvm_result *out);

/* MatchGray cvmatch.go:128-139.  istep/tstep: 1 = already gray (the
 * *image.Gray / YCbCr-Y views of toGray cvmatch.go:201-217), 4 = RGBA/NRGBA
 * bytes converted with rgbToGray's BT.601 weights (cvmatch.go:222-240) on
 * the device. */
int cvm_match_gray(const uint8_t *img, int64_t istride, int32_t iw, int32_t ih, int32_t istep, const uint8_t *tpl,
                   int64_t tstride, int32_t tw, int32_t th, int32_t tstep, cvm_result *out);

/* alphaConst cvmatch.go:146-168: *out = 1 iff every alpha byte equals the first. */
int cvm_alpha_const(const uint8_t *pix, int64_t stride, int32_t w, int32_t h, int32_t *out);

/* rgbToGray cvmatch.go:222-240: out is w*h bytes (stride w). */
int cvm_rgba_to_gray(const uint8_t *pix, int64_t stride, int32_t w, int32_t h, uint8_t *out);

/* statsCap impl.go:1118-1124: largest template area with exact int64 statistics. */
int64_t cvm_stats_cap(int32_t cn);

/* newGoPlan impl.go:246-295: the tile geometry the engine uses
 * (out = {dftW, dftH, blockW, blockH}); returns the model cost. */
int64_t cvm_plan(int32_t tw, int32_t th, int32_t rw, int32_t rh, int32_t *out);

/* ------------------------------------------------------------- Matcher -- */

/* NewMatcher matcher.go:45-53 (step 4) / NewGrayMatcher matcher.go:57-64
 * (step 1, tpl already gray).  The template bytes are copied; the caller's
 * buffer is not retained. */
int cvm_matcher_new(const uint8_t *tpl, int64_t stride, int32_t w, int32_t h, int32_t step, cvm_matcher **out);

/* (*Matcher).Find matcher.go:83-130: bit-identical to cvm_match (step 4) or
 * cvm_match_gray on gray inputs (step 1).  img has the matcher's step. */
int cvm_matcher_find(cvm_matcher *m, const uint8_t *img, int64_t stride, int32_t w, int32_t h, cvm_result *out);

/* Releases a matcher (the reference relies on the Go GC; the shim calls this
 * from a finalizer).  A fleet keeps its own reference to its members. */
void cvm_matcher_free(cvm_matcher *m);

/* ------------------------------------------------------- frame streams -- */
/* No reference counterpart: Find over a stream of captured frames (a
 * screenshot loop) with the upload of one frame overlapping the matching of
 * the previous one.  cvm_frames_new: up to depth (2..8) frames in flight.
 * cvm_frames_find submits Find (matcher.go:83-130; direct != 0: the direct
 * form of cvm_matcher_find_direct) of m on a host frame and returns once the
 * frame has been copied out of the caller's buffer, which may then be
 * reused; with depth frames already in flight it first completes the
 * oldest.  cvm_frames_result returns the oldest submitted frame's six-tuple
 * and its tag, waiting for it; CVM_E_BAD_ARGS when nothing is in flight.
 * Each result equals cvm_matcher_find (or _direct) on that frame, bit for
 * bit.  A stream holds a reference to each matcher while its frames are in
 * flight.  One stream is used by one thread at a time. */
typedef struct cvm_frames cvm_frames;
int cvm_frames_new(int32_t depth, cvm_frames **out);
int cvm_frames_find(cvm_frames *q, cvm_matcher *m, const uint8_t *img, int64_t stride, int32_t w, int32_t h,
                    int32_t direct, int64_t tag);
int cvm_frames_result(cvm_frames *q, int64_t *tag, cvm_result *out);
int32_t cvm_frames_pending(cvm_frames *q);
void cvm_frames_free(cvm_frames *q);

/* --------------------------------------------------------------- Fleet -- */

/* NewFleet fleet.go:85-100: members bucketed by shape dominance
 * (bucketize fleet.go:111-171).  The member list is copied. */
int cvm_fleet_new(cvm_matcher *const *ms, int32_t n, cvm_fleet **out);

/* (*Fleet).FindAll fleet.go:176-332: out receives n results in
 * construction order. */
int cvm_fleet_find_all(cvm_fleet *f, const uint8_t *img, int64_t stride, int32_t w, int32_t h, cvm_result *out);

void cvm_fleet_free(cvm_fleet *f);

/* -------------------------------------------------- every occurrence -- */
/* The step after MatchMap that README.md:82-90 describes and
 * TestFindAllOccurrences (cvmatch_test.go:320-381) verifies, run on the
 * device so the map never crosses PCIe: (x, y) is a peak unless
 * map[y*w+x] < thresh or an element within |dx| <= rx, |dy| <= ry (inside
 * the map) is strictly greater; tied maxima are all peaks.  Peaks are
 * written in raster order (the test's loop order): xy[2i], xy[2i+1] and
 * vals[i] for i < min(count, cap); *count receives the total, which may
 * exceed cap.  The test uses thresh 0.999 and rx, ry = half the template.
 *
 * cvm_match_peaks: matchU8 (impl.go:1126-1155) on host buffers, then the
 * peaks of its map; cn = 0 selects Match's RGBA alpha rule (step 4). */
int cvm_match_peaks(const uint8_t *img, int64_t istride, int32_t iw, int32_t ih, const uint8_t *tpl, int64_t tstride,
                    int32_t tw, int32_t th, int32_t cn, int32_t step, float thresh, int32_t rx, int32_t ry,
                    int32_t *xy, float *vals, int32_t cap, int32_t *count, cvm_result *out);
/* The peaks of a w x h map already in device memory; d_xy / d_vals are
 * device buffers, *count is written on the host when the call returns.
 * stream: a hipStream_t or NULL for the library's stream. */
int cvm_find_peaks_device(const float *d_map, int32_t w, int32_t h, float thresh, int32_t rx, int32_t ry,
                          int32_t *d_xy, float *d_vals, int32_t cap, int32_t *count, void *stream);

/* ------------------------------------------------- direct (mixed) form -- */
/* The same calls with the window x template sum computed exactly on int8
 * MFMA (bytes offset by 128, int32 accumulation) and every score evaluated
 * in FP64 from exact integers, then rounded once to float32: each score is
 * the FP64 oracle's (the TM_CCOEFF_NORMED formula with OpenCV's guard ladder
 * as in normOne impl.go:816-849, evaluated in binary64) rounded to float32.
 * Not bit-identical to the plain calls, which reproduce the reference's
 * float32 FFT rounding.  Applies when tw*th*cn < 131072 and tw, th <= 257,
 * on gray (cn = 1, step = 1) or RGBA (step = 4, cn = 3, 4 or 0 for the alpha
 * rule) pixels; other shapes and layouts return CVM_E_BAD_ARGS.  Arguments
 * and results as in the plain
 * counterparts: matchU8 impl.go:1126-1155, Find matcher.go:83-130, FindAll
 * fleet.go:176-332 (FindAll's members are not bucketed here: each result
 * equals that member's direct Find). */
int cvm_match_u8_direct(const uint8_t *img, int64_t istride, int32_t iw, int32_t ih, const uint8_t *tpl,
                        int64_t tstride, int32_t tw, int32_t th, int32_t cn, int32_t step, float *map,
                        cvm_result *out);
int cvm_matcher_find_direct(cvm_matcher *m, const uint8_t *img, int64_t stride, int32_t w, int32_t h,
                            cvm_result *out);
int cvm_fleet_find_all_direct(cvm_fleet *f, const uint8_t *img, int64_t stride, int32_t w, int32_t h,
                              cvm_result *out);

/* ------------------------------------------------------ the mixed form -- */
/* No reference counterpart: FindAll (fleet.go:176-332) with each member on
 * the form the autotuner picks for this frame geometry.  On the first call of
 * a geometry the tuner measures the crossover -- the smallest square template
 * side at which the plain FFT call on this frame is faster than the direct
 * one -- and, for every non-flat member at least that wide or tall, the
 * largest |FFT score - exact score| over its whole map on that frame; a
 * member within 1e-5 (CVMATCH_MIXED_GATE sets another bound) takes the FFT form (its result is then the plain Find's,
 * bit for bit), every other member the direct form (its result is the direct
 * Find's); members outside the direct form's shapes take the FFT form.
 * cvm_fleet_mixed_plan reports the last plan: routed[i] = 1 for members on
 * the FFT form, gate[i] their measured deviation (-1: not a candidate, -2:
 * outside the direct shapes), *crossover the side (0: no member routed by
 * speed).  n is the member count. */
int cvm_fleet_find_all_mixed(cvm_fleet *f, const uint8_t *img, int64_t stride, int32_t w, int32_t h,
                             cvm_result *out);
int cvm_fleet_find_all_mixed_device(cvm_fleet *f, const uint8_t *d_img, int64_t stride, int32_t w, int32_t h,
                                    cvm_result *out, void *stream);
int cvm_fleet_mixed_plan(cvm_fleet *f, int32_t *routed, float *gate, int32_t n, int32_t *crossover);

/* No reference counterpart: the autotuner's decisions so far as a JSON
 * object (key -> choice, source, candidate times in ms).  Writes at most cap
 * bytes (NUL included) and returns the full length.  The choices it makes
 * (waves per direct workgroup, scoring strip width, the mixed form's
 * crossover) change no output bit; CVMATCH_TUNE=0 turns measuring off,
 * CVMATCH_TUNE_FILE names its cache file, CVMATCH_DIRECT_WAVES,
 * CVMATCH_NORM_Q and CVMATCH_CROSSOVER force a choice. */
int cvm_tune_report(char *buf, int32_t cap);

/* ------------------------------------------------ device-resident calls -- */
/* No reference counterpart: the same pipeline on buffers already in device
 * memory (the benchmark's "inputs resident in HBM" case and callers that
 * keep frames on the GPU).  stream: a hipStream_t or NULL for the library's
 * stream.  cn = 0 selects Match's RGBA alpha rule (step must be 4). */
int cvm_match_u8_device(const uint8_t *d_img, int64_t istride, int32_t iw, int32_t ih, const uint8_t *d_tpl,
                        int64_t tstride, int32_t tw, int32_t th, int32_t cn, int32_t step, float *d_map,
                        cvm_result *out, void *stream);
/* The direct form of cvm_match_u8_device. */
int cvm_match_u8_direct_device(const uint8_t *d_img, int64_t istride, int32_t iw, int32_t ih, const uint8_t *d_tpl,
                               int64_t tstride, int32_t tw, int32_t th, int32_t cn, int32_t step, float *d_map,
                               cvm_result *out, void *stream);
/* No counterpart in the reference: cvm_match_u8_direct_device for callers
 * that cannot allocate device memory (code under HIP stream capture, serving
 * loops or harnesses that forbid hipMalloc).  Same arguments and results; it
 * allocates nothing: its scratch is one static device array of 64 MiB per
 * device (held until the process ends), which bounds the call to
 * direct-form shapes whose scratch fits it (a 3840x2160 RGBA frame with a
 * 128x128 template fits with or without a map); larger ones return
 * CVM_E_UNSUPPORTED.  stream NULL: the null stream.  Calls on one device
 * share the array and are serialised (each returns with its work done). */
int cvm_match_u8_direct_device_noalloc(const uint8_t *d_img, int64_t istride, int32_t iw, int32_t ih,
                                       const uint8_t *d_tpl, int64_t tstride, int32_t tw, int32_t th, int32_t cn,
                                       int32_t step, float *d_map, cvm_result *out, void *stream);

/* Find, FindAll and the gray conversion on frames already in device memory
 * (no reference counterpart): a caller that runs several fleets over one
 * captured frame -- one per region of interest, as the windowrules example
 * does (examples/windowrules/engine.go:276-310) -- uploads the frame once and
 * passes each region as a device view (pointer + row stride), converting to
 * gray once for all gray fleets.  Arguments and results as in
 * (*Matcher).Find matcher.go:83-130, (*Fleet).FindAll fleet.go:176-332 and
 * rgbToGray cvmatch.go:222-240; out is host memory, written on return. */
int cvm_matcher_find_device(cvm_matcher *m, const uint8_t *d_img, int64_t stride, int32_t w, int32_t h,
                            cvm_result *out, void *stream);
int cvm_fleet_find_all_device(cvm_fleet *f, const uint8_t *d_img, int64_t stride, int32_t w, int32_t h,
                              cvm_result *out, void *stream);
int cvm_fleet_find_all_direct_device(cvm_fleet *f, const uint8_t *d_img, int64_t stride, int32_t w, int32_t h,
                                     cvm_result *out, void *stream);
int cvm_rgba_to_gray_device(const uint8_t *d_pix, int64_t stride, int32_t w, int32_t h, uint8_t *d_out,
                            void *stream);

/* Device memory helpers for callers without their own allocator.  The two
 * copies are synchronous and take the staged host path described below. */
int cvm_device_alloc(size_t bytes, void **out);
int cvm_device_free(void *p);
int cvm_copy_h2d(void *dst, const void *src, size_t bytes);
int cvm_copy_d2h(void *dst, const void *src, size_t bytes);

/* ------------------------------------------------------- host transfers -- */
/* Every host-buffer entry point moves a pageable buffer of 1 MiB or more
 * through pinned staging chunks copied by host threads, so that the copy
 * overlaps the DMA of the previous chunk and runs at the pinned-memory PCIe
 * rate.  Memory from cvm_host_alloc is pinned already and skips the staging:
 * a caller that fills its frames there (a screenshot stream) moves them at
 * the DMA rate with no host copy.
 *
 * cvm_set_threads mirrors the reference's Threads variable (cvmatch.go:36-63)
 * and its CVMATCH_THREADS environment variable (read once): n <= 0 means
 * automatic (the CPUs this process may run on), other values are clamped to
 * [1, 16]; returns the resolved count, as threads() cvmatch.go:52-63 does.
 * Here it sizes the host copy threads only; no setting changes an output
 * bit.  cvm_threads returns the resolved count. */
int32_t cvm_set_threads(int32_t n);
int32_t cvm_threads(void);
/* Pinned host memory (no reference counterpart; Go frames live in the Go
 * heap).  cvm_host_free rejects pointers that cvm_host_alloc did not return. */
int cvm_host_alloc(size_t bytes, void **out);
int cvm_host_free(void *p);
/* Self-test of the host copy threads alone: copies bytes from src to dst
 * (host memory) split over `threads` threads; touches no GPU.  Tests only. */
int cvm_selftest_copy(void *dst, const void *src, size_t bytes, int32_t threads);
/* Self-test of the autotuner's locking (host only; tests): returns 2. */
int cvm_selftest_tuner(void);
int cvm_synchronize(void);
/* Device-to-device copy by a plain streaming kernel; also the known-size
 * kernel that calibrates the HBM traffic counters. */
int cvm_device_copy(void *dst, const void *src, size_t bytes);

/* ------------------------------------------------------------ profiling -- */
/* No reference counterpart.  When enabled, every pipeline stage is bracketed
 * by HIP events on the stream it runs on and its time accumulated per stage
 * name.  cvm_profile_read copies up to max entries; returns the entry count. */
int cvm_profile_enable(int32_t on);
int cvm_profile_reset(void);
int cvm_profile_read(char *names /* max*32 bytes */, double *total_ms, int64_t *count, int32_t max);

/* Self-test of the device arithmetic the bit contract relies on (binary64
 * sqrt/division rounding, int64->binary64->binary32 conversion).  Fills
 * out[i] for the n inputs; used by the tests only. */
int cvm_selftest_fp(const float *a, const float *b, const int64_t *ints, int32_t n, float *out_sqrt, float *out_div,
                    float *out_cvt);

/* Twiddle tables generated on the device (FP64), for parity tests:
 * tri receives the stage triplets (complex64, tri_len entries), brev n ints. */
int cvm_twiddles(int32_t n, float *tri, int32_t *brev, int32_t *tri_len);

#ifdef __cplusplus
}
#endif

#endif /* CVMATCH_HIP_H */
