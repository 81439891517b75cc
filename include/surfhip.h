/*
 * surfhip.h -- C ABI of libsurfhip, the MI355X (gfx950) analyze/synthesize
 * path of the 3-D surfacelet transforms.
 *
 * The entry points replace what the reference's device extension does today
 * (ext/SurfaceletsAdaptExt.jl:25 `adapt_workspace`, and the broadcasts plus
 * vendor-FFT calls it routes every analyze!/synthesize! through).  A host
 * binding (Julia `ccall`, Python ctypes) keeps the reference API
 * (src/transforms/ns_surfacelet.jl:88-173,504-641 and
 * src/transforms/surfacelet.jl:126-199,292-464) and calls these.
 *
 * Conventions
 * - Volumes are column-major n1 x n2 x n3 (axis 1 contiguous), as in the
 *   reference.  Real data is float32; complex data is interleaved float32
 *   pairs (ComplexF32) and runs as two real channels.
 * - NS coefficients are (n1, n2, n3, S) column-major: subband i (1-based) at
 *   element offset (i-1)*n1*n2*n3.  Decimated coefficients are one flat vector
 *   with subbands in subband_layout order (see surfhip_s_layout).
 * - Every pointer passed to analyze/synthesize/stream calls is a DEVICE
 *   pointer.  Calls enqueue work on the plan's HIP stream and return without
 *   synchronising; the caller synchronises the stream before reading results.
 * - A plan serves one call at a time (as the reference's workspaces:
 *   src/transforms/ns_surfacelet.jl:51-53, src/transforms/surfacelet.jl:40-42).
 * - Status codes: SURFHIP_OK, or an error whose message surfhip_last_error
 *   returns (thread-local).  ARG maps to the reference's ArgumentError, DIM to
 *   DimensionMismatch.
 */
#ifndef SURFHIP_H
#define SURFHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SURFHIP_OK = 0,
  SURFHIP_ERR_ARG = 1,     /* ArgumentError */
  SURFHIP_ERR_DIM = 2,     /* DimensionMismatch */
  SURFHIP_ERR_DEVICE = 3,  /* HIP runtime failure */
  SURFHIP_ERR_OOM = 4      /* device allocation failure */
};

enum { SURFHIP_F32 = 0, SURFHIP_C64 = 1 };

#define SURFHIP_MAX_SCALES 16

/* NSSurfacelet parameters: src/types/transforms.jl:86-113. */
typedef struct {
  int32_t scales;
  int32_t direction_levels[SURFHIP_MAX_SCALES]; /* coarsest scale first */
  int32_t bernstein_order;
  int32_t halfband_order;
  double transition;
  int32_t tight;
} surfhip_ns_params;

/* Surfacelet (decimated) parameters: src/types/transforms.jl:222-261.
 * levels[j][K][m]: checkerboard stages along axis m inside hourglass K at
 * scale j (coarsest first), diagonal zero -- the level matrix of
 * src/types/transforms.jl:270-284. */
typedef struct {
  int32_t scales;
  int32_t levels[SURFHIP_MAX_SCALES][3][3];
  double pyramid_ratio;     /* 1, 1.5 or 2 */
  int32_t checkerboard_order;
  int32_t diamond_length;
  double kaiser_beta;
  int32_t diamond_power;
} surfhip_s_params;

/* The directional filter bank alone (make_hourglass / make_ndfb:
 * src/directional/hourglass.jl:149-170, src/directional/ndfb.jl:91-104);
 * levels = 0 is the hourglass bank. */
typedef struct {
  int32_t levels;
  int32_t bernstein_order;
  int32_t halfband_order;
  double transition;
  int32_t tight;
} surfhip_ndfb_params;

typedef struct surfhip_ns_plan surfhip_ns_plan;
typedef struct surfhip_s_plan surfhip_s_plan;

/* Message of the last failed call on this thread; returns its full length. */
size_t surfhip_last_error(char *buf, size_t n);

/* Library version string, e.g. "surfhip 0.1 gfx950". */
const char *surfhip_version(void);

/* ------------------------------------------------------------ NSSurfacelet */

/* Subband count, host only: src/transforms/ns_surfacelet.jl:527
 * (nsubbands) and src/types/transforms.jl:157. */
int surfhip_ns_nsubbands(const surfhip_ns_params *p, int32_t *out);

/* make_workspace: src/transforms/ns_surfacelet.jl:88-123,165-215.  Builds
 * the one- and two-dimensional FP64 tables (filters, wedge planes, pyramid
 * profiles) on the host and uploads them rounded once to FP32; the
 * volume-sized group-factor tables are then built on the device in FP64
 * from those inputs and rounded once to FP32.  Allocates the plan's buffers
 * on the current device.  `stream` is a hipStream_t (NULL: the default
 * stream).  Errors as the reference: "volume size must be positive", "too
 * small for J scales"; SURFHIP_ERR_OOM when a device allocation fails. */
int surfhip_ns_plan_create(surfhip_ns_plan **out, const surfhip_ns_params *p,
                           int32_t n1, int32_t n2, int32_t n3, int32_t dtype,
                           void *stream);
int surfhip_ns_plan_destroy(surfhip_ns_plan *plan);
int surfhip_ns_set_stream(surfhip_ns_plan *plan, void *stream);

/* make_workspace(...; tune = true): the reference accepts `tune` and
 * ignores it (src/transforms/ns_surfacelet.jl:90,147; FFTW's own planner
 * flags play that role on the CPU).  Here it runs the plan's autotuner: for
 * sizes outside the factored kernels' grids it times the plan's own launch
 * sequence on its stream and keeps the fastest block shapes and radix orders
 * (synchronises the stream).  config (NULL or 6 int32): factored (1/0),
 * rows per axis-1 block, pencils per axis-2/3 block, and each axis's radix
 * order as decimal digit pairs (e.g. 804 = radix 8 then 4). */
int surfhip_ns_plan_tune(surfhip_ns_plan *plan, int32_t *config);

/* analyze!: src/transforms/ns_surfacelet.jl:551-566.  x: n1 n2 n3 volume,
 * coeffs: (n1, n2, n3, S). */
int surfhip_ns_analyze(surfhip_ns_plan *plan, const void *x, void *coeffs);

/* analyze! that allocates no device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * plan's analysis runs with its working spectra and slabs in a static device
 * array of 64 MiB per GPU instead of the plan's own buffers; the
 * coefficients equal surfhip_ns_analyze's bit for bit.  A plan whose
 * analysis needs more working memory (e.g. 128^3 on the factored kernels,
 * 179 MB) is refused with SURFHIP_ERR_ARG.  All calls on one GPU share that
 * array: calls that overlap in time on one GPU race.  The first call on a
 * device looks up the array's address (loading it); later calls allocate
 * nothing. */
int surfhip_ns_analyze_noalloc(surfhip_ns_plan *plan, const void *x, void *coeffs);

/* synthesize!: src/transforms/ns_surfacelet.jl:568-581. */
int surfhip_ns_synthesize(surfhip_ns_plan *plan, const void *coeffs, void *x);

/* Adjoints for reverse-mode rules (no counterpart functions in the
 * reference; a Julia caller would use them in a ChainRulesCore rrule).  Each
 * subband operator irfft . M_k . rfft is self-adjoint (M_k real and even), so
 *   surfhip_ns_adjoint:            x = A^T c, the accumulation with the analysis
 *                                  masks (= synthesize! for a tight bank:
 *                                  src/transforms/ns_surfacelet.jl:21-25);
 *   surfhip_ns_synthesize_adjoint: c = S^T x, the analysis with the synthesis
 *                                  masks (= analyze! for a tight bank). */
int surfhip_ns_adjoint(surfhip_ns_plan *plan, const void *coeffs, void *x);
int surfhip_ns_synthesize_adjoint(surfhip_ns_plan *plan, const void *x, void *coeffs);

/* Streaming analysis, analyze(f, t, x): src/transforms/ns_surfacelet.jl:
 * 474-479,620-624.  stream_begin computes the spectrum of x; each
 * stream_subband(i) then writes subband i (1-based, subband_layout order)
 * into `band` (one volume).  The reference calls f for i = 2..S, then 1. */
int surfhip_ns_stream_begin(surfhip_ns_plan *plan, const void *x);
int surfhip_ns_stream_subband(surfhip_ns_plan *plan, int32_t i, void *band);

/* Streaming synthesis, synthesize!(f, x, t): src/transforms/ns_surfacelet.jl:
 * 483-493,637-641.  inverse_begin clears the accumulator; accumulate adds
 * subband i; finish writes the volume. */
int surfhip_ns_stream_inverse_begin(surfhip_ns_plan *plan);
int surfhip_ns_stream_accumulate(surfhip_ns_plan *plan, int32_t i, const void *band);
int surfhip_ns_stream_finish(surfhip_ns_plan *plan, void *x);

/* Batched transforms of many volumes of uneven sizes (no counterpart in the
 * reference, which transforms one volume per call,
 * src/transforms/ns_surfacelet.jl:551-581, and has no device batching,
 * :147).  A batch holds one plan per (lane, volume size) for `nvol` volumes
 * (dims: n1 n2 n3 per volume) and spreads the volumes over `lanes` streams,
 * longest first onto the least-loaded lane.  surfhip_ns_batch_run with mode
 * 0 analyzes xs[k] into cs[k], mode 1 synthesizes cs[k] into ys[k], mode 2
 * does both, for every k; it is ordered after the batch stream's earlier work
 * and before its later work.  A run is recorded once as a hipGraph and
 * replayed with one launch while the mode and the pointers repeat; a batch
 * stream that is itself being captured gets direct launches.  Results equal
 * each volume's lone surfhip_ns_analyze / _synthesize on a plan of the same
 * configuration, bit for bit.  flags: SURFHIP_BATCH_TUNE runs each plan's
 * autotuner, SURFHIP_BATCH_NO_GRAPH always launches directly.
 * surfhip_ns_batch_info returns each volume's lane and whether a graph is
 * recorded. */
typedef struct surfhip_ns_batch surfhip_ns_batch;
enum { SURFHIP_BATCH_TUNE = 1, SURFHIP_BATCH_NO_GRAPH = 2 };
int surfhip_ns_batch_create(surfhip_ns_batch **out, const surfhip_ns_params *p, int32_t nvol,
                            const int32_t *dims, int32_t dtype, int32_t lanes, int32_t flags,
                            void *stream);
int surfhip_ns_batch_run(surfhip_ns_batch *b, int32_t mode, const void *const *xs,
                         void *const *cs, void *const *ys);
int surfhip_ns_batch_set_stream(surfhip_ns_batch *b, void *stream);
int surfhip_ns_batch_info(surfhip_ns_batch *b, int32_t *lane_of, int32_t *graph_recorded);
int surfhip_ns_batch_destroy(surfhip_ns_batch *b);

/* Host only: the analysis mask of subband i (1-based) on the rfft grid,
 * (n1/2+1, n2, n3) column-major, as the device evaluates it (the same FP32
 * tables and expression).  The inspection counterpart of hourglass_mask /
 * ndfb_mask (src/directional/hourglass.jl:264-270). */
int surfhip_ns_mask_host(const surfhip_ns_params *p, int32_t n1, int32_t n2,
                         int32_t n3, int32_t i, float *out);

/* make_ndfb(T, dims; levels, ...) / make_hourglass(T, dims; ...)
 * (src/directional/ndfb.jl:91-104, src/directional/hourglass.jl:149-170).
 * The bank runs on the NS kernels as a plan of one scale without the
 * pyramid: B = 1 and a zero lowpass, so subband 1 of the plan is zero and
 * subbands 2..S are the bank's 3 * 2^levels directional subbands in
 * ndfb_subband_index order (src/directional/ndfb.jl:184-190).
 * surfhip_ns_analyze / _synthesize / streaming / _adjoint serve it:
 * ndfb_decompose (:263-346) is the analysis, ndfb_reconstruct (:384-468) the
 * synthesis (a plain sum when not tight, since the masks are then a partition
 * of unity and the synthesis masks are 1). */
int surfhip_ndfb_plan_create(surfhip_ns_plan **out, const surfhip_ndfb_params *p,
                             int32_t n1, int32_t n2, int32_t n3, int32_t dtype,
                             void *stream);
/* ndfb_mask(fb, idx) (src/directional/ndfb.jl:215-220): the analysis mask of
 * directional subband idx (1-based) on the rfft grid, as the device evaluates
 * it; out has (n1/2+1) n2 n3 floats, axis 1 fastest.  Host only. */
int surfhip_ndfb_mask_host(const surfhip_ndfb_params *p, int32_t n1, int32_t n2,
                           int32_t n3, int32_t idx, float *out);

/* Group-factor tables (no counterpart in the reference: an inspection hook
 * for tests).  A plan builds its tables on the device in FP64, rounded once
 * to FP32; surfhip_ns_group_table copies table g (0-based; the lowpass is the
 * last) of the analysis (synthesis = 0) or synthesis masks to out, H floats
 * ordered [k1][k2][k3] with k3 fastest.  surfhip_ns_group_table_host builds
 * the same table on the host with the same code. */
int surfhip_ns_group_table(surfhip_ns_plan *plan, int32_t g, int32_t synthesis, float *out);
int surfhip_ns_group_table_host(const surfhip_ns_params *p, int32_t n1, int32_t n2,
                                int32_t n3, int32_t g, int32_t synthesis, float *out);

/* Per-kernel timing (no counterpart in the reference: a profiling hook for
 * benchmarks).  With timing on, every kernel launch of the plan is bracketed
 * by HIP events on the plan's stream.  surfhip_ns_timing synchronises the
 * stream and returns, per kernel class, the number of launches and their
 * summed duration in milliseconds since timing was last switched on; class
 * names are written comma-separated into `names`.  Classes of the factored
 * path: 0 pencil analysis, 1 plane analysis (subband write), 2 plane
 * synthesis (subband read), 3 pencil synthesis, 4 other.  Of the general
 * path: 0 axis-1 r2c (volume and subband read), 1 axis-1 c2r (subband and
 * volume write), 2 axes 2-3 forward, 3 axes 2-3 inverse, 4 other. */
#define SURFHIP_NS_TIMING_CLASSES 5
int surfhip_ns_set_timing(surfhip_ns_plan *plan, int32_t enable);
int surfhip_ns_timing(surfhip_ns_plan *plan, int64_t *launches, double *ms, char *names, size_t n);

/* ------------------------------------------------------ Surfacelet (dec.) */

/* Size check with the reference's message (including "the nearest length
 * that fits"): src/transforms/surfacelet.jl:230-269.  Host only. */
int surfhip_s_check_size(const surfhip_s_params *p, int32_t n1, int32_t n2, int32_t n3);

/* Subband count and total coefficient count, host only:
 * src/transforms/surfacelet.jl:275-290,310. */
int surfhip_s_counts(const surfhip_s_params *p, int32_t n1, int32_t n2,
                     int32_t n3, int32_t *nsub, int64_t *ncoef);

/* subband_layout sizes (3 per subband) and element offsets (1 per subband)
 * of the flat coefficient vector, host only. */
int surfhip_s_layout(const surfhip_s_params *p, int32_t n1, int32_t n2,
                     int32_t n3, int64_t *sizes3, int64_t *offsets, int32_t cap);

int surfhip_s_plan_create(surfhip_s_plan **out, const surfhip_s_params *p,
                          int32_t n1, int32_t n2, int32_t n3, int32_t dtype,
                          void *stream);
int surfhip_s_plan_destroy(surfhip_s_plan *plan);
int surfhip_s_set_stream(surfhip_s_plan *plan, void *stream);

/* analyze! / synthesize!: src/transforms/surfacelet.jl:446-464. */
int surfhip_s_analyze(surfhip_s_plan *plan, const void *x, void *coeffs);
int surfhip_s_synthesize(surfhip_s_plan *plan, const void *coeffs, void *x);

/* Integer index maps of a plan, read back from the device or computed there
 * by the same device functions the transform kernels use (no counterpart
 * function in the reference: an inspection hook so that the reference's
 * integer arithmetic can be checked bit for bit).  `level` is the pyramid
 * level and scale index, finest first (0-based); axis and hourglass K are
 * 0-based.
 *   kind 0: the box bins idx of level `level` along `axis`, 0-based
 *           (src/pyramid/fourier_pyramid.jl:68-85: [0:P; n-P:n-1]);
 *   kind 1: their lowpass bins idx_low ([0:P; n_low-P:n_low-1]);
 *   kind 2: for each lowpass bin l along `axis`, the box coordinates the
 *           split gathers from, two values per bin, -1 when absent (two
 *           when 2P = n_low, as src/pyramid/fourier_pyramid.jl:108-118
 *           assigns the later one);
 *   kind 3: the Hermitian mirror of hourglass `hourglass` at scale `level`
 *           along `axis` (src/directional/surfacelet_scale.jl:246-257,
 *           mod(1 - i, n) + 1): for each index j, the index it reads, -1 for
 *           j on the direct half of axis K;
 *   kind 4: the flat element offset of each block of hourglass `hourglass`
 *           at scale `level`, in column-major block order
 *           (src/directional/surfacelet_scale.jl:57-60,
 *           src/transforms/surfacelet.jl:78-81).
 * Writes min(cap, length) values to out (out may be NULL) and the map's
 * length to *count.  Synchronises the plan's stream for kinds 2 and 3. */
int surfhip_s_index_maps(surfhip_s_plan *plan, int32_t kind, int32_t level, int32_t hourglass,
                         int32_t axis, int64_t *out, int64_t cap, int64_t *count);

/* x = A^T c, the adjoint of analyze!.  The decimated transform is a frame but
 * not a tight one (src/transforms/surfacelet.jl:24-26), so this differs from
 * synthesize!; it is the backward of analyze in a reverse-mode rule (a
 * ChainRulesCore rrule on the Julia side).  No counterpart function in the
 * reference. */
int surfhip_s_adjoint(surfhip_s_plan *plan, const void *coeffs, void *x);

#ifdef __cplusplus
}
#endif
#endif /* SURFHIP_H */
