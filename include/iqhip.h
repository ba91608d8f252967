/*
 * iqhip.h -- C ABI of the MI355X (gfx950) image-quilting hot path.
 *
 * Drop-in boundary for ImageQuilting.jl's `iqsim` tile loop.  The reference
 * has no formal plugin API; its backend seam is three load-time function
 * constants (imfilter.jl:28 `imfilter_kernel`, utils.jl:61 `array_kernel`,
 * utils.jl:67 `view_kernel`) called from `fastdistance` (utils.jl:5-13) and
 * the tile loop (iqsim.jl:163-312).  Each entry point below replaces one
 * reference function, so each can be checked on its own, and
 * iq_plan_* / iq_simulate replace the whole loop.
 *
 * Conventions
 *   * Arrays are column-major (first axis fastest), exactly Julia's memory
 *     layout, and always described by three extents (pass 1 for missing
 *     trailing axes).
 *   * Indices are 0-based (Julia wrapper adds 1).
 *   * Host buffers are borrowed for the duration of a call only.
 *   * Every function returns an iq_status; iq_last_error() gives the text.
 *     No C++ exception crosses the ABI and no call waits forever.
 *   * One iq_ctx = one device + one HIP stream; calls on one context are not
 *     thread safe, distinct contexts are independent.
 */
#ifndef IQHIP_H
#define IQHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum iq_status {
  IQ_OK = 0,
  IQ_ERR_ARG = -1,          /* invalid argument (mirrors the @asserts of iqsim.jl:69-89) */
  IQ_ERR_HIP = -2,          /* HIP runtime error */
  IQ_ERR_NOCAND = -3,       /* every candidate pattern is disabled (iqsim.jl:237 / relaxation.jl:20-36) */
  IQ_ERR_NOTINDB = -4,      /* a forced pick is not in the pattern database */
  IQ_ERR_ITER = -5,         /* min-cut iteration guard reached */
  IQ_ERR_UNSUPPORTED = -6,  /* size beyond what the library handles (message names the limit) */
  IQ_ERR_NOMEM = -7
} iq_status;

typedef enum iq_dtype { IQ_F32 = 0, IQ_F64 = 1 } iq_dtype;

/* Element types of raw (unprepared) arrays: floats, and the integer /
 * categorical types the reference converts to Float64 (utils.jl:96-101). */
typedef enum iq_raw_dtype {
  IQ_RAW_F32 = 0, IQ_RAW_F64 = 1, IQ_RAW_U8 = 2, IQ_RAW_I8 = 3, IQ_RAW_U16 = 4,
  IQ_RAW_I16 = 5, IQ_RAW_U32 = 6, IQ_RAW_I32 = 7, IQ_RAW_U64 = 8, IQ_RAW_I64 = 9
} iq_raw_dtype;

typedef struct iq_ctx iq_ctx;
typedef struct iq_plan iq_plan;

/* ---- context --------------------------------------------------------- */
/* Replaces the load-time backend choice `CUDA.functional()`
 * (imfilter.jl:28, utils.jl:61, utils.jl:67) with a runtime handle. */
int iq_device_count(int* count);
int iq_ctx_create(int device, iq_ctx** out);
int iq_ctx_destroy(iq_ctx* ctx);
const char* iq_last_error(const iq_ctx* ctx); /* ctx may be NULL: last error of any context on this thread */
int iq_ctx_synchronize(iq_ctx* ctx);
/* Page-locked host memory on the context's device (no counterpart in the
 * reference): iq_plan_results copies into such a buffer at the full PCIe
 * rate (a pageable buffer goes through a staging copy).  iq_host_free
 * releases a buffer iq_host_alloc returned (NULL is ignored). */
int iq_host_alloc(iq_ctx* ctx, size_t bytes, void** out);
int iq_host_free(iq_ctx* ctx, void* p);

/* ---- per-function entry points (host buffers in, host buffers out) ---- */

/* imfilter_cpu / imfilter_gpu (imfilter.jl:5-26): valid cross-correlation
 * out[c] = sum_j img[c+j] * krn[j], out extents = idims - kdims + 1. */
int iq_imfilter(iq_ctx* ctx, const double* img, const int64_t idims[3], const double* krn,
                const int64_t kdims[3], double* out);

/* fastdistance (utils.jl:5-13): D[c] = |A2 - 2 AB + B2| over the voxels
 * where weights != 0 (weights may be NULL = all ones).  img is float or
 * double per `img_dtype`; D is FP64, exact to the project's FP64 definition
 * (fused-multiply-add chains in ascending column-major tile order). */
int iq_fastdistance(iq_ctx* ctx, const void* img, int img_dtype, const int64_t idims[3],
                    const double* kern, const uint8_t* weights, const int64_t kdims[3], double* out);

/* Tol-band pick (iqsim.jl:237, empty-auxiliary branch):
 * ascending indices i with D[i] <= (1 + tol) * minimum(D). */
int iq_band_pick(iq_ctx* ctx, const double* D, int64_t n, double tol, int64_t* out_idx,
                 int64_t* out_count);

/* relaxation (relaxation.jl:5-48): ascending pattern database for primary
 * distances D and `naux` auxiliary maps Ds[0..naux). */
int iq_relaxation(iq_ctx* ctx, const double* D, const double* const* Ds, int naux, int64_t n,
                  double tol, int64_t* out_idx, int64_t* out_count);

/* taumodel (taumodel.jl:5-45): unnormalised probabilities of `events`
 * (indices into the maps) from the primary map D and auxiliary maps Ds. */
int iq_taumodel(iq_ctx* ctx, const int64_t* events, int64_t nev, const double* D,
                const double* const* Ds, int naux, int64_t n, double* out_p);

/* Weighted draw (iqsim.jl:243, StatsBase direct method) with a caller-drawn
 * uniform u in [0,1): the first i whose cumulative weight reaches u * sum(p),
 * cumulative sums taken in fixed 64-element blocks over three levels. */
int iq_sample(iq_ctx* ctx, const double* p, int64_t n, double u, int64_t* out_index);

/* graphcut (graphcut.jl:5-84): mask over the slab A/B (extents dims, type
 * dtype) cut along 0-based axis `dim`: 1 where the voxel cannot reach the
 * sink in the residual graph of a maximum flow (keep A), 0 elsewhere. */
int iq_graphcut(iq_ctx* ctx, const void* A, const void* B, int dtype, const int64_t dims[3], int dim,
                uint8_t* out_mask);

/* Allocation-free variant of iq_graphcut.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  A, B and
 * out_mask are DEVICE pointers owned by the caller (A and B: slabs of extents
 * dims and element type dtype, column-major; out_mask: prod(dims) bytes),
 * their contents ready before the call (the work runs on the context's
 * stream; the call waits for it and returns the cut's status).  The library
 * allocates nothing: the status word and the node state of slabs beyond 1,536
 * interior nodes live in one static array per device of
 * 256 + 100 * IQ_NOALLOC_MAX_NODES bytes (about 1.6 MB), held until the
 * process ends.  Slabs with more than IQ_NOALLOC_MAX_NODES interior nodes
 * (voxels outside the first and last slices along dim) give
 * IQ_ERR_UNSUPPORTED.  Every call on one device shares that array: calls that
 * overlap on one device (two contexts, or two host threads) race. */
#define IQ_NOALLOC_MAX_NODES 16384
int iq_graphcut_noalloc(iq_ctx* ctx, const void* A, const void* B, int dtype, const int64_t dims[3], int dim,
                        uint8_t* out_mask);

/* ---- host preparation of raw arrays (no device work, no context) ------ */

/* prepare (utils.jl:96-101): raw array (extents dims, element type
 * raw_dtype; missing[i] != 0 marks a missing value, NULL = none) to the
 * element type the tile loop computes in: IQ_RAW_F32 stays FP32, every other
 * type becomes FP64 (out holds prod(dims) elements of that type); missing
 * and NaN values become 0. */
int iq_prepare(const void* raw, int raw_dtype, const uint8_t* missing, const int64_t dims[3], void* out);

/* Soft grid of one soft pair as the tile loop reads it (imagepreproc,
 * utils.jl:74-80, read through views at grid coordinates, iqsim.jl:222-227):
 * padarray(aux, Pad(:symmetric, 0, padsize - min(padsize, size(aux)))),
 * prepared, first padsize voxels per axis.  out: prod(padsize) FP64. */
int iq_prepare_soft(const void* aux, int raw_dtype, const uint8_t* missing, const int64_t dims[3],
                    const int64_t padsize[3], double* out);

/* finddisabled (utils.jl:115-129): disabled[c] = 1 when the training-image
 * window at candidate c (extents tilesize) holds a missing or NaN voxel;
 * disabled has tisize - tilesize + 1 extents. */
int iq_finddisabled(const void* raw_ti, int raw_dtype, const uint8_t* missing, const int64_t tisize[3],
                    const int64_t tilesize[3], uint8_t* disabled);

/* ---- whole simulation (iqsim.jl:163-300) ------------------------------ */
/* Everything the tile loop reads, prepared on the host exactly as
 * imagepreproc / finddisabled / findskipped / genpath do (utils.jl:69-204);
 * geometry as iqsim.jl:92-115. */
typedef struct iq_plan_desc {
  const void* ti;               /* prepared training image (NaN -> 0), tisize */
  int ti_dtype;                 /* IQ_F32 or IQ_F64 (element type of the prepared TI) */
  int64_t tisize[3];
  int64_t tilesize[3];
  int64_t ovlsize[3];
  int64_t spacing[3];
  int64_t ntiles[3];
  int64_t padsize[3];
  const uint8_t* disabled;      /* distsize = tisize - tilesize + 1 (finddisabled, utils.jl:115) */
  const int64_t* path;          /* npath tile linear indices (genpath, utils.jl:158) */
  int64_t npath;
  const uint8_t* skipped;       /* prod(ntiles) flags (findskipped, utils.jl:131) */
  const double* hardgrid;       /* padsize, NaN where no hard datum; NULL = no hard data */
  int naux;                     /* soft pairs (iqsim.jl:222-227) */
  const double* const* aux;     /* naux grids of padsize (padded + prepared) */
  const double* const* auxti;   /* naux training images of tisize (prepared) */
  double tol;
  int64_t nreal;
} iq_plan_desc;

/* Argument checks of iq_plan_create_multi without any device work (the
 * @asserts of iqsim.jl:69-89 as they reach the ABI and the geometry
 * identities of iqsim.jl:92-115): IQ_OK, or IQ_ERR_ARG with the reason in
 * iq_last_error(NULL).  Lets a binding check the descriptors it builds. */
int iq_plan_validate(const iq_plan_desc* descs, int ngrid);

/* Upload everything once (the TI stays resident in HBM). */
int iq_plan_create(iq_ctx* ctx, const iq_plan_desc* desc, iq_plan** out);

/* One plan over several simulation grids of different sizes (the nreal loop
 * of iqsim.jl:163 with a grid per group of realizations; every grid's
 * realizations advance in lockstep with all others).  descs[0..ngrid) share
 * the training image, ti_dtype, tisize, tilesize, ovlsize, spacing, tol,
 * naux and auxti (taken from descs[0], together with disabled); each has its
 * own ntiles, padsize, path, npath, skipped, hardgrid, aux and nreal.
 * Realizations are numbered grid by grid.  iq_plan_run / iq_plan_results
 * then take and give, realization after realization, npath uniforms /
 * forced picks / chosen picks and padsize grid voxels of that realization's
 * grid.  iq_plan_create(ctx, d, p) == iq_plan_create_multi(ctx, d, 1, p). */
int iq_plan_create_multi(iq_ctx* ctx, const iq_plan_desc* descs, int ngrid, iq_plan** out);
int iq_plan_destroy(iq_plan* plan);

/* Run all realizations.  uniforms: nreal x npath (row r holds realization
 * r's draws, indexed by path position).  forced: NULL, or nreal x npath
 * pattern indices to replay (-1 = draw).  Results stay on the device until
 * iq_plan_results; the call returns once the work is queued and checked. */
int iq_plan_run(iq_ctx* ctx, iq_plan* plan, const double* uniforms, const int64_t* forced);

/* Copy results: out = nreal x padsize in the TI element type (before the
 * hard-data overwrite and crop of iqsim.jl:291-303), chosen = nreal x npath
 * picked pattern indices (-1 for skipped tiles), cutgrid = nreal x padsize
 * boundary-cut values 0/1 (iqsim.jl:281; NULL to skip).  Any may be NULL. */
int iq_plan_results(iq_ctx* ctx, iq_plan* plan, void* out, int64_t* chosen, double* cutgrid);

/* Seconds of device time of the last iq_plan_run (HIP events on the
 * context's stream). */
int iq_plan_last_time(iq_plan* plan, double* seconds);

/* Live timing of the distance scan (the dominant kernel) inside
 * iq_plan_run: when enabled, every scan launch is bracketed by HIP events on
 * the context's stream.  iq_plan_scan_stats reports the launches of the last
 * run, their summed device time, their algorithmic FLOPs (2 * candidates *
 * masked voxels per query) and the FLOPs of the bf16 matrix instructions the
 * MFMA scan issues (0 for the VALU scan).  Diagnostic only; off by default. */
int iq_plan_profile(iq_plan* plan, int enable);
int iq_plan_scan_stats(iq_plan* plan, int64_t* launches, double* seconds, double* flop, double* mfma_flop);

/* Relative error bound of the MFMA distance scan's correlation term (no
 * counterpart in the reference; host only): |AB32 - AB| <= *gerr * sum|a||b|
 * for a mask of nstep k-steps of 16 voxels, split != 0 for FP32 data (three
 * bf16 products per voxel pair), 0 for bf16-exact data (one product).  The
 * scan's interval around each exact FP64 distance is built from it. */
int iq_scan_error_bound(int64_t nstep, int split, double* gerr);

/* Min-cut launches of the last run and their summed device time (HIP events
 * around each launch, recorded while profiling is enabled; no counterpart in
 * the reference).  Diagnostic only. */
int iq_plan_cut_time(iq_plan* plan, int64_t* launches, double* seconds);

/* Scan autotuning result of iq_plan_create (no counterpart in the reference):
 * wavefronts per MFMA-scan workgroup (2, 3 or 4) chosen for launches of one
 * and of two or more 32-query tiles (nw[0], nw[1]; 0 when the plan has no
 * MFMA scan), and the measured microseconds per launch us[tiles-1][0: 2
 * wavefronts, 1: 4] (0 where not measured; the 3-wavefront timings are not
 * reported).  Diagnostic only. */
int iq_plan_scan_tuning(iq_plan* plan, int nw[2], float us[4]);

/* Shortlist bank-scan launches of the last run (learned proposal; HIP
 * events around each launch while profiling is enabled), their summed device
 * time and their algorithmic HBM bytes (per launch: the bank rows and A2 maps
 * of each mask shape it scans, P x (128 + 12) bytes).  Diagnostic only. */
int iq_plan_bank_stats(iq_plan* plan, int64_t* launches, double* seconds, double* bytes);

/* Min-cut work of the last iq_plan_run (graphcut.jl:73 replacement):
 * push-relabel rounds, global-relabel passes (team-wide relaxation passes
 * of the distance labels) and cuts (register-resident kernel; slabs beyond
 * its size limit are not counted).  Diagnostic only. */
int iq_plan_cut_stats(iq_plan* plan, int64_t* rounds, int64_t* levels, int64_t* cuts);

/* Selection work of the last iq_plan_run (plain queries with an overlap):
 * number of queries, summed sizes of the band supersets that were rechecked
 * in FP64, and summed band sizes (iqsim.jl:237).  Diagnostic only. */
int iq_plan_select_stats(iq_plan* plan, int64_t* queries, int64_t* superset, int64_t* band);

/* ---- learned proposal (no counterpart in the reference) ----------------
 * For plans with many candidate patterns: a shortlist of the K best
 * approximate distances of each plain query replaces the full scan.  The
 * shortlist is reranked with exact FP64 distances, so a pick always comes
 * from exact distances, but a band member outside the shortlist is missed.
 * Sampled batches also run the full scan, which decides them and measures
 * the recall of the shortlist on the exact band.
 *
 * iq_plan_learn: one linear encoder (64 outputs) per overlap-mask shape of
 * the plan, learned from nsamples training-image windows by iters rounds of
 * block subspace iteration (seeded; iters = 0 leaves a random orthonormal
 * encoder), and the bank of every candidate's encoding (64 bf16, 128 bytes
 * per candidate and mask shape).
 * iq_plan_set_proposal: use the shortlist in iq_plan_run (enable), its size
 * K, a full-scan check every check_every-th batch (0: never), and return to
 * the full scan for the rest of a run once the cumulative recall of the
 * checks falls below min_recall: after at least 256 checked queries, or at
 * once when the recall is below 1 - 4 (1 - min_recall) (so min_recall = 1
 * falls back at the first band member missed).
 * iq_plan_set_check_sample: in a check batch, only `queries` queries (0:
 * all of them, the default), an evenly strided sample of the batch, go to
 * the full scan, which decides them and measures recall on their exact
 * bands; the shortlist decides the others.  No counterpart in the reference: the sampled recall check of
 * the learned proposal.
 * iq_plan_proposal_stats: for the last run, out[0] band members of the
 * checked queries, out[1] those inside the shortlist, out[2] checked
 * queries, out[3] queries with the whole band inside, out[4] 1 if the run
 * fell back to the full scan.
 * iq_plan_encoder: the encoder of mask shape `mask` (64 x V row-major over
 * the V masked voxels, whose 0-based column-major tile offsets, ascending,
 * go to tvox); any output may be NULL; *nmasks = number of mask shapes. */
int iq_plan_learn(iq_ctx* ctx, iq_plan* plan, int64_t nsamples, int iters, uint64_t seed);
int iq_plan_set_proposal(iq_plan* plan, int enable, int64_t K, int check_every, double min_recall);
int iq_plan_set_check_sample(iq_plan* plan, int queries);
int iq_plan_proposal_stats(iq_plan* plan, int64_t out[5]);
int iq_plan_encoder(iq_plan* plan, int mask, float* W, int32_t* tvox, int64_t* V, int* nmasks);

/* Convolutional encoder (no counterpart in the reference).  A network of
 * three 3x3x3 convolution layers over the training image, 1 -> 32 -> 32 -> 1
 * channels with ReLU after the first two and a residual output
 * f = x + conv3(...) (zero outside the image), feeds each mask shape's head
 * W (64 x V): e_c = W f[window c] for the bank, e_q = W q for a query.
 * iq_plan_learn_conv trains it on the GPU (forward and backward kernels,
 * Adam) from pairs of the plain method: it runs the full scan over every
 * realization (its own uniforms; the grids then hold that run) and takes
 * about `nqueries` plain tile steps, spread over the run, as they run: the
 * overlap values the step sees, its `topn` nearest candidates by exact FP64
 * distance (preselected by the full scan's lower bounds), `nhard` members
 * of its shortlist under the linear encoders that are not among those
 * nearest (the competitors the band must outrank) and `nrandom` random
 * candidates; the loss of a query is
 * KL(p* || p~) with p* ~ exp(-(D - min D)/tau) over the exact distances,
 * p~ ~ exp(-(dd - min dd)/tau) over the bank scan's approximate distances
 * dd = A2 - 2 e_c.e_q, tau the larger of the band width tol * min D and the
 * spread of its eight nearest exact distances.  The heads start from
 * iq_plan_learn's linear encoders (call it first) and are trained with the
 * network when lr_head > 0.  Every fifth query is held out: at the start
 * (where the network is the identity, i.e. the linear encoders alone) and
 * every few steps, the banks are rebuilt and the held-out tile steps are
 * shortlisted exactly as iq_plan_run does; the parameters with the highest
 * recall of the held-out exact bands (lower held-out loss on ties) are kept,
 * so training never leaves an encoder that recalls fewer held-out band
 * members than the linear encoders.  Afterwards every mask's bank holds the
 * encodings of f.  Parameter vector (float): w1[32][27], b1[32],
 * w2[32][32][27], b2[32], w3[32][27], b3, i.e. IQ_CONV_NPARAM values, tap
 * k = (dx+1) + 3(dy+1) + 9(dz+1); a mask's head W[64][V] follows when a
 * function takes one.
 * iq_plan_conv_stats: out[0] seconds harvesting pairs, out[1] seconds
 * training and rebuilding the banks, out[2] / out[3] mean loss of the first /
 * last ten steps, out[4] training pairs, out[5] steps, out[6] queries, out[7]
 * the step whose parameters were kept (0: the starting point), out[8] /
 * out[9] held-out loss at the start / of the kept parameters, out[10] /
 * out[11] held-out shortlist recall at the start / of the kept parameters.
 * iq_plan_conv_grad: loss and gradient (network, then head) for given
 * parameters on given pairs of mask shape `mask` (Q: nq x V values; cands,
 * D: nq x ns; tau: nq), without changing the plan.  iq_plan_bank_rows: bank
 * rows (64 values each) of the given candidates. */
#define IQ_CONV_NPARAM 29441
typedef struct iq_conv_opts {
  int64_t nqueries;  /* training queries (e.g. 640) */
  int topn;          /* exact nearest candidates per query (e.g. 256) */
  int nrandom;       /* random candidates per query (e.g. 64); topn + nhard + nrandom <= 4096 */
  int steps;         /* Adam steps (e.g. 200) */
  int batch;         /* queries per step (e.g. 16) */
  float lr;          /* learning rate of the network (e.g. 1e-3) */
  float lr_head;     /* learning rate of the heads (0: the linear encoders' heads stay) */
  uint64_t seed;
  int64_t val_k;     /* shortlist size of the held-out recall check (0: the plan's K, else 2048) */
  int nhard;         /* shortlist members of the linear encoders outside the nearest list, per query (e.g. 64) */
} iq_conv_opts;
int iq_plan_learn_conv(iq_ctx* ctx, iq_plan* plan, const iq_conv_opts* opts);
int iq_plan_conv_stats(iq_plan* plan, double out[12]);
int iq_plan_conv_params(iq_plan* plan, float* theta);
int iq_plan_conv_grad(iq_ctx* ctx, iq_plan* plan, int mask, const float* theta, int64_t nq, int64_t ns, const float* Q,
                      const int32_t* cands, const double* D, const double* tau, double* loss, float* grad);
int iq_plan_bank_rows(iq_ctx* ctx, iq_plan* plan, int mask, const int32_t* cands, int64_t n, float* out);

/* Convenience: create + run + results + destroy. */
int iq_simulate(iq_ctx* ctx, const iq_plan_desc* desc, const double* uniforms, const int64_t* forced,
                void* out, int64_t* chosen, double* cutgrid);

#ifdef __cplusplus
}
#endif

#endif /* IQHIP_H */
