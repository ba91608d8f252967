// iq_internal.h -- device-side data structures shared by the HIP kernels and
// the host driver of libiqhip.  Not part of the public ABI (include/iqhip.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define IQ_MAX_BOX 27       // disjoint boxes describing one overlap mask (3^3 cells at most)
#define IQ_MAX_SLAB 6       // prev/next slab per axis (iqsim.jl:252-275)
#define IQ_MAX_AUX 8        // soft-data pairs
#define IQ_SB 64            // block of the three-level cumulative sums of the draw
#define IQ_GR_FREQ 8        // push-relabel rounds between global relabels (slabs of <= IQ_GR_BIG interior nodes)
#ifndef IQ_GR_FREQ_BIG
#define IQ_GR_FREQ_BIG 24   // the same for larger slabs
#endif
#define IQ_GR_BIG 1024
#define IQ_PR_LIMIT 200000  // push-relabel round guard (never reached on valid input)
#define IQ_SORT_MAX 4096    // largest multi-valued band ranked in LDS
#define IQ_PR_NLDS 1536     // interior min-cut nodes kept in LDS (above: global scratch)

// error codes written by kernels into the per-run flag word (first error wins)
#define IQ_KERR_NOCAND 1
#define IQ_KERR_NOTINDB 2
#define IQ_KERR_ITER 3
#define IQ_KERR_SORTMAX 4
#define IQ_KERR_EMPTYBAND 5  // internal: a selection produced no pattern (never expected)

// One axis-aligned box inside the tile, [x0, x0+wx) x [y0, y0+wy) x [z0, z0+wz).
struct IqBox {
  int32_t x0, y0, z0, wx, wy, wz;
  int32_t wxp;   // wx rounded up to a multiple of 4 (zero-padded query row)
  int32_t kofs;  // start of this box's query values in the box-ordered query array
};

// One overlap-mask shape (a union of the prev/next slabs of iqsim.jl:188-205),
// stored as disjoint boxes for the FP32 scan and as the ascending column-major
// list of masked tile voxels for the exact FP64 chains.
struct IqMask {
  int32_t nbox;
  IqBox box[IQ_MAX_BOX];
  int32_t V;        // masked voxels
  int32_t kbox;     // length of the box-ordered, row-padded query array
  int32_t off;      // start of this mask's entries in the offset tables
  int32_t a2;       // index of the exact A2 map of this mask over the TI (-1: none)
};

// A cut slab inside the tile (iqsim.jl:258-274).
struct IqSlab {
  int32_t dim, side;  // side 0 = previous tile (keep graphcut mask), 1 = next tile (keep its negation)
  int32_t lo[3], sz[3];
};

// One path position of one simulation grid (shared by the grid's realizations).
struct IqStep {
  int32_t active;     // 0: skipped tile (iqsim.jl:174)
  int32_t origin[3];  // tile origin in the padded grid (iqsim.jl:180)
  int32_t mask;       // overlap mask index, -1 when nothing overlaps
  int32_t hard;       // hard-data points in the tile (iqsim.jl:213)
  int32_t hoff, hcnt; // hard voxels of the tile: entries [hoff, hoff+hcnt) of the hard tables
  int32_t nslab;
  int32_t ncut;       // slabs [0, ncut) have interior nodes and need a min cut; the others are two slices thick
  IqSlab slab[IQ_MAX_SLAB];
};

// One tile step of one realization inside a batch (a wave of independent tiles).
struct IqQuery {
  int32_t step;    // entry of the steps table (its grid's step0 + path position)
  int32_t real;    // realization
  int32_t slot;    // scratch slot
  int32_t multi;   // 1: relaxation / tau model over several sources
  double u;        // uniform of the draw
  int64_t forced;  // pick to replay, or -1
};

// MFMA scan: the K dimension is a list of 8-voxel chunks over the union of
// all overlap masks, grouped by box; a chunk runs along x (dir 0) or y (dir 1).
struct IqChunk {
  int16_t jx, jy, jz;  // first voxel (tile coordinates)
  int8_t dir, n;       // direction, valid voxels (0..8)
};
struct IqKBox {
  int32_t x0, y0, z0, wxr, wyr, wz;  // staged region of the box (x/y extents padded to the chunk runs)
  int32_t s0, s1;                    // k-steps [s0, s1) (two chunks per k-step)
};
#define IQ_MAX_KBOX 8
struct IqKTab {
  int32_t nbox, nstep;
  IqKBox box[IQ_MAX_KBOX];
};

// Learned proposal: per overlap-mask shape a linear encoder e(x) = W x of
// the masked values (IQ_EMB outputs) and the bank of encoded candidate
// windows; approximate distances A2 - 2 e(x_c).e(q) + B2 shortlist the K
// best candidates of a query, which the exact recheck then reranks.
#define IQ_EMB 64
struct IqLearn {
  int32_t on;                 // shortlist from the bank instead of the full scan
  int32_t K;                  // shortlist size
  const float* W;             // encoders: IQ_EMB x V of each mask, row-major, at wofs[mask]
  const float* WT;            // the same transposed (V x IQ_EMB per mask), for the query encodings
  const int64_t* wofs;        // per mask (-1: no encoder)
  const uint16_t* bank;       // P x IQ_EMB bf16 rows per encoded mask, at bofs[mask]
  const int64_t* bofs;
  float* eq;                  // per slot: the query's encoding
  uint4* bq;                  // B fragments of the query tiles, [tile][IQ_EMB / 16][64]
  uint32_t* brk;              // per slot: bracket lo, hi, bin shift
  uint4* binfo;               // per batch position: the bracket's top as a float, lo, hi, shift (bank scan columns)
  int4* stage;                // bank scan: per block, stage_cap listed entries (column, rank, candidate, key)
  int32_t stage_cap;
  int32_t* lcnt;              // per slot: candidates listed inside the bracket
  int2* list;                 // per slot: lcap (candidate, order key) pairs
  int64_t lcap;
  unsigned long long* sbits;  // per slot: shortlist bitmap (nw64 words)
  int32_t* slist;             // per slot: the shortlist (P capacity)
  int32_t* nslist;            // per slot: its length
  unsigned long long* stats;  // recall check: band members, members in the shortlist, queries, fully recalled
};

// Geometry shared by all kernels of a plan.
struct IqGeo {
  int32_t ti[3];     // training-image extents
  int32_t tile[3];
  int32_t dist[3];   // candidate extents = ti - tile + 1
  int64_t P;         // number of candidates
  int32_t tilevox;
};

// One realization: its simulation grid (realizations of one plan may have
// grids of different sizes) and where its state lives.
struct IqReal {
  int32_t pad[3];              // padded simulation grid (iqsim.jl:114)
  int32_t step0;               // first entry of its grid's steps
  int64_t sim_off;             // first voxel of its grid in sims / cuts
  int64_t log_off;             // pick-log entry of its path position 0
  const double* hard;          // hard-data grid (NaN: none), or null
  const double* const* aux;    // soft-data grids (padded), naux of them
};

// Everything a kernel needs, resident in device memory (read through the
// scalar cache).  Per-slot scratch arrays are laid out slot-major.
struct IqDev {
  IqGeo g;
  const void* ti;
  const float* ti32;
  int32_t ti_f32;
  int32_t naux;
  double tol;
  const uint8_t* disabled;
  const int32_t* enabled;
  int64_t nen;
  const IqStep* steps;
  const IqMask* masks;
  const int32_t* tvox;   // tile-linear offsets (ascending column-major) per mask, then the full tile
  const int32_t* toff;   // matching training-image offsets
  const int32_t* htvox;  // hard voxels per step
  const int32_t* htoff;
  const int32_t* rtvox;   // tvox entries as offsets into the tiled multi-map's staged region
  const int32_t* rhtvox;  // htvox entries likewise
  int32_t full_off;
  const double* const* a2maps;
  const float* const* a2f;  // FP32 copies of a2maps, NaN where the window is disabled
  const IqReal* reals;
  const double* const* auxti;
  const double* const* a2soft;
  void* sims;
  uint8_t* cuts;
  int64_t* log;
  int32_t* err;
  int64_t* rounds;
  // per-slot scratch
  double* kv;        int64_t kv_stride;    // (2 + naux) query vectors of tilevox
  float* kv32;       int64_t kv32_stride;  int32_t kv32_on;  // FP32 query boxes (VALU scan only)
  double* b2;                              // (2 + naux) per slot
  uint32_t* gmin_hi;
  unsigned long long* gmin64;
  float* lo32;                             // P per slot
  double* dexact;                          // P per slot
  unsigned long long* bitmap; int64_t nw64;
  int32_t* band;                           // P per slot
  int32_t* sup;                            // superset list, P per slot (then ranks of a large band)
  unsigned long long* rkey;                // 2P per slot: radix keys of a large band (then its weights)
  int32_t* nsup;
  double* sd;                              // IQ_SORT_MAX per slot (band distances, then weights)
  int32_t* si;                             // IQ_SORT_MAX per slot (ranks)
  double* s1;        int64_t s1_stride;
  int64_t* pick;
  int64_t* nband;
  uint8_t* keep;                           // IQ_MAX_SLAB * tilevox per slot
  char* cutscratch;  int64_t cutscratch_per;
  // multi-source selection (relaxation.jl / taumodel.jl), per slot
  double* maps;      int32_t nmap;         // nmap maps of P: 0 overlap, 1 hard, 2+a soft
  uint32_t* bits;    int64_t nw32;         // nmap bitmaps of P bits
  int32_t* ev;                             // pattern database (P)
  double* vals;                            // P
  // MFMA scan
  IqKTab ktab;
  const IqChunk* chunks;
  const uint8_t* maskbits;                 // per mask: tilevox membership bytes
  uint4* bh;                               // B fragments (hi), [ntile][nstep][64]
  uint4* bl;                               // B fragments (lo)
  IqLearn lrn;
  int32_t rc_small;  // exact recheck: a wavefront per member up to this superset size, a work-item per member above
  int32_t rc_lanes;  // 0: the exact recheck runs k_band_recheck_lanes (a lane per member), else k_band_recheck
};

// Order-preserving 32-bit key of a non-NaN float (-0 and +0 share a key).
__host__ __device__ inline uint32_t iq_f32_key(float x) {
  union { float f; uint32_t u; } c;
  c.f = x == 0.0f ? 0.0f : x;
  return (c.u >> 31) ? ~c.u : (c.u | 0x80000000u);
}

// global-address-space view of a device pointer: loads through it are
// global_load (counted by vmcnt only), not flat loads that also hold lgkmcnt
template <typename T>
__device__ inline const __attribute__((address_space(1))) T* iq_glob(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}
// constant-address-space view: uniform loads through it become scalar loads
template <typename T>
__device__ inline const __attribute__((address_space(4))) T* iq_const(const T* p) {
  return (const __attribute__((address_space(4))) T*)p;
}
template <typename T>
__device__ inline __attribute__((address_space(1))) T* iq_globw(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

// Push-relabel rounds between global relabels of a slab with n interior
// nodes: larger slabs do better with fewer relabels per round (1,536-node
// slabs: 1,499, 1,226, 1,188 and 1,194 us per launch of 362 cuts at 8, 16,
// 24 and 32 rounds; slabs of ~200 nodes gain nothing).
__host__ __device__ inline int iq_gr_freq(int64_t n) { return n > IQ_GR_BIG ? IQ_GR_FREQ_BIG : IQ_GR_FREQ; }

__host__ __device__ inline int64_t iq_lin3(const int32_t d[3], int64_t x, int64_t y, int64_t z) {
  return x + (int64_t)d[0] * (y + (int64_t)d[1] * z);
}

// float rounding of a non-negative double toward -inf / +inf (one ulp step
// of the non-negative float bit pattern when round-to-nearest overshot)
__device__ inline float iq_f32_down(double x) {
  float f = (float)x;
  if ((double)f > x) f = __uint_as_float(__float_as_uint(f) - 1u);
  return f;
}
__device__ inline float iq_f32_up(double x) {
  float f = (float)x;
  if ((double)f < x) f = __uint_as_float(__float_as_uint(f) + 1u);
  return f;
}

// broadcast of lane `l` (uniform) of a double held in VGPRs
__device__ inline double iq_readlane_f64(double v, int l) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}
