// iq_internal.h -- device-side data structures shared by the HIP kernels and
// the host driver of libiqhip.  Not part of the public ABI (include/iqhip.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define IQ_MAX_BOX 27       // disjoint boxes describing one overlap mask (3^3 cells at most)
#define IQ_MAX_SLAB 6       // prev/next slab per axis (iqsim.jl:252-275)
#define IQ_SB 64            // block of the three-level cumulative sums of the draw
#define IQ_GR_FREQ 8        // push-relabel rounds between global relabels
#define IQ_PR_LIMIT 200000  // push-relabel round guard (never reached on valid input)
#define IQ_SORT_MAX 4096    // largest multi-valued band ranked in LDS
#define IQ_PR_NLDS 1536     // interior min-cut nodes kept in LDS (above: global scratch)

// error codes written by kernels into the per-run flag word (first error wins)
#define IQ_KERR_NOCAND 1
#define IQ_KERR_NOTINDB 2
#define IQ_KERR_ITER 3
#define IQ_KERR_SORTMAX 4

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

// One path position.  Identical for every realization of a plan.
struct IqStep {
  int32_t active;     // 0: skipped tile (iqsim.jl:174)
  int32_t origin[3];  // tile origin in the padded grid (iqsim.jl:180)
  int32_t mask;       // overlap mask index, -1 when nothing overlaps
  int32_t hard;       // hard-data points in the tile (iqsim.jl:213)
  int32_t hoff, hcnt; // hard voxels of the tile: entries [hoff, hoff+hcnt) of the hard tables
  int32_t nslab;
  IqSlab slab[IQ_MAX_SLAB];
};

// Geometry shared by all kernels of a plan.
struct IqGeo {
  int32_t ti[3];     // training-image extents
  int32_t tile[3];
  int32_t pad[3];    // padded simulation grid
  int32_t dist[3];   // candidate extents = ti - tile + 1
  int64_t P;         // number of candidates
  int64_t padvox;
  int32_t tilevox;
};

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
