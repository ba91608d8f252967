// k_distance.hip -- masked sliding distances (fastdistance, utils.jl:5-13)
// and valid cross-correlation (imfilter.jl:5-26) for gfx950.
//
// Two evaluations of the same map:
//   * exact:  D[c] = |(A2 - 2 AB) + B2| with A2, AB, B2 fused-multiply-add
//             chains in FP64 over the masked tile voxels in ascending
//             column-major order;
//   * scan:   AB in FP32 over disjoint boxes with the training image staged
//             through LDS, A2 from the exact per-mask map, and a rigorous
//             error bound, giving an interval [lo, hi] that contains the exact
//             FP64 value.  The scan feeds the tol-band superset; members of
//             the superset are recomputed exactly.
// The per-step kernels serve a batch of queries (independent tile steps of a
// wave, all realizations) per launch; blockIdx.z (or .y) selects the query.
#include "iq_internal.h"

// ---------------------------------------------------------------------------
// imfilter: out[c] = sum_j img[c+j] * krn[j]   (FP64, fma chain, column-major j)
// ---------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(256)
k_imfilter(const double* __restrict__ img, int nx, int ny, int nz,
           const double* __restrict__ krn, int kx, int ky, int kz,
           double* __restrict__ out, int ox, int oy, int oz) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t P = (int64_t)ox * oy * oz;
  if (c >= P) return;
  int cx = (int)(c % ox), cy = (int)((c / ox) % oy), cz = (int)(c / ((int64_t)ox * oy));
  const double* base = img + cx + (int64_t)nx * (cy + (int64_t)ny * cz);
  double acc = 0.0;
  int j = 0;
  for (int z = 0; z < kz; z++)
    for (int y = 0; y < ky; y++) {
      const double* row = base + (int64_t)nx * (y + (int64_t)ny * z);
      for (int x = 0; x < kx; x++, j++) acc = __builtin_fma(row[x], krn[j], acc);
    }
  out[c] = acc;
}

// ---------------------------------------------------------------------------
// Exact FP64 map of one query over all candidates (standalone fastdistance and
// the per-mask A2 maps).  MODE 0: D with the A2 chain; MODE 2: the A2 chain
// alone.  toff: training-image offsets of the masked voxels in ascending
// column-major tile order; kv: query values.
// ---------------------------------------------------------------------------
template <typename T, int MODE>
__global__ void __launch_bounds__(256)
k_exact_map(const T* __restrict__ img, IqGeo g, const int32_t* __restrict__ toff,
            const double* __restrict__ kv, int V, const double* __restrict__ b2p,
            double* __restrict__ out) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.P) return;
  int cx = (int)(c % g.dist[0]), cy = (int)((c / g.dist[0]) % g.dist[1]);
  int cz = (int)(c / ((int64_t)g.dist[0] * g.dist[1]));
  const T* base = img + iq_lin3(g.ti, cx, cy, cz);
  double a2 = 0.0, ab = 0.0;
  for (int j = 0; j < V; j++) {
    double t = (double)base[toff[j]];
    a2 = __builtin_fma(t, t, a2);
    if (MODE == 0) ab = __builtin_fma(t, kv[j], ab);
  }
  if (MODE == 2) { out[c] = a2; return; }
  out[c] = __builtin_fabs((a2 - 2.0 * ab) + *b2p);
}

template __global__ void k_exact_map<float, 0>(const float*, IqGeo, const int32_t*, const double*, int, const double*, double*);
template __global__ void k_exact_map<float, 2>(const float*, IqGeo, const int32_t*, const double*, int, const double*, double*);
template __global__ void k_exact_map<double, 0>(const double*, IqGeo, const int32_t*, const double*, int, const double*, double*);
template __global__ void k_exact_map<double, 2>(const double*, IqGeo, const int32_t*, const double*, int, const double*, double*);

// Standalone query staging (iq_fastdistance): kv[j] = kern[tv[j]] and the
// exact B2 chain, one workgroup.
extern "C" __global__ void __launch_bounds__(256)
k_gather_b2(const double* __restrict__ kern, const int32_t* __restrict__ tv, int V, double* __restrict__ kv,
            double* __restrict__ b2) {
  for (int j = threadIdx.x; j < V; j += blockDim.x) kv[j] = kern[tv[j]];
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int j = 0; j < V; j++) s = __builtin_fma(kv[j], kv[j], s);
    *b2 = s;
  }
}

// ---------------------------------------------------------------------------
// Query staging, one workgroup per (query, source).  Source 0 is the overlap
// query (grid = the realization's simulation grid), 1 the hard-data query
// (grid = hard-data grid), 2+a the soft query of pair a (grid = aux[a]).
// Writes the query values at the masked voxels (ascending column-major), the
// box-ordered zero-padded FP32 copy for the scan, and the exact B2 chain; the
// overlap source also resets the query's reduction words.
// ---------------------------------------------------------------------------
#define PREP_LDS 8192

template <typename T>
__device__ __forceinline__ void query_prep_body(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, int qx,
                                                int src, double* kl) {
  const IqQuery q = qs[qx];
  const IqGeo& g = D->g;
  const IqStep& st = D->steps[q.step];
  const int64_t slot = q.slot;
  int V = 0;
  const int32_t* tv = nullptr;
  const IqMask* mask = nullptr;
  if (src == 0) {
    if (st.mask >= 0) {
      mask = &D->masks[st.mask];
      V = mask->V;
      tv = D->tvox + mask->off;
    }
    if (threadIdx.x == 0) {
      D->gmin_hi[slot] = 0x7f800000u;
      D->gmin64[slot] = 0x7ff0000000000000ull;
      D->nsup[slot] = 0;
    }
  } else if (src == 1) {
    if (!q.multi || !st.hard) return;
    V = st.hcnt;
    tv = D->htvox + st.hoff;
  } else {
    if (!q.multi) return;
    V = g.tilevox;
    tv = D->tvox + D->full_off;
  }
  double* kv = D->kv + slot * D->kv_stride + (int64_t)src * g.tilevox;
  const int ox = st.origin[0], oy = st.origin[1], oz = st.origin[2];
  const IqReal& R = D->reals[q.real];
  const T* sim = (const T*)D->sims + R.sim_off;
  const double* dgrid = src == 1 ? R.hard : (src >= 2 ? R.aux[src - 2] : nullptr);
  // four voxels per pass: offsets, then grid values, all loads in flight
  const auto tvg = iq_glob(tv);
  const int t0 = g.tile[0], t1 = g.tile[1];
  for (int j0 = threadIdx.x; j0 < V; j0 += 4 * blockDim.x) {
    int t[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int j = j0 + u * (int)blockDim.x;
      t[u] = tvg[j < V ? j : 0];
    }
    int64_t gi[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int x = t[u] % t0, y = (t[u] / t0) % t1, z = t[u] / (t0 * t1);
      gi[u] = iq_lin3(R.pad, ox + x, oy + y, oz + z);
    }
    double v[4];
    if (src == 0) {  // uniform per group
#pragma unroll
      for (int u = 0; u < 4; u++) v[u] = (double)sim[gi[u]];
    } else {
#pragma unroll
      for (int u = 0; u < 4; u++) v[u] = dgrid[gi[u]];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int j = j0 + u * (int)blockDim.x;
      if (j < V) {
        kv[j] = v[u];
        if (j < PREP_LDS) kl[j] = v[u];
      }
    }
  }
  if (src == 0 && mask && !q.multi && D->ti32 && D->kv32_on) {
    float* kv32 = D->kv32 + slot * D->kv32_stride;
    for (int b = 0; b < mask->nbox; b++) {
      const IqBox bx = mask->box[b];
      int n = bx.wxp * bx.wy * bx.wz;
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int x = i % bx.wxp, y = (i / bx.wxp) % bx.wy, z = i / (bx.wxp * bx.wy);
        float v = 0.0f;
        if (x < bx.wx) v = (float)sim[iq_lin3(R.pad, ox + bx.x0 + x, oy + bx.y0 + y, oz + bx.z0 + z)];
        kv32[bx.kofs + i] = v;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // B2 chain in ascending voxel order; the next eight values are read from
    // LDS while the current eight run through the fmas
    double s = 0.0;
    int j = 0;
    const int vl = V < PREP_LDS ? V : PREP_LDS;
    double a[8];
    if (vl >= 8) {
#pragma unroll
      for (int i = 0; i < 8; i++) a[i] = kl[i];
    }
    for (; j + 8 <= vl; j += 8) {
      const int jn = j + 16 <= vl ? j + 8 : j;
      double an[8];
#pragma unroll
      for (int i = 0; i < 8; i++) an[i] = kl[jn + i];
#pragma unroll
      for (int i = 0; i < 8; i++) s = __builtin_fma(a[i], a[i], s);
#pragma unroll
      for (int i = 0; i < 8; i++) a[i] = an[i];
    }
    for (; j < vl; j++) s = __builtin_fma(kl[j], kl[j], s);
    for (; j < V; j++) s = __builtin_fma(kv[j], kv[j], s);  // beyond the LDS stage
    D->b2[slot * (2 + D->naux) + src] = s;
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_query_prep(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  __shared__ double kl[PREP_LDS];
  query_prep_body<T>(D, qs, blockIdx.x, blockIdx.y, kl);
}

template __global__ void k_query_prep<float>(const IqDev*, const IqQuery*);
template __global__ void k_query_prep<double>(const IqDev*, const IqQuery*);

// ---------------------------------------------------------------------------
// FP32 scan.  Work-item (lx, ly, lz) of the group owns R consecutive
// candidates along x.  Per box, the needed training-image block is staged in
// LDS with an R-interleaved row layout (element x of a row lives at
// (x % R) * S + x / R) so that lanes read consecutive words, conflict free,
// while each loaded word feeds R fused multiply-adds.
// Grid: x, y candidate blocks; z = candidate z-block + nbz * query.
// ---------------------------------------------------------------------------
template <int R, int LX, int LY, int LZ>
__global__ void __launch_bounds__(LX * LY * LZ)
k_scan_f32(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, int nbz) {
  extern __shared__ float lds[];
  const IqQuery q = qs[blockIdx.z / nbz];
  const IqStep& st = D->steps[q.step];
  if (q.multi || st.mask < 0) return;
  const IqGeo& g = D->g;
  const IqMask* mask = &D->masks[st.mask];
  const int64_t slot = q.slot;
  const float* __restrict__ ti = D->ti32;
  const float* __restrict__ kv32 = D->kv32 + slot * D->kv32_stride;
  const int tid = threadIdx.x;
  const int lx = tid % LX, ly = (tid / LX) % LY, lz = tid / (LX * LY);
  const int cx0 = blockIdx.x * (LX * R), cy0 = blockIdx.y * LY, cz0 = (blockIdx.z % nbz) * LZ;
  float acc[R];
#pragma unroll
  for (int i = 0; i < R; i++) acc[i] = 0.0f;
  const int nbox = mask->nbox;
  for (int b = 0; b < nbox; b++) {
    const IqBox bx = mask->box[b];
    const int wxp = bx.wxp;
    const int S = LX + wxp / R;              // interleave stride
    const int RXL = R * LX + wxp - 1;        // logical row length
    const int RY = LY + bx.wy - 1, RZ = LZ + bx.wz - 1;
    const int rowst = R * S;
    const int nst = RXL * RY * RZ;
    const int gx0 = cx0 + bx.x0, gy0 = cy0 + bx.y0, gz0 = cz0 + bx.z0;
    for (int i = tid; i < nst; i += LX * LY * LZ) {
      int x = i % RXL, y = (i / RXL) % RY, z = i / (RXL * RY);
      int gx = gx0 + x, gy = gy0 + y, gz = gz0 + z;
      float v = 0.0f;
      if (gx < g.ti[0] && gy < g.ti[1] && gz < g.ti[2]) v = ti[iq_lin3(g.ti, gx, gy, gz)];
      lds[(z * RY + y) * rowst + (x % R) * S + x / R] = v;
    }
    __syncthreads();
    const float* kb = kv32 + bx.kofs;
    for (int jz = 0; jz < bx.wz; jz++) {
      for (int jy = 0; jy < bx.wy; jy++) {
        const float* rowp = lds + ((lz + jz) * RY + (ly + jy)) * rowst + lx;
        const float* kp = kb + (jz * bx.wy + jy) * wxp;
        for (int jx = 0; jx < wxp; jx += 4) {
          const float* rp = rowp + jx / R;
          float v[R + 3];
#pragma unroll
          for (int m = 0; m < R + 3; m++) v[m] = rp[(m % R) * S + m / R];
          const float k0 = kp[jx], k1 = kp[jx + 1], k2 = kp[jx + 2], k3 = kp[jx + 3];
#pragma unroll
          for (int i = 0; i < R; i++) {
            acc[i] = __builtin_fmaf(v[i], k0, acc[i]);
            acc[i] = __builtin_fmaf(v[i + 1], k1, acc[i]);
            acc[i] = __builtin_fmaf(v[i + 2], k2, acc[i]);
            acc[i] = __builtin_fmaf(v[i + 3], k3, acc[i]);
          }
        }
      }
    }
    __syncthreads();
  }
  // bound and publish
  const double* a2map = D->a2maps[mask->a2];
  const double b2 = D->b2[slot * (2 + D->naux)];
  const double V = (double)mask->V;
  const double u32 = 0x1p-24, u64 = 0x1p-53;
  const double g32 = V * u32 / (1.0 - V * u32), g64 = V * u64 / (1.0 - V * u64);
  float* lo_out = D->lo32 + slot * g.P;
  float lmin = __builtin_inff();
  const int cy = cy0 + ly, cz = cz0 + lz;
#pragma unroll
  for (int i = 0; i < R; i++) {
    const int cx = cx0 + R * lx + i;
    if (cx < g.dist[0] && cy < g.dist[1] && cz < g.dist[2]) {
      int64_t c = iq_lin3(g.dist, cx, cy, cz);
      float lo = __builtin_inff(), hi = __builtin_inff();
      if (!D->disabled[c]) {
        double a2 = a2map[c];
        double s = __builtin_sqrt(a2 * b2) * (1.0 + 0x1p-30);
        double d = __builtin_fabs((a2 - 2.0 * (double)acc[i]) + b2);
        double e = (2.0 * (g32 + g64) * s + 4.0 * u64 * (a2 + b2 + 2.0 * s) + 2.0 * V * 0x1p-148) * (1.0 + 0x1p-30)
                   + 4.0 * u64 * d;
        lo = iq_f32_down(d - e > 0.0 ? d - e : 0.0);
        hi = iq_f32_up(d + e);
      }
      lo_out[c] = lo;
      lmin = fminf(lmin, hi);
    }
  }
  // workgroup min of the upper bounds -> one atomic per group
  __shared__ float red[LX * LY * LZ / 64];
  for (int o = 32; o > 0; o >>= 1) lmin = fminf(lmin, __shfl_xor(lmin, o));
  if ((tid & 63) == 0) red[tid >> 6] = lmin;
  __syncthreads();
  if (tid == 0) {
    float m = red[0];
    for (int w = 1; w < LX * LY * LZ / 64; w++) m = fminf(m, red[w]);
    atomicMin(&D->gmin_hi[slot], __float_as_uint(m));  // non-negative floats order like their bits
  }
}

// ---------------------------------------------------------------------------
// FP32 scan with a 2-D register block (3-D candidate grids): a work-item owns
// R x-candidates times RY y-candidates at one z; the lanes of a wavefront span
// x (LX) and z (LZ) and the four wavefronts span y, so every loop over
// training-image rows is uniform across the wavefront.  One LDS row segment
// (R + 3 words) feeds up to 4 * R * RY fused multiply-adds: the row at offset
// ty serves candidate row iy with query row jy = ty - iy.  LDS rows use the
// R-interleaved layout of k_scan_f32 (conflict-free lane reads); the z-plane
// stride is padded to 8 words modulo 64 banks.
// Grid: x, y candidate blocks; z = candidate z-block + nbz * query.
// ---------------------------------------------------------------------------
template <int R, int RY, int LX, int LZ>
__global__ void __launch_bounds__(256)
k_scan2_f32(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, int nbz, int perz) {
  constexpr int LY = 256 / (LX * LZ);
  extern __shared__ float lds[];
  const IqQuery q = qs[blockIdx.z / nbz];
  const IqStep& st = D->steps[q.step];
  if (q.multi || st.mask < 0) return;
  const IqGeo& g = D->g;
  const IqMask* mask = &D->masks[st.mask];
  const int64_t slot = q.slot;
  const float* __restrict__ ti = D->ti32;
  const float* __restrict__ kv32 = D->kv32 + slot * D->kv32_stride;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int lx = lane % LX, lz = lane / LX, ly = tid >> 6;
  const int cx0 = blockIdx.x * (LX * R), cy0 = blockIdx.y * (LY * RY), cz0 = (blockIdx.z % nbz) * LZ;
  float acc[RY][R];
#pragma unroll
  for (int a = 0; a < RY; a++)
#pragma unroll
    for (int i = 0; i < R; i++) acc[a][i] = 0.0f;
  const int nbox = mask->nbox;
  for (int b = 0; b < nbox; b++) {
    const IqBox bx = mask->box[b];
    const int wxp = bx.wxp, wy = bx.wy, wz = bx.wz;
    const int S = LX + wxp / R;
    const int RXL = R * LX + wxp - 1;
    const int RYT = LY * RY + wy - 1, RZ = LZ + wz - 1;
    const int rowst = R * S;
    const int plane = ((RYT * rowst + 55) / 64) * 64 + 8;  // = 8 (mod 64)
    const int gx0 = cx0 + bx.x0, gy0 = cy0 + bx.y0, gz0 = cz0 + bx.z0;
    // perz: stage the LZ planes of one jz at a time (large tiles), else all RZ planes once
    const int nplanes = perz ? LZ : RZ;
    const int nst = RXL * RYT * nplanes;
    const float* kb = kv32 + bx.kofs;
    const int nty = RY + wy - 1;
    for (int jz = 0; jz < wz; jz++) {
      if (jz == 0 || perz) {
        if (jz > 0) __syncthreads();
        const int zoff = perz ? jz : 0;
        for (int i = tid; i < nst; i += 256) {
          int x = i % RXL, y = (i / RXL) % RYT, z = i / (RXL * RYT);
          int gx = gx0 + x, gy = gy0 + y, gz = gz0 + zoff + z;
          float v = 0.0f;
          if (gx < g.ti[0] && gy < g.ti[1] && gz < g.ti[2]) v = ti[iq_lin3(g.ti, gx, gy, gz)];
          lds[z * plane + y * rowst + (x % R) * S + x / R] = v;
        }
        __syncthreads();
      }
      const float* zp = lds + (lz + (perz ? 0 : jz)) * plane + lx + (ly * RY) * rowst;
      const float* kz = kb + jz * wy * wxp;
      for (int ty = 0; ty < nty; ty++) {
        const float* rowp = zp + ty * rowst;
        for (int jx = 0; jx < wxp; jx += 4) {
          const float* rp = rowp + jx / R;
          float v[R + 3];
#pragma unroll
          for (int m = 0; m < R + 3; m++) v[m] = rp[(m % R) * S + m / R];
#pragma unroll
          for (int a = 0; a < RY; a++) {
            const int jy = ty - a;
            if (jy >= 0 && jy < wy) {
              const float* kp = kz + jy * wxp + jx;
              const float k0 = kp[0], k1 = kp[1], k2 = kp[2], k3 = kp[3];
#pragma unroll
              for (int i = 0; i < R; i++) {
                acc[a][i] = __builtin_fmaf(v[i], k0, acc[a][i]);
                acc[a][i] = __builtin_fmaf(v[i + 1], k1, acc[a][i]);
                acc[a][i] = __builtin_fmaf(v[i + 2], k2, acc[a][i]);
                acc[a][i] = __builtin_fmaf(v[i + 3], k3, acc[a][i]);
              }
            }
          }
        }
      }
    }
    __syncthreads();
  }
  const double* a2map = D->a2maps[mask->a2];
  const double b2 = D->b2[slot * (2 + D->naux)];
  const double V = (double)mask->V;
  const double u32 = 0x1p-24, u64 = 0x1p-53;
  const double g32 = V * u32 / (1.0 - V * u32), g64 = V * u64 / (1.0 - V * u64);
  float* lo_out = D->lo32 + slot * g.P;
  float lmin = __builtin_inff();
  const int cz = cz0 + lz;
#pragma unroll
  for (int a = 0; a < RY; a++) {
    const int cy = cy0 + ly * RY + a;
#pragma unroll
    for (int i = 0; i < R; i++) {
      const int cx = cx0 + R * lx + i;
      if (cx < g.dist[0] && cy < g.dist[1] && cz < g.dist[2]) {
        int64_t c = iq_lin3(g.dist, cx, cy, cz);
        float lo = __builtin_inff(), hi = __builtin_inff();
        if (!D->disabled[c]) {
          double a2 = a2map[c];
          double sq = __builtin_sqrt(a2 * b2) * (1.0 + 0x1p-30);
          double d = __builtin_fabs((a2 - 2.0 * (double)acc[a][i]) + b2);
          double e = (2.0 * (g32 + g64) * sq + 4.0 * u64 * (a2 + b2 + 2.0 * sq) + 2.0 * V * 0x1p-148) * (1.0 + 0x1p-30)
                     + 4.0 * u64 * d;
          lo = iq_f32_down(d - e > 0.0 ? d - e : 0.0);
          hi = iq_f32_up(d + e);
        }
        lo_out[c] = lo;
        lmin = fminf(lmin, hi);
      }
    }
  }
  __shared__ float red[4];
  for (int o = 32; o > 0; o >>= 1) lmin = fminf(lmin, __shfl_xor(lmin, o));
  if (lane == 0) red[ly] = lmin;
  __syncthreads();
  if (tid == 0) {
    float m = fminf(fminf(red[0], red[1]), fminf(red[2], red[3]));
    atomicMin(&D->gmin_hi[slot], __float_as_uint(m));
  }
}

template __global__ void k_scan2_f32<4, 4, 8, 8>(const IqDev*, const IqQuery*, int, int);
template __global__ void k_scan2_f32<4, 4, 16, 4>(const IqDev*, const IqQuery*, int, int);

// ---------------------------------------------------------------------------
// MFMA scan (multi-query GEMM): AB[c, q] = sum_k A[c, k] * B[k, q] with
// A[c, k] = TI[c + voxel(k)] and B[k, q] = query q's value at voxel(k) (zero
// outside q's mask), on v_mfma_f32_32x32x16_bf16: 32 candidates (consecutive
// x) by 32 queries by 16 voxels per instruction, FP32 accumulation.  FP32
// data is split t = hi + lo (two bf16) and three products are accumulated
// (hi*hi, hi*lo, lo*hi); bf16-exact data (small integer codes) needs one.
// Each wavefront owns four 32-candidate rows (consecutive y) of one z, the
// group (two to four wavefronts, chosen per plan so that the groups of a
// launch pack the compute units evenly) 8 to 16 rows; the training-image block of each box
// is staged in LDS as words (hi | lo << 16), and A fragments are gathered
// from it along the chunk runs, so any run alignment works.
// Grid: x blocks (32), y blocks (4 x wavefronts), z = candidate z + dist_z * query tile.
// ---------------------------------------------------------------------------
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 iq_bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float iq_f32x16;
typedef unsigned int iq_u32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint32_t iq_bf16_rne(float f) {  // round-to-nearest-even bf16 bits of a finite float
  uint32_t u = __float_as_uint(f);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

__device__ inline uint32_t iq_split_word(float t) {  // hi | lo << 16
  uint32_t h = iq_bf16_rne(t);
  float r = t - __uint_as_float(h << 16);
  return h | (iq_bf16_rne(r) << 16);
}

// B fragments of a batch: lane l of (tile, step) holds the eight values of
// chunk 2*step + l/32 for query tile*32 + l%32, as bf16 hi and lo.
__device__ __forceinline__ void mfma_bprep_body(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, int nq,
                                                int step, int tile, int lane) {
  const int nstep = D->ktab.nstep;
  const int qi = tile * 32 + (lane & 31);
  uint32_t h[4] = {0u, 0u, 0u, 0u}, l[4] = {0u, 0u, 0u, 0u};
  if (qi < nq) {
    const IqQuery q = qs[qi];
    const IqStep& st = D->steps[q.step];
    if (!q.multi && st.mask >= 0) {
      const IqGeo& g = D->g;
      const IqChunk ch = D->chunks[2 * step + (lane >> 5)];
      const uint8_t* mb = D->maskbits + (int64_t)st.mask * g.tilevox;
      const IqReal& R = D->reals[q.real];
      if (ch.n > 0) {
        // mask bytes, then grid values of the chunk's (up to) eight voxels,
        // each batch of loads in flight together; positions past the chunk
        // repeat its last voxel and are dropped
        int64_t gi[8];
        uint8_t on[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const int ee = e < ch.n ? e : ch.n - 1;
          const int x = ch.jx + (ch.dir == 0 ? ee : 0), y = ch.jy + (ch.dir == 1 ? ee : 0), z = ch.jz;
          on[e] = mb[x + g.tile[0] * (y + g.tile[1] * z)];
          gi[e] = iq_lin3(R.pad, st.origin[0] + x, st.origin[1] + y, st.origin[2] + z);
        }
        float v[8];
        if (D->ti_f32) {
          const float* sim32 = (const float*)D->sims + R.sim_off;
#pragma unroll
          for (int e = 0; e < 8; e++) v[e] = sim32[gi[e]];
        } else {
          const double* sim64 = (const double*)D->sims + R.sim_off;
#pragma unroll
          for (int e = 0; e < 8; e++) v[e] = (float)sim64[gi[e]];
        }
#pragma unroll
        for (int e = 0; e < 8; e++) {
          if (e >= ch.n || !on[e]) continue;
          const uint32_t wv = iq_split_word(v[e]);
          h[e >> 1] |= (wv & 0xffffu) << (16 * (e & 1));
          l[e >> 1] |= (wv >> 16) << (16 * (e & 1));
        }
      }
    }
  }
  const int64_t o = ((int64_t)tile * nstep + step) * 64 + lane;
  D->bh[o] = make_uint4(h[0], h[1], h[2], h[3]);
  D->bl[o] = make_uint4(l[0], l[1], l[2], l[3]);
}

__global__ void __launch_bounds__(64)
k_mfma_bprep(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, int nq) {
  mfma_bprep_body(D, qs, nq, blockIdx.x, blockIdx.y, threadIdx.x);
}

// Query preparation and B fragments in one launch (both read the grid, neither
// reads what the other writes): blocks [0, nq) prepare one query each (the
// overlap source), the others hold four (k-step, query tile) items, one per
// wavefront.
template <typename T>
__global__ void __launch_bounds__(256)
k_prep_bprep(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, int nq, int nitems) {
  __shared__ double kl[PREP_LDS];
  if ((int)blockIdx.x < nq) {
    query_prep_body<T>(D, qs, blockIdx.x, 0, kl);
    return;
  }
  const int item = ((int)blockIdx.x - nq) * 4 + (int)(threadIdx.x >> 6);
  if (item >= nitems) return;
  const int nstep = D->ktab.nstep;
  mfma_bprep_body(D, qs, nq, item % nstep, item / nstep, threadIdx.x & 63);
}
template __global__ void k_prep_bprep<float>(const IqDev*, const IqQuery*, int, int);
template __global__ void k_prep_bprep<double>(const IqDev*, const IqQuery*, int, int);

// FP32 copy of an A2 map for the scan's interval bound; NaN marks a disabled
// window (an A2 that overflows FP32 becomes +inf and stays a candidate)
__global__ void __launch_bounds__(256)
k_a2_f32(const double* __restrict__ a2, const uint8_t* __restrict__ disabled, int64_t P, float* __restrict__ out) {
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < P; c += (int64_t)gridDim.x * 256)
    out[c] = disabled[c] ? __builtin_nanf("") : (float)a2[c];
}

template <bool SPLIT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_scan_mfma(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, int nq, double gerr,
            const int32_t* __restrict__ kofs_tab) {
  extern __shared__ uint32_t lw[];
  __shared__ float red[4][32];
  const IqGeo& g = D->g;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row = lane & 31, kb = lane >> 5;
  const int tile = blockIdx.z / g.dist[2];
  const int cz = blockIdx.z % g.dist[2];
  const int nw = (int)(blockDim.x >> 6), NT = (int)blockDim.x;  // 2 to 4 wavefronts, 4 rows each
  const int cx0 = blockIdx.x * 32, cy0 = blockIdx.y * 4 * nw;
  const bool wlive = cy0 + 4 * w < g.dist[1];  // uniform per wavefront
  const int nstep = D->ktab.nstep;
  const int t0 = g.ti[0], t1 = g.ti[1], t2 = g.ti[2];
  const int d0 = g.dist[0], d1 = g.dist[1];
  const auto ti = iq_glob(D->ti32);
  constexpr int SU = 8;
  iq_f32x16 acc[4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int v = 0; v < 16; v++) acc[i][v] = 0.0f;
  for (int b = 0; b < D->ktab.nbox; b++) {
    const IqKBox kbx = D->ktab.box[b];
    const int RX = 32 + kbx.wxr - 1, RY = 4 * nw + kbx.wyr - 1;
    const int n = RX * RY * kbx.wz;
    const int gx0 = cx0 + kbx.x0, gy0 = cy0 + kbx.y0, gz0 = cz + kbx.z0;
    // A box's chunks all run along x or all along y.  The block is staged so
    // that a run is contiguous: x-runs in the layout x + RX (y + RY z),
    // y-runs in y + RY (x + RX z); a wavefront's four rows then sit at a
    // uniform stride (RX, or 1) and the eight words of a run at immediate
    // offsets.  Element i = x + RX (y + RY z) of the walk; coordinates
    // advance incrementally (NT per element, SU elements per pass with all
    // their loads issued first); windows past the training image read as zero.
    const auto kofs = iq_glob(kofs_tab);
    const bool ydir = kbx.s0 < kbx.s1 && (kofs[2 * kbx.s0] >> 30) != 0;
    int x = (int)threadIdx.x % RX, y = ((int)threadIdx.x / RX) % RY, z = (int)threadIdx.x / (RX * RY);
    const int sx = NT % RX, sy = (NT / RX) % RY, sz = NT / (RX * RY);
    for (int i0 = threadIdx.x; i0 < n; i0 += NT * SU) {
      float v[SU];
      bool in[SU];
      int at[SU];
#pragma unroll
      for (int u = 0; u < SU; u++) {
        const int gx = gx0 + x, gy = gy0 + y, gz = gz0 + z;
        in[u] = gx < t0 && gy < t1 && gz < t2;
        v[u] = ti[in[u] ? gx + (int64_t)t0 * (gy + (int64_t)t1 * gz) : 0];
        at[u] = ydir ? y + RY * (x + RX * z) : i0 + NT * u;
        x += sx;
        y += sy;
        z += sz;
        if (x >= RX) { x -= RX; y++; }
        if (y >= RY) { y -= RY; z++; }
      }
#pragma unroll
      for (int u = 0; u < SU; u++)
        if (i0 + NT * u < n) lw[at[u]] = iq_split_word(in[u] ? v[u] : 0.0f);
    }
    __syncthreads();
    // B fragments and chunk offsets are loaded one k-step ahead
    const auto bhp = iq_glob((const iq_u32x4*)(D->bh + (int64_t)tile * nstep * 64 + lane));
    const auto blp = iq_glob((const iq_u32x4*)(D->bl + (int64_t)tile * nstep * 64 + lane));
    const int lane_off = ydir ? 4 * w + RY * row : 4 * w * RX + row;
    const int istr = ydir ? 1 : RX;
    iq_u32x4 bh_n = {0u, 0u, 0u, 0u}, bl_n = {0u, 0u, 0u, 0u};
    int ko_n = 0;
    if (kbx.s0 < kbx.s1) {
      bh_n = bhp[(int64_t)kbx.s0 * 64];
      if (SPLIT) bl_n = blp[(int64_t)kbx.s0 * 64];
      ko_n = kofs[2 * kbx.s0 + kb];
    }
    for (int step = kbx.s0; step < kbx.s1; step++) {
      const iq_bf16x8 bh = __builtin_bit_cast(iq_bf16x8, bh_n);
      iq_bf16x8 bl;
      if (SPLIT) bl = __builtin_bit_cast(iq_bf16x8, bl_n);
      const int ko = ko_n;
      if (step + 1 < kbx.s1) {
        bh_n = bhp[(int64_t)(step + 1) * 64];
        if (SPLIT) bl_n = blp[(int64_t)(step + 1) * 64];
        ko_n = kofs[2 * (step + 1) + kb];
      }
      const int base = (ko & 0x3fffffff) + lane_off;
      if (!wlive) continue;  // every candidate row of this wavefront is past the grid
      // all 32 LDS words of the four row blocks first, then the MFMAs in an
      // order where consecutive ones use different accumulators
      uint32_t wv[4][8];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int rb = base + i * istr;
#pragma unroll
        for (int e = 0; e < 8; e++) wv[i][e] = lw[rb + e];
      }
      iq_bf16x8 ah[4], al[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        uint4 h4, l4;
        h4.x = __builtin_amdgcn_perm(wv[i][1], wv[i][0], 0x05040100u);
        h4.y = __builtin_amdgcn_perm(wv[i][3], wv[i][2], 0x05040100u);
        h4.z = __builtin_amdgcn_perm(wv[i][5], wv[i][4], 0x05040100u);
        h4.w = __builtin_amdgcn_perm(wv[i][7], wv[i][6], 0x05040100u);
        ah[i] = __builtin_bit_cast(iq_bf16x8, h4);
        if (SPLIT) {
          l4.x = __builtin_amdgcn_perm(wv[i][1], wv[i][0], 0x07060302u);
          l4.y = __builtin_amdgcn_perm(wv[i][3], wv[i][2], 0x07060302u);
          l4.z = __builtin_amdgcn_perm(wv[i][5], wv[i][4], 0x07060302u);
          l4.w = __builtin_amdgcn_perm(wv[i][7], wv[i][6], 0x07060302u);
          al[i] = __builtin_bit_cast(iq_bf16x8, l4);
        }
      }
#pragma unroll
      for (int i = 0; i < 4; i++) acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh, acc[i], 0, 0, 0);
      if (SPLIT) {
#pragma unroll
        for (int i = 0; i < 4; i++) acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl, acc[i], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++) acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh, acc[i], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // epilogue: lane holds C[row'][col] for col = lane % 32 (its query)
  const int qi = tile * 32 + row;
  float lmin = __builtin_inff();
  bool live = false;
  int64_t slot = 0;
  if (qi < nq) {
    const IqQuery q = qs[qi];
    const IqStep& st = D->steps[q.step];
    live = !q.multi && st.mask >= 0;
    slot = q.slot;
    if (live) {
      // Interval around the oracle's FP64 distance, evaluated in FP32:
      // |AB32 - AB| <= gerr sum|a||b| and sum|a||b| <= S = sqrt(A2) sqrt(B2)
      // (square roots of operands clamped up to 2^-100, rounded up by
      // 2^-19); 8 u32 (A2 + B2 + 2S) covers the FP32 conversions and the
      // evaluation of d as well as the oracle's own FP64 rounding.  The main
      // term carries a 2^-10 margin for its own FP32 evaluation, the small
      // terms are doubled to cover the rounding of d -/+ e, and lo / hi are
      // pushed outward by 2^-20.  A2 = NaN marks a disabled window.
      const IqMask& m = D->masks[st.mask];
      const auto a2f = iq_glob(D->a2f[m.a2]);
      const float B = (float)D->b2[slot * (2 + D->naux)];
      const float sB = __builtin_amdgcn_sqrtf(fmaxf(B, 0x1p-100f)) * (1.0f + 0x1p-19f);
      const double V = (double)m.V;
      const double g64 = V * 0x1p-53 / (1.0 - V * 0x1p-53);
      const float gs2 = (float)((gerr + g64) * (1.0 + 0x1p-10)) * 2.0f * (1.0f + 0x1p-10f);
      const float u32 = 0x1p-24f;
      const float tiny2 = 2.0f * (float)(1e3 * V * 0x1p-125);
      // e = gs2 S + 2 (8 u32 ((A + B) + 2 S)) + tiny2, regrouped as
      // S (gs2 + 32 u32) + 16 u32 (A + B) + tiny2 and evaluated by two fmas:
      // fewer roundings than the plain sum, each within the margins above
      const float k1 = gs2 + 32.0f * u32, k2 = 16.0f * u32;
      const auto lo_out = iq_globw(D->lo32 + slot * g.P);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int cy = cy0 + 4 * w + i;
        if (cy >= d1) break;  // uniform per wavefront
        const int64_t rb = (int64_t)d0 * (cy + (int64_t)d1 * cz) + cx0 + 4 * kb;
        const int xl = d0 - cx0 - 4 * kb;  // candidates of this row left to the lane
        float A[16];
#pragma unroll
        for (int v = 0; v < 16; v++) {
          const int ox = 8 * (v >> 2) + (v & 3);
          A[v] = a2f[ox < xl ? rb + ox : rb - 4 * kb];
        }
#pragma unroll
        for (int v = 0; v < 16; v++) {
          const int ox = 8 * (v >> 2) + (v & 3);
          const float a = A[v];
          const float d = __builtin_fabsf(__builtin_fmaf(-2.0f, acc[i][v], a) + B);  // 2 acc is exact
          const float S = __builtin_amdgcn_sqrtf(fmaxf(a, 0x1p-100f)) * sB;
          const float e = __builtin_fmaf(S, k1, __builtin_fmaf(a + B, k2, tiny2));
          float lo = fmaxf(d - e, 0.0f) * (1.0f - 0x1p-20f);
          float hi = (d + e) * (1.0f + 0x1p-20f);
          if (a != a) lo = hi = __builtin_inff();
          if (ox < xl) {
            lo_out[rb + ox] = lo;
            lmin = fminf(lmin, hi);
          }
        }
      }
    }
  }
  lmin = fminf(lmin, __shfl_xor(lmin, 32));
  if (kb == 0) red[w][row] = lmin;
  __syncthreads();
  if (w == 0 && kb == 0 && live) {
    float m = red[0][row];
    for (int i = 1; i < nw; i++) m = fminf(m, red[i][row]);
    atomicMin(&D->gmin_hi[slot], __float_as_uint(m));
  }
}

template __global__ void k_scan_mfma<true>(const IqDev*, const IqQuery*, int, double, const int32_t*);
template __global__ void k_scan_mfma<false>(const IqDev*, const IqQuery*, int, double, const int32_t*);

#define IQ_SCAN_INST(R, LX, LY, LZ) \
  template __global__ void k_scan_f32<R, LX, LY, LZ>(const IqDev*, const IqQuery*, int);
IQ_SCAN_INST(4, 8, 8, 4)
IQ_SCAN_INST(4, 16, 4, 4)
IQ_SCAN_INST(2, 16, 8, 2)
IQ_SCAN_INST(4, 8, 16, 2)

// ---------------------------------------------------------------------------
// Band superset: lo <= T where T = round_up(fl(fl(1+tol) * U)) and U is the
// least upper bound of the scan, so every exact band member is selected.
// One bitmap bit per candidate (ballot words of 64 candidates) and an
// unordered list of the members.  Grid: (candidate blocks, query).
// ---------------------------------------------------------------------------
#define SEL_PT 8
__global__ void __launch_bounds__(256)
k_band_select(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  const IqQuery q = qs[blockIdx.y];
  const IqStep& st = D->steps[q.step];
  if (q.multi || st.mask < 0) return;
  const int64_t P = D->g.P;
  const int64_t slot = q.slot;
  const double U = (double)__uint_as_float(D->gmin_hi[slot]);
  const float Tf = iq_f32_up((1.0 + D->tol) * U);
  const auto lo = iq_glob(D->lo32 + slot * P);
  const int lane = threadIdx.x & 63;
  // a wavefront covers SEL_PT words of 64 consecutive candidates
  const int64_t c0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 64 * SEL_PT + lane;
  float v[SEL_PT];
#pragma unroll
  for (int j = 0; j < SEL_PT; j++) {
    const int64_t c = c0 + 64 * j;
    v[j] = lo[c < P ? c : 0];
  }
#pragma unroll
  for (int j = 0; j < SEL_PT; j++) {
    const int64_t c = c0 + 64 * j;
    const bool sel = c < P && v[j] <= Tf;
    const unsigned long long bal = __ballot(sel);
    if (lane == 0 && (c >> 6) < D->nw64) D->bitmap[slot * D->nw64 + (c >> 6)] = bal;
    if (bal) {
      int base = 0;
      if (lane == 0) base = atomicAdd(&D->nsup[slot], __popcll(bal));
      base = __shfl(base, 0);
      if (sel) D->sup[slot * P + base + __popcll(bal & ((1ull << lane) - 1ull))] = (int32_t)c;
    }
  }
}

// ---------------------------------------------------------------------------
// Exact recheck of the superset: the FMA chain over the masked voxels in
// ascending column-major order, so the FP64 value is the oracle's bit for
// bit.  Large supersets: one work-item per member, the query values and
// window offsets staged in LDS RC_CHUNK at a time for the whole workgroup.
// At most RC_SMALL members: one wavefront per member, window and query
// values staged by the wavefront, every lane running the same chain (short
// latency when the superset is small).  The exact minimum is reduced with an
// integer atomic on the non-negative double bits.
// Grid: (query, groups per query), 256 work-items per group.
// ---------------------------------------------------------------------------
#define RC_CHUNK 1024
#define RC_SMALL 512  // default of IqDev::rc_small
#define RC_WCH 512
#define RC_GROUPS 128  // member groups per query: 512 members in one pass, more loop
__global__ void __launch_bounds__(256)
k_band_recheck(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  __shared__ double kl[RC_CHUNK];
  __shared__ int32_t ol[RC_CHUNK];
  __shared__ double wl[4 * 2 * RC_WCH];
  __shared__ unsigned long long wbest[4];
  // grid (query, member group): the groups holding members of every query
  // are dispatched first, and groups past a query's superset leave at once
  const IqQuery q = qs[blockIdx.x];
  const int grp = blockIdx.y, ngrp = gridDim.y;
  const int n = D->nsup[q.slot];
  const int rcs = D->rc_small;
  if (grp * (n <= rcs ? 4 : 256) >= n) return;
  const IqStep& st = D->steps[q.step];
  if (q.multi || st.mask < 0) return;
  const IqGeo& g = D->g;
  const int64_t slot = q.slot;
  const IqMask& m = D->masks[st.mask];
  const int V = m.V;
  const int32_t* toff = D->toff + m.off;
  const double* kv = D->kv + slot * D->kv_stride;
  const double* a2map = D->a2maps[m.a2];
  const double b2 = D->b2[slot * (2 + D->naux)];
  double* dex = D->dexact + slot * g.P;
  unsigned long long best = 0x7ff0000000000000ull;
  if (n <= rcs) {  // few members: a wavefront per member, its chain fed from LDS
    // the wavefront stages RC_WCH window and query values per chunk and loads
    // the next chunk into registers while the chain runs over the current one
    double* tl = wl + (threadIdx.x >> 6) * (2 * RC_WCH);
    double* ql = tl + RC_WCH;
    const int lane = threadIdx.x & 63;
    const auto toffg = iq_glob(toff);
    const auto kvg = iq_glob(kv);
    const bool f32 = D->ti_f32;
    const auto ti32 = iq_glob((const float*)D->ti);
    const auto ti64 = iq_glob((const double*)D->ti);
    const int d0 = g.dist[0], d1 = g.dist[1];
    for (int e = grp * 4 + (int)(threadIdx.x >> 6); e < n; e += ngrp * 4) {
      const int32_t c = D->sup[slot * g.P + e];  // candidate indices fit 31 bits
      const int cx = c % d0, cy = (c / d0) % d1, cz = c / (d0 * d1);
      const int64_t base = iq_lin3(g.ti, cx, cy, cz);
      double rt[RC_WCH / 64], rq[RC_WCH / 64];
      // all offsets of a chunk first, then all window values (two latencies)
      auto fetch = [&](int j0) {
        int32_t o[RC_WCH / 64];
#pragma unroll
        for (int i = 0; i < RC_WCH / 64; i++) {
          const int k = j0 + lane + 64 * i;
          const int kc = k < V ? k : 0;
          o[i] = toffg[kc];
          rq[i] = kvg[kc];
        }
        if (f32) {  // uniform: the training image is read in its own element type
#pragma unroll
          for (int i = 0; i < RC_WCH / 64; i++) rt[i] = (double)ti32[base + o[i]];
        } else {
#pragma unroll
          for (int i = 0; i < RC_WCH / 64; i++) rt[i] = ti64[base + o[i]];
        }
#pragma unroll
        for (int i = 0; i < RC_WCH / 64; i++) {
          const bool ok = j0 + lane + 64 * i < V;
          rt[i] = ok ? rt[i] : 0.0;
          rq[i] = ok ? rq[i] : 0.0;
        }
      };
      fetch(0);
      double ab = 0.0;
      for (int j0 = 0; j0 < V; j0 += RC_WCH) {
        const int nj = V - j0 < RC_WCH ? V - j0 : RC_WCH;
#pragma unroll
        for (int i = 0; i < RC_WCH / 64; i++) {
          tl[lane + 64 * i] = rt[i];
          ql[lane + 64 * i] = rq[i];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (j0 + RC_WCH < V) fetch(j0 + RC_WCH);  // next chunk in flight during the chain
        // the ordered chain runs on lane 0 alone (the other lanes' LDS reads
        // would only repeat its values and take LDS bandwidth from other
        // wavefronts); eight LDS pairs of the next group in flight while the
        // fmas of the current group run
        if (lane == 0) {
          int k = 0;
          double t[8], u[8];
          if (nj >= 8) {
#pragma unroll
            for (int i = 0; i < 8; i++) { t[i] = tl[i]; u[i] = ql[i]; }
          }
          for (; k + 8 <= nj; k += 8) {
            double tn[8], un[8];
            const int kn = k + 16 <= nj ? k + 8 : k;
#pragma unroll
            for (int i = 0; i < 8; i++) { tn[i] = tl[kn + i]; un[i] = ql[kn + i]; }
#pragma unroll
            for (int i = 0; i < 8; i++) ab = __builtin_fma(t[i], u[i], ab);
#pragma unroll
            for (int i = 0; i < 8; i++) { t[i] = tn[i]; u[i] = un[i]; }
          }
          for (; k < nj; k++) ab = __builtin_fma(tl[k], ql[k], ab);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      if (lane == 0) {
        const double d = __builtin_fabs((a2map[c] - 2.0 * ab) + b2);
        dex[c] = d;
        const unsigned long long bits = (unsigned long long)__double_as_longlong(d);
        best = bits < best ? bits : best;
      }
    }
  }
  for (int e0 = grp * 256; n > rcs && e0 < n; e0 += ngrp * 256) {  // uniform over the group
    const int e = e0 + (int)threadIdx.x;
    const bool live = e < n;
    int64_t c = 0, base = 0;
    if (live) {
      c = D->sup[slot * g.P + e];
      const int cx = (int)(c % g.dist[0]), cy = (int)((c / g.dist[0]) % g.dist[1]);
      const int cz = (int)(c / ((int64_t)g.dist[0] * g.dist[1]));
      base = iq_lin3(g.ti, cx, cy, cz);
    }
    double ab = 0.0;
    for (int j0 = 0; j0 < V; j0 += RC_CHUNK) {
      const int nj = V - j0 < RC_CHUNK ? V - j0 : RC_CHUNK;
      __syncthreads();
      for (int k = threadIdx.x; k < nj; k += 256) {
        kl[k] = kv[j0 + k];
        ol[k] = toff[j0 + k];
      }
      __syncthreads();
      if (live) {  // 16 window loads in flight, then the ordered fmas
        int k = 0;
        if (D->ti_f32) {
          const float* ti = (const float*)D->ti + base;
          for (; k + 16 <= nj; k += 16) {
            float t[16];
#pragma unroll
            for (int i = 0; i < 16; i++) t[i] = ti[ol[k + i]];
#pragma unroll
            for (int i = 0; i < 16; i++) ab = __builtin_fma((double)t[i], kl[k + i], ab);
          }
          for (; k < nj; k++) ab = __builtin_fma((double)ti[ol[k]], kl[k], ab);
        } else {
          const double* ti = (const double*)D->ti + base;
          for (; k + 16 <= nj; k += 16) {
            double t[16];
#pragma unroll
            for (int i = 0; i < 16; i++) t[i] = ti[ol[k + i]];
#pragma unroll
            for (int i = 0; i < 16; i++) ab = __builtin_fma(t[i], kl[k + i], ab);
          }
          for (; k < nj; k++) ab = __builtin_fma(ti[ol[k]], kl[k], ab);
        }
      }
    }
    if (live) {
      const double d = __builtin_fabs((a2map[c] - 2.0 * ab) + b2);
      dex[c] = d;
      const unsigned long long bits = (unsigned long long)__double_as_longlong(d);
      best = bits < best ? bits : best;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long other = __shfl_xor(best, o);
    best = other < best ? other : best;
  }
  if ((threadIdx.x & 63) == 0) wbest[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long b = wbest[0];
    for (int i = 1; i < 4; i++) b = wbest[i] < b ? wbest[i] : b;
    if (b != 0x7ff0000000000000ull) atomicMin(&D->gmin64[slot], b);
  }
}

// Exact recheck with a lane per member: the same ordered FMA chain (so the
// same FP64 value), each member's chain run by one lane.  A wavefront serves
// RCL_MPW members; for every block of 64 masked voxels it loads the members'
// window values one member at a time (lane l: voxel l of that window, so
// each load is a few contiguous runs of the window), transposes them through
// LDS with the block's 64 query values (lane l loads value l), and lanes
// 0..RCL_MPW-1 then chain over their member's row, reading the query values
// from LDS (one address for all lanes).  The next block's loads are issued
// before the current block's chains run.  (Query values read by scalar
// loads inside the chain: 280 us per launch at the 2.9M-candidate workload;
// the scalar cache does not hold a query's 38 KB, and every 16 chain steps
// waited on a scalar load.)  Few members per wavefront keep
// one round of loads per block and many wavefronts per query.
// Grid: (query, RCL_GROUPS wavefronts), members beyond loop.
#define RCL_GROUPS 32
#define RCL_MPW 16
template <typename E>
__device__ __forceinline__ double rcl_chain(const IqDev* __restrict__ D, int64_t base, const int32_t* __restrict__ toff,
                                            const double* kvp, int V, E (*tile)[65], double* kt) {
  const int lane = threadIdx.x;
  const auto ti = iq_glob((const E*)D->ti);
  const auto kv = iq_glob(kvp);
  int64_t bj[RCL_MPW];
#pragma unroll
  for (int u = 0; u < RCL_MPW; u++)
    bj[u] = (int64_t)(uint32_t)__builtin_amdgcn_readlane((int)base, u) |
            ((int64_t)__builtin_amdgcn_readlane((int)(base >> 32), u) << 32);
  E t[RCL_MPW];
  double kq;
  {
    const int i = lane < V ? lane : 0;
    const int32_t o = toff[i];
#pragma unroll
    for (int u = 0; u < RCL_MPW; u++) t[u] = ti[bj[u] + o];
    kq = kv[i];
  }
  double ab = 0.0;
  for (int v0 = 0; v0 < V; v0 += 64) {
    const int nv = V - v0 < 64 ? V - v0 : 64;
#pragma unroll
    for (int u = 0; u < RCL_MPW; u++) tile[u][lane] = t[u];
    kt[lane] = kq;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int v1 = v0 + 64;
    if (v1 < V) {  // the next block's window and query values, in flight during the chains
      const int i = v1 + (v1 + lane < V ? lane : 0);
      const int32_t o = toff[i];
#pragma unroll
      for (int u = 0; u < RCL_MPW; u++) t[u] = ti[bj[u] + o];
      kq = kv[i];
    }
    if (lane < RCL_MPW) {
      const E* row = tile[lane];
      if (nv == 64) {
#pragma unroll 16
        for (int i = 0; i < 64; i++) ab = __builtin_fma((double)row[i], kt[i], ab);
      } else {
        for (int i = 0; i < nv; i++) ab = __builtin_fma((double)row[i], kt[i], ab);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  return ab;
}

__global__ void __launch_bounds__(64)
k_band_recheck_lanes(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  __shared__ double tile_mem[RCL_MPW * 65];  // the transposed block (doubles, or floats in its first half)
  __shared__ double kt[64];                  // the block's query values
  const IqQuery q = qs[blockIdx.x];
  const int n = D->nsup[q.slot];
  if ((int)blockIdx.y * RCL_MPW >= n) return;
  const IqStep& st = D->steps[q.step];
  if (q.multi || st.mask < 0) return;
  const IqGeo& g = D->g;
  const int64_t slot = q.slot;
  const IqMask& m = D->masks[st.mask];
  const int V = m.V;
  const int32_t* toff = D->toff + m.off;
  const double* kvp = D->kv + slot * D->kv_stride;
  const double* a2map = D->a2maps[m.a2];
  const double b2 = D->b2[slot * (2 + D->naux)];
  double* dex = D->dexact + slot * g.P;
  const int lane = threadIdx.x;
  unsigned long long best = 0x7ff0000000000000ull;
  for (int e0 = blockIdx.y * RCL_MPW; e0 < n; e0 += gridDim.y * RCL_MPW) {  // uniform over the wavefront
    const int e = e0 + lane;
    const bool live = lane < RCL_MPW && e < n;
    int64_t c = 0, base = 0;
    if (live) {
      c = D->sup[slot * g.P + e];
      const int cx = (int)(c % g.dist[0]), cy = (int)((c / g.dist[0]) % g.dist[1]);
      const int cz = (int)(c / ((int64_t)g.dist[0] * g.dist[1]));
      base = iq_lin3(g.ti, cx, cy, cz);
    }
    const double ab = D->ti_f32 ? rcl_chain<float>(D, base, toff, kvp, V, (float(*)[65])tile_mem, kt)
                                : rcl_chain<double>(D, base, toff, kvp, V, (double(*)[65])tile_mem, kt);
    if (live) {
      const double d = __builtin_fabs((a2map[c] - 2.0 * ab) + b2);
      dex[c] = d;
      const unsigned long long bits = (unsigned long long)__double_as_longlong(d);
      best = bits < best ? bits : best;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long other = __shfl_xor(best, o);
    best = other < best ? other : best;
  }
  if (lane == 0 && best != 0x7ff0000000000000ull) atomicMin(&D->gmin64[slot], best);
}

// ---------------------------------------------------------------------------
// Exact maps for the multi-source selection (relaxation): one work-item per
// candidate, grid (candidate blocks, query, source).  Source 0 uses the A2 map
// of its mask (or is identically 0 without a mask), the hard source chains
// A2, soft sources use the full-tile A2 map of their auxiliary image.
// Disabled candidates are +inf (iqsim.jl:207,217,226).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_multi_maps(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  const IqQuery q = qs[blockIdx.y];
  if (!q.multi) return;
  const int src = blockIdx.z;
  const IqStep& st = D->steps[q.step];
  if (src == 1 && !st.hard) return;
  const IqGeo& g = D->g;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.P) return;
  const int64_t slot = q.slot;
  double* out = D->maps + (slot * D->nmap + src) * g.P;
  if (D->disabled[c]) { out[c] = __builtin_inf(); return; }
  int V;
  const int32_t* toff;
  const double* a2map = nullptr;
  const double* img64 = nullptr;
  if (src == 0) {
    if (st.mask < 0) { out[c] = 0.0; return; }
    const IqMask& m = D->masks[st.mask];
    V = m.V; toff = D->toff + m.off; a2map = D->a2maps[m.a2];
  } else if (src == 1) {
    V = st.hcnt; toff = D->htoff + st.hoff;
  } else {
    V = g.tilevox; toff = D->toff + D->full_off; a2map = D->a2soft[src - 2]; img64 = D->auxti[src - 2];
  }
  const double* kv = D->kv + slot * D->kv_stride + (int64_t)src * g.tilevox;
  const double b2 = D->b2[slot * (2 + D->naux) + src];
  int cx = (int)(c % g.dist[0]), cy = (int)((c / g.dist[0]) % g.dist[1]);
  int cz = (int)(c / ((int64_t)g.dist[0] * g.dist[1]));
  const int64_t base = iq_lin3(g.ti, cx, cy, cz);
  double a2 = 0.0, ab = 0.0;
  if (img64) {
    for (int j = 0; j < V; j++) ab = __builtin_fma(img64[base + toff[j]], kv[j], ab);
  } else if (D->ti_f32) {
    const float* ti = (const float*)D->ti;
    for (int j = 0; j < V; j++) {
      double t = (double)ti[base + toff[j]];
      if (!a2map) a2 = __builtin_fma(t, t, a2);
      ab = __builtin_fma(t, kv[j], ab);
    }
  } else {
    const double* ti = (const double*)D->ti;
    for (int j = 0; j < V; j++) {
      double t = ti[base + toff[j]];
      if (!a2map) a2 = __builtin_fma(t, t, a2);
      ab = __builtin_fma(t, kv[j], ab);
    }
  }
  if (a2map) a2 = a2map[c];
  out[c] = __builtin_fabs((a2 - 2.0 * ab) + b2);
}

// Tiled form of k_multi_maps: a workgroup of MX x MY work-items owns MX x MY
// candidates at one z (32 x 16 where the region fits, else 16 x 16: rows of
// 32 consecutive candidates keep a half-wavefront's 8-byte reads on distinct
// LDS banks, and 512-wide groups hold more wavefronts per compute unit); the training-image (or auxiliary-TI) region their windows cover is
// staged in LDS as FP64, so every work-item's chain (same ascending order,
// same operations as k_multi_maps) reads LDS instead of gathering from HBM.
// The query values and the voxels' region offsets (prepared per plan) are
// the same for every work-item and come through scalar loads.  Grid: (x
// blocks * y blocks * dist z, query, source); LDS = the region (the host
// checks it fits).
template <int MT_X, int MT_Y>
__global__ void __launch_bounds__(MT_X * MT_Y)
k_multi_maps_tiled(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  extern __shared__ double mreg[];
  const IqQuery q = qs[blockIdx.y];
  if (!q.multi) return;
  const int src = blockIdx.z;
  const IqStep& st = D->steps[q.step];
  if (src == 1 && !st.hard) return;
  const IqGeo& g = D->g;
  const int nbx = (g.dist[0] + MT_X - 1) / MT_X, nby = (g.dist[1] + MT_Y - 1) / MT_Y;
  const int bx = blockIdx.x % nbx, by = (blockIdx.x / nbx) % nby, cz = blockIdx.x / (nbx * nby);
  const int64_t slot = q.slot;
  double* out = D->maps + (slot * D->nmap + src) * g.P;
  int V;
  const int32_t* rtv;
  const double* a2map = nullptr;
  const double* img64 = nullptr;
  if (src == 0) {
    if (st.mask < 0) {
      const int cx = bx * MT_X + (threadIdx.x % MT_X), cy = by * MT_Y + (threadIdx.x / MT_X);
      if (cx < g.dist[0] && cy < g.dist[1]) {
        const int64_t c = iq_lin3(g.dist, cx, cy, cz);
        out[c] = D->disabled[c] ? __builtin_inf() : 0.0;
      }
      return;
    }
    const IqMask& m = D->masks[st.mask];
    V = m.V; rtv = D->rtvox + m.off; a2map = D->a2maps[m.a2];
  } else if (src == 1) {
    V = st.hcnt; rtv = D->rhtvox + st.hoff;
  } else {
    V = g.tilevox; rtv = D->rtvox + D->full_off; a2map = D->a2soft[src - 2]; img64 = D->auxti[src - 2];
  }
  const int RX = MT_X + g.tile[0] - 1, RY = MT_Y + g.tile[1] - 1, RZ = g.tile[2];
  const int nreg = RX * RY * RZ;
  double* reg = mreg;
  const int gx0 = bx * MT_X, gy0 = by * MT_Y;
  for (int i = threadIdx.x; i < nreg; i += blockDim.x) {
    const int x = i % RX, y = (i / RX) % RY, z = i / (RX * RY);
    const int gx = gx0 + x, gy = gy0 + y, gz = cz + z;
    double v = 0.0;
    if (gx < g.ti[0] && gy < g.ti[1]) {
      const int64_t o = iq_lin3(g.ti, gx, gy, gz);
      v = img64 ? img64[o] : (D->ti_f32 ? (double)((const float*)D->ti)[o] : ((const double*)D->ti)[o]);
    }
    reg[i] = v;
  }
  // query values and region offsets are the same for every work-item:
  // constant-address loads (scalar unit, no LDS traffic)
  const auto kq = iq_const(D->kv + slot * D->kv_stride + (int64_t)src * g.tilevox);
  const auto ro = iq_const(rtv);
  __syncthreads();
  const int lx = threadIdx.x % MT_X, ly = threadIdx.x / MT_X;
  const int cx = gx0 + lx, cy = gy0 + ly;
  if (cx >= g.dist[0] || cy >= g.dist[1]) return;
  const int64_t c = iq_lin3(g.dist, cx, cy, cz);
  if (D->disabled[c]) { out[c] = __builtin_inf(); return; }
  const double b2 = D->b2[slot * (2 + D->naux) + src];
  const double* r0 = reg + lx + RX * ly;
  double a2 = 0.0, ab = 0.0;
  int j = 0;
  if (!a2map) {
    for (; j + 8 <= V; j += 8) {
      double t[8], u[8];
#pragma unroll
      for (int i = 0; i < 8; i++) { t[i] = r0[ro[j + i]]; u[i] = kq[j + i]; }
#pragma unroll
      for (int i = 0; i < 8; i++) { a2 = __builtin_fma(t[i], t[i], a2); ab = __builtin_fma(t[i], u[i], ab); }
    }
    for (; j < V; j++) { const double t = r0[ro[j]]; a2 = __builtin_fma(t, t, a2); ab = __builtin_fma(t, kq[j], ab); }
  } else {
    for (; j + 8 <= V; j += 8) {
      double t[8], u[8];
#pragma unroll
      for (int i = 0; i < 8; i++) { t[i] = r0[ro[j + i]]; u[i] = kq[j + i]; }
#pragma unroll
      for (int i = 0; i < 8; i++) ab = __builtin_fma(t[i], u[i], ab);
    }
    for (; j < V; j++) ab = __builtin_fma(r0[ro[j]], kq[j], ab);
    a2 = a2map[c];
  }
  out[c] = __builtin_fabs((a2 - 2.0 * ab) + b2);
}

// Plain queries whose training image is not exactly representable in FP32:
// the overlap map is computed exactly for every candidate, all finite
// candidates form the superset.
__global__ void __launch_bounds__(256)
k_band_exact_all(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  const IqQuery q = qs[blockIdx.y];
  const IqStep& st = D->steps[q.step];
  if (q.multi || st.mask < 0) return;
  const IqGeo& g = D->g;
  const int64_t slot = q.slot;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const IqMask& m = D->masks[st.mask];
  bool in = false;
  double d = __builtin_inf();
  if (c < g.P && !D->disabled[c]) {
    const int32_t* toff = D->toff + m.off;
    const double* kv = D->kv + slot * D->kv_stride;
    int cx = (int)(c % g.dist[0]), cy = (int)((c / g.dist[0]) % g.dist[1]);
    int cz = (int)(c / ((int64_t)g.dist[0] * g.dist[1]));
    const int64_t base = iq_lin3(g.ti, cx, cy, cz);
    const double* ti = (const double*)D->ti;
    double ab = 0.0;
    for (int j = 0; j < m.V; j++) ab = __builtin_fma(ti[base + toff[j]], kv[j], ab);
    d = __builtin_fabs((D->a2maps[m.a2][c] - 2.0 * ab) + D->b2[slot * (2 + D->naux)]);
    D->dexact[slot * g.P + c] = d;
    in = true;
  }
  const unsigned long long bal = __ballot(in);
  if ((threadIdx.x & 63) == 0 && (c >> 6) < D->nw64) D->bitmap[slot * D->nw64 + (c >> 6)] = bal;
  unsigned long long bits = in ? (unsigned long long)__double_as_longlong(d) : 0x7ff0000000000000ull;
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long other = __shfl_xor(bits, o);
    bits = other < bits ? other : bits;
  }
  if ((threadIdx.x & 63) == 0 && bal) atomicMin(&D->gmin64[slot], bits);
}
