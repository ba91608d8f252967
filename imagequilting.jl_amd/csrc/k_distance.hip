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
__global__ void __launch_bounds__(256)
k_query_prep(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  __shared__ double kl[PREP_LDS];
  const IqQuery q = qs[blockIdx.x];
  const int src = blockIdx.y;
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
  const T* sim = (const T*)D->sims + (int64_t)q.real * g.padvox;
  const double* dgrid = src == 1 ? D->hardgrid : (src >= 2 ? D->aux[src - 2] : nullptr);
  for (int j = threadIdx.x; j < V; j += blockDim.x) {
    int t = tv[j];
    int x = t % g.tile[0], y = (t / g.tile[0]) % g.tile[1], z = t / (g.tile[0] * g.tile[1]);
    int64_t gi = iq_lin3(g.pad, ox + x, oy + y, oz + z);
    double v = src == 0 ? (double)sim[gi] : dgrid[gi];
    kv[j] = v;
    if (j < PREP_LDS) kl[j] = v;
  }
  if (src == 0 && mask && !q.multi && D->ti32) {
    float* kv32 = D->kv32 + slot * D->kv32_stride;
    for (int b = 0; b < mask->nbox; b++) {
      const IqBox bx = mask->box[b];
      int n = bx.wxp * bx.wy * bx.wz;
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int x = i % bx.wxp, y = (i / bx.wxp) % bx.wy, z = i / (bx.wxp * bx.wy);
        float v = 0.0f;
        if (x < bx.wx) v = (float)sim[iq_lin3(g.pad, ox + bx.x0 + x, oy + bx.y0 + y, oz + bx.z0 + z)];
        kv32[bx.kofs + i] = v;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    int j = 0;
    const int vl = V < PREP_LDS ? V : PREP_LDS;
    for (; j + 8 <= vl; j += 8) {
      double a0 = kl[j], a1 = kl[j + 1], a2 = kl[j + 2], a3 = kl[j + 3];
      double a4 = kl[j + 4], a5 = kl[j + 5], a6 = kl[j + 6], a7 = kl[j + 7];
      s = __builtin_fma(a0, a0, s); s = __builtin_fma(a1, a1, s);
      s = __builtin_fma(a2, a2, s); s = __builtin_fma(a3, a3, s);
      s = __builtin_fma(a4, a4, s); s = __builtin_fma(a5, a5, s);
      s = __builtin_fma(a6, a6, s); s = __builtin_fma(a7, a7, s);
    }
    for (; j < vl; j++) s = __builtin_fma(kl[j], kl[j], s);
    for (; j < V; j++) s = __builtin_fma(kv[j], kv[j], s);  // beyond the LDS stage
    D->b2[slot * (2 + D->naux) + src] = s;
  }
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
__global__ void __launch_bounds__(256)
k_band_select(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  const IqQuery q = qs[blockIdx.y];
  const IqStep& st = D->steps[q.step];
  if (q.multi || st.mask < 0) return;
  const IqGeo& g = D->g;
  const int64_t slot = q.slot;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double U = (double)__uint_as_float(D->gmin_hi[slot]);
  const float Tf = iq_f32_up((1.0 + D->tol) * U);
  bool sel = false;
  if (c < g.P) sel = D->lo32[slot * g.P + c] <= Tf;
  const unsigned long long bal = __ballot(sel);
  const int lane = threadIdx.x & 63;
  if (lane == 0 && (c >> 6) < D->nw64) D->bitmap[slot * D->nw64 + (c >> 6)] = bal;
  if (bal) {
    int base = 0;
    if (lane == 0) base = atomicAdd(&D->nsup[slot], __popcll(bal));
    base = __shfl(base, 0);
    if (sel) D->sup[slot * g.P + base + __popcll(bal & ((1ull << lane) - 1ull))] = (int32_t)c;
  }
}

// ---------------------------------------------------------------------------
// Exact recheck of the superset: one wavefront per member.  The masked window
// values and the query values are staged in LDS 512 at a time by the whole
// wavefront; every lane then runs the same ascending fma chain (broadcast LDS
// reads), so the FP64 value is the oracle's bit for bit.  The exact minimum
// is reduced with an integer atomic on the non-negative double bits.
// Grid: (groups per query, query); 4 wavefronts per group.
// ---------------------------------------------------------------------------
#define RC_CHUNK 512
__global__ void __launch_bounds__(256)
k_band_recheck(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  __shared__ double tl[4][RC_CHUNK];
  __shared__ double kl[4][RC_CHUNK];
  const IqQuery q = qs[blockIdx.y];
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
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nwave = gridDim.x * 4;
  const int n = D->nsup[slot];
  double* dex = D->dexact + slot * g.P;
  unsigned long long best = 0x7ff0000000000000ull;
  for (int e = blockIdx.x * 4 + w; e < n; e += nwave) {
    const int64_t c = D->sup[slot * g.P + e];
    int cx = (int)(c % g.dist[0]), cy = (int)((c / g.dist[0]) % g.dist[1]);
    int cz = (int)(c / ((int64_t)g.dist[0] * g.dist[1]));
    const int64_t base = iq_lin3(g.ti, cx, cy, cz);
    double ab = 0.0;
    for (int j0 = 0; j0 < V; j0 += RC_CHUNK) {
      const int nj = V - j0 < RC_CHUNK ? V - j0 : RC_CHUNK;
      for (int k = lane; k < nj; k += 64) {
        tl[w][k] = D->ti_f32 ? (double)((const float*)D->ti)[base + toff[j0 + k]]
                             : ((const double*)D->ti)[base + toff[j0 + k]];
        kl[w][k] = kv[j0 + k];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int k = 0; k < nj; k++) ab = __builtin_fma(tl[w][k], kl[w][k], ab);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    const double d = __builtin_fabs((a2map[c] - 2.0 * ab) + b2);
    if (lane == 0) dex[c] = d;
    unsigned long long bits = (unsigned long long)__double_as_longlong(d);
    best = bits < best ? bits : best;
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
