// k_distance.hip -- masked sliding distances (fastdistance, utils.jl:5-13)
// and valid cross-correlation (imfilter.jl:5-26) for gfx950.
//
// Two evaluations of the same map:
//   * exact:  D[c] = |(A2 - 2 AB) + B2| with A2, AB, B2 fused-multiply-add
//             chains in FP64 over the masked tile voxels in ascending
//             column-major order (one work-item per candidate);
//   * scan:   AB in FP32 over disjoint boxes with the training image staged
//             through LDS, A2 from the exact per-mask map, and a rigorous
//             error bound, giving an interval [lo, hi] that contains the exact
//             FP64 value.  The scan feeds the tol-band superset; members of
//             the superset are recomputed exactly.
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
// Exact FP64 maps.  MODE 0: D with the A2 chain; MODE 1: D with a precomputed
// A2 map; MODE 2: the A2 chain alone.  toff: training-image linear offsets of
// the masked voxels (ascending column-major tile order); kv: query values.
// ---------------------------------------------------------------------------
template <typename T, int MODE>
__global__ void __launch_bounds__(256)
k_exact_map(const T* __restrict__ img, IqGeo g, const int32_t* __restrict__ toff,
            const double* __restrict__ kv, int V, const double* __restrict__ b2p,
            const double* __restrict__ a2map, const uint8_t* __restrict__ disabled,
            double* __restrict__ out) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.P) return;
  if (MODE != 2 && disabled && disabled[c]) { out[c] = __builtin_inf(); return; }
  int cx = (int)(c % g.dist[0]), cy = (int)((c / g.dist[0]) % g.dist[1]);
  int cz = (int)(c / ((int64_t)g.dist[0] * g.dist[1]));
  const T* base = img + iq_lin3(g.ti, cx, cy, cz);
  double a2 = 0.0, ab = 0.0;
  for (int j = 0; j < V; j++) {
    double t = (double)base[toff[j]];
    if (MODE != 1) a2 = __builtin_fma(t, t, a2);
    if (MODE != 2) ab = __builtin_fma(t, kv[j], ab);
  }
  if (MODE == 2) { out[c] = a2; return; }
  if (MODE == 1) a2 = a2map[c];
  double b2 = *b2p;
  out[c] = __builtin_fabs((a2 - 2.0 * ab) + b2);
}

template __global__ void k_exact_map<float, 0>(const float*, IqGeo, const int32_t*, const double*, int, const double*, const double*, const uint8_t*, double*);
template __global__ void k_exact_map<float, 1>(const float*, IqGeo, const int32_t*, const double*, int, const double*, const double*, const uint8_t*, double*);
template __global__ void k_exact_map<float, 2>(const float*, IqGeo, const int32_t*, const double*, int, const double*, const double*, const uint8_t*, double*);
template __global__ void k_exact_map<double, 0>(const double*, IqGeo, const int32_t*, const double*, int, const double*, const double*, const uint8_t*, double*);
template __global__ void k_exact_map<double, 1>(const double*, IqGeo, const int32_t*, const double*, int, const double*, const double*, const uint8_t*, double*);
template __global__ void k_exact_map<double, 2>(const double*, IqGeo, const int32_t*, const double*, int, const double*, const double*, const uint8_t*, double*);

// ---------------------------------------------------------------------------
// Query staging (one workgroup): gather the query values at the masked voxels
// of a tile (ascending column-major) into kv, the box-ordered zero-padded FP32
// copy into kv32 (when a mask is given), the exact B2 chain into *b2, and reset
// the per-step reduction words.
// ---------------------------------------------------------------------------
template <typename G>
__global__ void __launch_bounds__(256)
k_query_prep(const G* __restrict__ grid, IqGeo g, const int32_t* __restrict__ origin3,
             const int32_t* __restrict__ tvox, int V, const IqMask* __restrict__ mask,
             double* __restrict__ kv, float* __restrict__ kv32, double* __restrict__ b2,
             uint32_t* __restrict__ gmin_hi, unsigned long long* __restrict__ gmin64) {
  const int ox = origin3[0], oy = origin3[1], oz = origin3[2];
  for (int j = threadIdx.x; j < V; j += blockDim.x) {
    int t = tvox[j];
    int x = t % g.tile[0], y = (t / g.tile[0]) % g.tile[1], z = t / (g.tile[0] * g.tile[1]);
    kv[j] = (double)grid[iq_lin3(g.pad, ox + x, oy + y, oz + z)];
  }
  if (mask && kv32) {
    for (int b = 0; b < mask->nbox; b++) {
      const IqBox bx = mask->box[b];
      int n = bx.wxp * bx.wy * bx.wz;
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int x = i % bx.wxp, y = (i / bx.wxp) % bx.wy, z = i / (bx.wxp * bx.wy);
        float v = 0.0f;
        if (x < bx.wx) v = (float)grid[iq_lin3(g.pad, ox + bx.x0 + x, oy + bx.y0 + y, oz + bx.z0 + z)];
        kv32[bx.kofs + i] = v;
      }
    }
  }
  if (threadIdx.x == 0) {
    if (gmin_hi) *gmin_hi = 0x7f800000u;                 // +inf
    if (gmin64) *gmin64 = 0x7ff0000000000000ull;         // +inf
  }
  __syncthreads();  // kv complete (the chain reads it in order)
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int j = 0; j < V; j++) s = __builtin_fma(kv[j], kv[j], s);
    *b2 = s;
  }
}

template __global__ void k_query_prep<float>(const float*, IqGeo, const int32_t*, const int32_t*, int, const IqMask*, double*, float*, double*, uint32_t*, unsigned long long*);
template __global__ void k_query_prep<double>(const double*, IqGeo, const int32_t*, const int32_t*, int, const IqMask*, double*, float*, double*, uint32_t*, unsigned long long*);

// ---------------------------------------------------------------------------
// FP32 scan.  Work-item (lx, ly, lz) of a 256-wide group owns R consecutive
// candidates along x.  Per box, the needed training-image block is staged in
// LDS with an R-interleaved row layout (element x of a row lives at
// (x % R) * S + x / R) so that lanes read consecutive words, conflict free,
// while each loaded word feeds R fused multiply-adds.
// ---------------------------------------------------------------------------
template <int R, int LX, int LY, int LZ>
__global__ void __launch_bounds__(LX * LY * LZ)
k_scan_f32(const float* __restrict__ ti, IqGeo g, const IqMask* __restrict__ mask,
           const float* __restrict__ kv32, const double* __restrict__ b2p,
           const double* __restrict__ a2map, const uint8_t* __restrict__ disabled,
           float* __restrict__ lo_out, uint32_t* __restrict__ gmin_hi) {
  extern __shared__ float lds[];
  const int tid = threadIdx.x;
  const int lx = tid % LX, ly = (tid / LX) % LY, lz = tid / (LX * LY);
  const int cx0 = blockIdx.x * (LX * R), cy0 = blockIdx.y * LY, cz0 = blockIdx.z * LZ;
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
  const double b2 = *b2p;
  const double V = (double)mask->V;
  const double u32 = 5.9604644775390625e-08, u64 = 1.1102230246251565e-16;
  const double g32 = V * u32 / (1.0 - V * u32), g64 = V * u64 / (1.0 - V * u64);
  float lmin = __builtin_inff();
  const int cy = cy0 + ly, cz = cz0 + lz;
#pragma unroll
  for (int i = 0; i < R; i++) {
    const int cx = cx0 + R * lx + i;
    if (cx < g.dist[0] && cy < g.dist[1] && cz < g.dist[2]) {
      int64_t c = iq_lin3(g.dist, cx, cy, cz);
      float lo = __builtin_inff(), hi = __builtin_inff();
      if (!disabled[c]) {
        double a2 = a2map[c];
        double s = __builtin_sqrt(a2 * b2) * (1.0 + 0x1p-40);
        double d = __builtin_fabs((a2 - 2.0 * (double)acc[i]) + b2);
        double e = 2.0 * (g32 + g64) * s + 4.0 * u64 * (a2 + b2 + 2.0 * s) + 2.0 * V * 0x1p-148;
        e = e * (1.0 + 0x1p-30);
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
    atomicMin(gmin_hi, __float_as_uint(m));  // non-negative floats order like their bits
  }
}

template __global__ void k_scan_f32<4, 8, 8, 4>(const float*, IqGeo, const IqMask*, const float*, const double*, const double*, const uint8_t*, float*, uint32_t*);
template __global__ void k_scan_f32<4, 16, 4, 4>(const float*, IqGeo, const IqMask*, const float*, const double*, const double*, const uint8_t*, float*, uint32_t*);
template __global__ void k_scan_f32<2, 16, 8, 2>(const float*, IqGeo, const IqMask*, const float*, const double*, const double*, const uint8_t*, float*, uint32_t*);
template __global__ void k_scan_f32<4, 8, 16, 2>(const float*, IqGeo, const IqMask*, const float*, const double*, const double*, const uint8_t*, float*, uint32_t*);

// ---------------------------------------------------------------------------
// Band superset + exact recheck.  T = round_up(fl(fl(1+tol) * U)) where U is
// the least upper bound of the scan; every exact band member has lo <= T.
// Members get their exact FP64 value (AB chain, A2 from the map) and the exact
// minimum is reduced with an integer atomic on the non-negative double bits.
// One bit per candidate is written (ballot words of 64 candidates).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
k_band_superset(const T* __restrict__ img, IqGeo g, const float* __restrict__ lo,
                const uint32_t* __restrict__ gmin_hi, double tol,
                const int32_t* __restrict__ toff, const double* __restrict__ kv, int V,
                const double* __restrict__ b2p, const double* __restrict__ a2map,
                double* __restrict__ dexact, unsigned long long* __restrict__ bitmap,
                unsigned long long* __restrict__ gmin64) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double U = (double)__uint_as_float(*gmin_hi);
  const float Tf = iq_f32_up((1.0 + tol) * U);
  bool sel = false;
  if (c < g.P) sel = lo[c] <= Tf;
  unsigned long long bal = __ballot(sel);
  if ((threadIdx.x & 63) == 0 && (c >> 6) < (g.P + 63) / 64) bitmap[c >> 6] = bal;
  double d = __builtin_inf();
  if (sel) {
    int cx = (int)(c % g.dist[0]), cy = (int)((c / g.dist[0]) % g.dist[1]);
    int cz = (int)(c / ((int64_t)g.dist[0] * g.dist[1]));
    const T* base = img + iq_lin3(g.ti, cx, cy, cz);
    double ab = 0.0;
    for (int j = 0; j < V; j++) ab = __builtin_fma((double)base[toff[j]], kv[j], ab);
    d = __builtin_fabs((a2map[c] - 2.0 * ab) + *b2p);
    dexact[c] = d;
  }
  if (bal) {
    unsigned long long bits = __double_as_longlong(d);
    for (int o = 32; o > 0; o >>= 1) {
      unsigned long long other = __shfl_xor(bits, o);
      bits = other < bits ? other : bits;
    }
    if ((threadIdx.x & 63) == 0) atomicMin(gmin64, bits);
  }
}

template __global__ void k_band_superset<float>(const float*, IqGeo, const float*, const uint32_t*, double, const int32_t*, const double*, int, const double*, const double*, double*, unsigned long long*, unsigned long long*);
template __global__ void k_band_superset<double>(const double*, IqGeo, const float*, const uint32_t*, double, const int32_t*, const double*, int, const double*, const double*, double*, unsigned long long*, unsigned long long*);

// Exact maps where the training image is not exactly representable in FP32:
// every candidate is a superset member.
extern "C" __global__ void __launch_bounds__(256)
k_band_all(const double* __restrict__ d, IqGeo g, unsigned long long* __restrict__ bitmap,
           unsigned long long* __restrict__ gmin64) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool in = c < g.P && !__builtin_isinf(c < g.P ? d[c] : 0.0);
  unsigned long long bal = __ballot(in);
  if ((threadIdx.x & 63) == 0 && (c >> 6) < (g.P + 63) / 64) bitmap[c >> 6] = bal;
  unsigned long long bits = in ? (unsigned long long)__double_as_longlong(d[c]) : 0x7ff0000000000000ull;
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long other = __shfl_xor(bits, o);
    bits = other < bits ? other : bits;
  }
  if ((threadIdx.x & 63) == 0 && bal) atomicMin(gmin64, bits);
}
