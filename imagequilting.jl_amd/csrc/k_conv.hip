// k_conv.hip -- convolutional encoder of the learned proposal: forward and
// backward kernels of three 3x3x3 convolution layers over the training image
// (1 -> 32 -> 32 -> 1 channels, residual output f = x + conv3(...)), the
// window embeddings e_c = W f[window c] and e_q = W q, the ranking loss that
// trains them on (query, exact distance) pairs of the full scan, and Adam.
// No counterpart in the reference (it scans every candidate, iqsim.jl:206).
//
// Maps are channel planes: plane c of a map lives at c * N, N = training-image
// voxels in column-major order.  Tap k = (dx + 1) + 3 (dy + 1) + 9 (dz + 1),
// offsets d_k in {-1, 0, 1}^3; values outside the training image are zero.
// Parameters (float): w1[32][27], b1[32], w2[32][32][27], b2[32], w3[32][27],
// b3 (CV_NW of them), then one head W[64][V] per overlap-mask shape.
#include "iq_internal.h"

#define CV_C 32
#define CV_W1 0
#define CV_B1 (CV_C * 27)
#define CV_W2 (CV_B1 + CV_C)
#define CV_B2 (CV_W2 + CV_C * CV_C * 27)
#define CV_W3 (CV_B2 + CV_C)
#define CV_B3 (CV_W3 + CV_C * 27)
#define CV_NW (CV_B3 + 1)  // 29,441 convolution parameters

struct CvDims {
  int X, Y, Z;
  int64_t N;
};

__device__ inline void cv_coords(const CvDims& d, int64_t p, int& x, int& y, int& z) {
  x = (int)(p % d.X);
  y = (int)((p / d.X) % d.Y);
  z = (int)(p / ((int64_t)d.X * d.Y));
}

// One input plane -> 32 output planes, a work-item per voxel.
//   FWD (conv 1):           out[c][p] = relu(b[c] + sum_k w[c][k] in[p + d_k])
//   BWD (input gradient of conv 3 through the relu of h2):
//                           out[c][p] = [mask[c][p] > 0] sum_k w[c][k] in[p - d_k]
template <bool BWD>
__global__ void __launch_bounds__(256)
k_cv_in(CvDims d, const float* __restrict__ in, const float* __restrict__ w, const float* __restrict__ b,
        const float* __restrict__ mask, float* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= d.N) return;
  int x, y, z;
  cv_coords(d, p, x, y, z);
  float v[27];
#pragma unroll
  for (int k = 0; k < 27; k++) {
    int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
    if (BWD) { dx = -dx; dy = -dy; dz = -dz; }
    const int xx = x + dx, yy = y + dy, zz = z + dz;
    const bool in_ = xx >= 0 && xx < d.X && yy >= 0 && yy < d.Y && zz >= 0 && zz < d.Z;
    v[k] = in_ ? in[xx + (int64_t)d.X * (yy + (int64_t)d.Y * zz)] : 0.0f;
  }
  const auto wc = iq_const(w);
  for (int c = 0; c < CV_C; c++) {
    float acc = BWD ? 0.0f : iq_const(b)[c];
#pragma unroll
    for (int k = 0; k < 27; k++) acc = __builtin_fmaf(wc[c * 27 + k], v[k], acc);
    if (BWD) out[c * d.N + p] = mask[c * d.N + p] > 0.0f ? acc : 0.0f;
    else out[c * d.N + p] = fmaxf(acc, 0.0f);
  }
}

// 32 input planes -> one output plane (conv 3 forward, residual):
//   f[p] = x[p] + b3 + sum_c sum_k w3[c][k] h2[c][p + d_k]
__global__ void __launch_bounds__(256)
k_cv_out(CvDims d, const float* __restrict__ h, const float* __restrict__ w, const float* __restrict__ b,
         const float* __restrict__ x0, float* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= d.N) return;
  int x, y, z;
  cv_coords(d, p, x, y, z);
  int64_t o[27];
  uint32_t ok = 0;
#pragma unroll
  for (int k = 0; k < 27; k++) {
    const int xx = x + k % 3 - 1, yy = y + (k / 3) % 3 - 1, zz = z + k / 9 - 1;
    const bool in_ = xx >= 0 && xx < d.X && yy >= 0 && yy < d.Y && zz >= 0 && zz < d.Z;
    ok |= in_ ? (1u << k) : 0u;
    o[k] = in_ ? xx + (int64_t)d.X * (yy + (int64_t)d.Y * zz) : 0;
  }
  const auto wc = iq_const(w);
  float acc = iq_const(b)[0];
  for (int c = 0; c < CV_C; c++) {
    const float* hc = h + c * d.N;
    float v[27];
#pragma unroll
    for (int k = 0; k < 27; k++) v[k] = hc[o[k]];
#pragma unroll
    for (int k = 0; k < 27; k++) acc = __builtin_fmaf(wc[c * 27 + k], (ok >> k) & 1u ? v[k] : 0.0f, acc);
  }
  out[p] = x0[p] + acc;
}

// Weights of conv 2 rearranged for k_cv_mid: wr[ci][k][co], co fastest.
//   FWD: wr[ci][k][co] = w2[co][ci][k]
//   BWD (input gradient: channels swapped, taps mirrored, d_{26-k} = -d_k):
//        wr[ci][k][co] = w2[ci][co][26 - k]
template <bool BWD>
__global__ void k_cv_wprep(const float* __restrict__ w2, float* __restrict__ wr) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= CV_C * CV_C * 27) return;
  const int co = i % CV_C, k = (i / CV_C) % 27, ci = i / (CV_C * 27);
  wr[i] = BWD ? w2[(ci * CV_C + co) * 27 + 26 - k] : w2[(co * CV_C + ci) * 27 + k];
}

// 32 -> 32 planes through LDS tiles: a group owns 16 x 16 x 2 output voxels
// (a work-item: one (x, y) and two z, 32 channels each), stages eight input
// planes at a time with a one-voxel halo, and takes the weights of each
// (input plane, tap) as 32 uniform (scalar) loads.
//   FWD: out[co][p] = relu(b[co] + sum_ci sum_k wr[ci][k][co] in[ci][p + d_k])
//   BWD: out[co][p] = [mask[co][p] > 0] sum_ci sum_k wr[ci][k][co] in[ci][p + d_k]
#define CVM_TX 16
#define CVM_TY 16
#define CVM_CH 8
template <bool BWD>
__global__ void __launch_bounds__(256)
k_cv_mid(CvDims d, const float* __restrict__ in, const float* __restrict__ wr, const float* __restrict__ b,
         const float* __restrict__ mask, float* __restrict__ out) {
  __shared__ float tile[CVM_CH][4][CVM_TY + 2][CVM_TX + 2];
  const int tx = threadIdx.x % CVM_TX, ty = threadIdx.x / CVM_TX;
  const int nbx = (d.X + CVM_TX - 1) / CVM_TX, nby = (d.Y + CVM_TY - 1) / CVM_TY;
  int bid = blockIdx.x;
  const int bx = bid % nbx;
  bid /= nbx;
  const int by = bid % nby, bz = bid / nby;
  const int x0 = bx * CVM_TX, y0 = by * CVM_TY, z0 = bz * 2;
  float acc0[CV_C], acc1[CV_C];
#pragma unroll
  for (int c = 0; c < CV_C; c++) { acc0[c] = 0.0f; acc1[c] = 0.0f; }
  const auto wg = iq_const(wr);
  for (int c0 = 0; c0 < CV_C; c0 += CVM_CH) {
    __syncthreads();
    for (int i = threadIdx.x; i < CVM_CH * 4 * (CVM_TY + 2) * (CVM_TX + 2); i += 256) {
      const int xx = i % (CVM_TX + 2);
      int r = i / (CVM_TX + 2);
      const int yy = r % (CVM_TY + 2);
      r /= (CVM_TY + 2);
      const int zz = r % 4, cc = r / 4;
      const int gx = x0 + xx - 1, gy = y0 + yy - 1, gz = z0 + zz - 1;
      float v = 0.0f;
      if (gx >= 0 && gx < d.X && gy >= 0 && gy < d.Y && gz >= 0 && gz < d.Z)
        v = in[(int64_t)(c0 + cc) * d.N + gx + (int64_t)d.X * (gy + (int64_t)d.Y * gz)];
      tile[cc][zz][yy][xx] = v;
    }
    __syncthreads();
    for (int cc = 0; cc < CVM_CH; cc++) {
      const int ci = c0 + cc;
#pragma unroll
      for (int dy = 0; dy < 3; dy++) {
#pragma unroll
        for (int dx = 0; dx < 3; dx++) {
          float v[4];
#pragma unroll
          for (int zz = 0; zz < 4; zz++) v[zz] = tile[cc][zz][ty + dy][tx + dx];
#pragma unroll
          for (int dz = 0; dz < 3; dz++) {
            const int k = dx + 3 * dy + 9 * dz;
            const auto wk = wg + (ci * 27 + k) * CV_C;
#pragma unroll
            for (int co = 0; co < CV_C; co++) {
              const float wv = wk[co];
              acc0[co] = __builtin_fmaf(wv, v[dz], acc0[co]);
              acc1[co] = __builtin_fmaf(wv, v[dz + 1], acc1[co]);
            }
          }
        }
      }
    }
  }
  const int x = x0 + tx, y = y0 + ty;
  if (x >= d.X || y >= d.Y) return;
#pragma unroll
  for (int zi = 0; zi < 2; zi++) {
    const int z = z0 + zi;
    if (z >= d.Z) break;
    const int64_t p = x + (int64_t)d.X * (y + (int64_t)d.Y * z);
#pragma unroll
    for (int co = 0; co < CV_C; co++) {
      const float a = zi == 0 ? acc0[co] : acc1[co];
      if (BWD) out[co * d.N + p] = mask[co * d.N + p] > 0.0f ? a : 0.0f;
      else out[co * d.N + p] = fmaxf(a + iq_const(b)[co], 0.0f);
    }
  }
}

// Weight gradient of a convolution out = conv(A, w) given the output
// gradient G: dw[a][b][k] = sum_p G[a][p] A[b][p + d_k], and the bias
// gradient db[a] = sum_p G[a][p], as per-group partial sums (a group walks
// 8 x 8 x 2 voxel tiles staged in LDS; the partials are summed in a fixed
// order by k_cv_wreduce, so the result does not depend on scheduling).
// A work-item owns AG output planes a, one input plane b and a share (one of
// VG) of each tile's voxels.  Partial row of a group: NGA * NGB * 27 weights,
// then NGA biases.
#define CVW_T 8
#define CVW_PL (4 * (CVW_T + 2) * (CVW_T + 2) + 1)  // staged plane stride (odd: conflict free across b)
template <int NGA, int NGB, int AG, int VG>
__global__ void __launch_bounds__(256)
k_cv_wgrad(CvDims d, const float* __restrict__ G, const float* __restrict__ A, float* __restrict__ part) {
  static_assert((NGA / AG) * NGB * VG == 256, "work-item mapping");
  constexpr int SA = NGB * CVW_PL > 256 * 28 ? NGB * CVW_PL : 256 * 28;  // staged planes, then the VG reduction
  __shared__ float sa[SA];
  __shared__ float sg[NGA][CVW_T * CVW_T * 2];
  const int t = threadIdx.x;
  const int b = t % NGB;
  const int ag = (t / NGB) % (NGA / AG);
  const int vg = t / (NGB * (NGA / AG));
  float acc[AG][27];
  float bacc[AG];
#pragma unroll
  for (int i = 0; i < AG; i++) {
    bacc[i] = 0.0f;
#pragma unroll
    for (int k = 0; k < 27; k++) acc[i][k] = 0.0f;
  }
  const int ntx = (d.X + CVW_T - 1) / CVW_T, nty = (d.Y + CVW_T - 1) / CVW_T, ntz = (d.Z + 1) / 2;
  const int64_t ntiles = (int64_t)ntx * nty * ntz;
  for (int64_t tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
    const int x0 = (int)(tl % ntx) * CVW_T, y0 = (int)((tl / ntx) % nty) * CVW_T, z0 = (int)(tl / ((int64_t)ntx * nty)) * 2;
    __syncthreads();
    for (int i = t; i < NGB * 4 * (CVW_T + 2) * (CVW_T + 2); i += 256) {
      const int xx = i % (CVW_T + 2);
      int r = i / (CVW_T + 2);
      const int yy = r % (CVW_T + 2);
      r /= (CVW_T + 2);
      const int zz = r % 4, cb = r / 4;
      const int gx = x0 + xx - 1, gy = y0 + yy - 1, gz = z0 + zz - 1;
      float v = 0.0f;
      if (gx >= 0 && gx < d.X && gy >= 0 && gy < d.Y && gz >= 0 && gz < d.Z)
        v = A[(int64_t)cb * d.N + gx + (int64_t)d.X * (gy + (int64_t)d.Y * gz)];
      sa[cb * CVW_PL + (zz * (CVW_T + 2) + yy) * (CVW_T + 2) + xx] = v;
    }
    for (int i = t; i < NGA * CVW_T * CVW_T * 2; i += 256) {
      const int xx = i % CVW_T, yy = (i / CVW_T) % CVW_T, zz = (i / (CVW_T * CVW_T)) % 2, ca = i / (CVW_T * CVW_T * 2);
      const int gx = x0 + xx, gy = y0 + yy, gz = z0 + zz;
      float v = 0.0f;
      if (gx < d.X && gy < d.Y && gz < d.Z) v = G[(int64_t)ca * d.N + gx + (int64_t)d.X * (gy + (int64_t)d.Y * gz)];
      sg[ca][(zz * CVW_T + yy) * CVW_T + xx] = v;
    }
    __syncthreads();
    const float* ab = sa + b * CVW_PL;
    for (int j = vg; j < CVW_T * CVW_T; j += VG) {  // (x, y) of the tile; both z
      const int xx = j % CVW_T, yy = j / CVW_T;
      float g0[AG], g1[AG];
#pragma unroll
      for (int i = 0; i < AG; i++) {
        g0[i] = sg[ag * AG + i][yy * CVW_T + xx];
        g1[i] = sg[ag * AG + i][CVW_T * CVW_T + yy * CVW_T + xx];
        if (b == 0) bacc[i] += g0[i] + g1[i];
      }
#pragma unroll
      for (int dy = 0; dy < 3; dy++)
#pragma unroll
        for (int dx = 0; dx < 3; dx++) {
          float v[4];
#pragma unroll
          for (int zz = 0; zz < 4; zz++) v[zz] = ab[(zz * (CVW_T + 2) + yy + dy) * (CVW_T + 2) + xx + dx];
#pragma unroll
          for (int dz = 0; dz < 3; dz++)
#pragma unroll
            for (int i = 0; i < AG; i++)
              acc[i][dx + 3 * dy + 9 * dz] =
                  __builtin_fmaf(g1[i], v[dz + 1], __builtin_fmaf(g0[i], v[dz], acc[i][dx + 3 * dy + 9 * dz]));
        }
    }
  }
  // sum the VG shares (LDS reused), then one partial row per group
  float* row = part + (int64_t)blockIdx.x * (NGA * NGB * 27 + NGA);
  __syncthreads();
  float* red = sa;  // VG x (AG * 28) floats per (ag, b) pair
  static_assert(VG == 1 || AG == 1, "the VG reduction holds one output plane per work-item");
  if (VG == 1) {
#pragma unroll
    for (int i = 0; i < AG; i++) {
      const int a = ag * AG + i;
#pragma unroll
      for (int k = 0; k < 27; k++) row[(a * NGB + b) * 27 + k] = acc[i][k];
      if (b == 0) row[NGA * NGB * 27 + a] = bacc[i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < AG; i++) {
#pragma unroll
      for (int k = 0; k < 27; k++) red[t * 28 + i * 28 + k] = acc[i][k];
      red[t * 28 + i * 28 + 27] = bacc[i];
    }
    __syncthreads();
    if (vg == 0) {
#pragma unroll
      for (int i = 0; i < AG; i++) {
        const int a = ag * AG + i;
        for (int k = 0; k < 28; k++) {
          float s = 0.0f;
          for (int u = 0; u < VG; u++) s += red[(t + u * NGB * (NGA / AG)) * 28 + i * 28 + k];
          if (k < 27) row[(a * NGB + b) * 27 + k] = s;
          else if (b == 0) row[NGA * NGB * 27 + a] = s;
        }
      }
    }
  }
}

// grad[j] = sum over groups of part[group][j], accumulated in FP64 in group order.
__global__ void __launch_bounds__(256)
k_cv_wreduce(const float* __restrict__ part, int ngroups, int n, float* __restrict__ grad) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  double s = 0.0;
  for (int gi = 0; gi < ngroups; gi++) s += (double)part[(int64_t)gi * n + j];
  grad[j] = (float)s;
}

// Rows of the embedding GEMMs: windows then queries.  F[r][v] = f[window of
// cands[r] at mask voxel v] for r < nw, the query values Q[r - nw][v] after.
__global__ void __launch_bounds__(256)
k_cv_gather(const IqDev* __restrict__ D, int mask, const float* __restrict__ f, const int32_t* __restrict__ cands,
            int nw, const float* __restrict__ Q, float* __restrict__ F) {
  const IqGeo& g = D->g;
  const IqMask& m = D->masks[mask];
  const int r = blockIdx.x;
  float* row = F + (int64_t)r * m.V;
  if (r >= nw) {
    const float* q = Q + (int64_t)(r - nw) * m.V;
    for (int v = threadIdx.x; v < m.V; v += 256) row[v] = q[v];
    return;
  }
  const int64_t c = cands[r];
  const int cx = (int)(c % g.dist[0]), cy = (int)((c / g.dist[0]) % g.dist[1]);
  const int cz = (int)(c / ((int64_t)g.dist[0] * g.dist[1]));
  const int64_t base = iq_lin3(g.ti, cx, cy, cz);
  const int32_t* toff = D->toff + m.off;
  for (int v = threadIdx.x; v < m.V; v += 256) row[v] = f[base + toff[v]];
}

// Ranking loss of query i (one group each) over its ns candidates
// (E rows i*ns .. i*ns+ns-1, the query's encoding at row nq*ns + i):
//   dd_c = A2_c - 2 e_c.e_q,  p~ = softmax(-(dd - min dd) / tau),
//   p* = softmax(-(D - min D) / tau),  loss_i = KL(p* || p~) / nq_total,
// and the gradients G[c] = dL/de_c = -2 dL/ddd_c e_q,
// G[nq*ns + i] = dL/de_q = -2 sum_c dL/ddd_c e_c  (dL/ddd_c = (p*_c - p~_c) / tau / nq_total).
#define CVL_MAXS 4096
__global__ void __launch_bounds__(256)
k_cv_rankloss(int nq, int ns, const float* __restrict__ E, const double* __restrict__ A2c,
              const double* __restrict__ Dt, const double* __restrict__ tau, double inv_nq, float* __restrict__ G,
              double* __restrict__ lossq) {
  __shared__ float eq[IQ_EMB];
  __shared__ double dd[CVL_MAXS];
  __shared__ double gd[CVL_MAXS];
  __shared__ double red[4][256];
  const int i = blockIdx.x, t = threadIdx.x;
  const float* Ei = E + (int64_t)i * ns * IQ_EMB;
  if (t < IQ_EMB) eq[t] = E[((int64_t)nq * ns + i) * IQ_EMB + t];
  __syncthreads();
  double mn = __builtin_inf(), mt = __builtin_inf();
  for (int c = t; c < ns; c += 256) {
    const float* e = Ei + (int64_t)c * IQ_EMB;
    float s = 0.0f;
    for (int j = 0; j < IQ_EMB; j++) s = __builtin_fmaf(e[j], eq[j], s);
    dd[c] = A2c[(int64_t)i * ns + c] - 2.0 * (double)s;
    mn = fmin(mn, dd[c]);
    mt = fmin(mt, Dt[(int64_t)i * ns + c]);
  }
  red[0][t] = mn;
  red[1][t] = mt;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) { red[0][t] = fmin(red[0][t], red[0][t + o]); red[1][t] = fmin(red[1][t], red[1][t + o]); }
    __syncthreads();
  }
  mn = red[0][0];
  mt = red[1][0];
  const double ti = tau[i];
  __syncthreads();
  double zp = 0.0, zt = 0.0;
  for (int c = t; c < ns; c += 256) {
    zp += exp(-(dd[c] - mn) / ti);
    zt += exp(-(Dt[(int64_t)i * ns + c] - mt) / ti);
  }
  red[0][t] = zp;
  red[1][t] = zt;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) { red[0][t] += red[0][t + o]; red[1][t] += red[1][t + o]; }
    __syncthreads();
  }
  const double lzp = log(red[0][0]), lzt = log(red[1][0]);
  __syncthreads();
  double ls = 0.0;
  for (int c = t; c < ns; c += 256) {
    const double lp = -(dd[c] - mn) / ti - lzp;
    const double lt = -(Dt[(int64_t)i * ns + c] - mt) / ti - lzt;
    const double pt = exp(lt);
    ls += pt * (lt - lp);
    gd[c] = (pt - exp(lp)) / ti * inv_nq;
  }
  red[0][t] = ls;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[0][t] += red[0][t + o];
    __syncthreads();
  }
  if (t == 0) lossq[i] = red[0][0] * inv_nq;
  // dL/de_c
  float* Gi = G + (int64_t)i * ns * IQ_EMB;
  for (int e = t; e < ns * IQ_EMB; e += 256) {
    const int c = e / IQ_EMB, j = e % IQ_EMB;
    Gi[e] = (float)(-2.0 * gd[c] * (double)eq[j]);
  }
  // dL/de_q: four partial sums per component, summed in order
  const int j = t % IQ_EMB, part = t / IQ_EMB;
  double s = 0.0;
  for (int c = part; c < ns; c += 4) s += gd[c] * (double)Ei[(int64_t)c * IQ_EMB + j];
  red[part][j] = s;
  __syncthreads();
  if (t < IQ_EMB) G[((int64_t)nq * ns + i) * IQ_EMB + t] = (float)(-2.0 * ((red[0][t] + red[1][t]) + (red[2][t] + red[3][t])));
}

// Largest |U| (float bits of a non-negative value order as integers).
__global__ void __launch_bounds__(256)
k_cv_absmax(const float* __restrict__ U, int64_t n, unsigned* __restrict__ amax) {
  unsigned m = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    m = max(m, __float_as_uint(__builtin_fabsf(U[i])));
  for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o));
  if ((threadIdx.x & 63) == 0) atomicMax(amax, m);
}

// Fixed-point scale 2^e of the window scatter: |sum| <= nw * amax < 2^62.
__device__ inline double cv_fix_scale(unsigned amax_bits, int nw) {
  const float a = __uint_as_float(amax_bits);
  if (!(a > 0.0f)) return 1.0;
  int ex;
  __builtin_frexp((double)a * (double)nw, &ex);  // a * nw < 2^ex
  return __builtin_ldexp(1.0, 61 - ex);
}

// gf[window of cands[r] at mask voxel v] += U[r][v] as 64-bit fixed point
// (integer sums do not depend on the order of the atomics).
__global__ void __launch_bounds__(256)
k_cv_scatter(const IqDev* __restrict__ D, int mask, const int32_t* __restrict__ cands, const float* __restrict__ U,
             int nw, const unsigned* __restrict__ amax, unsigned long long* __restrict__ acc) {
  const IqGeo& g = D->g;
  const IqMask& m = D->masks[mask];
  const int r = blockIdx.x;
  const int64_t c = cands[r];
  const int cx = (int)(c % g.dist[0]), cy = (int)((c / g.dist[0]) % g.dist[1]);
  const int cz = (int)(c / ((int64_t)g.dist[0] * g.dist[1]));
  const int64_t base = iq_lin3(g.ti, cx, cy, cz);
  const int32_t* toff = D->toff + m.off;
  const double sc = cv_fix_scale(*amax, nw);
  const float* u = U + (int64_t)r * m.V;
  for (int v = threadIdx.x; v < m.V; v += 256) {
    const long long q = __builtin_llrint((double)u[v] * sc);
    atomicAdd(acc + base + toff[v], (unsigned long long)q);
  }
}

__global__ void __launch_bounds__(256)
k_cv_unfix(const unsigned long long* __restrict__ acc, int64_t n, const unsigned* __restrict__ amax, int nw,
           float* __restrict__ out) {
  const double sc = cv_fix_scale(*amax, nw);
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    out[i] = (float)((double)(long long)acc[i] / sc);
}

// Sum of the per-query losses in query order.
__global__ void k_cv_losssum(const double* __restrict__ lossq, int nq, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < nq; i++) s += lossq[i];
  *out = s;
}

// Adam (bias-corrected moments; c1 = 1 - beta1^t, c2 = 1 - beta2^t).
__global__ void __launch_bounds__(256)
k_cv_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
          float lr, float b1, float b2, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (__builtin_sqrtf(vi / c2) + 1e-8f);
}

// ---------------------------------------------------------------------------
// Training pairs from the full scan: per slot, the N smallest lower bounds
// lo32 of the scan (order keys of non-negative floats; +inf marks a
// disabled window) by two 11-bit histogram passes, then every candidate
// whose key is at or below the threshold, listed (at most cap), and the
// exact FP64 distance of each listed or random candidate.
// ---------------------------------------------------------------------------
#define CVT_BINS 2048
// pass 0: bins of key >> 20; pass 1: bins of (key >> 9) & 2047 among keys in bin sel[slot]
template <int PASS>
__global__ void __launch_bounds__(256)
k_cv_tophist(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, int32_t* __restrict__ hist,
             const uint32_t* __restrict__ sel) {
  __shared__ int32_t h[CVT_BINS];
  const IqQuery q = qs[blockIdx.y];
  const int64_t slot = q.slot;
  const int64_t P = D->g.P;
  for (int i = threadIdx.x; i < CVT_BINS; i += 256) h[i] = 0;
  __syncthreads();
  const uint32_t b0 = PASS == 1 ? sel[slot * 4] : 0u;
  const auto lo = iq_glob(D->lo32 + slot * P);
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < P; c += (int64_t)gridDim.x * 256) {
    const uint32_t k = __float_as_uint(lo[c]);
    if (k >= 0x7f800000u) continue;
    if (PASS == 0) atomicAdd(&h[k >> 20], 1);
    else if ((k >> 20) == b0) atomicAdd(&h[(k >> 9) & 2047], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < CVT_BINS; i += 256)
    if (h[i]) atomicAdd(&hist[slot * CVT_BINS + i], h[i]);
}

// sel[slot]: [0] bin of pass 0 where the count reaches N, [1] count below it,
// [2] the key threshold once pass 1 is known.  Grid: one group per query.
template <int PASS>
__global__ void k_cv_topsel(const IqQuery* __restrict__ qs, const int32_t* __restrict__ hist, int N,
                            uint32_t* __restrict__ sel) {
  if (threadIdx.x != 0) return;
  const int64_t slot = qs[blockIdx.x].slot;
  const int32_t* h = hist + slot * CVT_BINS;
  uint32_t* s = sel + slot * 4;
  int64_t run = PASS == 0 ? 0 : (int64_t)s[1];
  uint32_t b = CVT_BINS - 1;
  for (int i = 0; i < CVT_BINS; i++) {
    if (run + h[i] >= N) { b = (uint32_t)i; break; }
    run += h[i];
  }
  if (PASS == 0) { s[0] = b; s[1] = (uint32_t)run; }
  else s[2] = (s[0] << 20) | (b << 9) | 511u;
}

// list[slot][0..cap) = candidates with key <= threshold (order unspecified).
__global__ void __launch_bounds__(256)
k_cv_topcollect(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, const uint32_t* __restrict__ sel,
                int32_t* __restrict__ list, int32_t* __restrict__ cnt, int cap) {
  const IqQuery q = qs[blockIdx.y];
  const int64_t slot = q.slot;
  const int64_t P = D->g.P;
  const uint32_t th = sel[slot * 4 + 2];
  const auto lo = iq_glob(D->lo32 + slot * P);
  const int lane = threadIdx.x & 63;
  for (int64_t c0 = (int64_t)blockIdx.x * 256; c0 < P; c0 += (int64_t)gridDim.x * 256) {
    const int64_t c = c0 + threadIdx.x;
    const uint32_t k = c < P ? __float_as_uint(lo[c]) : 0xffffffffu;
    const bool take = k < 0x7f800000u && k <= th;  // +inf: disabled window
    const unsigned long long bal = __ballot(take);
    if (!bal) continue;
    int base = 0;
    if (lane == 0) base = atomicAdd(&cnt[slot], __popcll(bal));
    base = __shfl(base, 0);
    const int at = base + __popcll(bal & ((1ull << lane) - 1ull));
    if (take && at < cap) list[slot * cap + at] = (int32_t)c;
  }
}

// Exact distance |A2 - 2 AB + B2| (FP64) of candidate cands[i][j] for the
// query in slot slots[i]: a wavefront per pair.  Grid: (ns, nq) of 64.
__global__ void __launch_bounds__(64)
k_cv_exact(const IqDev* __restrict__ D, int mask, const int32_t* __restrict__ slots, const int32_t* __restrict__ cands,
           int ns, double* __restrict__ out) {
  const IqGeo& g = D->g;
  const IqMask& m = D->masks[mask];
  const int i = blockIdx.y, j = blockIdx.x, l = threadIdx.x;
  const int64_t slot = slots[i];
  const int64_t c = cands[(int64_t)i * ns + j];
  const int cx = (int)(c % g.dist[0]), cy = (int)((c / g.dist[0]) % g.dist[1]);
  const int cz = (int)(c / ((int64_t)g.dist[0] * g.dist[1]));
  const int64_t base = iq_lin3(g.ti, cx, cy, cz);
  const int32_t* toff = D->toff + m.off;
  const double* kv = D->kv + slot * D->kv_stride;
  double s = 0.0;
  for (int v = l; v < m.V; v += 64) {
    const double t = D->ti_f32 ? (double)((const float*)D->ti)[base + toff[v]] : ((const double*)D->ti)[base + toff[v]];
    s = __builtin_fma(t, kv[v], s);
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (l == 0) {
    const double a2 = D->a2maps[m.a2][c];
    out[(int64_t)i * ns + j] = __builtin_fabs((a2 - 2.0 * s) + D->b2[slot * (2 + D->naux)]);
  }
}

// A2 of listed candidates (FP64 maps of the mask).
__global__ void __launch_bounds__(256)
k_cv_a2(const IqDev* __restrict__ D, int mask, const int32_t* __restrict__ cands, int64_t n, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = D->a2maps[D->masks[mask].a2][cands[i]];
}

// Band members of each held-out query that its shortlist holds:
// hits += #{b in band[i][0..nband[i]) : the slot's shortlist bitmap has b}.
// Grid: nq groups.
__global__ void __launch_bounds__(256)
k_cv_bandhits(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, const int32_t* __restrict__ band,
              const int32_t* __restrict__ nband, int stride, unsigned long long* __restrict__ hits) {
  const int i = blockIdx.x;
  const int64_t slot = qs[i].slot;
  const unsigned long long* sb = D->lrn.sbits + slot * D->nw64;
  int h = 0;
  for (int k = threadIdx.x; k < nband[i]; k += 256) {
    const int32_t c = band[(int64_t)i * stride + k];
    h += (int)((sb[c >> 6] >> (c & 63)) & 1ull);
  }
  for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o);
  if ((threadIdx.x & 63) == 0 && h) atomicAdd(hits, (unsigned long long)h);
}
