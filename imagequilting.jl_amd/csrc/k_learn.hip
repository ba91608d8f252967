// k_learn.hip -- learned proposal for the distance scan.
//
// Per overlap-mask shape m, a linear encoder e(x) = W_m x maps the masked
// tile values (V_m of them, ascending column-major) to IQ_EMB numbers.  W_m
// is learned from the training image's own windows by block subspace
// iteration (its rows converge to the leading principal directions of the
// masked windows, so e(x).e(q) approximates the correlation term x.q).  The
// bank holds e(x_c) for every candidate c.  A query's approximate distances
// A2[c] - 2 e(x_c).e(q) + B2 select its K best candidates; those go through
// the exact FP64 recheck, so only the shortlist, never a distance, is
// approximate.  Sampled batches also run the full scan, which measures the
// recall of the shortlist on the exact band.
#include "iq_internal.h"

#define LG_T 64  // sgemm tile
#define LG_K 16  // sgemm k-chunk

// Training windows: X[i][v] = TI[window of candidate cands[i]] at the mask's
// voxels.  Grid: n blocks.
__global__ void __launch_bounds__(256)
k_learn_gather(const IqDev* __restrict__ D, int mask, const int32_t* __restrict__ cands, float* __restrict__ X) {
  const IqGeo& g = D->g;
  const IqMask& m = D->masks[mask];
  const int64_t c = cands[blockIdx.x];
  const int cx = (int)(c % g.dist[0]), cy = (int)((c / g.dist[0]) % g.dist[1]);
  const int cz = (int)(c / ((int64_t)g.dist[0] * g.dist[1]));
  const int64_t base = iq_lin3(g.ti, cx, cy, cz);
  const int32_t* toff = D->toff + m.off;
  float* x = X + (int64_t)blockIdx.x * m.V;
  for (int v = threadIdx.x; v < m.V; v += blockDim.x)
    x[v] = D->ti_f32 ? ((const float*)D->ti)[base + toff[v]] : (float)((const double*)D->ti)[base + toff[v]];
}

// C[m][n] = sum_k A(m,k) B(n,k) with A(m,k) = A[m*sam + k*sak], B(n,k) =
// B[n*sbn + k*sbk] (FP32).  64 x 64 tiles, 256 threads of 4 x 4 outputs.
__global__ void __launch_bounds__(256)
k_sgemm(int M, int N, int K, const float* __restrict__ A, int64_t sam, int64_t sak, const float* __restrict__ B,
        int64_t sbn, int64_t sbk, float* __restrict__ C, int64_t ldc) {
  __shared__ float As[LG_K][LG_T + 1], Bs[LG_K][LG_T + 1];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int m0 = blockIdx.y * LG_T, n0 = blockIdx.x * LG_T;
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0.0f;
  for (int k0 = 0; k0 < K; k0 += LG_K) {
    for (int i = threadIdx.x; i < LG_K * LG_T; i += 256) {
      int kk, mm;
      if (sak == 1) { kk = i % LG_K; mm = i / LG_K; } else { kk = i / LG_T; mm = i % LG_T; }
      const int m = m0 + mm, k = k0 + kk;
      As[kk][mm] = (m < M && k < K) ? A[m * sam + k * sak] : 0.0f;
      int kb, nn;
      if (sbk == 1) { kb = i % LG_K; nn = i / LG_K; } else { kb = i / LG_T; nn = i % LG_T; }
      const int n = n0 + nn, k2 = k0 + kb;
      Bs[kb][nn] = (n < N && k2 < K) ? B[n * sbn + k2 * sbk] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < LG_K; kk++) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) { a[i] = As[kk][ty + 16 * i]; b[i] = Bs[kk][tx + 16 * i]; }
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = __builtin_fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int m = m0 + ty + 16 * i, n = n0 + tx + 16 * j;
      if (m < M && n < N) C[(int64_t)m * ldc + n] = acc[i][j];
    }
}

// Bank of one mask: bank[j * P + c] = sum_v W[j][v] * TI[window c at voxel v]
// (one work-item per candidate; W staged through LDS 64 voxels at a time).
__global__ void __launch_bounds__(256)
k_learn_bank(const IqDev* __restrict__ D, int mask, const float* __restrict__ W, float* __restrict__ bank) {
  __shared__ float wc[64][IQ_EMB];
  const IqGeo& g = D->g;
  const IqMask& m = D->masks[mask];
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = c < g.P;
  int64_t base = 0;
  if (live) {
    const int cx = (int)(c % g.dist[0]), cy = (int)((c / g.dist[0]) % g.dist[1]);
    const int cz = (int)(c / ((int64_t)g.dist[0] * g.dist[1]));
    base = iq_lin3(g.ti, cx, cy, cz);
  }
  const int32_t* toff = D->toff + m.off;
  float acc[IQ_EMB];
#pragma unroll
  for (int j = 0; j < IQ_EMB; j++) acc[j] = 0.0f;
  for (int v0 = 0; v0 < m.V; v0 += 64) {
    const int nv = m.V - v0 < 64 ? m.V - v0 : 64;
    for (int i = threadIdx.x; i < 64 * IQ_EMB; i += blockDim.x) {
      const int vv = i % 64, j = i / 64;
      wc[vv][j] = vv < nv ? W[(int64_t)j * m.V + v0 + vv] : 0.0f;
    }
    __syncthreads();
    if (live) {
      for (int vv = 0; vv < nv; vv++) {
        const int64_t o = base + toff[v0 + vv];
        const float t = D->ti_f32 ? ((const float*)D->ti)[o] : (float)((const double*)D->ti)[o];
#pragma unroll
        for (int j = 0; j < IQ_EMB; j++) acc[j] = __builtin_fmaf(wc[vv][j], t, acc[j]);
      }
    }
    __syncthreads();
  }
  if (live)
#pragma unroll
    for (int j = 0; j < IQ_EMB; j++) bank[(int64_t)j * g.P + c] = acc[j];
}

__device__ inline bool learn_applies(const IqDev* D, const IqQuery& q, const IqStep& st) {
  return !q.multi && st.mask >= 0 && D->lrn.wofs[st.mask] >= 0;
}

// Query encodings eq[slot][j] = sum_v W[j][v] * q_v.  Grid: nq blocks of 256.
__global__ void __launch_bounds__(256)
k_learn_embed(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  __shared__ float part[4][IQ_EMB];
  const IqQuery q = qs[blockIdx.x];
  const IqStep& st = D->steps[q.step];
  if (!learn_applies(D, q, st)) return;
  const IqMask& m = D->masks[st.mask];
  const float* W = D->lrn.W + D->lrn.wofs[st.mask];
  const double* kv = D->kv + (int64_t)q.slot * D->kv_stride;
  const int j = threadIdx.x & 63, p = threadIdx.x >> 6;
  float s = 0.0f;
  for (int v = p; v < m.V; v += 4) s = __builtin_fmaf(W[(int64_t)j * m.V + v], (float)kv[v], s);
  part[p][j] = s;
  __syncthreads();
  if (p == 0) D->lrn.eq[(int64_t)q.slot * IQ_EMB + j] = (part[0][j] + part[1][j]) + (part[2][j] + part[3][j]);
}

// Approximate distances of every candidate for a group of LQG queries:
// lsc[slot][c] = A2[c] - 2 e(x_c).e(q) + B2 (+inf for disabled windows).
// Grid: (candidate blocks of 256, query groups).
#define LQG 16
__global__ void __launch_bounds__(256)
k_learn_scores(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, int nq) {
  __shared__ float eqs[LQG][IQ_EMB];
  __shared__ int qm[LQG];
  const IqGeo& g = D->g;
  const int q0 = blockIdx.y * LQG;
  for (int i = threadIdx.x; i < LQG * IQ_EMB; i += blockDim.x) {
    const int k = i / IQ_EMB, j = i % IQ_EMB;
    float v = 0.0f;
    if (q0 + k < nq) v = D->lrn.eq[(int64_t)qs[q0 + k].slot * IQ_EMB + j];
    eqs[k][j] = v;
  }
  if (threadIdx.x < LQG) {
    int mk = -1;
    if (q0 + (int)threadIdx.x < nq) {
      const IqQuery q = qs[q0 + threadIdx.x];
      const IqStep& st = D->steps[q.step];
      if (learn_applies(D, q, st)) mk = st.mask;
    }
    qm[threadIdx.x] = mk;
  }
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.P) return;
  const bool off = D->disabled[c];
  float e[IQ_EMB];
  int cur = -1;
  double a2 = 0.0;
  for (int k = 0; k < LQG && q0 + k < nq; k++) {
    const int mk = qm[k];
    if (mk < 0) continue;
    const IqQuery q = qs[q0 + k];
    float* out = D->lrn.lsc + (int64_t)q.slot * g.P;
    if (off) { out[c] = __builtin_inff(); continue; }
    if (mk != cur) {
      const float* bk = D->lrn.bank + D->lrn.bofs[mk];
#pragma unroll
      for (int j = 0; j < IQ_EMB; j++) e[j] = bk[(int64_t)j * g.P + c];
      a2 = D->a2maps[D->masks[mk].a2][c];
      cur = mk;
    }
    float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f, d3 = 0.0f;
#pragma unroll
    for (int j = 0; j < IQ_EMB; j += 4) {
      d0 = __builtin_fmaf(e[j], eqs[k][j], d0);
      d1 = __builtin_fmaf(e[j + 1], eqs[k][j + 1], d1);
      d2 = __builtin_fmaf(e[j + 2], eqs[k][j + 2], d2);
      d3 = __builtin_fmaf(e[j + 3], eqs[k][j + 3], d3);
    }
    const double b2 = D->b2[(int64_t)q.slot * (2 + D->naux)];
    out[c] = (float)((a2 - 2.0 * (double)((d0 + d1) + (d2 + d3))) + b2);
  }
}

// Shortlist threshold of a query: an order key at or above its K-th
// smallest approximate distance.  A strided sample of LKS keys (sorted in
// LDS) brackets the K-th smallest between the sample minimum and a quantile
// at four times K/P; one pass over the P keys histograms only that bracket
// (2048 bins), so few work-items touch the histogram, and the threshold is
// the upper edge of the bin where the count reaches K (the bracket's top when
// the sample underestimated it).  Grid: nq blocks of 1024.
#define LKS 4096
__global__ void __launch_bounds__(1024)
k_learn_kth(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  __shared__ uint32_t smp[LKS];
  __shared__ int hist[2048];
  const IqQuery q = qs[blockIdx.x];
  const IqStep& st = D->steps[q.step];
  if (!learn_applies(D, q, st)) return;
  const IqGeo& g = D->g;
  const float* s = D->lrn.lsc + (int64_t)q.slot * g.P;
  const int64_t K = D->lrn.K < D->nen ? D->lrn.K : D->nen;
  for (int i = threadIdx.x; i < LKS; i += blockDim.x) smp[i] = iq_f32_key(s[(int64_t)i * g.P / LKS]);
  for (int i = threadIdx.x; i < 2048; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  for (int k = 2; k <= LKS; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < LKS; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const uint32_t a = smp[i], b = smp[ixj];
          if ((a > b) == ((i & k) == 0)) { smp[i] = b; smp[ixj] = a; }
        }
      }
      __syncthreads();
    }
  const uint32_t lo = smp[0];
  int64_t qi = (4 * K * LKS) / g.P + 1;
  const uint32_t hi = qi >= LKS ? 0xffffffffu : smp[qi];
  const uint32_t span = hi - lo;
  int shift = 0;
  while (shift < 32 && (span >> shift) >= 2048u) shift++;
  for (int64_t c = threadIdx.x; c < g.P; c += blockDim.x) {
    const uint32_t k = iq_f32_key(s[c]);
    if (k <= hi) atomicAdd(&hist[k <= lo ? 0 : (k - lo) >> shift], 1);
  }
  __syncthreads();
  // fresh shortlist and band bitmap for this query
  unsigned long long* bm = D->bitmap + (int64_t)q.slot * D->nw64;
  for (int64_t w = threadIdx.x; w < D->nw64; w += blockDim.x) bm[w] = 0ull;
  if (threadIdx.x == 0) {
    D->lrn.nslist[q.slot] = 0;
    int64_t run = 0;
    uint32_t thr = hi;
    for (int b = 0; b < 2048; b++) {
      run += hist[b];
      if (run >= K) {
        const uint64_t top = (uint64_t)lo + (((uint64_t)b + 1) << shift) - 1;
        thr = top < (uint64_t)hi ? (uint32_t)top : hi;
        break;
      }
    }
    D->lrn.thr[q.slot] = thr;
  }
}

// FP32 bounds of the shortlist: one wavefront per member sums a.q over the
// masked voxels in FP32 (lanes take strided voxels, then a shuffle tree) with
// the bound |AB32 - AB| <= gamma_V sum|a||b| <= gamma_V sqrt(A2 B2), giving an
// interval [lo, hi] around the exact FP64 distance as in the full scan's
// epilogue; lo goes to lo32, the least hi to the query's reduction word.
// Grid: (groups, query), 4 wavefronts per group.
__global__ void __launch_bounds__(256)
k_short_bound(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  extern __shared__ float qv[];  // the query values, V floats
  const IqQuery q = qs[blockIdx.y];
  const IqStep& st = D->steps[q.step];
  if (!learn_applies(D, q, st)) return;
  const IqGeo& g = D->g;
  const int64_t slot = q.slot;
  const IqMask& m = D->masks[st.mask];
  const int V = m.V;
  const int32_t* toff = D->toff + m.off;
  const double* kv = D->kv + slot * D->kv_stride;
  for (int v = threadIdx.x; v < V; v += blockDim.x) qv[v] = (float)kv[v];
  __syncthreads();
  const double* a2map = D->a2maps[m.a2];
  const double b2 = D->b2[slot * (2 + D->naux)];
  const double u32 = 0x1p-24, u64 = 0x1p-53;
  const double gerr = 1.02 * ((double)V * u32 / (1.0 - (double)V * u32));
  const double g64 = (double)V * u64 / (1.0 - (double)V * u64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n = D->lrn.nslist[slot];
  const int32_t* sl = D->lrn.slist + slot * g.P;
  float* lo_out = D->lo32 + slot * g.P;
  float lmin = __builtin_inff();
  for (int e = blockIdx.x * 4 + w; e < n; e += gridDim.x * 4) {
    const int64_t c = sl[e];
    const int cx = (int)(c % g.dist[0]), cy = (int)((c / g.dist[0]) % g.dist[1]);
    const int cz = (int)(c / ((int64_t)g.dist[0] * g.dist[1]));
    const int64_t base = iq_lin3(g.ti, cx, cy, cz);
    float acc = 0.0f;
    for (int v = lane; v < V; v += 64) {
      const float t = D->ti_f32 ? ((const float*)D->ti)[base + toff[v]] : (float)((const double*)D->ti)[base + toff[v]];
      acc = __builtin_fmaf(t, qv[v], acc);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    const double a2 = a2map[c];
    const double sq = __builtin_sqrt(a2 * b2) * (1.0 + 0x1p-30);
    const double d = __builtin_fabs((a2 - 2.0 * (double)acc) + b2);
    const double er = (2.0 * (gerr + g64) * sq + 4.0 * u64 * (a2 + b2 + 2.0 * sq) + 1e3 * V * 0x1p-125) * (1.0 + 0x1p-30) +
                      4.0 * u64 * d;
    if (lane == 0) lo_out[c] = iq_f32_down(d - er > 0.0 ? d - er : 0.0);
    lmin = fminf(lmin, iq_f32_up(d + er));
  }
  if (lane == 0 && lmin < __builtin_inff()) atomicMin(&D->gmin_hi[slot], __float_as_uint(lmin));
}

// Superset of the band inside the shortlist: members with lo <= T, T =
// round_up(fl(fl(1+tol) * U)), U the least upper bound; into the superset
// list and the band bitmap.  Grid: (groups, query).
__global__ void __launch_bounds__(256)
k_short_select(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  const IqQuery q = qs[blockIdx.y];
  const IqStep& st = D->steps[q.step];
  if (!learn_applies(D, q, st)) return;
  const IqGeo& g = D->g;
  const int64_t slot = q.slot;
  const double U = (double)__uint_as_float(D->gmin_hi[slot]);
  const float Tf = iq_f32_up((1.0 + D->tol) * U);
  const int n = D->lrn.nslist[slot];
  const int32_t* sl = D->lrn.slist + slot * g.P;
  const int lane = threadIdx.x & 63;
  for (int e0 = blockIdx.x * blockDim.x; e0 < n; e0 += gridDim.x * blockDim.x) {
    const int e = e0 + (int)threadIdx.x;
    int64_t c = 0;
    bool sel = false;
    if (e < n) {
      c = sl[e];
      sel = D->lo32[slot * g.P + c] <= Tf;
    }
    const unsigned long long bal = __ballot(sel);
    if (bal) {
      int base = 0;
      if (lane == 0) base = atomicAdd(&D->nsup[slot], __popcll(bal));
      base = __shfl(base, 0);
      if (sel) {
        D->sup[slot * g.P + base + __popcll(bal & ((1ull << lane) - 1ull))] = (int32_t)c;
        atomicOr(&D->bitmap[slot * D->nw64 + (c >> 6)], 1ull << (c & 63));
      }
    }
  }
}

// Recall of the shortlist on the exact band of a full-scan batch:
// stats += (band members, members inside the shortlist, queries, queries
// with every member inside).  Grid: nq blocks of 256.
__global__ void __launch_bounds__(256)
k_learn_recall(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  __shared__ int hits[4];
  const IqQuery q = qs[blockIdx.x];
  const IqStep& st = D->steps[q.step];
  if (!learn_applies(D, q, st)) return;
  const IqGeo& g = D->g;
  const int64_t slot = q.slot;
  const int64_t nb = D->nband[slot];
  const int32_t* band = D->band + slot * g.P;
  const float* s = D->lrn.lsc + slot * g.P;
  const uint32_t thr = D->lrn.thr[slot];
  int h = 0;
  for (int64_t i = threadIdx.x; i < nb; i += blockDim.x) h += iq_f32_key(s[band[i]]) <= thr;
  for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o);
  if ((threadIdx.x & 63) == 0) hits[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int64_t tot = (int64_t)hits[0] + hits[1] + hits[2] + hits[3];
    atomicAdd(&D->lrn.stats[0], (unsigned long long)nb);
    atomicAdd(&D->lrn.stats[1], (unsigned long long)tot);
    atomicAdd(&D->lrn.stats[2], 1ull);
    if (tot == nb) atomicAdd(&D->lrn.stats[3], 1ull);
  }
}
