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

// Bank of one mask: bank[c * IQ_EMB + j] = bf16(sum_v W[j][v] * X[window c
// at voxel v]) with X the training image, or the convolutional encoder's
// output map `src` (training-image extents) when given (one work-item per
// candidate; W staged through LDS 64 voxels at a time; rows of 64 bf16 are
// the A operand of the MFMA shortlist scan).
__global__ void __launch_bounds__(256)
k_learn_bank(const IqDev* __restrict__ D, int mask, const float* __restrict__ W, const float* __restrict__ src,
             uint16_t* __restrict__ bank) {
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
        const float t = src ? src[o] : D->ti_f32 ? ((const float*)D->ti)[o] : (float)((const double*)D->ti)[o];
#pragma unroll
        for (int j = 0; j < IQ_EMB; j++) acc[j] = __builtin_fmaf(wc[vv][j], t, acc[j]);
      }
    }
    __syncthreads();
  }
  if (live) {
    uint4* row = (uint4*)(bank + c * IQ_EMB);
#pragma unroll
    for (int j = 0; j < IQ_EMB; j += 8) {
      uint4 w;
      w.x = iq_bf16_rne(acc[j]) | (iq_bf16_rne(acc[j + 1]) << 16);
      w.y = iq_bf16_rne(acc[j + 2]) | (iq_bf16_rne(acc[j + 3]) << 16);
      w.z = iq_bf16_rne(acc[j + 4]) | (iq_bf16_rne(acc[j + 5]) << 16);
      w.w = iq_bf16_rne(acc[j + 6]) | (iq_bf16_rne(acc[j + 7]) << 16);
      row[j / 8] = w;
    }
  }
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

__device__ inline float bf16_val(uint32_t b) { return __uint_as_float(b << 16); }

// B fragments of the query tiles (32 queries of one mask each): lane l of
// (tile, k-step ks) holds bf16(e_q[16 ks + 8 (l / 32) + e]), e < 8, of query
// tile.q0 + l % 32.  Grid: ntiles blocks of 64.
__global__ void __launch_bounds__(64)
k_learn_bfrag(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, const int4* __restrict__ tiles) {
  const int4 t = tiles[blockIdx.x];
  const int l = threadIdx.x;
  const float* eq = (l % 32) < t.z ? D->lrn.eq + (int64_t)qs[t.y + l % 32].slot * IQ_EMB : nullptr;
  for (int ks = 0; ks < IQ_EMB / 16; ks++) {
    uint4 w = {0u, 0u, 0u, 0u};
    if (eq) {
      const float* e = eq + 16 * ks + 8 * (l / 32);
      w.x = iq_bf16_rne(e[0]) | (iq_bf16_rne(e[1]) << 16);
      w.y = iq_bf16_rne(e[2]) | (iq_bf16_rne(e[3]) << 16);
      w.z = iq_bf16_rne(e[4]) | (iq_bf16_rne(e[5]) << 16);
      w.w = iq_bf16_rne(e[6]) | (iq_bf16_rne(e[7]) << 16);
    }
    D->lrn.bq[((int64_t)blockIdx.x * (IQ_EMB / 16) + ks) * 64 + l] = w;
  }
}

// Approximate distance of candidate c for a query, in FP32 as the MFMA bank
// scan forms it: A2[c] - 2 sum_j bf16(E[c][j]) bf16(e_q[j]) + B2 (A2 from
// the FP32 copy of the mask's A2 map, NaN for a disabled window).
__device__ inline float learn_score(const IqDev* D, int mask, int64_t c, const float* eq, float b2) {
  const uint16_t* row = D->lrn.bank + D->lrn.bofs[mask] + c * IQ_EMB;
  float acc = 0.0f;
  for (int j = 0; j < IQ_EMB; j++) acc = __builtin_fmaf(bf16_val(row[j]), bf16_val(iq_bf16_rne(eq[j])), acc);
  return __builtin_fmaf(-2.0f, acc, D->a2f[D->masks[mask].a2][c]) + b2;
}

// Per query: a bracket [lo, hi] of order keys around the K-th smallest
// approximate distance from a strided sample of LKS candidates (sorted in
// LDS; hi = the sample quantile at three times K/P), the bin width of a
// 2048-bin histogram over the bracket, fresh per-query state (candidate
// list, shortlist, bitmaps), and the bank scan's per-column word of the
// query at its batch position (FP32 B2, lo, hi, shift).  Grid: nq blocks of
// 1024.
#define LKS 4096
__global__ void __launch_bounds__(1024)
k_learn_bracket(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  __shared__ uint32_t smp[LKS];
  __shared__ unsigned long long best;  // (key, sample) of the sample minimum
  const IqQuery q = qs[blockIdx.x];
  const IqStep& st = D->steps[q.step];
  if (!learn_applies(D, q, st)) return;
  const IqGeo& g = D->g;
  const int64_t slot = q.slot;
  const float* eq = D->lrn.eq + slot * IQ_EMB;
  const float b2 = (float)D->b2[slot * (2 + D->naux)];
  const int64_t K = D->lrn.K < D->nen ? D->lrn.K : D->nen;
  if (threadIdx.x == 0) best = ~0ull;
  __syncthreads();
  for (int i = threadIdx.x; i < LKS; i += blockDim.x) {
    const int64_t c = (int64_t)i * g.P / LKS;
    smp[i] = D->disabled[c] ? 0xffffffffu : iq_f32_key(learn_score(D, st.mask, c, eq, b2));
    atomicMin(&best, ((unsigned long long)smp[i] << 32) | (unsigned)i);
  }
  unsigned long long* bm = D->bitmap + slot * D->nw64;
  unsigned long long* sb = D->lrn.sbits + slot * D->nw64;
  for (int64_t w = threadIdx.x; w < D->nw64; w += blockDim.x) { bm[w] = 0ull; sb[w] = 0ull; }
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
  if (threadIdx.x == 0) {
    const uint32_t lo = smp[0];
    const int64_t qi = (3 * K * LKS) / g.P + 1;
    const uint32_t hi = qi >= LKS ? 0xfffffffeu : smp[qi];
    const uint32_t span = hi - lo;
    uint32_t shift = 0;
    while (shift < 32 && (span >> shift) >= 2048u) shift++;
    uint32_t* b = D->lrn.brk + slot * 4;
    b[0] = lo; b[1] = hi; b[2] = shift;
    D->lrn.binfo[blockIdx.x] = make_uint4(__float_as_uint(b2), lo, hi, shift);
    // seed of the shortlist should no candidate fall inside the bracket: the
    // sample minimum, or the first enabled candidate when every sampled
    // window is disabled
    b[3] = (uint32_t)((best >> 32) != 0xffffffffull ? (int64_t)(best & 0xffffffffull) * g.P / LKS
                                                      : (int64_t)D->enabled[0]);
    D->lrn.lcnt[slot] = 0;
    D->lrn.nslist[slot] = 0;
  }
}

// Shortlist scan on MFMA (v_mfma_f32_32x32x16_bf16): approximate distances
// of 32 candidates x 32 queries of a tile per wavefront from the bf16 bank
// rows (A) and the tile's B fragments, in FP32 (A2 from the mask's FP32 map,
// whose NaN marks a disabled window); the few inside a query's bracket are
// appended to its candidate list with their order keys (one reservation per
// lane and hit set; the histogram is built from the list afterwards).  Each
// wavefront owns 32 consecutive candidates and walks all tiles (tiles of one
// mask are adjacent, so bank rows load once per mask); a column's query word
// (B2, bracket) is one load at its batch position.
// Grid: ceil(P / 128) blocks of 4 wavefronts.
__global__ void __launch_bounds__(256)
k_learn_mscan(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, const int4* __restrict__ tiles, int ntiles) {
  const IqGeo& g = D->g;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t c0 = (int64_t)blockIdx.x * 128 + w * 32;
  const int kb = lane >> 5;
  iq_bf16x8 a[IQ_EMB / 16];
  float a2[16];
  uint32_t offm = 0;  // candidate rows of this lane that are disabled or past P
  int cur = -1;
  const int64_t LC = D->lrn.lcap;
  for (int t = 0; t < ntiles; t++) {
    const int4 tl = tiles[t];
    if (tl.x != cur) {
      cur = tl.x;
      const uint16_t* bank = D->lrn.bank + D->lrn.bofs[cur];
      const int64_t ca = c0 + (lane & 31);
#pragma unroll
      for (int ks = 0; ks < IQ_EMB / 16; ks++) {
        uint4 v = {0u, 0u, 0u, 0u};
        if (ca < g.P) v = *(const uint4*)(bank + ca * IQ_EMB + 16 * ks + 8 * kb);
        a[ks] = __builtin_bit_cast(iq_bf16x8, v);
      }
      // A2 of the lane's 16 candidate rows (all 16 loads in flight) and
      // whether each is live
      const auto a2f = iq_glob(D->a2f[D->masks[cur].a2]);
      float fl[16];
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int64_t c = c0 + 8 * (v >> 2) + 4 * kb + (v & 3);
        fl[v] = a2f[c < g.P ? c : g.P - 1];
      }
      offm = 0;
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int64_t c = c0 + 8 * (v >> 2) + 4 * kb + (v & 3);
        const bool ok = c < g.P && fl[v] == fl[v];
        a2[v] = ok ? fl[v] : 0.0f;
        offm |= ok ? 0u : (1u << v);
      }
    }
    // the column's query word is fetched before the MFMAs run
    const int col = lane & 31;
    const bool qlive = col < tl.z;
    const uint4 bi = D->lrn.binfo[tl.y + (qlive ? col : 0)];
    const float b2 = __uint_as_float(bi.x);
    const uint32_t hi = bi.z;
    iq_f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; v++) acc[v] = 0.0f;
    const uint4* bq = D->lrn.bq + (int64_t)t * (IQ_EMB / 16) * 64 + lane;
#pragma unroll
    for (int ks = 0; ks < IQ_EMB / 16; ks++)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ks], __builtin_bit_cast(iq_bf16x8, bq[ks * 64]), acc, 0, 0, 0);
    if (!qlive) continue;
    uint32_t keys[16];
    uint32_t in = 0;
#pragma unroll
    for (int v = 0; v < 16; v++) {
      keys[v] = iq_f32_key(__builtin_fmaf(-2.0f, acc[v], a2[v]) + b2);
      in |= (!(offm & (1u << v)) && keys[v] <= hi) ? (1u << v) : 0u;
    }
    if (!in) continue;
    const int64_t slot = qs[tl.y + col].slot;
    int idx = atomicAdd(&D->lrn.lcnt[slot], __builtin_popcount(in));  // one reservation per lane
#pragma unroll
    for (int v = 0; v < 16; v++) {
      if (!(in & (1u << v))) continue;
      if (idx < LC) D->lrn.list[slot * LC + idx] = make_int2((int)(c0 + 8 * (v >> 2) + 4 * kb + (v & 3)), (int)keys[v]);
      idx++;
    }
  }
}

// Per query: a 2048-bin histogram (in LDS) of the listed order keys over the
// bracket, the threshold key where its count reaches K (the bracket's top
// when the sample underestimated it), then the shortlist: the listed
// candidates at or below it (list and bitmap).  Grid: nq blocks of 1024.
__global__ void __launch_bounds__(1024)
k_learn_pick(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  __shared__ uint32_t thr;
  __shared__ int32_t hist[2048];
  const IqQuery q = qs[blockIdx.x];
  const IqStep& st = D->steps[q.step];
  if (!learn_applies(D, q, st)) return;
  const IqGeo& g = D->g;
  const int64_t slot = q.slot;
  const int64_t K = D->lrn.K < D->nen ? D->lrn.K : D->nen;
  const int64_t LC = D->lrn.lcap;
  const int64_t n = D->lrn.lcnt[slot] < LC ? D->lrn.lcnt[slot] : LC;
  const int2* list = D->lrn.list + slot * LC;
  {
    const uint32_t* br = D->lrn.brk + slot * 4;
    const uint32_t lo = br[0], shift = br[2];
    for (int i = threadIdx.x; i < 2048; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
      const uint32_t k = (uint32_t)list[e].y;
      atomicAdd(&hist[k <= lo ? 0 : (k - lo) >> shift], 1);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const uint32_t* br = D->lrn.brk + slot * 4;
    const uint32_t lo = br[0], hi = br[1], shift = br[2];
    int64_t run = 0;
    uint32_t th = hi;
    for (int b = 0; b < 2048; b++) {
      run += hist[b];
      if (run >= K) {
        const uint64_t top = (uint64_t)lo + (((uint64_t)b + 1) << shift) - 1;
        th = top < (uint64_t)hi ? (uint32_t)top : hi;
        break;
      }
    }
    thr = th;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (int64_t e0 = 0; e0 < n; e0 += blockDim.x) {
    const int64_t e = e0 + threadIdx.x;
    int2 it = make_int2(0, 0);
    bool sel = false;
    if (e < n) {
      it = list[e];
      sel = (uint32_t)it.y <= thr;
    }
    const unsigned long long bal = __ballot(sel);
    if (bal) {
      int base = 0;
      if (lane == 0) base = atomicAdd(&D->lrn.nslist[slot], __popcll(bal));
      base = __shfl(base, 0);
      if (sel) {
        D->lrn.slist[slot * g.P + base + __popcll(bal & ((1ull << lane) - 1ull))] = it.x;
        atomicOr(&D->lrn.sbits[slot * D->nw64 + (it.x >> 6)], 1ull << (it.x & 63));
      }
    }
  }
  // an empty shortlist (no approximate distance fell inside the bracket,
  // e.g. ties resolved one ulp apart) takes the seed: the exact recheck and
  // the pick then run on it instead of finding no candidate at all
  __syncthreads();
  if (threadIdx.x == 0 && D->lrn.nslist[slot] == 0) {
    const int64_t c = D->lrn.brk[slot * 4 + 3];
    D->lrn.slist[slot * g.P] = (int32_t)c;
    D->lrn.nslist[slot] = 1;
    atomicOr(&D->lrn.sbits[slot * D->nw64 + (c >> 6)], 1ull << (c & 63));
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
  extern __shared__ float qv[];  // the query values (V floats), then the window offsets (V ints)
  const IqQuery q = qs[blockIdx.y];
  const IqStep& st = D->steps[q.step];
  if (!learn_applies(D, q, st)) return;
  const IqGeo& g = D->g;
  const int64_t slot = q.slot;
  const IqMask& m = D->masks[st.mask];
  const int V = m.V;
  const int32_t* toff = D->toff + m.off;
  const double* kv = D->kv + slot * D->kv_stride;
  int32_t* tl = (int32_t*)(qv + V);
  for (int v = threadIdx.x; v < V; v += blockDim.x) {
    qv[v] = (float)kv[v];
    tl[v] = toff[v];
  }
  __syncthreads();
  const bool f32 = D->ti_f32;
  const auto ti32 = iq_glob((const float*)D->ti);
  const auto ti64 = iq_glob((const double*)D->ti);
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
    // each lane sums its strided voxels in ascending order, sixteen window
    // values in flight at a time
    float acc = 0.0f;
    int v = lane;
    if (f32) {
      for (; v + 64 * 15 < V; v += 64 * 16) {
        float t[16];
#pragma unroll
        for (int u = 0; u < 16; u++) t[u] = ti32[base + tl[v + 64 * u]];
#pragma unroll
        for (int u = 0; u < 16; u++) acc = __builtin_fmaf(t[u], qv[v + 64 * u], acc);
      }
      for (; v < V; v += 64) acc = __builtin_fmaf(ti32[base + tl[v]], qv[v], acc);
    } else {
      for (; v + 64 * 15 < V; v += 64 * 16) {
        float t[16];
#pragma unroll
        for (int u = 0; u < 16; u++) t[u] = (float)ti64[base + tl[v + 64 * u]];
#pragma unroll
        for (int u = 0; u < 16; u++) acc = __builtin_fmaf(t[u], qv[v + 64 * u], acc);
      }
      for (; v < V; v += 64) acc = __builtin_fmaf((float)ti64[base + tl[v]], qv[v], acc);
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
  const unsigned long long* sb = D->lrn.sbits + slot * D->nw64;
  int h = 0;
  for (int64_t i = threadIdx.x; i < nb; i += blockDim.x) h += (int)((sb[band[i] >> 6] >> (band[i] & 63)) & 1ull);
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
