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

// Query encodings eq[slot][j] = sum_v W[j][v] * q_v (from the transposed
// heads: a wavefront's loads of one voxel's 64 weights are contiguous).
// Sixteen wavefronts per query take every sixteenth voxel; their partial
// sums are added in a fixed order.  Grid: nq blocks of 1024.
__global__ void __launch_bounds__(1024)
k_learn_embed(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  __shared__ float part[16][IQ_EMB];
  const IqQuery q = qs[blockIdx.x];
  const IqStep& st = D->steps[q.step];
  if (!learn_applies(D, q, st)) return;
  const IqMask& m = D->masks[st.mask];
  const float* WT = D->lrn.WT + D->lrn.wofs[st.mask];  // V x 64: a voxel's 64 weights side by side
  const double* kv = D->kv + (int64_t)q.slot * D->kv_stride;
  const int j = threadIdx.x & 63, p = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const auto kvs = iq_const(kv);  // the same voxel for the whole wavefront: scalar loads
  float s = 0.0f;
  int v = p;
  for (; v + 16 * 7 < m.V; v += 16 * 8) {  // eight voxels' loads in flight
    float wv[8];
    double qv[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      wv[u] = WT[(int64_t)(v + 16 * u) * IQ_EMB + j];
      qv[u] = kvs[v + 16 * u];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) s = __builtin_fmaf(wv[u], (float)qv[u], s);
  }
  for (; v < m.V; v += 16) s = __builtin_fmaf(WT[(int64_t)v * IQ_EMB + j], (float)kvs[v], s);
  part[threadIdx.x >> 6][j] = s;
  __syncthreads();
  if (p == 0) {
    float t = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) t += part[i][j];
    D->lrn.eq[(int64_t)q.slot * IQ_EMB + j] = t;
  }
}

// WT[v][j] = W[j][v] for one mask (64 x V heads, row-major).  Grid: V blocks of 64.
__global__ void __launch_bounds__(64)
k_learn_transpose(const float* __restrict__ W, int V, float* __restrict__ WT) {
  WT[(int64_t)blockIdx.x * IQ_EMB + threadIdx.x] = W[(int64_t)threadIdx.x * V + blockIdx.x];
}

__device__ inline float bf16_val(uint32_t b) { return __uint_as_float(b << 16); }

// B fragments of the query tiles (32 queries of one mask each): lane l of
// (tile, k-step ks) holds bf16(-2 e_q[16 ks + 8 (l / 32) + e]), e < 8, of
// query tile.q0 + l % 32 (the factor -2 of the score is exact in bf16).
// Grid: ntiles blocks of 64.
__global__ void __launch_bounds__(64)
k_learn_bfrag(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, const int4* __restrict__ tiles) {
  const int4 t = tiles[blockIdx.x];
  const int l = threadIdx.x;
  const float* eq = (l % 32) < t.z ? D->lrn.eq + (int64_t)qs[t.y + l % 32].slot * IQ_EMB : nullptr;
  for (int ks = 0; ks < IQ_EMB / 16; ks++) {
    uint4 w = {0u, 0u, 0u, 0u};
    if (eq) {
      const float* e = eq + 16 * ks + 8 * (l / 32);
      w.x = iq_bf16_rne(-2.0f * e[0]) | (iq_bf16_rne(-2.0f * e[1]) << 16);
      w.y = iq_bf16_rne(-2.0f * e[2]) | (iq_bf16_rne(-2.0f * e[3]) << 16);
      w.z = iq_bf16_rne(-2.0f * e[4]) | (iq_bf16_rne(-2.0f * e[5]) << 16);
      w.w = iq_bf16_rne(-2.0f * e[6]) | (iq_bf16_rne(-2.0f * e[7]) << 16);
    }
    D->lrn.bq[((int64_t)blockIdx.x * (IQ_EMB / 16) + ks) * 64 + l] = w;
  }
}

// Approximate score of candidate c for a query, in FP32: A2[c] + sum_j
// bf16(E[c][j]) bf16(-2 e_q[j]) (A2 from the FP32 copy of the mask's A2 map,
// NaN for a disabled window), the terms added in order of j.  The exact
// distance is A2 - 2 AB + B2: B2 is the same for every candidate of a
// query, so it is left out of the order.

// the float whose order key is k (iq_f32_key's inverse on non-NaN keys)
__device__ inline float key_float(uint32_t k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k); }

// Per query: a bracket [lo, hi] of order keys around the K-th smallest
// approximate distance from a strided sample of LKS candidates (lo = the
// sample minimum, hi = the sample quantile at three times K/P, selected in
// LDS), the bin width of a
// 2048-bin histogram over the bracket, fresh per-query state (candidate
// list, shortlist, bitmaps), and the bank scan's per-column word of the
// query at its batch position (the bracket's top as a float, lo, hi,
// shift).  Grid: nq blocks of 1024.
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
  const int64_t K = D->lrn.K < D->nen ? D->lrn.K : D->nen;
  // the query's bf16(-2 e_q) as floats, once per block
  __shared__ float qe[IQ_EMB];
  if (threadIdx.x < IQ_EMB) qe[threadIdx.x] = bf16_val(iq_bf16_rne(-2.0f * eq[threadIdx.x]));
  if (threadIdx.x == 0) best = ~0ull;
  __syncthreads();
  const uint16_t* bank = D->lrn.bank + D->lrn.bofs[st.mask];
  const float* a2f = D->a2f[D->masks[st.mask].a2];
  unsigned long long mine = ~0ull;
  for (int i = threadIdx.x; i < LKS; i += blockDim.x) {
    const int64_t c = (int64_t)i * g.P / LKS;
    // the approximate score, the row in eight 16-byte loads
    const auto row = iq_glob((const iq_u32x4*)(bank + c * IQ_EMB));
    iq_u32x4 rw[IQ_EMB / 8];
#pragma unroll
    for (int k = 0; k < IQ_EMB / 8; k++) rw[k] = row[k];
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < IQ_EMB / 8; k++) {
      const uint32_t wv[4] = {rw[k].x, rw[k].y, rw[k].z, rw[k].w};
#pragma unroll
      for (int h = 0; h < 4; h++) {
        acc = __builtin_fmaf(bf16_val(wv[h] & 0xffffu), qe[8 * k + 2 * h], acc);
        acc = __builtin_fmaf(bf16_val(wv[h] >> 16), qe[8 * k + 2 * h + 1], acc);
      }
    }
    const float sc = acc + a2f[c];
    smp[i] = D->disabled[c] ? 0xffffffffu : iq_f32_key(sc);
    const unsigned long long kv = ((unsigned long long)smp[i] << 32) | (unsigned)i;
    mine = kv < mine ? kv : mine;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long other = __shfl_xor(mine, o);
    mine = other < mine ? other : mine;
  }
  if ((threadIdx.x & 63) == 0) atomicMin(&best, mine);  // one LDS atomic per wavefront
  unsigned long long* bm = D->bitmap + slot * D->nw64;
  unsigned long long* sb = D->lrn.sbits + slot * D->nw64;
  for (int64_t w = threadIdx.x; w < D->nw64; w += blockDim.x) { bm[w] = 0ull; sb[w] = 0ull; }
  // the sample's qi-th smallest key (0-based) by radix selection, eight bits
  // a pass from the top: a histogram of the keys that share the digits fixed
  // so far, then the bin that holds the remaining rank
  const int64_t qi = (3 * K * LKS) / g.P + 1;
  __shared__ uint32_t rhist[256];
  __shared__ uint32_t rsel[2];  // the digits fixed so far, the rank left inside them
  if (threadIdx.x == 0) { rsel[0] = 0u; rsel[1] = (uint32_t)(qi < LKS ? qi : 0); }
  if (qi < LKS) {
    for (int sh = 24; sh >= 0; sh -= 8) {
      for (int i = threadIdx.x; i < 256; i += blockDim.x) rhist[i] = 0u;
      __syncthreads();
      const uint32_t pre = rsel[0];
      const uint32_t pmask = sh == 24 ? 0u : (0xffffffffu << (sh + 8));
      for (int i = threadIdx.x; i < LKS; i += blockDim.x) {
        const uint32_t k = smp[i];
        if ((k & pmask) == pre) atomicAdd(&rhist[(k >> sh) & 255u], 1u);
      }
      __syncthreads();
      if (threadIdx.x < 64) {  // one wavefront: four bins a lane, a wavefront prefix sum
        const int l = threadIdx.x;
        uint32_t c[4];
#pragma unroll
        for (int u = 0; u < 4; u++) c[u] = rhist[4 * l + u];
        const uint32_t sum = c[0] + c[1] + c[2] + c[3];
        uint32_t inc = sum;
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t y = __shfl_up(inc, o);
          if (l >= o) inc += y;
        }
        const uint32_t r = rsel[1];
        uint32_t acc = inc - sum;
        if (acc <= r && r < inc) {  // one lane
          for (int u = 0; u < 4; u++) {
            if (r < acc + c[u]) {
              rsel[0] = pre | ((uint32_t)(4 * l + u) << sh);
              rsel[1] = r - acc;
              break;
            }
            acc += c[u];
          }
        }
      }
      __syncthreads();
    }
  } else {
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const uint32_t lo = (uint32_t)(best >> 32);
    const uint32_t hi = qi >= LKS ? 0xfffffffeu : rsel[0];
    const uint32_t span = hi - lo;
    uint32_t shift = 0;
    while (shift < 32 && (span >> shift) >= 2048u) shift++;
    uint32_t* b = D->lrn.brk + slot * 4;
    b[0] = lo; b[1] = hi; b[2] = shift;
    // every key from that of +inf up (NaN keys) admits every non-NaN score
    const float hif = hi >= 0xff800000u ? __builtin_inff() : key_float(hi);
    D->lrn.binfo[blockIdx.x] = make_uint4(__float_as_uint(hif), lo, hi, shift);
    // seed of the shortlist should no candidate fall inside the bracket: the
    // sample minimum, or the first enabled candidate when every sampled
    // window is disabled
    b[3] = (uint32_t)((best >> 32) != 0xffffffffull ? (int64_t)(best & 0xffffffffull) * g.P / LKS
                                                      : (int64_t)D->enabled[0]);
    D->lrn.lcnt[slot] = 0;
    D->lrn.nslist[slot] = 0;
  }
}

// Shortlist scan on MFMA (v_mfma_f32_32x32x16_bf16): approximate scores of
// 32 candidates x 32 queries of a tile per MFMA chain, in FP32: the chain
// starts from the candidates' A2 (the FP32 map; NaN for a disabled window or
// a pad row past P, which no comparison admits) and adds bf16 bank rows times
// the tile's B fragments (-2 e_q).  The few scores inside a query's bracket
// (score <= its top as a float) go to its candidate list with their order
// keys, in two steps: during the scan a block stages them in its own region
// of global memory (the rank within the column and the staging slot from LDS
// counters), then reserves each column's share of the query's list with one
// global atomic and copies them there (entries beyond the staging region go
// to the list directly).  Global atomics on a few hundred list counters from
// every wavefront of the chip were this kernel's bound: 668 us per launch
// against 168 us for the scan alone at the 2.9M-candidate workload.
// Persistent blocks of 8 wavefronts stage up to LMS_TMAX tiles' B
// fragments, bracket tops and query slots in LDS once, then each wavefront
// walks chunks of 32 consecutive candidates, strided over the grid; tiles of
// one mask are adjacent, so a chunk's bank rows and A2 load once per mask.
// Grid: the blocks resident at once (at most LMS_BLOCKS_PER_CU per CU).
#define LMS_TMAX 8
#define LMS_WAVES 8
#define LMS_BLOCKS_PER_CU 4
#define LMS_STAGE 4096
__global__ void __launch_bounds__(64 * LMS_WAVES) __attribute__((amdgpu_waves_per_eu(6)))
k_learn_mscan(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, const int4* __restrict__ tiles, int ntiles) {
  __shared__ uint4 sbq[LMS_TMAX * (IQ_EMB / 16) * 64];
  __shared__ float shi[LMS_TMAX * 32];
  __shared__ int64_t sslot[LMS_TMAX * 32];
  __shared__ int scnt[LMS_TMAX * 32];   // entries staged per column (ranks)
  __shared__ int sbase[LMS_TMAX * 32];  // the column's reserved start in the query's list
  __shared__ int snent;                 // entries staged by the block
  __shared__ int smask[LMS_TMAX];
  __shared__ const uint16_t* sbank[LMS_TMAX];
  __shared__ const float* sa2[LMS_TMAX];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // the chunk loop runs on scalar registers
  const int kb = lane >> 5, col = lane & 31;
  const int64_t LC = D->lrn.lcap;
  const int SC = D->lrn.stage_cap;
  int4* stage = D->lrn.stage + (int64_t)blockIdx.x * SC;
  const int64_t nchunk = (D->g.P + 31) / 32;
  const float qnan = __builtin_nanf("");
  for (int t0 = 0; t0 < ntiles; t0 += LMS_TMAX) {
    const int nt = ntiles - t0 < LMS_TMAX ? ntiles - t0 : LMS_TMAX;
    __syncthreads();  // the previous group's tiles and staged entries are no longer read
    for (int i = threadIdx.x; i < nt * (IQ_EMB / 16) * 64; i += blockDim.x)
      sbq[i] = D->lrn.bq[(int64_t)t0 * (IQ_EMB / 16) * 64 + i];
    for (int i = threadIdx.x; i < nt * 32; i += blockDim.x) {
      const int4 tl = tiles[t0 + i / 32];
      const bool live = i % 32 < tl.z;
      shi[i] = live ? __uint_as_float(D->lrn.binfo[tl.y + i % 32].x) : qnan;
      sslot[i] = live ? qs[tl.y + i % 32].slot : 0;
      scnt[i] = 0;
    }
    if (threadIdx.x < nt) {
      const int mk = tiles[t0 + threadIdx.x].x;
      smask[threadIdx.x] = mk;
      sbank[threadIdx.x] = D->lrn.bank + D->lrn.bofs[mk];
      sa2[threadIdx.x] = D->a2f[D->masks[mk].a2];
    }
    if (threadIdx.x == 0) snent = 0;
    __syncthreads();
    for (int64_t ch = (int64_t)blockIdx.x * LMS_WAVES + w; ch < nchunk; ch += (int64_t)gridDim.x * LMS_WAVES) {
      const int64_t c0 = ch * 32;
      iq_bf16x8 a[IQ_EMB / 16];
      iq_f32x16 a2;
      int cur = -1;
      for (int t = 0; t < nt; t++) {
        if (smask[t] != cur) {
          cur = smask[t];
          // bank rows and A2 map are padded to whole chunks (pad A2 NaN)
          const auto bank = iq_glob((const iq_u32x4*)(sbank[t] + (c0 + col) * IQ_EMB + 8 * kb));
#pragma unroll
          for (int ks = 0; ks < IQ_EMB / 16; ks++) a[ks] = __builtin_bit_cast(iq_bf16x8, bank[2 * ks]);
          // rows 8 i + 4 kb .. + 3 of the chunk: four 16-byte loads (the map
          // is 16-byte aligned, c0 a multiple of 32)
          const auto a2f = iq_glob((const iq_u32x4*)(sa2[t] + c0 + 4 * kb));
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const iq_u32x4 f = a2f[2 * i];
            a2[4 * i] = __uint_as_float(f.x);
            a2[4 * i + 1] = __uint_as_float(f.y);
            a2[4 * i + 2] = __uint_as_float(f.z);
            a2[4 * i + 3] = __uint_as_float(f.w);
          }
        }
        const float hif = shi[t * 32 + col];
        const uint4* bq = sbq + t * (IQ_EMB / 16) * 64 + lane;
        iq_f32x16 acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], __builtin_bit_cast(iq_bf16x8, bq[0]), a2, 0, 0, 0);
#pragma unroll
        for (int ks = 1; ks < IQ_EMB / 16; ks++)
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ks], __builtin_bit_cast(iq_bf16x8, bq[ks * 64]), acc, 0, 0, 0);
        uint32_t in = 0;
#pragma unroll
        for (int v = 0; v < 16; v++) in |= (acc[v] <= hif) ? (1u << v) : 0u;
        if (in) {
          const int cc = t * 32 + col;
          const int nh = __builtin_popcount(in);
          const int e0 = atomicAdd(&snent, nh);
          const int nfit = e0 >= SC ? 0 : (SC - e0 < nh ? SC - e0 : nh);
          const int r0 = nfit ? atomicAdd(&scnt[cc], nfit) : 0;
          int didx = 0;
          if (nfit < nh) didx = atomicAdd(&D->lrn.lcnt[sslot[cc]], nh - nfit);  // past the staging region
          // the row's index is formed here (the empty asm keeps the
          // compiler from holding all sixteen across the loop)
          int cand = (int)c0 + 4 * kb;
          asm volatile("" : "+v"(cand));
          int k = 0;
#pragma unroll
          for (int v = 0; v < 16; v++) {
            if (!(in & (1u << v))) continue;
            const int cv = cand + 8 * (v >> 2) + (v & 3);
            const int key = (int)iq_f32_key(acc[v]);
            if (k < nfit) {
              stage[e0 + k] = make_int4(cc, r0 + k, cv, key);
            } else {
              const int idx = didx + (k - nfit);
              if (idx < LC) D->lrn.list[sslot[cc] * LC + idx] = make_int2(cv, key);
            }
            k++;
          }
        }
      }
    }
    // each column's staged entries: one reservation in its query's list
    __syncthreads();
    for (int i = threadIdx.x; i < nt * 32; i += blockDim.x)
      if (scnt[i] > 0) sbase[i] = atomicAdd(&D->lrn.lcnt[sslot[i]], scnt[i]);
    __syncthreads();
    const int ne = snent < SC ? snent : SC;
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
      const int4 it = stage[e];
      const int idx = sbase[it.x] + it.y;
      if (idx < LC) D->lrn.list[sslot[it.x] * LC + idx] = make_int2(it.z, it.w);
    }
  }
}

// Exclusive prefix sum of one int per work-item over a block of 1024
// (wavefront scans, then the 16 wavefront totals); ends with a barrier.
__device__ inline int block_exclusive_sum_1024(int x, int* wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(inc, o);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int before = 0;
  for (int i = 0; i < w; i++) before += wsum[i];
  __syncthreads();
  return before + inc - x;
}

#define LPK_SORT 4096
#define LPK_UNROLL 4
// Per query: a 2048-bin histogram (in LDS) of the listed order keys over the
// bracket, the threshold key where its count reaches K (the bracket's top
// when the sample underestimated it), then the shortlist: the listed
// candidates at or below it (list and bitmap), in ascending candidate order
// when it fits LPK_SORT.  The block owns its query's shortlist, so members
// take their positions from an LDS counter; list entries are read
// LPK_UNROLL per work-item at a time.  Grid: nq blocks of 1024.
__global__ void __launch_bounds__(1024)
k_learn_pick(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  __shared__ uint32_t thr;
  __shared__ int32_t hist[2048];
  __shared__ int wsum[16];
  __shared__ int32_t skey[LPK_SORT];
  __shared__ int scount;
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
    if (threadIdx.x == 0) scount = 0;
    __syncthreads();
    for (int64_t e0 = threadIdx.x; e0 < n; e0 += LPK_UNROLL * blockDim.x) {
      uint32_t k[LPK_UNROLL];
#pragma unroll
      for (int u = 0; u < LPK_UNROLL; u++) {
        const int64_t e = e0 + u * blockDim.x;
        k[u] = e < n ? (uint32_t)list[e].y : 0u;
      }
#pragma unroll
      for (int u = 0; u < LPK_UNROLL; u++)
        if (e0 + u * blockDim.x < n) atomicAdd(&hist[k[u] <= lo ? 0 : (k[u] - lo) >> shift], 1);
    }
    __syncthreads();
  }
  {
    // the first bin whose running count reaches K: each work-item holds two
    // bins, a block-wide prefix sum gives the count before them
    const uint32_t* br = D->lrn.brk + slot * 4;
    const uint32_t lo = br[0], hi = br[1], shift = br[2];
    if (threadIdx.x == 0) thr = hi;
    const int b0 = 2 * (int)threadIdx.x;
    const int h0 = hist[b0], h1 = hist[b0 + 1];
    const int ex = block_exclusive_sum_1024(h0 + h1, wsum);
    int b = -1;
    if (ex < K && ex + h0 >= K) b = b0;
    else if (ex + h0 < K && ex + h0 + h1 >= K) b = b0 + 1;
    if (b >= 0) {  // one work-item at most
      const uint64_t top = (uint64_t)lo + (((uint64_t)b + 1) << shift) - 1;
      thr = top < (uint64_t)hi ? (uint32_t)top : hi;
    }
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  int32_t* sl = D->lrn.slist + slot * g.P;
  for (int64_t e0 = 0; e0 < n; e0 += LPK_UNROLL * blockDim.x) {
    int2 it[LPK_UNROLL];
#pragma unroll
    for (int u = 0; u < LPK_UNROLL; u++) {
      const int64_t e = e0 + u * blockDim.x + threadIdx.x;
      it[u] = e < n ? list[e] : make_int2(-1, -1);  // candidate -1: past the list
    }
#pragma unroll
    for (int u = 0; u < LPK_UNROLL; u++) {
      const bool sel = it[u].x >= 0 && (uint32_t)it[u].y <= thr;
      const unsigned long long bal = __ballot(sel);
      if (!bal) continue;
      int base = 0;
      if (lane == 0) base = atomicAdd(&scount, __popcll(bal));
      base = __shfl(base, 0);
      if (sel) {
        const int idx = base + __popcll(bal & ((1ull << lane) - 1ull));
        sl[idx] = it[u].x;
        if (idx < LPK_SORT) skey[idx] = it[u].x;
        atomicOr(&D->lrn.sbits[slot * D->nw64 + (it[u].x >> 6)], 1ull << (it[u].x & 63));
      }
    }
  }
  __syncthreads();
  int ns = scount;
  // an empty shortlist (no approximate distance fell inside the bracket,
  // e.g. ties resolved one ulp apart) takes the seed: the exact recheck and
  // the pick then run on it instead of finding no candidate at all
  if (ns == 0) {
    if (threadIdx.x == 0) {
      const int64_t c = D->lrn.brk[slot * 4 + 3];
      sl[0] = (int32_t)c;
      atomicOr(&D->lrn.sbits[slot * D->nw64 + (c >> 6)], 1ull << (c & 63));
    }
    ns = 1;
  }
  if (threadIdx.x == 0) D->lrn.nslist[slot] = ns;
  // the shortlist in ascending candidate order (bitonic sort in LDS), so the
  // bound kernel's neighbouring wavefronts read overlapping windows
  if (ns > 1 && ns <= LPK_SORT) {
    int N = 2;
    while (N < ns) N <<= 1;
    for (int i = ns + (int)threadIdx.x; i < N; i += blockDim.x) skey[i] = 0x7fffffff;
    __syncthreads();
    for (int k = 2; k <= N; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < N; i += blockDim.x) {
          const int ixj = i ^ j;
          if (ixj > i) {
            const int32_t a = skey[i], b = skey[ixj];
            if ((a > b) == ((i & k) == 0)) { skey[i] = b; skey[ixj] = a; }
          }
        }
        __syncthreads();
      }
    for (int i = threadIdx.x; i < ns; i += blockDim.x) sl[i] = skey[i];
  }
}

// FP32 sum a.q of one candidate window over the masked voxels: each lane
// sums its strided voxels in ascending order, eight window values in flight
// at a time (sixteen measured slower), then a shuffle tree.
template <typename E>
__device__ __forceinline__ float bound_window_sum(const __attribute__((address_space(1))) E* ti, int64_t base,
                                                  const int32_t* tl, const float* qv, int V, int lane) {
  float acc = 0.0f;
  int v = lane;
  for (; v + 64 * 7 < V; v += 64 * 8) {
    float t[8];
#pragma unroll
    for (int u = 0; u < 8; u++) t[u] = (float)ti[base + tl[v + 64 * u]];
#pragma unroll
    for (int u = 0; u < 8; u++) acc = __builtin_fmaf(t[u], qv[v + 64 * u], acc);
  }
  for (; v < V; v += 64) acc = __builtin_fmaf((float)ti[base + tl[v]], qv[v], acc);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  return acc;
}

// four floats at a 4-byte aligned address: one 16-byte global load (the
// hardware takes dword-aligned multi-dword loads)
typedef float iq_f32x4u __attribute__((ext_vector_type(4), aligned(4)));

// the first candidate of c's unit: same row, x rounded down to a multiple of 4
__device__ inline int64_t bound_unit(int64_t c, int d0) { return c - ((c % d0) & 3); }

// FP32 bounds of the shortlist: a wavefront sums a.q over the masked voxels
// in FP32 with the bound |AB32 - AB| <= gamma_V sum|a||b| <= gamma_V
// sqrt(A2 B2), giving an interval [lo, hi] around the exact FP64 distance as
// in the full scan's epilogue; lo goes to lo32, the least hi to the query's
// reduction word.  Members whose candidates share a row and an aligned block
// of four x positions form a unit: their windows are one window shifted by
// 0..3 voxels in x, so a 16-byte load at each masked voxel of the block's
// first window feeds all of them, and the lines a unit brings through L2 are
// shared (a unit whose block runs past the row's last candidate, and FP64
// training images, take their members one window at a time).  The shortlist
// is sorted, so a unit's members are adjacent in it.
// Grid: (groups, query), 8 wavefronts per group (the staged query takes
// 8 V bytes of LDS per group, so larger groups keep more wavefronts per CU);
// a group lists the unit starts of its range LSB_WIN entries at a time in
// LDS, and a wavefront takes one unit at a time.
#define LSB_WIN 512
__global__ void __launch_bounds__(LSB_WIN)
k_short_bound(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  extern __shared__ float qv[];  // the query values (V floats), then the window offsets (V ints)
  __shared__ int ustart[LSB_WIN + 1];
  __shared__ int wcnt[LSB_WIN / 64];
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
  const int d0 = g.dist[0];
  float lmin = __builtin_inff();
  // group b takes a contiguous range of the (sorted) shortlist; ranges are
  // dealt so that the groups of one XCD (b mod 8 when groups are dispatched
  // round robin) cover neighbouring ranges, and its L2 sees nearby windows
  const int G = gridDim.x;
  const int r = (G % 8 == 0) ? ((int)blockIdx.x % 8) * (G / 8) + (int)blockIdx.x / 8 : (int)blockIdx.x;
  const int ebeg = (int)((int64_t)n * r / G), eend = (int)((int64_t)n * (r + 1) / G);
  for (int wb = ebeg; wb < eend; wb += LSB_WIN) {
    const int we = eend - wb < LSB_WIN ? eend : wb + LSB_WIN;
    // unit starts of [wb, we), in order (the window's first entry starts one)
    const int e = wb + (int)threadIdx.x;
    bool first = false;
    if (e < we) first = e == wb || bound_unit(sl[e], d0) != bound_unit(sl[e - 1], d0);
    const unsigned long long bal = __ballot(first);
    if (lane == 0) wcnt[w] = __popcll(bal);
    __syncthreads();
    int before = 0, nu = 0;
#pragma unroll
    for (int i = 0; i < LSB_WIN / 64; i++) {
      before += i < w ? wcnt[i] : 0;
      nu += wcnt[i];
    }
    if (first) ustart[before + __popcll(bal & ((1ull << lane) - 1ull))] = e;
    if (threadIdx.x == 0) ustart[nu] = we;
    __syncthreads();
    for (int k = w; k < nu; k += LSB_WIN / 64) {
      const int s0 = ustart[k], s1 = ustart[k + 1];
      const int64_t c0 = bound_unit(sl[s0], d0);
      int jm = 0;  // members j of the unit: candidates c0 + j
      for (int i = s0; i < s1; i++) jm |= 1 << (int)(sl[i] - c0);
      const int x0 = (int)(c0 % d0), cy = (int)((c0 / d0) % g.dist[1]);
      const int cz = (int)(c0 / ((int64_t)d0 * g.dist[1]));
      const int64_t base = iq_lin3(g.ti, x0, cy, cz);
      float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      if (f32 && x0 + 3 < d0) {
        int v = lane;
        for (; v + 64 * 3 < V; v += 64 * 4) {
          iq_f32x4u t[4];
#pragma unroll
          for (int u = 0; u < 4; u++) t[u] = *iq_glob((const iq_f32x4u*)((const float*)D->ti + base + tl[v + 64 * u]));
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const float qq = qv[v + 64 * u];
            acc[0] = __builtin_fmaf(t[u].x, qq, acc[0]);
            acc[1] = __builtin_fmaf(t[u].y, qq, acc[1]);
            acc[2] = __builtin_fmaf(t[u].z, qq, acc[2]);
            acc[3] = __builtin_fmaf(t[u].w, qq, acc[3]);
          }
        }
        for (; v < V; v += 64) {
          const iq_f32x4u t = *iq_glob((const iq_f32x4u*)((const float*)D->ti + base + tl[v]));
          const float qq = qv[v];
          acc[0] = __builtin_fmaf(t.x, qq, acc[0]);
          acc[1] = __builtin_fmaf(t.y, qq, acc[1]);
          acc[2] = __builtin_fmaf(t.z, qq, acc[2]);
          acc[3] = __builtin_fmaf(t.w, qq, acc[3]);
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (jm & (1 << j))
            for (int o = 32; o > 0; o >>= 1) acc[j] += __shfl_xor(acc[j], o);
      } else {
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (jm & (1 << j))
            acc[j] = f32 ? bound_window_sum(ti32, base + j, tl, qv, V, lane) : bound_window_sum(ti64, base + j, tl, qv, V, lane);
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (!(jm & (1 << j))) continue;
        const int64_t c = c0 + j;
        const double a2 = a2map[c];
        const double sq = __builtin_sqrt(a2 * b2) * (1.0 + 0x1p-30);
        const double d = __builtin_fabs((a2 - 2.0 * (double)acc[j]) + b2);
        const double er = (2.0 * (gerr + g64) * sq + 4.0 * u64 * (a2 + b2 + 2.0 * sq) + 1e3 * V * 0x1p-125) * (1.0 + 0x1p-30) +
                          4.0 * u64 * d;
        if (lane == 0) lo_out[c] = iq_f32_down(d - er > 0.0 ? d - er : 0.0);
        lmin = fminf(lmin, iq_f32_up(d + er));
      }
    }
    __syncthreads();  // the next window's unit starts overwrite these
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
