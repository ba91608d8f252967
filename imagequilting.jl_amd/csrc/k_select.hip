// k_select.hip -- candidate selection of one tile step (iqsim.jl:237-243):
//   * tol-band pick  findall(D .<= (1 + tol) * minimum(D))   (iqsim.jl:237)
//   * relaxation     soft/hard pattern database              (relaxation.jl:5-48)
//   * taumodel       dense ranks -> tau-model weights         (taumodel.jl:5-45)
//   * weighted draw  with a caller-supplied uniform          (iqsim.jl:243)
// One 1024-wide workgroup per query.  Every floating-point step uses the
// reference's operator order (the library is compiled with
// -ffp-contract=off), so results are bit-identical to the FP64 oracle.
#include "iq_internal.h"

#define FB 1024
#define NW (FB / 64)
#define S2MAX 2048

struct SelScratch {
  int64_t i64[NW + 4];
  double f64[NW + 4];
  int64_t result;
  int64_t n;
  int64_t aux;
  int32_t flag;
};

// ---------------------------------------------------------------------------
// block primitives (blockDim.x == FB)
// ---------------------------------------------------------------------------
__device__ int64_t blk_excl_scan(int64_t v, int64_t* total, SelScratch& sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    int64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh.i64[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t run = 0;
    for (int i = 0; i < NW; i++) { int64_t t = sh.i64[i]; sh.i64[i] = run; run += t; }
    sh.i64[NW] = run;
  }
  __syncthreads();
  int64_t r = sh.i64[w] + x - v;
  if (total) *total = sh.i64[NW];
  __syncthreads();
  return r;
}

__device__ int64_t blk_sum_i64(int64_t v, SelScratch& sh) {
  int64_t t;
  blk_excl_scan(v, &t, sh);
  return t;
}

__device__ double blk_max_f64(double v, SelScratch& sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) sh.f64[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sh.f64[0];
    for (int i = 1; i < NW; i++) m = fmax(m, sh.f64[i]);
    sh.f64[NW] = m;
  }
  __syncthreads();
  double r = sh.f64[NW];
  __syncthreads();
  return r;
}

__device__ double blk_min_f64(double v, SelScratch& sh) { return -blk_max_f64(-v, sh); }

// Ascending bitonic sort of sv[0..m) in LDS, m a power of two (padded with +inf).
__device__ void blk_bitonic(double* sv, int m) {
  for (int k = 2; k <= m; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < m; i += FB) {
        int ixj = i ^ j;
        if (ixj > i) {
          double a = sv[i], b = sv[ixj];
          bool up = (i & k) == 0;
          if ((a > b) == up) { sv[i] = b; sv[ixj] = a; }
        }
      }
      __syncthreads();
    }
  }
}

__device__ unsigned long long blk_or_u64(unsigned long long v, SelScratch& sh) {
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) sh.i64[threadIdx.x >> 6] = (int64_t)v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long r = 0;
    for (int i = 0; i < NW; i++) r |= (unsigned long long)sh.i64[i];
    sh.i64[NW] = (int64_t)r;
  }
  __syncthreads();
  unsigned long long r = (unsigned long long)sh.i64[NW];
  __syncthreads();
  return r;
}

// Lanes of the wavefront whose 8-bit digit equals this lane's (valid lanes only).
__device__ inline unsigned long long digit_peers(bool valid, int d) {
  unsigned long long peers = __ballot(valid);
  for (int b = 0; b < 8; b++) {
    const unsigned long long m = __ballot((d >> b) & 1);
    peers &= ((d >> b) & 1) ? m : ~m;
  }
  return peers;
}

// Stable ascending LSD radix sort (8-bit digits) of n >= 1 keys in global
// memory by one workgroup.  Each wavefront owns a contiguous segment of the
// keys (a multiple of 64); inside a 64-key tile, lanes with equal digits are
// ordered by lane through ballots.  Counts live in cnt (NW * 256 ints of LDS,
// digit-major), so the scatter order is (digit, wavefront, position).
// Digits on which all keys agree are skipped.  Returns the buffer holding the
// sorted keys (a or b).
__device__ unsigned long long* blk_radix_sort(unsigned long long* a, unsigned long long* b, int64_t n, int32_t* cnt,
                                              SelScratch& sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long k0 = a[0];
  unsigned long long diff = 0;
  for (int64_t i = threadIdx.x; i < n; i += FB) diff |= a[i] ^ k0;
  diff = blk_or_u64(diff, sh);
  const int64_t seg = (((n + NW - 1) / NW) + 63) & ~(int64_t)63;
  const int64_t s0 = (int64_t)w * seg < n ? (int64_t)w * seg : n;
  const int64_t s1 = s0 + seg < n ? s0 + seg : n;
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int sh8 = 0; sh8 < 64; sh8 += 8) {
    if (((diff >> sh8) & 0xffull) == 0) continue;
    for (int i = threadIdx.x; i < NW * 256; i += FB) cnt[i] = 0;
    __syncthreads();
    for (int64_t t = s0; t < s1; t += 64) {
      const int64_t i = t + lane;
      const bool v = i < s1;
      const int d = v ? (int)((a[i] >> sh8) & 0xffull) : 0;
      const unsigned long long peers = digit_peers(v, d);
      if (v && (peers & lt) == 0) cnt[d * NW + w] += __popcll(peers);
    }
    __syncthreads();
    {
      const int e0 = threadIdx.x * 4;  // NW * 256 == 4 * FB entries
      const int c0 = cnt[e0], c1 = cnt[e0 + 1], c2 = cnt[e0 + 2], c3 = cnt[e0 + 3];
      const int ex = (int)blk_excl_scan(c0 + c1 + c2 + c3, nullptr, sh);
      cnt[e0] = ex;
      cnt[e0 + 1] = ex + c0;
      cnt[e0 + 2] = ex + c0 + c1;
      cnt[e0 + 3] = ex + c0 + c1 + c2;
    }
    __syncthreads();
    for (int64_t t = s0; t < s1; t += 64) {
      const int64_t i = t + lane;
      const bool v = i < s1;
      const unsigned long long key = v ? a[i] : 0ull;
      const int d = (int)((key >> sh8) & 0xffull);
      const unsigned long long peers = digit_peers(v, d);
      if (v) b[cnt[d * NW + w] + __popcll(peers & lt)] = key;
      if (v && (peers & lt) == 0) cnt[d * NW + w] += __popcll(peers);
    }
    __syncthreads();
    unsigned long long* x = a;
    a = b;
    b = x;
  }
  return a;
}

// Order-preserving integer key of a non-NaN double (-0 and +0 share a key).
__device__ inline unsigned long long f64_key(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x == 0.0 ? 0.0 : x);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}

// Dense ranks of any number of non-NaN values val(i), i < n: sort a copy of
// their order-preserving keys, compact the distinct ones, binary-search each
// value.  keys: 2n words of global scratch; cnt: NW * 256 ints of LDS.
template <class F>
__device__ void blk_dense_rank_big(F val, int64_t n, int32_t* rank, unsigned long long* keys, int32_t* cnt,
                                   SelScratch& sh) {
  unsigned long long* ka = keys;
  unsigned long long* kb = keys + n;
  for (int64_t i = threadIdx.x; i < n; i += FB) ka[i] = f64_key(val(i));
  __syncthreads();
  unsigned long long* s = blk_radix_sort(ka, kb, n, cnt, sh);
  unsigned long long* u = s == ka ? kb : ka;
  const int64_t per = (n + FB - 1) / FB;
  const int64_t i0 = threadIdx.x * per < n ? threadIdx.x * per : n;
  const int64_t i1 = i0 + per < n ? i0 + per : n;
  int64_t local = 0;
  for (int64_t i = i0; i < i1; i++) local += (i == 0 || s[i] != s[i - 1]);
  int64_t nd;
  int64_t off = blk_excl_scan(local, &nd, sh);
  for (int64_t i = i0; i < i1; i++)
    if (i == 0 || s[i] != s[i - 1]) u[off++] = s[i];
  __syncthreads();
  for (int64_t i = threadIdx.x; i < n; i += FB) {
    const unsigned long long v = f64_key(val(i));
    int64_t lo = 0, hi = nd;  // first distinct value >= v
    while (lo < hi) { int64_t mid = (lo + hi) >> 1; if (u[mid] < v) lo = mid + 1; else hi = mid; }
    rank[i] = (int32_t)(lo + 1);
  }
  __syncthreads();
}

// Dense ranks (ties share a rank, taumodel.jl:21-31) of vals[0..n).  When all
// values are equal every rank is 1 and *constant is set (rank is not
// written).  Up to IQ_SORT_MAX values are ranked in LDS (sv, dp); more go
// through blk_dense_rank_big with the global scratch `keys` (2n words; the
// LDS of sv then holds the radix counts).  keys == nullptr: false for more
// than IQ_SORT_MAX values with several distinct ones.
__device__ bool blk_dense_rank(const double* vals, int64_t n, int32_t* rank, bool* constant, double* sv,
                               int32_t* dp, SelScratch& sh, unsigned long long* keys = nullptr) {
  double lmin = __builtin_inf(), lmax = -__builtin_inf();
  for (int64_t i = threadIdx.x; i < n; i += FB) { lmin = fmin(lmin, vals[i]); lmax = fmax(lmax, vals[i]); }
  double mn = blk_min_f64(lmin, sh), mx = blk_max_f64(lmax, sh);
  *constant = mn == mx;
  if (*constant) return true;
  if (n > IQ_SORT_MAX) {
    if (!keys) return false;
    blk_dense_rank_big([&](int64_t i) { return vals[i]; }, n, rank, keys, (int32_t*)sv, sh);
    return true;
  }
  int m = 1;
  while (m < n) m <<= 1;
  for (int i = threadIdx.x; i < m; i += FB) sv[i] = i < n ? vals[i] : __builtin_inf();
  __syncthreads();
  blk_bitonic(sv, m);
  int64_t local = 0;
  const int per = (m + FB - 1) / FB;
  const int i0 = threadIdx.x * per;
  for (int i = i0; i < i0 + per && i < m; i++) local += (i == 0 || sv[i] != sv[i - 1]);
  int64_t ex = blk_excl_scan(local, nullptr, sh);
  for (int i = i0; i < i0 + per && i < m; i++) {
    ex += (i == 0 || sv[i] != sv[i - 1]);
    dp[i] = (int32_t)ex;
  }
  __syncthreads();
  for (int64_t i = threadIdx.x; i < n; i += FB) {
    double v = vals[i];
    int lo = 0, hi = m;  // first position with sv >= v
    while (lo < hi) { int mid = (lo + hi) >> 1; if (sv[mid] < v) lo = mid + 1; else hi = mid; }
    rank[i] = dp[lo];
  }
  __syncthreads();
  return true;
}

// Sequential sum of n consecutive weights starting at i0 (p == nullptr:
// every weight is w); loads are issued eight at a time, zero padding is exact.
__device__ inline double seq_sum(const double* p, double w, int64_t i0, int64_t n) {
  double s = 0.0;
  if (!p) {
    for (int64_t i = 0; i < n; i++) s += w;
    return s;
  }
  int64_t i = 0;
  for (; i + 8 <= n; i += 8) {
    double v0 = p[i0 + i], v1 = p[i0 + i + 1], v2 = p[i0 + i + 2], v3 = p[i0 + i + 3];
    double v4 = p[i0 + i + 4], v5 = p[i0 + i + 5], v6 = p[i0 + i + 6], v7 = p[i0 + i + 7];
    s += v0; s += v1; s += v2; s += v3; s += v4; s += v5; s += v6; s += v7;
  }
  for (; i < n; i++) s += p[i0 + i];
  return s;
}

// Three-level blocked cumulative draw (the oracle's summation order: blocks of
// IQ_SB at two levels, a sequential top level).  p == nullptr: all weights w.
__device__ int64_t blk_sample(const double* p, double w, int64_t n, double u, double* s1, double* s2,
                              SelScratch& sh) {
  const int64_t n1 = (n + IQ_SB - 1) / IQ_SB, n2 = (n1 + IQ_SB - 1) / IQ_SB;
  for (int64_t k = threadIdx.x; k < n1; k += FB) {
    int64_t e = (k + 1) * IQ_SB < n ? (k + 1) * IQ_SB : n;
    s1[k] = seq_sum(p, w, k * IQ_SB, e - k * IQ_SB);
  }
  __syncthreads();
  for (int64_t m = threadIdx.x; m < n2; m += FB) {
    int64_t e = (m + 1) * IQ_SB < n1 ? (m + 1) * IQ_SB : n1;
    s2[m] = seq_sum(s1, 0.0, m * IQ_SB, e - m * IQ_SB);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double W = 0.0;
    for (int64_t m = 0; m < n2; m++) W += s2[m];
    double t = u * W;
    double acc = 0.0;
    int64_t m = 0;
    for (;; m++) {
      double nx = acc + s2[m];
      if (nx >= t || m == n2 - 1) break;
      acc = nx;
    }
    int64_t k = m * IQ_SB, kend = (m + 1) * IQ_SB < n1 ? (m + 1) * IQ_SB : n1;
    for (;; k++) {
      double nx = acc + s1[k];
      if (nx >= t || k == kend - 1) break;
      acc = nx;
    }
    int64_t i = k * IQ_SB, iend = (k + 1) * IQ_SB < n ? (k + 1) * IQ_SB : n;
    for (;; i++) {
      double nx = acc + (p ? p[i] : w);
      if (nx >= t || i == iend - 1) break;
      acc = nx;
    }
    sh.result = i;
  }
  __syncthreads();
  int64_t r = sh.result;
  __syncthreads();
  return r;
}

// Position of `key` in the ascending list idx[0..n) (or -1).
__device__ int64_t blk_find(const int32_t* idx, int64_t n, int64_t key, SelScratch& sh) {
  if (threadIdx.x == 0) {
    int64_t lo = 0, hi = n;
    while (lo < hi) { int64_t mid = (lo + hi) >> 1; if (idx[mid] < key) lo = mid + 1; else hi = mid; }
    sh.result = (lo < n && idx[lo] == key) ? lo : -1;
  }
  __syncthreads();
  int64_t r = sh.result;
  __syncthreads();
  return r;
}

__device__ void set_err(int32_t* err, int code) { atomicCAS(err, 0, code); }

// Tau-model factor of one more source (taumodel.jl:34-44): p[i] *= X_i / x0
// with P = (n - r + 1) / sum(n - r + 1), X = (1 - P) / P.  A constant
// source (all ranks 1) gives X == x0 exactly, a neutral factor of 1.0.
__device__ void tau_accumulate(int64_t n, const int32_t* rank, bool constant, double* p, SelScratch& sh) {
  if (constant) return;  // p * 1.0 == p
  int64_t ls = 0;
  for (int64_t i = threadIdx.x; i < n; i += FB) ls += (n - rank[i]) + 1;
  const double S = (double)blk_sum_i64(ls, sh);  // exact integer sum
  const double nd = (double)n;
  const double x0 = (1.0 - 1.0 / nd) / (1.0 / nd);
  for (int64_t i = threadIdx.x; i < n; i += FB) {
    double P = ((nd - (double)rank[i]) + 1.0) / S;
    double X = (1.0 - P) / P;
    p[i] = p[i] * (X / x0);
  }
  __syncthreads();
}

__device__ inline double tau_weight(int64_t n, double prodv) {
  const double nd = (double)n;
  const double x0 = (1.0 - 1.0 / nd) / (1.0 / nd);
  double x = x0 * prodv;
  return 1.0 / (1.0 + x);
}

__device__ void tau_finish(int64_t n, double* p) {
  for (int64_t i = threadIdx.x; i < n; i += FB) p[i] = tau_weight(n, p[i]);
  __syncthreads();
}

// Weight shared by every event when all sources are constant (all ranks 1).
__device__ inline double tau_constant_weight(int64_t n) {
  const double nd = (double)n;
  const double S = nd * nd;  // exact for n < 2^26
  const double x0 = (1.0 - 1.0 / nd) / (1.0 / nd);
  const double P = ((nd - 1.0) + 1.0) / S;
  const double X = (1.0 - P) / P;
  return tau_weight(n, 1.0 * (X / x0));
}

struct SelOut {
  int64_t* pick;  // current pick of this query (consumed by the cut / paste)
  int64_t* log;   // pick log entry of this path position
  int32_t* err;   // run error word
  int64_t* nband; // size of the pattern database
};

// Draw (or replay `forced`) from db[0..n) with weights p (or constant w).
__device__ void publish_pick(const int32_t* db, int64_t n, const double* p, double w, double u, int64_t forced,
                             double* s1, double* s2, SelOut o, SelScratch& sh) {
  int64_t i;
  if (forced >= 0) {
    i = blk_find(db, n, forced, sh);
    if (i < 0) {
      if (threadIdx.x == 0) { set_err(o.err, IQ_KERR_NOTINDB); *o.pick = -2; *o.log = forced; }
      return;
    }
  } else {
    i = blk_sample(p, w, n, u, s1, s2, sh);
  }
  if (threadIdx.x == 0) {
    *o.pick = db[i];
    *o.log = db[i];
    if (o.nband) *o.nband = n;
  }
}

// Tau weights of n events over nsrc sources (values of source j at event i are
// src_vals(j, i), staged into vals) -> p (or constant).  rank and p hold n
// entries; keys (2n words, or nullptr) is the scratch of the large-n ranker.
// False when n exceeds the LDS ranker and no keys are given.
template <class F>
__device__ bool tau_weights(int64_t n, int nsrc, F src_vals, double* vals, int32_t* rank, double* p, bool* allconst,
                            double* sv, int32_t* dp, SelScratch& sh, unsigned long long* keys = nullptr) {
  *allconst = true;
  if (n == 1) return true;
  bool started = false;
  for (int j = 0; j < nsrc; j++) {
    for (int64_t i = threadIdx.x; i < n; i += FB) vals[i] = src_vals(j, i);
    __syncthreads();
    bool c;
    if (!blk_dense_rank(vals, n, rank, &c, sv, dp, sh, keys)) return false;
    if (!c) {
      if (!started) {
        for (int64_t i = threadIdx.x; i < n; i += FB) p[i] = 1.0;
        __syncthreads();
        started = true;
      }
      tau_accumulate(n, rank, false, p, sh);
      *allconst = false;
    }
  }
  if (started) tau_finish(n, p);
  return true;
}

// ---------------------------------------------------------------------------
// Tol-band finalize (plain queries): exact min -> threshold -> ordered
// compaction of the band -> dense ranks -> weights -> draw.  Queries without
// an overlap (first tile, ...) have D == 0 on every enabled candidate, so the
// band is the enabled list and every rank is 1.
// ---------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(FB)
k_band_finalize(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs, int use_list) {
  __shared__ SelScratch sh;
  __shared__ double sv[IQ_SORT_MAX];
  __shared__ int32_t dp[IQ_SORT_MAX];
  __shared__ double s2[S2MAX];
  const IqQuery q = qs[blockIdx.x];
  if (q.multi) return;
  const IqGeo& g = D->g;
  const IqStep& st = D->steps[q.step];
  const int64_t slot = q.slot;
  SelOut o{D->pick + slot, D->log + D->reals[q.real].log_off + (q.step - D->reals[q.real].step0), D->err,
           D->nband + slot};
  double* s1 = D->s1 + slot * D->s1_stride;
  double* sd = D->sd + slot * IQ_SORT_MAX;
  int32_t* si = D->si + slot * IQ_SORT_MAX;
  if (st.mask < 0) {
    const int64_t nen = D->nen;
    if (nen == 0) {
      if (threadIdx.x == 0) { set_err(D->err, IQ_KERR_NOCAND); *o.pick = -2; }
      return;
    }
    const double w = nen == 1 ? 1.0 : tau_constant_weight(nen);
    publish_pick(D->enabled, nen, nullptr, w, q.u, q.forced, s1, s2, o, sh);
    return;
  }
  const double m = __longlong_as_double((long long)D->gmin64[slot]);
  if (__builtin_isinf(m)) {
    if (threadIdx.x == 0) { set_err(D->err, IQ_KERR_NOCAND); *o.pick = -2; }
    return;
  }
  const double thr = (1.0 + D->tol) * m;  // iqsim.jl:237
  int32_t* band = D->band + slot * g.P;
  const int64_t nsup = D->nsup[slot];
  int64_t nb;
  double dmax;
  bool small, inlds = false;
  // band, distances, ranks and weights of a list-built band stay in LDS
  // (upper halves of sv / dp; the LDS ranker uses at most their first FB)
  static_assert(FB <= IQ_SORT_MAX / 4, "LDS band regions overlap the ranker's");
  static_assert((FB + IQ_SB - 1) / IQ_SB <= S2MAX / 2, "draw scratch of a list-built band");
  int32_t* lband = dp + IQ_SORT_MAX / 2;
  double* ldist = sv + IQ_SORT_MAX / 2;
  if (use_list && nsup <= FB) {
    // usual case: the superset list (unordered) holds every candidate of the
    // band; each member's place in ascending candidate order is the number
    // of band members with a smaller index.  (use_list = 0: the bitmap was
    // filled without a list, by the exhaustive exact map.)
    const auto dexact = iq_glob(D->dexact + slot * g.P);
    int32_t c = 0;
    double d = __builtin_inf();
    if ((int64_t)threadIdx.x < nsup) {
      c = D->sup[slot * g.P + threadIdx.x];
      d = dexact[c];
    }
    const bool inb = (int64_t)threadIdx.x < nsup && d <= thr;
    dp[threadIdx.x] = inb ? c : 0x7fffffff;
    blk_excl_scan(inb ? 1 : 0, &nb, sh);
    dmax = blk_max_f64(inb ? d : -__builtin_inf(), sh);
    small = true;
    inlds = true;
    if (inb) {
      int r = 0;
      for (int j = 0; j < (int)nsup; j++) r += dp[j] < c;
      band[r] = c;
      lband[r] = c;
      ldist[r] = d;
    }
  } else {
    const unsigned long long* bitmap = D->bitmap + slot * D->nw64;
    const double* dexact = D->dexact + slot * g.P;
    const int64_t nw = D->nw64;
    const int64_t per = (nw + FB - 1) / FB;
    const int64_t w0 = threadIdx.x * per, w1 = (w0 + per < nw) ? w0 + per : nw;
    int64_t cnt = 0;
    double lmax = -__builtin_inf();
    for (int64_t w = w0; w < w1; w++) {
      unsigned long long b = bitmap[w];
      while (b) {
        int bit = __builtin_ctzll(b);
        b &= b - 1;
        double d = dexact[w * 64 + bit];
        if (d <= thr) { cnt++; lmax = fmax(lmax, d); }
      }
    }
    int64_t off = blk_excl_scan(cnt, &nb, sh);
    dmax = blk_max_f64(lmax, sh);
    small = nb <= IQ_SORT_MAX;
    for (int64_t w = w0; w < w1; w++) {
      unsigned long long b = bitmap[w];
      while (b) {
        int bit = __builtin_ctzll(b);
        b &= b - 1;
        int64_t c = w * 64 + bit;
        double d = dexact[c];
        if (d <= thr) {
          band[off] = (int32_t)c;
          if (small) sd[off] = d;
          off++;
        }
      }
    }
  }
  if (threadIdx.x == 0 && D->rounds) {  // work counters: superset and band sizes
    atomicAdd((unsigned long long*)D->rounds + 3, 1ull);
    atomicAdd((unsigned long long*)D->rounds + 4, (unsigned long long)nsup);
    atomicAdd((unsigned long long*)D->rounds + 5, (unsigned long long)nb);
  }
  if (nb == 0) {  // the exact minimum is always in the band; never sample an empty list
    if (threadIdx.x == 0) { set_err(D->err, IQ_KERR_EMPTYBAND); *o.pick = -2; }
    return;
  }
  __syncthreads();
  const int32_t* db = inlds ? lband : band;
  double* ps1 = inlds ? s2 + S2MAX / 2 : s1;  // the top level of the draw needs few entries of s2
  if (nb == 1) {
    publish_pick(db, 1, nullptr, 1.0, q.u, q.forced, ps1, s2, o, sh);  // taumodel.jl:9
    return;
  }
  if (dmax == m) {  // every band distance equal: all ranks 1
    publish_pick(db, nb, nullptr, tau_constant_weight(nb), q.u, q.forced, ps1, s2, o, sh);
    return;
  }
  double* p = inlds ? sv + 3 * (IQ_SORT_MAX / 4) : sd;
  int32_t* rk = inlds ? dp + 3 * (IQ_SORT_MAX / 4) : si;
  if (small) {
    bool c;
    blk_dense_rank(inlds ? ldist : sd, nb, rk, &c, sv, dp, sh);
  } else {
    // large band: radix-sorted ranks; ranks go to the (consumed) superset
    // list, weights over the key scratch once the ranks are known
    unsigned long long* keys = D->rkey + slot * 2 * g.P;
    const double* dexact = D->dexact + slot * g.P;
    rk = D->sup + slot * g.P;
    blk_dense_rank_big([&](int64_t i) { return dexact[band[i]]; }, nb, rk, keys, (int32_t*)sv, sh);
    p = (double*)keys;
  }
  for (int64_t i = threadIdx.x; i < nb; i += FB) p[i] = 1.0;
  __syncthreads();
  tau_accumulate(nb, rk, false, p, sh);
  tau_finish(nb, p);
  publish_pick(db, nb, p, 0.0, q.u, q.forced, ps1, s2, o, sh);
}

// ---------------------------------------------------------------------------
// Relaxation (relaxation.jl:5-39) by order statistics.
// maps[0] is the primary map D, maps[1..nsrc) the auxiliary maps Ds.  Every
// set the loop builds is a prefix of a map's stable ascending order (value,
// then index: Julia's Perm ordering for partialsortperm, relaxation.jl:12,27),
// and a prefix of length m is { i : (bits(v_i), i) <= the m-th smallest
// pair }: non-negative doubles and +inf order like their IEEE bit patterns.
// So each set needs one order statistic (radix selection, blk_kth_key) and
// one pass that writes it as a bitmap (bits: nsrc maps of nwd words); the
// running intersection is the AND of the bitmaps.
// ---------------------------------------------------------------------------
#define KS_CAP 512          // bins this small are ranked pairwise in LDS
#define KS_LDS_WORDS 1792   // 64-bit LDS words: histogram 2048 x u32, KS_CAP values, KS_CAP indices

__device__ inline uint64_t f64_bits(double x) { return (uint64_t)__double_as_longlong(x); }

// Bin of the running 1-based rank r over nb histogram counts; r becomes the
// rank inside that bin and cnt its count.
__device__ __forceinline__ int blk_find_bin(const uint32_t* hist, int nb, int64_t& r, int64_t& cnt, SelScratch& sh) {
  const int per = (nb + FB - 1) / FB;
  const int b0 = threadIdx.x * per;
  int64_t c = 0;
  for (int k = 0; k < per; k++)
    if (b0 + k < nb) c += hist[b0 + k];
  const int64_t ex = blk_excl_scan(c, nullptr, sh);
  if (ex < r && r <= ex + c) {
    int64_t acc = ex;
    for (int k = 0; k < per; k++) {
      const int64_t h = hist[b0 + k];
      if (r <= acc + h) { sh.result = b0 + k; sh.n = r - acc; sh.aux = h; break; }
      acc += h;
    }
  }
  __syncthreads();
  const int bin = (int)sh.result;
  r = sh.n;
  cnt = sh.aux;
  __syncthreads();
  return bin;
}

// The m-th smallest (1 <= m <= P) pair (bits(map[i]), i) over i in [0, P).
// Radix digits of 11 bits over the value bits, each a histogram pass over
// the map restricted to the prefix fixed so far, until the bin holding rank
// m has at most KS_CAP members; those are compacted into LDS and ranked
// pairwise.  When the whole value is fixed with more than KS_CAP ties, the
// m-th is found by index order directly (a counting pass per wavefront
// segment, then a walk of the one segment that holds it).
__device__ __forceinline__ void blk_kth_key(const double* __restrict__ map, int64_t P, int64_t m, uint64_t* lds, SelScratch& sh,
                            uint64_t* tv, int64_t* ti) {
  uint32_t* hist = (uint32_t*)lds;
  uint64_t* cv = lds + 1024;
  uint32_t* ci = (uint32_t*)(lds + 1024 + KS_CAP);
  const auto* mg = iq_glob(map);
  const int lane = threadIdx.x & 63;
  uint64_t pv = 0, hm = 0;  // prefix value and mask of the fixed top bits
  int known = 0;
  int64_t r = m, cnt = P;
  while (known < 64 && cnt > KS_CAP) {
    const int w = 64 - known >= 11 ? 11 : 64 - known;
    const int sft = 64 - known - w;
    const uint32_t dm = (1u << w) - 1u;
    for (int b = threadIdx.x; b < (1 << w); b += FB) hist[b] = 0u;
    __syncthreads();
    for (int64_t i0 = threadIdx.x; i0 < P; i0 += 8 * FB) {
      uint64_t v[8];
#pragma unroll
      for (int k = 0; k < 8; k++) v[k] = i0 + k * FB < P ? f64_bits(mg[i0 + k * FB]) : 0ull;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const bool mt = i0 + k * FB < P && (v[k] & hm) == pv;
        const uint32_t d = (uint32_t)(v[k] >> sft) & dm;
        // lanes sharing the first matching lane's digit add once, through it
        const uint64_t act = __ballot(mt);
        if (act) {
          const int lead = __ffsll((unsigned long long)act) - 1;
          const uint32_t ld = __shfl(d, lead);
          const uint64_t same = __ballot(mt && d == ld);
          if (lane == lead) atomicAdd(&hist[ld], (uint32_t)__popcll(same));
          if (mt && d != ld) atomicAdd(&hist[d], 1u);
        }
      }
    }
    __syncthreads();
    const int bin = blk_find_bin(hist, 1 << w, r, cnt, sh);
    pv |= (uint64_t)bin << sft;
    hm |= (uint64_t)dm << sft;
    known += w;
  }
  if (cnt > KS_CAP) {
    // value fixed (pv), rank r among its ties in index order
    const int wv = threadIdx.x >> 6;
    const int64_t seg = (P + NW * 64 - 1) / (NW * 64) * 64;
    const int64_t a = wv * seg, e = a + seg < P ? a + seg : P;
    int64_t c = 0;
    for (int64_t base = a; base < e; base += 64) {
      const int64_t i = base + lane;
      c += __popcll(__ballot(i < e && f64_bits(mg[i]) == pv));
    }
    const int64_t ex = blk_excl_scan(lane == 0 ? c : 0, nullptr, sh);
    const int64_t pre = __shfl(ex, 0);
    if (pre < r && r <= pre + c) {
      int64_t need = r - pre;
      for (int64_t base = a; base < e; base += 64) {
        const int64_t i = base + lane;
        uint64_t b = __ballot(i < e && f64_bits(mg[i]) == pv);
        const int k = __popcll(b);
        if (need <= k) {
          for (int64_t t = 1; t < need; t++) b &= b - 1;
          if (lane == 0) sh.result = base + __ffsll((unsigned long long)b) - 1;
          break;
        }
        need -= k;
      }
    }
    __syncthreads();
    *tv = pv;
    *ti = sh.result;
    __syncthreads();
    return;
  }
  // compact the bin's members into LDS (any order), then rank them pairwise
  if (threadIdx.x == 0) sh.flag = 0;
  __syncthreads();
  for (int64_t i0 = threadIdx.x; i0 < P; i0 += 8 * FB) {
    uint64_t v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = i0 + k * FB < P ? f64_bits(mg[i0 + k * FB]) : 0ull;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const bool mt = i0 + k * FB < P && (v[k] & hm) == pv;
      const uint64_t b = __ballot(mt);
      if (b) {
        const int lead = __ffsll((unsigned long long)b) - 1;
        int base = 0;
        if (lane == lead) base = atomicAdd(&sh.flag, __popcll(b));
        base = __shfl(base, lead);
        if (mt) {
          const int slot = base + __popcll(b & ((1ull << lane) - 1ull));
          cv[slot] = v[k];
          ci[slot] = (uint32_t)(i0 + k * FB);
        }
      }
    }
  }
  __syncthreads();
  const int n = sh.flag;
  if ((int)threadIdx.x < n) {
    const uint64_t myv = cv[threadIdx.x];
    const uint32_t myi = ci[threadIdx.x];
    int rank = 0;
    for (int f = 0; f < n; f++) {
      const uint64_t fv = cv[f];
      rank += (fv < myv || (fv == myv && ci[f] < myi)) ? 1 : 0;
    }
    if (rank == r - 1) { sh.result = myi; sh.aux = (int64_t)myv; }
  }
  __syncthreads();
  *tv = (uint64_t)sh.aux;
  *ti = sh.result;
  __syncthreads();
}

// bits[w] for the prefix set {i : (bits(map[i]), i) <= (tv, ti)}: one
// wavefront per 64 consecutive candidates, its ballot is two words.
__device__ __forceinline__ void blk_prefix_bits(const double* __restrict__ map, int64_t P, uint64_t tv, int64_t ti,
                                uint32_t* __restrict__ bits) {
  const auto* mg = iq_glob(map);
  const int lane = threadIdx.x & 63;
  const int64_t nwd = (P + 31) / 32;
  // eight loads in flight per work-item
  for (int64_t base0 = (int64_t)(threadIdx.x >> 6) * 64; base0 < P; base0 += 8 * FB) {
    uint64_t v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int64_t i = base0 + k * FB + lane;
      v[k] = i < P ? f64_bits(mg[i]) : 0ull;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int64_t base = base0 + k * FB, i = base + lane;
      const uint64_t b = __ballot(i < P && (v[k] < tv || (v[k] == tv && i <= ti)));
      if (lane == 0 && base < P) {
        const int64_t w = base >> 5;
        bits[w] = (uint32_t)b;
        if (w + 1 < nwd) bits[w + 1] = (uint32_t)(b >> 32);
      }
    }
  }
}

// Returns the size of the database written to ev (ascending), -1 if every
// candidate is disabled.  lds: KS_LDS_WORDS 64-bit words.
__device__ __forceinline__ int64_t relax_db(int64_t P, int nsrc, const double* const* maps, double tol, uint32_t* bits, int32_t* ev,
                            uint64_t* lds, SelScratch& sh) {
  __shared__ int64_t lenj[IQ_MAX_AUX + 2];
  const int64_t nwd = (P + 31) / 32;
  // enabled patterns = finite entries of D; all zero when none is finite and nonzero (relaxation.jl:7-11)
  const auto* dg = iq_glob(maps[0]);
  int64_t cf = 0, cz = 0;
  for (int64_t i0 = threadIdx.x; i0 < P; i0 += 8 * FB) {  // eight loads in flight
    double x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = i0 + k * FB < P ? dg[i0 + k * FB] : __builtin_inf();
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const bool fin = !__builtin_isinf(x[k]);
      cf += fin;
      cz += fin && x[k] != 0.0;
    }
  }
  const int64_t nen = blk_sum_i64(cf, sh);
  const int64_t nnz = blk_sum_i64(cz, sh);
  if (nen == 0) return -1;  // relaxation.jl would never terminate
  const bool allzero = nnz == 0;
  const int64_t dbsize = allzero ? nen : (int64_t)ceil(tol * (double)nen);
  if (dbsize <= 0) return -1;  // tol outside (0, 1]: the loop would never end
  if (threadIdx.x < IQ_MAX_AUX + 2) lenj[threadIdx.x] = 0;
  {
    uint64_t tv;
    int64_t ti;
    blk_kth_key(maps[0], P, dbsize, lds, sh, &tv, &ti);
    blk_prefix_bits(maps[0], P, tv, ti, bits);
  }
  __syncthreads();
  double frac = 0.1 * ((double)dbsize / (double)nen);
  int64_t cnt = dbsize;
  for (;;) {
    const int64_t softdbsize = (int64_t)ceil(frac * (double)nen);
    cnt = dbsize;
    for (int j = 1; j < nsrc; j++) {
      // the prefix only grows (frac does); an unchanged length keeps its bitmap
      if (softdbsize != lenj[j]) {
        uint64_t tv = 0;
        int64_t ti = -1;  // empty prefix
        if (softdbsize > 0) blk_kth_key(maps[j], P, softdbsize < P ? softdbsize : P, lds, sh, &tv, &ti);
        blk_prefix_bits(maps[j], P, tv, ti, bits + j * nwd);
        __syncthreads();
        if (threadIdx.x == 0) lenj[j] = softdbsize;
      }
      int64_t lc = 0;
      for (int64_t w = threadIdx.x; w < nwd; w += FB) {
        uint32_t x = bits[w];
        for (int k = 1; k <= j; k++) x &= bits[k * nwd + w];
        lc += __builtin_popcount(x);
      }
      cnt = blk_sum_i64(lc, sh);
      if (cnt == 0) break;
    }
    if (cnt != 0) break;
    frac = fmin(frac + 0.1, 1.0);
  }
  // ordered compaction of the pattern database
  const int64_t per = (nwd + FB - 1) / FB;
  const int64_t w0 = threadIdx.x * per, w1 = (w0 + per < nwd) ? w0 + per : nwd;
  int64_t lc = 0;
  for (int64_t w = w0; w < w1; w++) {
    uint32_t x = bits[w];
    for (int k = 1; k < nsrc; k++) x &= bits[k * nwd + w];
    lc += __builtin_popcount(x);
  }
  int64_t nev;
  int64_t off = blk_excl_scan(lc, &nev, sh);
  for (int64_t w = w0; w < w1; w++) {
    uint32_t x = bits[w];
    for (int k = 1; k < nsrc; k++) x &= bits[k * nwd + w];
    while (x) {
      int b = __builtin_ctz(x);
      x &= x - 1;
      ev[off++] = (int32_t)(w * 32 + b);
    }
  }
  __syncthreads();
  return nev;
}

// Multi-source finalize: relaxation + tau model + draw (iqsim.jl:230-243).
// Source order: hard tile -> [hard, overlap, soft...], else [overlap, soft...].
extern "C" __global__ void __launch_bounds__(FB)
k_relax_finalize(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  __shared__ SelScratch sh;
  __shared__ double sv[IQ_SORT_MAX];
  __shared__ int32_t dp[IQ_SORT_MAX];
  __shared__ double s2[S2MAX];
  __shared__ const double* mp[IQ_MAX_AUX + 2];
  const IqQuery q = qs[blockIdx.x];
  if (!q.multi) return;
  const IqGeo& g = D->g;
  const IqStep& st = D->steps[q.step];
  const int64_t slot = q.slot;
  SelOut o{D->pick + slot, D->log + D->reals[q.real].log_off + (q.step - D->reals[q.real].step0), D->err,
           D->nband + slot};
  int nsrc = 0;
  if (threadIdx.x == 0) {
    int k = 0;
    if (st.hard) mp[k++] = D->maps + (slot * D->nmap + 1) * g.P;
    mp[k++] = D->maps + (slot * D->nmap + 0) * g.P;
    for (int a = 0; a < D->naux; a++) {
      mp[k++] = D->maps + (slot * D->nmap + 2 + a) * g.P;
    }
    sh.flag = k;
  }
  __syncthreads();
  nsrc = sh.flag;
  __syncthreads();
  int32_t* ev = D->ev + slot * g.P;
  const int64_t nev = relax_db(g.P, nsrc, mp, D->tol, D->bits + slot * D->nmap * D->nw32, ev, (uint64_t*)sv, sh);
  if (nev < 0) {
    if (threadIdx.x == 0) { set_err(D->err, IQ_KERR_NOCAND); *o.pick = -2; }
    return;
  }
  double* vals = D->vals + slot * g.P;
  const bool small = nev <= IQ_SORT_MAX;  // else ranks and weights in P-sized scratch
  double* p = small ? D->sd + slot * IQ_SORT_MAX : D->dexact + slot * g.P;
  int32_t* si = small ? D->si + slot * IQ_SORT_MAX : D->sup + slot * g.P;
  bool allconst;
  auto srcv = [&](int j, int64_t i) { return mp[j][ev[i]]; };
  if (!tau_weights(nev, nsrc, srcv, vals, si, p, &allconst, sv, dp, sh, D->rkey + slot * 2 * g.P)) {
    if (threadIdx.x == 0) { set_err(D->err, IQ_KERR_SORTMAX); *o.pick = -2; }
    return;
  }
  double* s1 = D->s1 + slot * D->s1_stride;
  if (nev == 1) publish_pick(ev, 1, nullptr, 1.0, q.u, q.forced, s1, s2, o, sh);
  else if (allconst) publish_pick(ev, nev, nullptr, tau_constant_weight(nev), q.u, q.forced, s1, s2, o, sh);
  else publish_pick(ev, nev, p, 0.0, q.u, q.forced, s1, s2, o, sh);
}

// ---------------------------------------------------------------------------
// Standalone entry points (iq_band_pick / iq_relaxation / iq_taumodel / iq_sample)
// ---------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(FB)
k_band_pick_host(const double* __restrict__ Dm, int64_t n, double tol, int32_t* __restrict__ out,
                 int64_t* __restrict__ count) {
  __shared__ SelScratch sh;
  double lm = __builtin_inf();
  for (int64_t i = threadIdx.x; i < n; i += FB) lm = fmin(lm, Dm[i]);
  const double m = blk_min_f64(lm, sh);
  const double thr = (1.0 + tol) * m;
  const int64_t per = (n + FB - 1) / FB;
  const int64_t i0 = threadIdx.x * per, i1 = (i0 + per < n) ? i0 + per : n;
  int64_t c = 0;
  for (int64_t i = i0; i < i1; i++) c += Dm[i] <= thr;
  int64_t tot;
  int64_t off = blk_excl_scan(c, &tot, sh);
  for (int64_t i = i0; i < i1; i++)
    if (Dm[i] <= thr) out[off++] = (int32_t)i;
  if (threadIdx.x == 0) *count = tot;
}

extern "C" __global__ void __launch_bounds__(FB)
k_relax_host(int64_t P, int nsrc, const double* const* __restrict__ maps, double tol, uint32_t* __restrict__ bits,
             int32_t* __restrict__ ev, int64_t* __restrict__ count, int32_t* __restrict__ err) {
  __shared__ SelScratch sh;
  __shared__ uint64_t ks[KS_LDS_WORDS];
  const int64_t n = relax_db(P, nsrc, maps, tol, bits, ev, ks, sh);
  if (threadIdx.x == 0) {
    if (n < 0) set_err(err, IQ_KERR_NOCAND);
    *count = n < 0 ? 0 : n;
  }
}

extern "C" __global__ void __launch_bounds__(FB)
k_taumodel_host(const int32_t* __restrict__ ev, int64_t nev, int nsrc, const double* const* __restrict__ maps,
                double* __restrict__ vals, int32_t* __restrict__ rank, double* __restrict__ p,
                unsigned long long* __restrict__ keys, int32_t* __restrict__ err) {
  __shared__ SelScratch sh;
  __shared__ double sv[IQ_SORT_MAX];
  __shared__ int32_t dp[IQ_SORT_MAX];
  if (nev == 1) { if (threadIdx.x == 0) p[0] = 1.0; return; }  // taumodel.jl:9
  for (int64_t i = threadIdx.x; i < nev; i += FB) p[i] = 1.0;
  __syncthreads();
  for (int j = 0; j < nsrc; j++) {
    for (int64_t i = threadIdx.x; i < nev; i += FB) vals[i] = maps[j][ev[i]];
    __syncthreads();
    bool c;
    if (!blk_dense_rank(vals, nev, rank, &c, sv, dp, sh, keys)) {
      if (threadIdx.x == 0) set_err(err, IQ_KERR_SORTMAX);
      return;
    }
    tau_accumulate(nev, rank, c, p, sh);
  }
  tau_finish(nev, p);
}

extern "C" __global__ void __launch_bounds__(FB)
k_sample_host(const double* __restrict__ p, int64_t n, double u, double* __restrict__ s1, int64_t* __restrict__ out) {
  __shared__ SelScratch sh;
  __shared__ double s2[S2MAX];
  int64_t i = blk_sample(p, 0.0, n, u, s1, s2, sh);
  if (threadIdx.x == 0) *out = i;
}
