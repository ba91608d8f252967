/*
 * iqref.c -- FP64 CPU restatement of the ImageQuilting.jl `iqsim` hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP library is
 * checked against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the product (libiqhip.so) never links or calls
 * it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Arrays are column-major (first axis
 * fastest), always viewed as 3-D (missing trailing axes have extent 1), and
 * every index is 0-based (the reference is 1-based).
 *
 * Arithmetic contract (what "exact" means everywhere in this project):
 *   * a masked sliding distance is
 *         D[c] = | (A2[c] - 2*AB[c]) + B2 |                 (utils.jl:12)
 *     where A2, AB and B2 are fused-multiply-add chains over the masked tile
 *     voxels in ascending column-major order:
 *         A2 = fma(t_j, t_j, A2), AB = fma(t_j, k_j, AB), B2 = fma(k_j, k_j, B2)
 *     (the reference evaluates the two correlations by FFT, imfilter.jl:5-26;
 *     direct summation in a fixed order replaces the FFT's rounding noise);
 *   * every other floating-point step follows the reference's operator order
 *     with no contraction (compile with -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_OK 0
#define ORA_ERR_NOCAND (-3)   /* every candidate disabled (relaxation.jl:7-36 would loop forever) */
#define ORA_ERR_NOTINDB (-4)  /* a forced choice is not in the pattern database */
#define ORA_ERR_ITER (-5)     /* push-relabel iteration guard hit */
#define ORA_ERR_ARG (-6)

/* ------------------------------------------------------------------ */
/* small helpers                                                        */
/* ------------------------------------------------------------------ */
static inline int64_t lin3(const int64_t d[3], int64_t i, int64_t j, int64_t k) {
  return i + d[0] * (j + d[1] * k);
}

int ora_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void ora_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* A5  imfilter_cpu / imfilter_gpu  (imfilter.jl:5-26)                  */
/* valid-region cross-correlation out[c] = sum_j img[c+j] * krn[j]      */
/* ------------------------------------------------------------------ */
void ora_imfilter(const double* img, const int64_t idims[3], const double* krn,
                  const int64_t kdims[3], double* out) {
  int64_t od[3] = {idims[0] - kdims[0] + 1, idims[1] - kdims[1] + 1, idims[2] - kdims[2] + 1};
  int64_t P = od[0] * od[1] * od[2];
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < P; c++) {
    int64_t cx = c % od[0], cy = (c / od[0]) % od[1], cz = c / (od[0] * od[1]);
    double acc = 0.0;
    for (int64_t jz = 0; jz < kdims[2]; jz++)
      for (int64_t jy = 0; jy < kdims[1]; jy++)
        for (int64_t jx = 0; jx < kdims[0]; jx++)
          acc = fma(img[lin3(idims, cx + jx, cy + jy, cz + jz)], krn[lin3(kdims, jx, jy, jz)], acc);
    out[c] = acc;
  }
}

/* ------------------------------------------------------------------ */
/* A4  fastdistance  (utils.jl:5-13)                                    */
/* masked offsets are visited in ascending column-major tile order      */
/* ------------------------------------------------------------------ */
static int64_t masked_offsets(const uint8_t* w, const int64_t kd[3], int64_t* jx, int64_t* jy, int64_t* jz) {
  int64_t n = 0;
  for (int64_t z = 0; z < kd[2]; z++)
    for (int64_t y = 0; y < kd[1]; y++)
      for (int64_t x = 0; x < kd[0]; x++)
        if (!w || w[lin3(kd, x, y, z)]) { jx[n] = x; jy[n] = y; jz[n] = z; n++; }
  return n;
}

/* kern values at the masked offsets; B2 as an fma chain (utils.jl:10) */
static double b2_of(const double* kv, int64_t V) {
  double b2 = 0.0;
  for (int64_t j = 0; j < V; j++) b2 = fma(kv[j], kv[j], b2);
  return b2;
}

/* D over all candidates for one masked query.  Shared by ora_fastdistance and
 * the simulate loop.  kv[j] = kern value at the j-th masked offset. */
static void distance_map(const double* img, const int64_t idims[3], const int64_t kdims[3],
                         const int64_t* jx, const int64_t* jy, const int64_t* jz,
                         const double* kv, int64_t V, double* out) {
  int64_t od[3] = {idims[0] - kdims[0] + 1, idims[1] - kdims[1] + 1, idims[2] - kdims[2] + 1};
  int64_t P = od[0] * od[1] * od[2];
  double b2 = b2_of(kv, V);
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (V > 0 ? V : 1));
  for (int64_t j = 0; j < V; j++) off[j] = lin3(idims, jx[j], jy[j], jz[j]);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < P; c++) {
    int64_t cx = c % od[0], cy = (c / od[0]) % od[1], cz = c / (od[0] * od[1]);
    const double* base = img + lin3(idims, cx, cy, cz);
    double a2 = 0.0, ab = 0.0;
    for (int64_t j = 0; j < V; j++) {
      double t = base[off[j]];
      a2 = fma(t, t, a2);
      ab = fma(t, kv[j], ab);
    }
    out[c] = fabs((a2 - 2.0 * ab) + b2);
  }
  free(off);
}

void ora_fastdistance(const double* img, const int64_t idims[3], const double* kern,
                      const uint8_t* weights, const int64_t kdims[3], double* out) {
  int64_t nk = kdims[0] * kdims[1] * kdims[2];
  int64_t* jx = (int64_t*)malloc(sizeof(int64_t) * nk * 3);
  int64_t *jy = jx + nk, *jz = jx + 2 * nk;
  double* kv = (double*)malloc(sizeof(double) * nk);
  int64_t V = masked_offsets(weights, kdims, jx, jy, jz);
  for (int64_t j = 0; j < V; j++) kv[j] = kern[lin3(kdims, jx[j], jy[j], jz[j])];
  distance_map(img, idims, kdims, jx, jy, jz, kv, V, out);
  free(kv);
  free(jx);
}

/* ------------------------------------------------------------------ */
/* A10  tol-band pick  (iqsim.jl:237, empty Ds branch)                  */
/*   patterndb = findall(vec(D .<= (1 + tol) * minimum(D)))             */
/* ------------------------------------------------------------------ */
int64_t ora_band(const double* D, int64_t n, double tol, int64_t* out) {
  double m = INFINITY;
  for (int64_t i = 0; i < n; i++)
    if (D[i] < m) m = D[i];
  double thr = (1.0 + tol) * m;
  int64_t k = 0;
  for (int64_t i = 0; i < n; i++)
    if (D[i] <= thr) out[k++] = i;
  return k;
}

/* ------------------------------------------------------------------ */
/* A11  relaxation  (relaxation.jl:5-48), restated literally            */
/* ------------------------------------------------------------------ */
typedef struct { double v; int64_t i; } vi_t;

static int cmp_vi(const void* a, const void* b) {
  const vi_t* x = (const vi_t*)a;
  const vi_t* y = (const vi_t*)b;
  if (x->v < y->v) return -1;
  if (x->v > y->v) return 1;
  return (x->i < y->i) ? -1 : (x->i > y->i);
}

/* partialsortperm(v, 1:k): indices of the k smallest by (value, index)
 * (Julia's Perm ordering breaks ties by index). */
static void partialsortperm(const double* v, int64_t n, int64_t k, int64_t* out, vi_t* scratch) {
  for (int64_t i = 0; i < n; i++) { scratch[i].v = v[i]; scratch[i].i = i; }
  qsort(scratch, (size_t)n, sizeof(vi_t), cmp_vi);
  for (int64_t i = 0; i < k; i++) out[i] = scratch[i].i;
}

/* fastintersect (relaxation.jl:41-48): ascending indices present in both */
static int64_t fastintersect(const int64_t* A, int64_t na, const int64_t* B, int64_t nb, int64_t nbits,
                             uint8_t* bitsA, uint8_t* bitsB, int64_t* out) {
  memset(bitsA, 0, (size_t)nbits);
  memset(bitsB, 0, (size_t)nbits);
  for (int64_t i = 0; i < na; i++) bitsA[A[i]] = 1;
  for (int64_t i = 0; i < nb; i++) bitsB[B[i]] = 1;
  int64_t k = 0;
  for (int64_t i = 0; i < nbits; i++)
    if (bitsA[i] & bitsB[i]) out[k++] = i;
  return k;
}

int64_t ora_relaxation(const double* D, const double* const* Ds, int naux, int64_t n, double tol,
                       int64_t* out) {
  /* relaxation.jl:7-11 */
  int64_t npat = 0;
  int allzero = 1;
  for (int64_t i = 0; i < n; i++)
    if (!isinf(D[i])) {
      npat++;
      if (D[i] != 0.0) allzero = 0;
    }
  if (npat == 0) return ORA_ERR_NOCAND; /* the reference never terminates here */
  int64_t dbsize = allzero ? npat : (int64_t)ceil(tol * (double)npat);
  vi_t* scratch = (vi_t*)malloc(sizeof(vi_t) * n);
  int64_t* overlapdb = (int64_t*)malloc(sizeof(int64_t) * (dbsize > 0 ? dbsize : 1));
  partialsortperm(D, n, dbsize, overlapdb, scratch);

  /* relaxation.jl:14-18 */
  int64_t** softdb = (int64_t**)calloc((size_t)(naux > 0 ? naux : 1), sizeof(int64_t*));
  int64_t* softlen = (int64_t*)calloc((size_t)(naux > 0 ? naux : 1), sizeof(int64_t));
  double** softdist = (double**)calloc((size_t)(naux > 0 ? naux : 1), sizeof(double*));
  for (int a = 0; a < naux; a++) {
    softdb[a] = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
    softdist[a] = (double*)malloc(sizeof(double) * n);
    memcpy(softdist[a], Ds[a], sizeof(double) * n);
  }
  int64_t* pat = (int64_t*)malloc(sizeof(int64_t) * (dbsize > 0 ? dbsize : 1));
  int64_t* tmp = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
  uint8_t* bA = (uint8_t*)malloc((size_t)n);
  uint8_t* bB = (uint8_t*)malloc((size_t)n);
  int64_t npat_db = 0;

  /* relaxation.jl:19-36 */
  double frac = 0.1 * ((double)dbsize / (double)npat);
  for (;;) {
    int64_t softdbsize = (int64_t)ceil(frac * (double)npat);
    memcpy(pat, overlapdb, sizeof(int64_t) * dbsize);
    npat_db = dbsize;
    for (int a = 0; a < naux; a++) {
      for (int64_t i = 0; i < softlen[a]; i++) softdist[a][softdb[a][i]] = INFINITY;
      int64_t grow = softdbsize - softlen[a];
      if (grow > 0) {
        partialsortperm(softdist[a], n, grow, tmp, scratch);
        memcpy(softdb[a] + softlen[a], tmp, sizeof(int64_t) * grow);
        softlen[a] += grow;
      }
      npat_db = fastintersect(pat, npat_db, softdb[a], softlen[a], n, bA, bB, tmp);
      memcpy(pat, tmp, sizeof(int64_t) * npat_db);
      if (npat_db == 0) break;
    }
    if (npat_db != 0) break;
    frac = fmin(frac + 0.1, 1.0);
  }
  /* with no auxiliary source the loop above returns overlapdb in sort order;
   * the reference only reaches relaxation with naux >= 1 (iqsim.jl:237) */
  memcpy(out, pat, sizeof(int64_t) * npat_db);

  for (int a = 0; a < naux; a++) { free(softdb[a]); free(softdist[a]); }
  free(softdb); free(softlen); free(softdist);
  free(pat); free(tmp); free(bA); free(bB); free(overlapdb); free(scratch);
  return npat_db;
}

/* ------------------------------------------------------------------ */
/* A12  taumodel  (taumodel.jl:5-45)                                    */
/* D1 / Ds are full candidate maps; events index into them.             */
/* ------------------------------------------------------------------ */
void ora_taumodel(const int64_t* events, int64_t nev, const double* D1, const double* const* Ds,
                  int naux, double* p) {
  if (nev == 1) { p[0] = 1.0; return; } /* taumodel.jl:9 */
  int ns = 1 + naux;
  double* R = (double*)malloc(sizeof(double) * nev * ns);
  vi_t* sc = (vi_t*)malloc(sizeof(vi_t) * nev);
  for (int j = 0; j < ns; j++) {
    const double* src = (j == 0) ? D1 : Ds[j - 1];
    /* sortperm (stable) then dense ranks, taumodel.jl:21-31 */
    for (int64_t i = 0; i < nev; i++) { sc[i].v = src[events[i]]; sc[i].i = i; }
    qsort(sc, (size_t)nev, sizeof(vi_t), cmp_vi);
    double r = 0.0, prev = -INFINITY;
    for (int64_t q = 0; q < nev; q++) {
      if (sc[q].v > prev) r += 1.0;
      prev = sc[q].v;
      R[sc[q].i + nev * j] = r;
    }
  }
  /* taumodel.jl:34-44 */
  double nd = (double)nev;
  double x0 = (1.0 - 1.0 / nd) / (1.0 / nd);
  for (int64_t i = 0; i < nev; i++) p[i] = 1.0;
  for (int j = 0; j < ns; j++) {
    double S = 0.0;
    for (int64_t i = 0; i < nev; i++) S += (nd - R[i + nev * j]) + 1.0;
    for (int64_t i = 0; i < nev; i++) {
      double P = ((nd - R[i + nev * j]) + 1.0) / S;
      double X = (1.0 - P) / P;
      p[i] = p[i] * (X / x0); /* prod(X / x0, dims=2), left to right from 1.0 */
    }
  }
  for (int64_t i = 0; i < nev; i++) {
    double x = x0 * p[i];
    p[i] = 1.0 / (1.0 + x);
  }
  free(R);
  free(sc);
}

/* ------------------------------------------------------------------ */
/* A13  weighted draw  (iqsim.jl:243, StatsBase direct method)          */
/* Three-level blocked cumulative sums (block 64) fix the summation    */
/* order: t = u * W; the first index whose cumulative weight reaches t. */
/* ------------------------------------------------------------------ */
#define ORA_SB 64
int64_t ora_sample(const double* p, int64_t n, double u) {
  if (n <= 0) return -1;
  int64_t n1 = (n + ORA_SB - 1) / ORA_SB, n2 = (n1 + ORA_SB - 1) / ORA_SB;
  double* s1 = (double*)malloc(sizeof(double) * n1);
  double* s2 = (double*)malloc(sizeof(double) * n2);
  for (int64_t k = 0; k < n1; k++) {
    double s = 0.0;
    for (int64_t i = k * ORA_SB; i < n && i < (k + 1) * ORA_SB; i++) s += p[i];
    s1[k] = s;
  }
  for (int64_t m = 0; m < n2; m++) {
    double s = 0.0;
    for (int64_t k = m * ORA_SB; k < n1 && k < (m + 1) * ORA_SB; k++) s += s1[k];
    s2[m] = s;
  }
  double W = 0.0;
  for (int64_t m = 0; m < n2; m++) W += s2[m];
  double t = u * W;
  /* descend */
  double acc = 0.0;
  int64_t m = 0;
  for (;; m++) {
    double nx = acc + s2[m];
    if (nx >= t || m == n2 - 1) break;
    acc = nx;
  }
  int64_t k = m * ORA_SB, kend = (m + 1) * ORA_SB < n1 ? (m + 1) * ORA_SB : n1;
  for (;; k++) {
    double nx = acc + s1[k];
    if (nx >= t || k == kend - 1) break;
    acc = nx;
  }
  int64_t i = k * ORA_SB, iend = (k + 1) * ORA_SB < n ? (k + 1) * ORA_SB : n;
  for (;; i++) {
    double nx = acc + p[i];
    if (nx >= t || i == iend - 1) break;
    acc = nx;
  }
  free(s1);
  free(s2);
  return i;
}

/* ------------------------------------------------------------------ */
/* A16  graphcut  (graphcut.jl:5-84)                                    */
/* ------------------------------------------------------------------ */
/* Edge capacity between slab voxels u and v = u + e_d (graphcut.jl:22-54);
 * `wok` says whether w = v + e_d is inside the slab.  For Float32 slabs
 * the differences and the four-term gradient sum are Float32 and `+ eps()`
 * promotes to Float64 (graphcut.jl:52). */
static double edge_cap(const double* A, const double* B, int is_f32, int64_t u, int64_t v, int64_t w, int wok) {
  if (is_f32) {
    float Au = (float)A[u], Av = (float)A[v], Bu = (float)B[u], Bv = (float)B[v];
    float Du = fabsf(Au - Bu), Dv = fabsf(Av - Bv);
    float gAu = fabsf(Av - Au), gBu = fabsf(Bv - Bu);
    float gAv = wok ? fabsf((float)A[w] - Av) : gAu;
    float gBv = wok ? fabsf((float)B[w] - Bv) : gBu;
    float num = Du + Dv;
    float g = ((gAu + gAv) + gBu) + gBv;
    return (double)num / ((double)g + DBL_EPSILON);
  } else {
    double Du = fabs(A[u] - B[u]), Dv = fabs(A[v] - B[v]);
    double gAu = fabs(A[v] - A[u]), gBu = fabs(B[v] - B[u]);
    double gAv = wok ? fabs(A[w] - A[v]) : gAu;
    double gBv = wok ? fabs(B[w] - B[v]) : gBu;
    return (Du + Dv) / ((((gAu + gAv) + gBu) + gBv) + DBL_EPSILON);
  }
}

/* Capacity of the slab edge (p, p + e_d), p given by its slab linear index. */
static double slab_cap(const double* A, const double* B, int is_f32, const int64_t s[3], int64_t pu, int d) {
  int64_t c[3] = {pu % s[0], (pu / s[0]) % s[1], pu / (s[0] * s[1])};
  int64_t st = (d == 0) ? 1 : (d == 1 ? s[0] : s[0] * s[1]);
  int wok = (c[d] + 2) < s[d];
  return edge_cap(A, B, is_f32, pu, pu + st, wok ? pu + 2 * st : 0, wok);
}

/*
 * Minimum cut with the reference's labelling: mask = V \ {voxels that can
 * reach the sink in the residual graph of a maximum flow} (graphcut.jl:73-81;
 * the BK labels 0 "free" and 1 "source tree" map to true).  That set does not
 * depend on which maximum flow is found.
 *
 * The first slice along `dim` is tied to the source and the last slice to the
 * sink with infinite capacity (graphcut.jl:56-70), so both slices are
 * contracted into the terminals and a preflow push-relabel runs on the
 * interior slices.  The schedule is synchronous and deterministic (relabel,
 * push, gather phases; global relabel every `gr_freq` rounds; gr_freq <= 0
 * takes the kernel's rule: 8 rounds, 24 above 1,024 interior nodes), the
 * same schedule the HIP kernel executes, so the two agree bit for bit.
 */
int ora_graphcut(const double* A, const double* B, int is_f32, const int64_t s[3], int dim, int gr_freq,
                 uint8_t* mask, int64_t* iters_out) {
  int64_t nvox = s[0] * s[1] * s[2];
  int64_t L = s[dim];
  if (iters_out) *iters_out = 0;
  if (L < 2) return ORA_ERR_ARG;
  for (int64_t p = 0; p < nvox; p++) {
    int64_t c[3] = {p % s[0], (p / s[0]) % s[1], p / (s[0] * s[1])};
    mask[p] = (c[dim] == L - 1) ? 0 : 1;
  }
  if (L == 2) return ORA_OK; /* no interior: every slice-0 -> slice-1 edge is saturated */
  int64_t si[3] = {s[0], s[1], s[2]};
  si[dim] = L - 2;
  int64_t n = si[0] * si[1] * si[2];
  const int64_t HMAX = n + 1;
  if (gr_freq <= 0) gr_freq = n > 1024 ? 24 : 8;
  int64_t sst[3] = {1, s[0], s[0] * s[1]};
  int64_t ist[3] = {1, si[0], si[0] * si[1]};
  double* e = (double*)calloc((size_t)n, sizeof(double));
  double* rt = (double*)calloc((size_t)n, sizeof(double));
  int64_t* h = (int64_t*)malloc(sizeof(int64_t) * n);
  int64_t* hn = (int64_t*)malloc(sizeof(int64_t) * n);
  double* rf = (double*)calloc((size_t)(3 * n), sizeof(double));
  double* rb = (double*)calloc((size_t)(3 * n), sizeof(double));
  double* pr = (double*)calloc((size_t)(3 * n), sizeof(double));
  uint8_t* hasp = (uint8_t*)calloc((size_t)(3 * n), 1); /* plus-edge exists */
  uint8_t* hast = (uint8_t*)calloc((size_t)n, 1);
  int rc = ORA_OK;

  /* build: interior node u <-> slab voxel p with coordinate[dim] + 1 */
  for (int64_t u = 0; u < n; u++) {
    int64_t ic[3] = {u % si[0], (u / si[0]) % si[1], u / (si[0] * si[1])};
    int64_t pc[3] = {ic[0], ic[1], ic[2]};
    pc[dim] += 1;
    int64_t p = pc[0] + s[0] * (pc[1] + s[1] * pc[2]);
    for (int d = 0; d < 3; d++) {
      if (ic[d] + 1 < si[d]) {
        double c = slab_cap(A, B, is_f32, s, p, d);
        hasp[3 * u + d] = 1;
        rf[3 * u + d] = c;
        rb[3 * u + d] = c;
      }
    }
    if (pc[dim] == 1) e[u] += slab_cap(A, B, is_f32, s, p - sst[dim], dim); /* saturated source edge */
    if (pc[dim] == L - 2) { hast[u] = 1; rt[u] = slab_cap(A, B, is_f32, s, p, dim); }
  }

  int64_t total = 0;
  const int64_t LIMIT = 200000; /* same guard as the HIP kernel */
  for (;;) {
    /* global relabel: exact residual distance to the sink */
    for (int64_t u = 0; u < n; u++) h[u] = (hast[u] && rt[u] > 0.0) ? 1 : HMAX;
    for (int64_t lev = 1;; lev++) {
      int changed = 0;
      for (int64_t u = 0; u < n; u++) {
        if (h[u] != HMAX) continue;
        int64_t ic[3] = {u % si[0], (u / si[0]) % si[1], u / (si[0] * si[1])};
        for (int d = 0; d < 3 && h[u] == HMAX; d++) {
          if (ic[d] > 0) {
            int64_t v = u - ist[d];
            if (rb[3 * v + d] > 0.0 && h[v] == lev) h[u] = lev + 1;
          }
          if (h[u] == HMAX && hasp[3 * u + d]) {
            int64_t v = u + ist[d];
            if (rf[3 * u + d] > 0.0 && h[v] == lev) h[u] = lev + 1;
          }
        }
        if (h[u] == lev + 1) changed = 1;
      }
      if (!changed) break;
    }
    int any = 0;
    for (int64_t u = 0; u < n; u++)
      if (e[u] > 0.0 && h[u] < HMAX) { any = 1; break; }
    if (!any) break;
    if (total > LIMIT) { rc = ORA_ERR_ITER; break; }

    for (int it = 0; it < gr_freq; it++) {
      /* relabel phase */
      for (int64_t u = 0; u < n; u++) {
        hn[u] = h[u];
        if (!(e[u] > 0.0 && h[u] < HMAX)) continue;
        int64_t ic[3] = {u % si[0], (u / si[0]) % si[1], u / (si[0] * si[1])};
        int adm = 0;
        int64_t mh = HMAX;
        if (hast[u] && rt[u] > 0.0) { if (h[u] == 1) adm = 1; mh = 0; }
        for (int d = 0; d < 3; d++) {
          if (ic[d] > 0) {
            int64_t v = u - ist[d];
            if (rb[3 * v + d] > 0.0) { if (h[u] == h[v] + 1) adm = 1; if (h[v] < mh) mh = h[v]; }
          }
          if (hasp[3 * u + d]) {
            int64_t v = u + ist[d];
            if (rf[3 * u + d] > 0.0) { if (h[u] == h[v] + 1) adm = 1; if (h[v] < mh) mh = h[v]; }
          }
        }
        if (!adm) hn[u] = (mh + 1 < HMAX) ? mh + 1 : HMAX;
      }
      memcpy(h, hn, sizeof(int64_t) * n);
      /* push phase */
      for (int64_t u = 0; u < n; u++) {
        if (!(e[u] > 0.0 && h[u] < HMAX)) continue;
        int64_t ic[3] = {u % si[0], (u / si[0]) % si[1], u / (si[0] * si[1])};
        double ex = e[u];
        if (hast[u] && rt[u] > 0.0 && h[u] == 1) {
          double dl = ex < rt[u] ? ex : rt[u];
          ex -= dl;
          rt[u] -= dl;
        }
        for (int d = 0; d < 3 && ex > 0.0; d++) {
          if (ic[d] > 0) {
            int64_t v = u - ist[d];
            if (h[u] == h[v] + 1 && rb[3 * v + d] > 0.0) {
              double dl = ex < rb[3 * v + d] ? ex : rb[3 * v + d];
              ex -= dl;
              rb[3 * v + d] -= dl;
              rf[3 * v + d] += dl;
              pr[3 * v + d] = -dl;
            }
          }
          if (ex > 0.0 && hasp[3 * u + d]) {
            int64_t v = u + ist[d];
            if (h[u] == h[v] + 1 && rf[3 * u + d] > 0.0) {
              double dl = ex < rf[3 * u + d] ? ex : rf[3 * u + d];
              ex -= dl;
              rf[3 * u + d] -= dl;
              rb[3 * u + d] += dl;
              pr[3 * u + d] = dl;
            }
          }
        }
        e[u] = ex;
      }
      /* gather phase */
      for (int64_t u = 0; u < n; u++) {
        int64_t ic[3] = {u % si[0], (u / si[0]) % si[1], u / (si[0] * si[1])};
        double ex = e[u];
        for (int d = 0; d < 3; d++) {
          if (ic[d] > 0) {
            int64_t v = u - ist[d];
            if (pr[3 * v + d] > 0.0) ex += pr[3 * v + d];
          }
          if (hasp[3 * u + d] && pr[3 * u + d] < 0.0) ex += -pr[3 * u + d];
        }
        e[u] = ex;
      }
      memset(pr, 0, sizeof(double) * 3 * n);
      total++;
      int act = 0;
      for (int64_t u = 0; u < n; u++)
        if (e[u] > 0.0 && h[u] < HMAX) { act = 1; break; }
      if (!act) break;
    }
  }

  for (int64_t u = 0; u < n; u++) {
    int64_t ic[3] = {u % si[0], (u / si[0]) % si[1], u / (si[0] * si[1])};
    int64_t pc[3] = {ic[0], ic[1], ic[2]};
    pc[dim] += 1;
    mask[pc[0] + s[0] * (pc[1] + s[1] * pc[2])] = (h[u] < HMAX) ? 0 : 1;
  }
  if (iters_out) *iters_out = total;
  free(e); free(rt); free(h); free(hn); free(rf); free(rb); free(pr); free(hasp); free(hast);
  return rc;
}

/*
 * Independent cross-check of ora_graphcut: Edmonds-Karp max-flow on the
 * uncontracted graph of graphcut.jl:18-70 (terminal edges of capacity +Inf),
 * then the complement of the residual co-reachability of the sink.
 */
int ora_graphcut_ek(const double* A, const double* B, int is_f32, const int64_t s[3], int dim, uint8_t* mask) {
  int64_t nvox = s[0] * s[1] * s[2];
  int64_t N = nvox + 2, S = nvox, T = nvox + 1;
  /* adjacency: each voxel has <= 6 lattice arcs + terminal arcs */
  int64_t maxarcs = nvox * 6 + 2 * nvox + 4;
  int64_t* head = (int64_t*)malloc(sizeof(int64_t) * N);
  int64_t* nxt = (int64_t*)malloc(sizeof(int64_t) * maxarcs * 2);
  int64_t* to = (int64_t*)malloc(sizeof(int64_t) * maxarcs * 2);
  double* cap = (double*)malloc(sizeof(double) * maxarcs * 2);
  int64_t na = 0;
  for (int64_t i = 0; i < N; i++) head[i] = -1;
#define ADD_ARC(a, b, c1, c2)                                   \
  do {                                                          \
    to[na] = (b); cap[na] = (c1); nxt[na] = head[a]; head[a] = na; na++; \
    to[na] = (a); cap[na] = (c2); nxt[na] = head[b]; head[b] = na; na++; \
  } while (0)
  int64_t st[3] = {1, s[0], s[0] * s[1]};
  for (int d = 0; d < 3; d++) {
    for (int64_t p = 0; p < nvox; p++) {
      int64_t c[3] = {p % s[0], (p / s[0]) % s[1], p / (s[0] * s[1])};
      if (c[d] + 1 >= s[d]) continue;
      double cc = slab_cap(A, B, is_f32, s, p, d);
      ADD_ARC(p, p + st[d], cc, cc);
    }
  }
  for (int64_t p = 0; p < nvox; p++) {
    int64_t c[3] = {p % s[0], (p / s[0]) % s[1], p / (s[0] * s[1])};
    if (c[dim] == 0) ADD_ARC(S, p, INFINITY, 0.0);
    if (c[dim] == s[dim] - 1) ADD_ARC(p, T, INFINITY, 0.0);
  }
#undef ADD_ARC
  int64_t* q = (int64_t*)malloc(sizeof(int64_t) * N);
  int64_t* par = (int64_t*)malloc(sizeof(int64_t) * N);
  for (;;) {
    for (int64_t i = 0; i < N; i++) par[i] = -2;
    par[S] = -1;
    int64_t qh = 0, qt = 0;
    q[qt++] = S;
    while (qh < qt && par[T] == -2) {
      int64_t x = q[qh++];
      for (int64_t a = head[x]; a >= 0; a = nxt[a])
        if (cap[a] > 0.0 && par[to[a]] == -2) { par[to[a]] = a; q[qt++] = to[a]; }
    }
    if (par[T] == -2) break;
    double f = INFINITY;
    for (int64_t x = T; x != S; x = to[par[x] ^ 1]) if (cap[par[x]] < f) f = cap[par[x]];
    if (isinf(f)) break; /* an all-infinite path cannot occur for a slab with >= 2 slices */
    for (int64_t x = T; x != S; x = to[par[x] ^ 1]) {
      int64_t a = par[x];
      cap[a] -= f;
      if (!isinf(cap[a ^ 1])) cap[a ^ 1] += f;
    }
  }
  /* co-reachability of T: reverse BFS over arcs with residual capacity */
  uint8_t* co = (uint8_t*)calloc((size_t)N, 1);
  int64_t qh = 0, qt = 0;
  q[qt++] = T;
  co[T] = 1;
  while (qh < qt) {
    int64_t x = q[qh++];
    for (int64_t a = head[x]; a >= 0; a = nxt[a]) {
      /* arc a goes x -> to[a]; its twin a^1 goes to[a] -> x */
      int64_t y = to[a];
      if (!co[y] && cap[a ^ 1] > 0.0) { co[y] = 1; q[qt++] = y; }
    }
  }
  for (int64_t p = 0; p < nvox; p++) mask[p] = co[p] ? 0 : 1;
  free(co); free(q); free(par); free(head); free(nxt); free(to); free(cap);
  return ORA_OK;
}

/* ------------------------------------------------------------------ */
/* A1  tile-step driver  (iqsim.jl:163-300)                             */
/* ------------------------------------------------------------------ */
typedef struct {
  int64_t tisize[3], tilesize[3], ovlsize[3], spacing[3], ntiles[3], padsize[3], distsize[3];
  const double* ti;          /* prepared TI (NaN -> 0), tisize */
  int ti_is_f32;             /* element type of the prepared TI is Float32 */
  const uint8_t* disabled;   /* distsize */
  const int64_t* path;       /* tile linear indices (0-based) */
  int64_t npath;
  const uint8_t* skipped;    /* prod(ntiles) */
  const double* hardgrid;    /* padsize, NaN where no hard datum, or NULL */
  int naux;
  const double* const* aux;    /* naux arrays of padsize (prepared, padded) */
  const double* const* auxti;  /* naux arrays of tisize (prepared) */
  double tol;
  int64_t nreal;
  const double* uniforms;    /* nreal x npath */
  const int64_t* forced;     /* nreal x npath, -1 = draw from uniforms; or NULL */
  int gr_freq;
  double* out;               /* nreal x padsize */
  int64_t* chosen;           /* nreal x npath, -1 for skipped */
  double* cutgrid;           /* nreal x padsize or NULL (debug) */
  int64_t* cut_iters;        /* total push-relabel rounds, or NULL */
  int64_t* nband;            /* nreal x npath pattern-database sizes (0 for skipped), or NULL */
  uint8_t* step_cut;         /* nreal x npath x tilevox cut masks per step (1 = keep the grid), or NULL */
  double* step_before;       /* nreal x npath x tilevox grid values of the tile before the paste, or NULL */
  int mutate;                /* mistakes planted on purpose by tests of the checks themselves:
                                0 none, 1 next slabs not negated, 2 prev slabs one voxel further in */
} ora_sim_args;

static void tile_offsets(const uint8_t* m, const int64_t ts[3], int64_t* jx, int64_t* jy, int64_t* jz, int64_t* V) {
  *V = masked_offsets(m, ts, jx, jy, jz);
}

int ora_simulate(const ora_sim_args* a) {
  const int64_t* ts = a->tilesize;
  int64_t tvox = ts[0] * ts[1] * ts[2];
  int64_t P = a->distsize[0] * a->distsize[1] * a->distsize[2];
  int64_t padvox = a->padsize[0] * a->padsize[1] * a->padsize[2];
  int64_t nt = a->ntiles[0] * a->ntiles[1] * a->ntiles[2];
  int nsrc_max = 2 + a->naux;
  uint8_t* ovlmask = (uint8_t*)malloc((size_t)tvox);
  uint8_t* hardmask = (uint8_t*)malloc((size_t)tvox);
  uint8_t* cutmask = (uint8_t*)malloc((size_t)tvox);
  uint8_t* pasted = (uint8_t*)malloc((size_t)nt);
  int64_t* jx = (int64_t*)malloc(sizeof(int64_t) * tvox * 3);
  int64_t *jy = jx + tvox, *jz = jx + 2 * tvox;
  double* kv = (double*)malloc(sizeof(double) * tvox);
  double** maps = (double**)malloc(sizeof(double*) * nsrc_max);
  for (int i = 0; i < nsrc_max; i++) maps[i] = (double*)malloc(sizeof(double) * P);
  double* ovldist = maps[0];
  double* harddist = maps[1];
  double** softdists = maps + 2;
  int64_t* db = (int64_t*)malloc(sizeof(int64_t) * P);
  double* pw = (double*)malloc(sizeof(double) * P);
  double* Aslab = (double*)malloc(sizeof(double) * tvox);
  double* Bslab = (double*)malloc(sizeof(double) * tvox);
  uint8_t* smask = (uint8_t*)malloc((size_t)tvox);
  const double** Dsp = (const double**)malloc(sizeof(double*) * (1 + a->naux));
  int rc = ORA_OK;
  int64_t cut_total = 0;

  for (int64_t r = 0; r < a->nreal && rc == ORA_OK; r++) {
    double* sim = a->out + r * padvox;
    double* cg = a->cutgrid ? a->cutgrid + r * padvox : NULL;
    memset(sim, 0, sizeof(double) * padvox);                       /* iqsim.jl:165 */
    if (cg) memset(cg, 0, sizeof(double) * padvox);                 /* iqsim.jl:166 */
    memset(pasted, 0, (size_t)nt);                                  /* iqsim.jl:169 */
    for (int64_t sidx = 0; sidx < a->npath; sidx++) {
      int64_t ind = a->path[sidx];
      a->chosen[r * a->npath + sidx] = -1;
      if (a->skipped[ind]) continue;                                /* iqsim.jl:174 */
      int64_t ti3[3] = {ind % a->ntiles[0], (ind / a->ntiles[0]) % a->ntiles[1], ind / (a->ntiles[0] * a->ntiles[1])};
      int64_t st[3];
      for (int d = 0; d < 3; d++) st[d] = ti3[d] * a->spacing[d];  /* iqsim.jl:180 */

      /* overlap mask, iqsim.jl:188-205 */
      memset(ovlmask, 0, (size_t)tvox);
      int prevp[3] = {0, 0, 0}, nextp[3] = {0, 0, 0};
      for (int d = 0; d < 3; d++) {
        if (a->ovlsize[d] <= 1) continue;
        int64_t nb[3] = {ti3[0], ti3[1], ti3[2]};
        nb[d] = ti3[d] - 1;
        if (nb[d] >= 0 && pasted[lin3(a->ntiles, nb[0], nb[1], nb[2])]) prevp[d] = 1;
        nb[d] = ti3[d] + 1;
        if (nb[d] < a->ntiles[d] && pasted[lin3(a->ntiles, nb[0], nb[1], nb[2])]) nextp[d] = 1;
        for (int64_t z = 0; z < ts[2]; z++)
          for (int64_t y = 0; y < ts[1]; y++)
            for (int64_t x = 0; x < ts[0]; x++) {
              int64_t c[3] = {x, y, z};
              if ((prevp[d] && c[d] < a->ovlsize[d]) || (nextp[d] && c[d] >= a->spacing[d]))
                ovlmask[lin3(ts, x, y, z)] = 1;
            }
      }
      /* ovldist (iqsim.jl:206-207) */
      int64_t V;
      tile_offsets(ovlmask, ts, jx, jy, jz, &V);
      for (int64_t j = 0; j < V; j++) kv[j] = sim[lin3(a->padsize, st[0] + jx[j], st[1] + jy[j], st[2] + jz[j])];
      distance_map(a->ti, a->tisize, ts, jx, jy, jz, kv, V, ovldist);
      for (int64_t c = 0; c < P; c++) if (a->disabled[c]) ovldist[c] = INFINITY;

      /* hard distance (iqsim.jl:210-219) */
      int hardtile = 0;
      if (a->hardgrid) {
        int64_t nh = 0;
        for (int64_t z = 0; z < ts[2]; z++)
          for (int64_t y = 0; y < ts[1]; y++)
            for (int64_t x = 0; x < ts[0]; x++) {
              double hv = a->hardgrid[lin3(a->padsize, st[0] + x, st[1] + y, st[2] + z)];
              hardmask[lin3(ts, x, y, z)] = !isnan(hv);
              if (!isnan(hv)) nh++;
            }
        if (nh > 0) {
          tile_offsets(hardmask, ts, jx, jy, jz, &V);
          for (int64_t j = 0; j < V; j++) kv[j] = a->hardgrid[lin3(a->padsize, st[0] + jx[j], st[1] + jy[j], st[2] + jz[j])];
          distance_map(a->ti, a->tisize, ts, jx, jy, jz, kv, V, harddist);
          for (int64_t c = 0; c < P; c++) if (a->disabled[c]) harddist[c] = INFINITY;
          hardtile = 1;
        }
      }
      /* soft distances (iqsim.jl:222-227), all-ones weights */
      for (int s = 0; s < a->naux; s++) {
        tile_offsets(NULL, ts, jx, jy, jz, &V);
        for (int64_t j = 0; j < V; j++) kv[j] = a->aux[s][lin3(a->padsize, st[0] + jx[j], st[1] + jy[j], st[2] + jz[j])];
        distance_map(a->auxti[s], a->tisize, ts, jx, jy, jz, kv, V, softdists[s]);
        for (int64_t c = 0; c < P; c++) if (a->disabled[c]) softdists[s][c] = INFINITY;
      }
      /* main and auxiliary distances (iqsim.jl:230-234) */
      const double* D;
      int nds = 0;
      if (hardtile) {
        D = harddist;
        Dsp[nds++] = ovldist;
        for (int s = 0; s < a->naux; s++) Dsp[nds++] = softdists[s];
      } else {
        D = ovldist;
        for (int s = 0; s < a->naux; s++) Dsp[nds++] = softdists[s];
      }
      /* pattern database (iqsim.jl:237) */
      int64_t ndb;
      if (nds == 0) {
        int64_t nen = 0;
        for (int64_t c = 0; c < P; c++) nen += !isinf(D[c]);
        if (nen == 0) { rc = ORA_ERR_NOCAND; break; }
        ndb = ora_band(D, P, a->tol, db);
      } else {
        ndb = ora_relaxation(D, Dsp, nds, P, a->tol, db);
        if (ndb < 0) { rc = (int)ndb; break; }
      }
      /* probabilities and the draw (iqsim.jl:240-243) */
      ora_taumodel(db, ndb, D, Dsp, nds, pw);
      int64_t pick;
      int64_t forced = a->forced ? a->forced[r * a->npath + sidx] : -1;
      if (forced >= 0) {
        pick = -1;
        for (int64_t i = 0; i < ndb; i++) if (db[i] == forced) { pick = i; break; }
        if (pick < 0) { rc = ORA_ERR_NOTINDB; a->chosen[r * a->npath + sidx] = forced; break; }
      } else {
        pick = ora_sample(pw, ndb, a->uniforms[r * a->npath + sidx]);
      }
      int64_t rind = db[pick];
      a->chosen[r * a->npath + sidx] = rind;
      if (a->nband) a->nband[r * a->npath + sidx] = ndb;
      int64_t cs[3] = {rind % a->distsize[0], (rind / a->distsize[0]) % a->distsize[1], rind / (a->distsize[0] * a->distsize[1])};

      /* boundary cut mask (iqsim.jl:252-275) */
      memset(cutmask, 0, (size_t)tvox);
      for (int d = 0; d < 3; d++) {
        for (int side = 0; side < 2; side++) {
          if (!(side == 0 ? prevp[d] : nextp[d])) continue;
          int64_t lo[3] = {0, 0, 0}, sl[3] = {ts[0], ts[1], ts[2]};
          if (side == 0) sl[d] = a->ovlsize[d];
          else { lo[d] = a->spacing[d]; sl[d] = ts[d] - a->spacing[d]; }
          if (a->mutate == 2 && side == 0 && sl[d] < ts[d]) lo[d] = 1;
          for (int64_t z = 0; z < sl[2]; z++)
            for (int64_t y = 0; y < sl[1]; y++)
              for (int64_t x = 0; x < sl[0]; x++) {
                int64_t q = lin3(sl, x, y, z);
                Aslab[q] = sim[lin3(a->padsize, st[0] + lo[0] + x, st[1] + lo[1] + y, st[2] + lo[2] + z)];
                Bslab[q] = a->ti[lin3(a->tisize, cs[0] + lo[0] + x, cs[1] + lo[1] + y, cs[2] + lo[2] + z)];
              }
          int64_t it = 0;
          int grc = ora_graphcut(Aslab, Bslab, a->ti_is_f32, sl, d, a->gr_freq, smask, &it);
          cut_total += it;
          if (grc != ORA_OK) { rc = grc; break; }
          for (int64_t z = 0; z < sl[2]; z++)
            for (int64_t y = 0; y < sl[1]; y++)
              for (int64_t x = 0; x < sl[0]; x++) {
                uint8_t mv = smask[lin3(sl, x, y, z)];
                if (side == 1 && a->mutate != 1) mv = !mv;
                cutmask[lin3(ts, lo[0] + x, lo[1] + y, lo[2] + z)] |= mv;
              }
        }
        if (rc != ORA_OK) break;
      }
      if (rc != ORA_OK) break;
      /* paste (iqsim.jl:278) and debug cut (iqsim.jl:281) */
      for (int64_t z = 0; z < ts[2]; z++)
        for (int64_t y = 0; y < ts[1]; y++)
          for (int64_t x = 0; x < ts[0]; x++) {
            int64_t g = lin3(a->padsize, st[0] + x, st[1] + y, st[2] + z);
            uint8_t cm = cutmask[lin3(ts, x, y, z)];
            const int64_t lg = (r * a->npath + sidx) * tvox + lin3(ts, x, y, z);
            if (a->step_cut) a->step_cut[lg] = cm;
            if (a->step_before) a->step_before[lg] = sim[g];
            if (!cm) sim[g] = a->ti[lin3(a->tisize, cs[0] + x, cs[1] + y, cs[2] + z)];
            if (cg) cg[g] = (double)cm;
          }
      pasted[ind] = 1;                                              /* iqsim.jl:284 */
    }
  }
  if (a->cut_iters) *a->cut_iters = cut_total;
  for (int i = 0; i < nsrc_max; i++) free(maps[i]);
  free(maps); free(ovlmask); free(hardmask); free(cutmask); free(pasted); free(jx); free(kv);
  free(db); free(pw); free(Aslab); free(Bslab); free(smask); free(Dsp);
  return rc;
}
