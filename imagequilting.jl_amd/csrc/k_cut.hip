// k_cut.hip -- boundary cut (graphcut.jl:5-84) and paste (iqsim.jl:252-281).
//
// One 256-wide workgroup per overlap slab.  The first slice along the cut
// axis is tied to the source and the last to the sink with infinite capacity
// (graphcut.jl:56-70), so both are contracted into the terminals and a
// preflow push-relabel runs on the interior slices, its state resident in LDS
// (global scratch beyond IQ_PR_NLDS interior nodes).  The schedule is
// synchronous: a relabel phase, a push phase (only the higher endpoint of an
// edge can push on it, so every residual pair has one writer) and a gather
// phase that adds incoming flow in a fixed neighbour order; a global relabel
// (exact residual BFS distance to the sink) runs every iq_gr_freq(n) rounds
// (8, or 24 above 1,024 interior nodes) and decides termination.  The result is deterministic and matches the oracle's
// restatement bit for bit.  Cut mask = voxels that cannot reach the sink in
// the residual graph (graphcut.jl:79-81).
#include "iq_internal.h"

#define CB 256

template <typename T>
__device__ inline double cut_cap(const T* A, const T* B, int64_t u, int64_t v, int64_t w, bool wok);

template <>
__device__ inline double cut_cap<float>(const float* A, const float* B, int64_t u, int64_t v, int64_t w, bool wok) {
  float Au = A[u], Av = A[v], Bu = B[u], Bv = B[v];
  float Du = __builtin_fabsf(Au - Bu), Dv = __builtin_fabsf(Av - Bv);
  float gAu = __builtin_fabsf(Av - Au), gBu = __builtin_fabsf(Bv - Bu);
  float gAv = wok ? __builtin_fabsf(A[w] - Av) : gAu;
  float gBv = wok ? __builtin_fabsf(B[w] - Bv) : gBu;
  float num = Du + Dv;
  float g = ((gAu + gAv) + gBu) + gBv;
  return (double)num / ((double)g + 0x1p-52);
}

template <>
__device__ inline double cut_cap<double>(const double* A, const double* B, int64_t u, int64_t v, int64_t w, bool wok) {
  double Du = __builtin_fabs(A[u] - B[u]), Dv = __builtin_fabs(A[v] - B[v]);
  double gAu = __builtin_fabs(A[v] - A[u]), gBu = __builtin_fabs(B[v] - B[u]);
  double gAv = wok ? __builtin_fabs(A[w] - A[v]) : gAu;
  double gBv = wok ? __builtin_fabs(B[w] - B[v]) : gBu;
  return (Du + Dv) / ((((gAu + gAv) + gBu) + gBv) + 0x1p-52);
}

// Slab view: element (x,y,z) of the slab lives at base + x*sx + y*sy + z*sz.
template <typename T>
struct SlabView {
  const T* base;
  int64_t st[3];
  __device__ int64_t at(int x, int y, int z) const { return x * st[0] + y * st[1] + z * st[2]; }
};

template <typename T>
__device__ double slab_cap(const SlabView<T>& A, const SlabView<T>& B, const int s[3], int c0, int c1, int c2, int d) {
  int c[3] = {c0, c1, c2};
  int cv[3] = {c0, c1, c2};
  cv[d] += 1;
  int cw[3] = {c0, c1, c2};
  cw[d] += 2;
  bool wok = cw[d] < s[d];
  int64_t u = A.at(c[0], c[1], c[2]), v = A.at(cv[0], cv[1], cv[2]);
  int64_t w = wok ? A.at(cw[0], cw[1], cw[2]) : 0;
  // A and B share strides only when both live in the same array; use per-view offsets
  int64_t ub = B.at(c[0], c[1], c[2]), vb = B.at(cv[0], cv[1], cv[2]);
  int64_t wb = wok ? B.at(cw[0], cw[1], cw[2]) : 0;
  // evaluate with the element order of graphcut.jl:37-52
  T Au = A.base[u], Av = A.base[v], Bu = B.base[ub], Bv = B.base[vb];
  T Aw = wok ? A.base[w] : (T)0, Bw = wok ? B.base[wb] : (T)0;
  T a3[3] = {Au, Av, Aw}, b3[3] = {Bu, Bv, Bw};
  return cut_cap<T>(a3, b3, 0, 1, 2, wok);
}

struct CutState {
  double *e, *rt, *rf, *rb, *pr;
  int32_t *h, *hn, *fl;
};

__device__ inline void cut_bind(CutState& S, char* mem, int n) {
  S.e = (double*)mem;
  S.rt = S.e + n;
  S.rf = S.rt + n;
  S.rb = S.rf + 3 * n;
  S.pr = S.rb + 3 * n;
  S.h = (int32_t*)(S.pr + 3 * n);
  S.hn = S.h + n;
  S.fl = S.hn + n;
}

// bytes of state per interior node
#define CUT_NODE_BYTES (8 * 11 + 4 * 3)

// Minimum cut of one slab (graphcut.jl:5-84) by the whole workgroup: writes
// keep = mask XOR side for every slab voxel into out (slab column-major).
// A and B are strided views (the simulation grid and the chosen training-image
// window); S points at LDS or global scratch sized for the interior nodes.
template <typename T>
__device__ void cut_slab(const SlabView<T>& A, const SlabView<T>& B, const int s[3], int dim, uint8_t side,
                         uint8_t* out, char* mem, int32_t* err, int64_t* rounds) {
  const int L = s[dim];
  const int nvox = s[0] * s[1] * s[2];
  // terminal slices
  for (int p = threadIdx.x; p < nvox; p += CB) {
    int c[3] = {p % s[0], (p / s[0]) % s[1], p / (s[0] * s[1])};
    uint8_t m = (c[dim] == L - 1) ? 0 : 1;
    out[p] = m ^ side;
  }
  if (L <= 2) return;
  int si[3] = {s[0], s[1], s[2]};
  si[dim] = L - 2;
  const int n = si[0] * si[1] * si[2];
  const int HMAX = n + 1;
  const int GR = iq_gr_freq(n);
  const int ist[3] = {1, si[0], si[0] * si[1]};
  CutState S;
  cut_bind(S, mem, n);

  // build: capacities of graphcut.jl:22-54 on interior edges; saturated source edges
  for (int u = threadIdx.x; u < n; u += CB) {
    int ic[3] = {u % si[0], (u / si[0]) % si[1], u / (si[0] * si[1])};
    int pc[3] = {ic[0], ic[1], ic[2]};
    pc[dim] += 1;
    int f = 0;
    double e = 0.0, rt = 0.0;
    for (int d = 0; d < 3; d++) {
      double c = 0.0;
      if (ic[d] + 1 < si[d]) { f |= 1 << d; c = slab_cap<T>(A, B, s, pc[0], pc[1], pc[2], d); }
      S.rf[3 * u + d] = c;
      S.rb[3 * u + d] = c;
      S.pr[3 * u + d] = 0.0;
    }
    if (pc[dim] == 1) {
      int q[3] = {pc[0], pc[1], pc[2]};
      q[dim] -= 1;
      e = e + slab_cap<T>(A, B, s, q[0], q[1], q[2], dim);
    }
    if (pc[dim] == L - 2) { f |= 8; rt = slab_cap<T>(A, B, s, pc[0], pc[1], pc[2], dim); }
    S.e[u] = e;
    S.rt[u] = rt;
    S.fl[u] = f;
  }
  __syncthreads();

  int64_t total = 0;
  for (;;) {
    // global relabel
    for (int u = threadIdx.x; u < n; u += CB) S.h[u] = ((S.fl[u] & 8) && S.rt[u] > 0.0) ? 1 : HMAX;
    __syncthreads();
    for (int lev = 1;; lev++) {
      int ch = 0;
      for (int u = threadIdx.x; u < n; u += CB) {
        if (S.h[u] != HMAX) continue;
        int ic[3] = {u % si[0], (u / si[0]) % si[1], u / (si[0] * si[1])};
        bool hit = false;
        for (int d = 0; d < 3 && !hit; d++) {
          if (ic[d] > 0) {
            int v = u - ist[d];
            if (S.rb[3 * v + d] > 0.0 && S.h[v] == lev) hit = true;
          }
          if (!hit && (S.fl[u] & (1 << d))) {
            int v = u + ist[d];
            if (S.rf[3 * u + d] > 0.0 && S.h[v] == lev) hit = true;
          }
        }
        if (hit) { S.h[u] = lev + 1; ch = 1; }
      }
      if (!__syncthreads_or(ch)) break;
    }
    int act = 0;
    for (int u = threadIdx.x; u < n; u += CB) act |= (S.e[u] > 0.0 && S.h[u] < HMAX);
    if (!__syncthreads_or(act)) break;
    if (total > IQ_PR_LIMIT) {
      if (threadIdx.x == 0) atomicCAS(err, 0, IQ_KERR_ITER);
      break;
    }
    for (int it = 0; it < GR; it++) {
      // relabel phase (also clears the push records of the previous round)
      for (int u = threadIdx.x; u < n; u += CB) {
        S.pr[3 * u] = 0.0; S.pr[3 * u + 1] = 0.0; S.pr[3 * u + 2] = 0.0;
        int hu = S.h[u];
        int nh = hu;
        if (S.e[u] > 0.0 && hu < HMAX) {
          int ic[3] = {u % si[0], (u / si[0]) % si[1], u / (si[0] * si[1])};
          bool adm = false;
          int mh = HMAX;
          if ((S.fl[u] & 8) && S.rt[u] > 0.0) { if (hu == 1) adm = true; mh = 0; }
          for (int d = 0; d < 3; d++) {
            if (ic[d] > 0) {
              int v = u - ist[d];
              if (S.rb[3 * v + d] > 0.0) { int hv = S.h[v]; if (hu == hv + 1) adm = true; if (hv < mh) mh = hv; }
            }
            if (S.fl[u] & (1 << d)) {
              int v = u + ist[d];
              if (S.rf[3 * u + d] > 0.0) { int hv = S.h[v]; if (hu == hv + 1) adm = true; if (hv < mh) mh = hv; }
            }
          }
          if (!adm) nh = (mh + 1 < HMAX) ? mh + 1 : HMAX;
        }
        S.hn[u] = nh;
      }
      __syncthreads();
      for (int u = threadIdx.x; u < n; u += CB) S.h[u] = S.hn[u];
      __syncthreads();
      // push phase
      for (int u = threadIdx.x; u < n; u += CB) {
        double ex = S.e[u];
        int hu = S.h[u];
        if (!(ex > 0.0 && hu < HMAX)) continue;
        int ic[3] = {u % si[0], (u / si[0]) % si[1], u / (si[0] * si[1])};
        if ((S.fl[u] & 8) && hu == 1 && S.rt[u] > 0.0) {
          double r = S.rt[u];
          double dl = ex < r ? ex : r;
          ex -= dl;
          S.rt[u] = r - dl;
        }
        for (int d = 0; d < 3 && ex > 0.0; d++) {
          if (ic[d] > 0) {
            int v = u - ist[d];
            if (hu == S.h[v] + 1) {
              double r = S.rb[3 * v + d];
              if (r > 0.0) {
                double dl = ex < r ? ex : r;
                ex -= dl;
                S.rb[3 * v + d] = r - dl;
                S.rf[3 * v + d] += dl;
                S.pr[3 * v + d] = -dl;
              }
            }
          }
          if (ex > 0.0 && (S.fl[u] & (1 << d))) {
            int v = u + ist[d];
            if (hu == S.h[v] + 1) {
              double r = S.rf[3 * u + d];
              if (r > 0.0) {
                double dl = ex < r ? ex : r;
                ex -= dl;
                S.rf[3 * u + d] = r - dl;
                S.rb[3 * u + d] += dl;
                S.pr[3 * u + d] = dl;
              }
            }
          }
        }
        S.e[u] = ex;
      }
      __syncthreads();
      // gather phase
      int a2 = 0;
      for (int u = threadIdx.x; u < n; u += CB) {
        int ic[3] = {u % si[0], (u / si[0]) % si[1], u / (si[0] * si[1])};
        double ex = S.e[u];
        for (int d = 0; d < 3; d++) {
          if (ic[d] > 0) {
            double q = S.pr[3 * (u - ist[d]) + d];
            if (q > 0.0) ex += q;
          }
          if (S.fl[u] & (1 << d)) {
            double q = S.pr[3 * u + d];
            if (q < 0.0) ex += -q;
          }
        }
        S.e[u] = ex;
        a2 |= (ex > 0.0 && S.h[u] < HMAX);
      }
      total++;
      if (!__syncthreads_or(a2)) break;
    }
  }
  for (int u = threadIdx.x; u < n; u += CB) {
    int ic[3] = {u % si[0], (u / si[0]) % si[1], u / (si[0] * si[1])};
    int pc[3] = {ic[0], ic[1], ic[2]};
    pc[dim] += 1;
    uint8_t m = (S.h[u] < HMAX) ? 0 : 1;
    out[pc[0] + s[0] * (pc[1] + s[1] * pc[2])] = m ^ side;
  }
  if (threadIdx.x == 0 && rounds) {
    atomicAdd((unsigned long long*)rounds, (unsigned long long)total);
    atomicAdd((unsigned long long*)rounds + 2, 1ull);
  }
}

// Register-resident form of cut_slab for slabs of at most NPT * TEAM
// interior nodes: same schedule, same arithmetic, bit-identical result.  A
// work-item owns nodes u = tid + k * TEAM; their coordinates, excess, sink
// residual, flags, height and the residuals of their six outgoing edges
// (towards the minus and the plus neighbour per axis) stay in registers:
// only the owner reads them.  A push from u to w lowers u's outgoing
// residual at once and leaves the amount in w's inbox (LDS, one slot per
// direction a push can come from, so each slot has one writer); w adds it
// to its excess and to its own outgoing residual on that edge at the start
// of the next round, in the same fixed order as the reference schedule's
// gather.  LDS holds only the inboxes (6 doubles per node) and the heights
// (double buffered, 16-bit): 52 bytes per node, so a 1,536-node slab takes
// 78 KB and two slabs fit in one compute unit.  A round costs two barriers.
// Team synchronisation of the register-resident cut: a 64-wide team is one
// wavefront (its LDS accesses complete in order; the fences only stop the
// compiler from moving them across the step), wider teams use the
// workgroup barrier.
template <int TEAM>
__device__ inline void team_sync() {
  if constexpr (TEAM == 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// OR of a predicate over the team: wavefront ballots, then (wider teams)
// one flag per wavefront in LDS behind a single barrier; the flags alternate
// between two buffers so the next call cannot overwrite a flag still unread.
template <int TEAM>
__device__ inline bool team_or(int p, int* tflag, int& par) {
  if constexpr (TEAM == 64) {
    team_sync<64>();
    return __any(p);
  } else {
    const bool w = __any(p);
    int* f = tflag + par * (TEAM / 64);
    if ((threadIdx.x & 63) == 0) f[threadIdx.x >> 6] = w ? 1 : 0;
    __syncthreads();
    int r = 0;
#pragma unroll
    for (int i = 0; i < TEAM / 64; i++) r |= f[i];
    par ^= 1;
    return r != 0;
  }
}

#define CUT_FAST_NODE_BYTES (8 * 6 + 2 * 2)

template <typename T, int NPT, int TEAM>
__device__ __forceinline__ void cut_slab_fast(const SlabView<T>& A, const SlabView<T>& B, const int s[3], int dim, uint8_t side,
                              uint8_t* out, char* mem, int32_t* err, int64_t* rounds) {
  const int L = s[dim];
  const int nvox = s[0] * s[1] * s[2];
  for (int p = threadIdx.x; p < nvox; p += TEAM) {
    int c[3] = {p % s[0], (p / s[0]) % s[1], p / (s[0] * s[1])};
    uint8_t m = (c[dim] == L - 1) ? 0 : 1;
    out[p] = m ^ side;
  }
  if (L <= 2) return;
  int si[3] = {s[0], s[1], s[2]};
  si[dim] = L - 2;
  const int n = si[0] * si[1] * si[2];
  const int HMAX = n + 1;
  const int GR = iq_gr_freq(n);
  const int ist[3] = {1, si[0], si[0] * si[1]};
  __shared__ int tflag[2 * (TEAM / 64)];
  int par = 0;
  // inbox of node w: [6 w + d] the amount its minus neighbour along d pushed
  // up to it, [6 w + 3 + d] what its plus neighbour pushed down
  double* inb = (double*)mem;
  int16_t* h = (int16_t*)(inb + 6 * n);  // committed heights
  int16_t* hn = h + n;                   // next heights
  // per owned node: flags (bit d: plus edge d, bit 3: sink edge, bits 4+d: minus edge d),
  // excess, sink residual, outgoing residuals towards the minus (om) and plus
  // (op) neighbours, height; neighbour indices are formed where used
  // (clamped to the node itself where the edge is absent)
  int fl[NPT];
  int uu[NPT];
  double e[NPT], rt[NPT], om[NPT][3], op[NPT][3];
  int hr[NPT];
#define CUT_VM(k, d) ((fl[k] & (16 << (d))) ? uu[k] - ist[d] : uu[k])
#define CUT_VP(k, d) ((fl[k] & (1 << (d))) ? uu[k] + ist[d] : uu[k])
#pragma unroll
  for (int k = 0; k < NPT; k++) {
    const int u = threadIdx.x + k * TEAM;
    uu[k] = u;
    fl[k] = 0;
    e[k] = 0.0;
    rt[k] = 0.0;
    hr[k] = HMAX;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      om[k][d] = 0.0;
      op[k][d] = 0.0;
    }
    if (u < n) {
      int ic[3] = {u % si[0], (u / si[0]) % si[1], u / (si[0] * si[1])};
      int pc[3] = {ic[0], ic[1], ic[2]};
      pc[dim] += 1;
      int f = 0;
      for (int d = 0; d < 3; d++) {
        // an edge's capacity is the same both ways; each end evaluates it
        if (ic[d] + 1 < si[d]) { f |= 1 << d; op[k][d] = slab_cap<T>(A, B, s, pc[0], pc[1], pc[2], d); }
        if (ic[d] > 0) {
          int q[3] = {pc[0], pc[1], pc[2]};
          q[d] -= 1;
          f |= 16 << d;
          om[k][d] = slab_cap<T>(A, B, s, q[0], q[1], q[2], d);
        }
        inb[6 * u + d] = 0.0;
        inb[6 * u + 3 + d] = 0.0;
      }
      if (pc[dim] == 1) {
        int q[3] = {pc[0], pc[1], pc[2]};
        q[dim] -= 1;
        e[k] = e[k] + slab_cap<T>(A, B, s, q[0], q[1], q[2], dim);
      }
      if (pc[dim] == L - 2) { f |= 8; rt[k] = slab_cap<T>(A, B, s, pc[0], pc[1], pc[2], dim); }
      fl[k] = f;
    }
  }
  team_sync<TEAM>();

  int64_t total = 0, levels = 0;
  for (;;) {
    // global relabel (exact residual distance to the sink)
#pragma unroll
    for (int k = 0; k < NPT; k++)
      if (uu[k] < n) h[uu[k]] = (int16_t)(((fl[k] & 8) && rt[k] > 0.0) ? 1 : HMAX);
    team_sync<TEAM>();
    // The labels are the residual hop distances to the sink (+1): the least
    // fixed point of h(u) = min over residual edges u->v of h(v) + 1 from
    // these starting values, which relaxing in any order reaches as well as
    // a level-by-level BFS does (every label written is the length of a real
    // path, and labels only decrease).
    for (;;) {
      // Each wavefront relaxes its own nodes until they hold still (no
      // barrier; labels of other wavefronts' nodes read meanwhile are valid
      // upper bounds), then one team OR asks whether any wavefront lowered a
      // label; a pass in which none did ends the relabel.  Labels strictly
      // decrease, so both loops end.
      int ch = 0;
      for (;;) {
        int lc = 0;
#pragma unroll
        for (int k = 0; k < NPT; k++) {
          const int u = uu[k];
          if (u >= n) continue;
          const int hu = h[u];
          int hm[3], hp[3];
#pragma unroll
          for (int d = 0; d < 3; d++) { hm[d] = h[CUT_VM(k, d)]; hp[d] = h[CUT_VP(k, d)]; }
          int best = hu;
#pragma unroll
          for (int d = 0; d < 3; d++) {
            if ((fl[k] & (16 << d)) && om[k][d] > 0.0 && hm[d] + 1 < best) best = hm[d] + 1;
            if ((fl[k] & (1 << d)) && op[k][d] > 0.0 && hp[d] + 1 < best) best = hp[d] + 1;
          }
          if (best < hu) { h[u] = (int16_t)best; lc = 1; }
        }
        if (!__any(lc)) break;
        ch = 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      levels++;
      if (!team_or<TEAM>(ch, tflag, par)) break;
    }
    int act = 0;
#pragma unroll
    for (int k = 0; k < NPT; k++)
      if (uu[k] < n) {
        hr[k] = h[uu[k]];
        act |= (e[k] > 0.0 && hr[k] < HMAX);
      }
    if (!team_or<TEAM>(act, tflag, par)) break;
    if (total > IQ_PR_LIMIT) {
      if (threadIdx.x == 0) atomicCAS(err, 0, IQ_KERR_ITER);
      break;
    }
    for (int it = 0; it <= GR; it++) {
      // phase A: take in the previous round's pushes, then relabel
      int a2 = 0;
      int hnx[NPT];
#pragma unroll
      for (int k = 0; k < NPT; k++) {
        const int u = uu[k];
        hnx[k] = hr[k];
        if (u >= n) continue;
        const int hu = hr[k];
        if (it > 0) {
          double im[3], ip[3];
#pragma unroll
          for (int d = 0; d < 3; d++) {
            im[d] = inb[6 * u + d];
            ip[d] = inb[6 * u + 3 + d];
          }
          double ex = e[k];
#pragma unroll
          for (int d = 0; d < 3; d++) {
            if ((fl[k] & (16 << d)) && im[d] > 0.0) {
              ex += im[d];
              om[k][d] = om[k][d] + im[d];
              inb[6 * u + d] = 0.0;
            }
            if ((fl[k] & (1 << d)) && ip[d] > 0.0) {
              ex += ip[d];
              op[k][d] = op[k][d] + ip[d];
              inb[6 * u + 3 + d] = 0.0;
            }
          }
          e[k] = ex;
        }
        int nh = hu;
        const bool active = e[k] > 0.0 && hu < HMAX;
        a2 |= active;
        if (active && it < GR) {
          int hm[3], hp[3];
#pragma unroll
          for (int d = 0; d < 3; d++) {
            hm[d] = h[CUT_VM(k, d)];
            hp[d] = h[CUT_VP(k, d)];
          }
          bool adm = false;
          int mh = HMAX;
          if ((fl[k] & 8) && rt[k] > 0.0) { if (hu == 1) adm = true; mh = 0; }
#pragma unroll
          for (int d = 0; d < 3; d++) {
            if ((fl[k] & (16 << d)) && om[k][d] > 0.0) { if (hu == hm[d] + 1) adm = true; if (hm[d] < mh) mh = hm[d]; }
            if ((fl[k] & (1 << d)) && op[k][d] > 0.0) { if (hu == hp[d] + 1) adm = true; if (hp[d] < mh) mh = hp[d]; }
          }
          if (!adm) nh = (mh + 1 < HMAX) ? mh + 1 : HMAX;
        }
        hn[u] = (int16_t)nh;
        hnx[k] = nh;
      }
      if (it > 0) {
        total++;
        if (!team_or<TEAM>(a2, tflag, par)) break;
      } else {
        team_sync<TEAM>();
      }
      if (it == GR) break;
      { int16_t* t = h; h = hn; hn = t; }
#pragma unroll
      for (int k = 0; k < NPT; k++) hr[k] = hnx[k];
      // phase B: push with the new heights (an edge has one pusher per
      // round, its higher end: the residuals it reads are its own)
#pragma unroll
      for (int k = 0; k < NPT; k++) {
        const int u = uu[k];
        if (u >= n) continue;
        double ex = e[k];
        const int hu = hr[k];
        if (!(ex > 0.0 && hu < HMAX)) continue;
        int hm[3], hp[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          hm[d] = h[CUT_VM(k, d)];
          hp[d] = h[CUT_VP(k, d)];
        }
        if ((fl[k] & 8) && hu == 1 && rt[k] > 0.0) {
          double r = rt[k];
          double dl = ex < r ? ex : r;
          ex -= dl;
          rt[k] = r - dl;
        }
#pragma unroll
        for (int d = 0; d < 3; d++) {
          if (ex > 0.0 && (fl[k] & (16 << d)) && hu == hm[d] + 1) {
            double r = om[k][d];
            if (r > 0.0) {
              double dl = ex < r ? ex : r;
              ex -= dl;
              om[k][d] = r - dl;
              inb[6 * (u - ist[d]) + 3 + d] = dl;
            }
          }
          if (ex > 0.0 && (fl[k] & (1 << d)) && hu == hp[d] + 1) {
            double r = op[k][d];
            if (r > 0.0) {
              double dl = ex < r ? ex : r;
              ex -= dl;
              op[k][d] = r - dl;
              inb[6 * (u + ist[d]) + d] = dl;
            }
          }
        }
        e[k] = ex;
      }
      team_sync<TEAM>();
    }
  }
#pragma unroll
  for (int k = 0; k < NPT; k++) {
    const int u = uu[k];
    if (u >= n) continue;
    int ic[3] = {u % si[0], (u / si[0]) % si[1], u / (si[0] * si[1])};
    int pc[3] = {ic[0], ic[1], ic[2]};
    pc[dim] += 1;
    uint8_t m = (h[u] < HMAX) ? 0 : 1;
    out[pc[0] + s[0] * (pc[1] + s[1] * pc[2])] = m ^ side;
  }
  if (threadIdx.x == 0 && rounds) {
    atomicAdd((unsigned long long*)rounds, (unsigned long long)total);
    atomicAdd((unsigned long long*)rounds + 1, (unsigned long long)levels);
    atomicAdd((unsigned long long*)rounds + 2, 1ull);
  }
#undef CUT_VM
#undef CUT_VP
}

template <typename T>
__device__ void cut_dispatch(const SlabView<T>& A, const SlabView<T>& B, const int s[3], int dim, uint8_t side,
                             uint8_t* out, char* lds, char* gmem, int32_t* err, int64_t* rounds) {
  const int n = s[0] * s[1] * s[2] / s[dim] * (s[dim] > 2 ? s[dim] - 2 : 0);
  if (n <= CB) cut_slab_fast<T, 1, CB>(A, B, s, dim, side, out, lds, err, rounds);
  else if (n <= 2 * CB) cut_slab_fast<T, 2, CB>(A, B, s, dim, side, out, lds, err, rounds);
  else if (n <= 4 * CB) cut_slab_fast<T, 4, CB>(A, B, s, dim, side, out, lds, err, rounds);
  else if (n <= IQ_PR_NLDS) cut_slab<T>(A, B, s, dim, side, out, lds, err, rounds);
  else cut_slab<T>(A, B, s, dim, side, out, gmem, err, rounds);
}


// Batched cut: grid (slab, query).  A = the realization's grid over the
// tile, B = the picked training-image window (iqsim.jl:244-273).
template <typename T>
__global__ void __launch_bounds__(CB)
k_graphcut(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  extern __shared__ char cut_lds[];
  const IqQuery q = qs[blockIdx.y];
  const IqStep& st = D->steps[q.step];
  if ((int)blockIdx.x >= st.ncut) return;
  const int64_t slot = q.slot;
  const int64_t pk = D->pick[slot];
  if (pk < 0) return;
  const IqGeo& g = D->g;
  const IqSlab sl = st.slab[blockIdx.x];
  const int s[3] = {sl.sz[0], sl.sz[1], sl.sz[2]};
  const int rc[3] = {(int)(pk % g.dist[0]), (int)((pk / g.dist[0]) % g.dist[1]),
                     (int)(pk / ((int64_t)g.dist[0] * g.dist[1]))};
  SlabView<T> A, B;
  const IqReal& R = D->reals[q.real];
  A.base = (const T*)D->sims + R.sim_off +
           iq_lin3(R.pad, st.origin[0] + sl.lo[0], st.origin[1] + sl.lo[1], st.origin[2] + sl.lo[2]);
  A.st[0] = 1; A.st[1] = R.pad[0]; A.st[2] = (int64_t)R.pad[0] * R.pad[1];
  B.base = (const T*)D->ti + iq_lin3(g.ti, rc[0] + sl.lo[0], rc[1] + sl.lo[1], rc[2] + sl.lo[2]);
  B.st[0] = 1; B.st[1] = g.ti[0]; B.st[2] = (int64_t)g.ti[0] * g.ti[1];
  uint8_t* out = D->keep + (slot * IQ_MAX_SLAB + blockIdx.x) * g.tilevox;
  char* gmem = D->cutscratch ? D->cutscratch + (slot * IQ_MAX_SLAB + blockIdx.x) * D->cutscratch_per : nullptr;
  cut_dispatch<T>(A, B, s, sl.dim, (uint8_t)sl.side, out, cut_lds, gmem, D->err, D->rounds);
}

template __global__ void k_graphcut<float>(const IqDev*, const IqQuery*);
template __global__ void k_graphcut<double>(const IqDev*, const IqQuery*);

// Plans whose slabs all have at most CUT_WAVE_MAX interior nodes: one
// wavefront per slab, up to two register-resident nodes per lane, no
// workgroup barriers in the push-relabel rounds.  (A round costs about the
// per-node work times the nodes per lane, so beyond two nodes per lane the
// 256-wide team with one node per lane is faster despite its barriers.)
#define CUT_WAVE_MAX (2 * 64)
template <typename T>
__device__ __forceinline__ void cut_dispatch_wave(const SlabView<T>& A, const SlabView<T>& B, const int s[3], int dim, uint8_t side,
                                  uint8_t* out, char* lds, int32_t* err, int64_t* rounds) {
  const int n = s[0] * s[1] * s[2] / s[dim] * (s[dim] > 2 ? s[dim] - 2 : 0);
  if (n <= 64) cut_slab_fast<T, 1, 64>(A, B, s, dim, side, out, lds, err, rounds);
  else cut_slab_fast<T, 2, 64>(A, B, s, dim, side, out, lds, err, rounds);
}

template <typename T>
__global__ void __launch_bounds__(64)
k_graphcut_wave(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  extern __shared__ char cut_lds[];
  const IqQuery q = qs[blockIdx.y];
  const IqStep& st = D->steps[q.step];
  if ((int)blockIdx.x >= st.ncut) return;
  const int64_t slot = q.slot;
  const int64_t pk = D->pick[slot];
  if (pk < 0) return;
  const IqGeo& g = D->g;
  const IqSlab sl = st.slab[blockIdx.x];
  const int s[3] = {sl.sz[0], sl.sz[1], sl.sz[2]};
  const int rc[3] = {(int)(pk % g.dist[0]), (int)((pk / g.dist[0]) % g.dist[1]),
                     (int)(pk / ((int64_t)g.dist[0] * g.dist[1]))};
  SlabView<T> A, B;
  const IqReal& R = D->reals[q.real];
  A.base = (const T*)D->sims + R.sim_off +
           iq_lin3(R.pad, st.origin[0] + sl.lo[0], st.origin[1] + sl.lo[1], st.origin[2] + sl.lo[2]);
  A.st[0] = 1; A.st[1] = R.pad[0]; A.st[2] = (int64_t)R.pad[0] * R.pad[1];
  B.base = (const T*)D->ti + iq_lin3(g.ti, rc[0] + sl.lo[0], rc[1] + sl.lo[1], rc[2] + sl.lo[2]);
  B.st[0] = 1; B.st[1] = g.ti[0]; B.st[2] = (int64_t)g.ti[0] * g.ti[1];
  uint8_t* out = D->keep + (slot * IQ_MAX_SLAB + blockIdx.x) * g.tilevox;
  cut_dispatch_wave<T>(A, B, s, sl.dim, (uint8_t)sl.side, out, cut_lds, D->err, D->rounds);
}

template __global__ void k_graphcut_wave<float>(const IqDev*, const IqQuery*);
template __global__ void k_graphcut_wave<double>(const IqDev*, const IqQuery*);

// Plans whose largest slab has CUT_BIG_MIN..CUT_BIG_MAX interior nodes: TEAM
// (1,024, or 512) work-items per slab, up to two (three) register-resident
// nodes each (node state of a 1,536-node slab fills the 160 KB of LDS, so one
// slab per compute unit runs at a time and more lanes per slab shorten each
// round: config 5's 1,536-node cuts take 1,469 us per launch with 1,024
// work-items against 1,780 us with 512).
#define CB_BIG 512
#define CUT_BIG_MAX 1536
template <typename T, int TEAM>
__global__ void __launch_bounds__(TEAM)
k_graphcut_big(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  extern __shared__ char cut_lds[];
  const IqQuery q = qs[blockIdx.y];
  const IqStep& st = D->steps[q.step];
  if ((int)blockIdx.x >= st.ncut) return;
  const int64_t slot = q.slot;
  const int64_t pk = D->pick[slot];
  if (pk < 0) return;
  const IqGeo& g = D->g;
  const IqSlab sl = st.slab[blockIdx.x];
  const int s[3] = {sl.sz[0], sl.sz[1], sl.sz[2]};
  const int rc[3] = {(int)(pk % g.dist[0]), (int)((pk / g.dist[0]) % g.dist[1]),
                     (int)(pk / ((int64_t)g.dist[0] * g.dist[1]))};
  SlabView<T> A, B;
  const IqReal& R = D->reals[q.real];
  A.base = (const T*)D->sims + R.sim_off +
           iq_lin3(R.pad, st.origin[0] + sl.lo[0], st.origin[1] + sl.lo[1], st.origin[2] + sl.lo[2]);
  A.st[0] = 1; A.st[1] = R.pad[0]; A.st[2] = (int64_t)R.pad[0] * R.pad[1];
  B.base = (const T*)D->ti + iq_lin3(g.ti, rc[0] + sl.lo[0], rc[1] + sl.lo[1], rc[2] + sl.lo[2]);
  B.st[0] = 1; B.st[1] = g.ti[0]; B.st[2] = (int64_t)g.ti[0] * g.ti[1];
  uint8_t* out = D->keep + (slot * IQ_MAX_SLAB + blockIdx.x) * g.tilevox;
  const int n = s[0] * s[1] * s[2] / s[sl.dim] * (s[sl.dim] > 2 ? s[sl.dim] - 2 : 0);
  static_assert(3 * TEAM >= CUT_BIG_MAX, "team too small for the largest slab");
  if (n <= TEAM) cut_slab_fast<T, 1, TEAM>(A, B, s, sl.dim, (uint8_t)sl.side, out, cut_lds, D->err, D->rounds);
  else if (n <= 2 * TEAM) cut_slab_fast<T, 2, TEAM>(A, B, s, sl.dim, (uint8_t)sl.side, out, cut_lds, D->err, D->rounds);
  else if constexpr (2 * TEAM < CUT_BIG_MAX)
    cut_slab_fast<T, 3, TEAM>(A, B, s, sl.dim, (uint8_t)sl.side, out, cut_lds, D->err, D->rounds);
}

template __global__ void k_graphcut_big<float, 512>(const IqDev*, const IqQuery*);
template __global__ void k_graphcut_big<double, 512>(const IqDev*, const IqQuery*);
template __global__ void k_graphcut_big<float, 1024>(const IqDev*, const IqQuery*);
template __global__ void k_graphcut_big<double, 1024>(const IqDev*, const IqQuery*);

// Standalone cut of one slab pair (iq_graphcut): A and B both column-major of extents s.
template <typename T>
__global__ void __launch_bounds__(CB)
k_graphcut_host(const T* __restrict__ a, const T* __restrict__ b, int s0, int s1, int s2, int dim,
                uint8_t* __restrict__ out, char* __restrict__ gmem, int32_t* __restrict__ err) {
  extern __shared__ char cut_lds[];
  const int s[3] = {s0, s1, s2};
  SlabView<T> A, B;
  A.base = a; A.st[0] = 1; A.st[1] = s0; A.st[2] = (int64_t)s0 * s1;
  B.base = b; B.st[0] = 1; B.st[1] = s0; B.st[2] = (int64_t)s0 * s1;
  cut_dispatch<T>(A, B, s, dim, (uint8_t)0, out, cut_lds, gmem, err, nullptr);
}

template __global__ void k_graphcut_host<float>(const float*, const float*, int, int, int, int, uint8_t*, char*, int32_t*);
template __global__ void k_graphcut_host<double>(const double*, const double*, int, int, int, int, uint8_t*, char*, int32_t*);

template <typename T>
__global__ void __launch_bounds__(64)
k_graphcut_host_wave(const T* __restrict__ a, const T* __restrict__ b, int s0, int s1, int s2, int dim,
                     uint8_t* __restrict__ out, int32_t* __restrict__ err) {
  extern __shared__ char cut_lds[];
  const int s[3] = {s0, s1, s2};
  SlabView<T> A, B;
  A.base = a; A.st[0] = 1; A.st[1] = s0; A.st[2] = (int64_t)s0 * s1;
  B.base = b; B.st[0] = 1; B.st[1] = s0; B.st[2] = (int64_t)s0 * s1;
  cut_dispatch_wave<T>(A, B, s, dim, (uint8_t)0, out, cut_lds, err, nullptr);
}

template __global__ void k_graphcut_host_wave<float>(const float*, const float*, int, int, int, int, uint8_t*, int32_t*);
template __global__ void k_graphcut_host_wave<double>(const double*, const double*, int, int, int, int, uint8_t*, int32_t*);

// Paste (iqsim.jl:278): voxels whose OR-ed keep flag is 0 take the picked
// training-image window; the keep flags are the debug cut grid (iqsim.jl:281).
// Grid: (tile blocks, query).
template <typename T>
__global__ void __launch_bounds__(256)
k_paste(const IqDev* __restrict__ D, const IqQuery* __restrict__ qs) {
  const IqQuery q = qs[blockIdx.y];
  const int64_t slot = q.slot;
  const int64_t pk = D->pick[slot];
  if (pk < 0) return;
  const IqGeo& g = D->g;
  const IqStep& st = D->steps[q.step];
  const int rc[3] = {(int)(pk % g.dist[0]), (int)((pk / g.dist[0]) % g.dist[1]),
                     (int)(pk / ((int64_t)g.dist[0] * g.dist[1]))};
  const int ns = st.nslab;
  const IqReal& R = D->reals[q.real];
  T* sim = (T*)D->sims + R.sim_off;
  uint8_t* cut = D->cuts + R.sim_off;
  const T* ti = (const T*)D->ti;
  const uint8_t* keep = D->keep + slot * IQ_MAX_SLAB * g.tilevox;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < g.tilevox; j += gridDim.x * blockDim.x) {
    int x = j % g.tile[0], y = (j / g.tile[0]) % g.tile[1], z = j / (g.tile[0] * g.tile[1]);
    uint8_t k = 0;
    for (int s = 0; s < ns; s++) {
      const IqSlab& sl = st.slab[s];
      int a = x - sl.lo[0], b = y - sl.lo[1], c = z - sl.lo[2];
      if (a >= 0 && b >= 0 && c >= 0 && a < sl.sz[0] && b < sl.sz[1] && c < sl.sz[2]) {
        if (s < st.ncut) {
          k |= keep[(int64_t)s * g.tilevox + a + sl.sz[0] * (b + sl.sz[1] * c)];
        } else {
          // two slices, no interior: the slice next to the source keeps the
          // grid, the other takes the window (negated for a next slab), as
          // the cut kernels write for such a slab
          const int cd = sl.dim == 0 ? a : (sl.dim == 1 ? b : c);
          k |= (uint8_t)((cd == sl.sz[sl.dim] - 1 ? 0 : 1) ^ sl.side);
        }
      }
    }
    int64_t gi = iq_lin3(R.pad, st.origin[0] + x, st.origin[1] + y, st.origin[2] + z);
    if (!k) sim[gi] = ti[iq_lin3(g.ti, rc[0] + x, rc[1] + y, rc[2] + z)];
    cut[gi] = k;
  }
}

template __global__ void k_paste<float>(const IqDev*, const IqQuery*);
template __global__ void k_paste<double>(const IqDev*, const IqQuery*);

// Fill helpers (no memset nodes: the run may be captured in a graph)
template <typename T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
template __global__ void k_fill<float>(float*, int64_t, float);
template __global__ void k_fill<double>(double*, int64_t, double);
template __global__ void k_fill<uint8_t>(uint8_t*, int64_t, uint8_t);
template __global__ void k_fill<int64_t>(int64_t*, int64_t, int64_t);
template __global__ void k_fill<int32_t>(int32_t*, int64_t, int32_t);
