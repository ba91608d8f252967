// iq_api.hip -- host driver and C ABI (include/iqhip.h) of libiqhip.
//
// One translation unit holds the host code and the kernels (k_*.hip are
// included below).
#include "../../include/iqhip.h"
#include "iq_internal.h"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "k_distance.hip"
#include "k_select.hip"
#include "k_cut.hip"
#include "k_learn.hip"
#include "k_conv.hip"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
namespace {
thread_local std::string g_last_error;

struct IqError {
  int code;
  std::string msg;
};

#define HC(x)                                                                               \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) throw IqError{IQ_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)}; \
  } while (0)
#define HCL()                                                                               \
  do {                                                                                      \
    hipError_t e_ = hipGetLastError();                                                      \
    if (e_ != hipSuccess) throw IqError{IQ_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)}; \
  } while (0)

[[noreturn]] void bad(int code, const std::string& m) { throw IqError{code, m}; }
}  // namespace

struct iq_ctx {
  int device = 0;
  int ncu = 256;  // compute units of the device
  hipStream_t stream = nullptr;
  std::string err;
};

namespace {
// LDS bytes of one slab's cut of n interior nodes: register-resident form
// (at most 4 CB nodes, or any slab of a plan run by the large-slab kernel),
// LDS-resident form (up to IQ_PR_NLDS), else none (global scratch)
size_t cut_lds_bytes(int64_t n, bool big) {
  if (n <= 4 * CB || (big && n <= CUT_BIG_MAX)) return (size_t)n * CUT_FAST_NODE_BYTES;
  if (n <= IQ_PR_NLDS) return (size_t)n * CUT_NODE_BYTES;
  return 0;
}

int fail(const iq_ctx* ctx, int code, const std::string& msg) {
  g_last_error = msg;
  if (ctx) const_cast<iq_ctx*>(ctx)->err = msg;
  return code;
}

template <class F>
int guard(iq_ctx* ctx, F&& f) {
  try {
    if (ctx) HC(hipSetDevice(ctx->device));
    f();
    return IQ_OK;
  } catch (const IqError& e) {
    return fail(ctx, e.code, e.msg);
  } catch (const std::bad_alloc&) {
    return fail(ctx, IQ_ERR_NOMEM, "host allocation failed");
  } catch (...) {
    return fail(ctx, IQ_ERR_ARG, "unexpected error");
  }
}

// device allocations owned by one object, freed together
struct DMem {
  std::vector<void*> ptrs;
  template <class T>
  T* alloc(size_t n) {
    void* p = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) bad(IQ_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    ptrs.push_back(p);
    return (T*)p;
  }
  template <class T>
  T* upload(const T* h, size_t n, hipStream_t s) {
    T* d = alloc<T>(n);
    if (n) HC(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    return d;
  }
  void release() {
    for (void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
  }
  ~DMem() { release(); }
};

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void check_kernel_error(int32_t code) {
  switch (code) {
    case 0: return;
    case IQ_KERR_NOCAND: bad(IQ_ERR_NOCAND, "every candidate pattern is disabled (all training-image windows hold inactive voxels)");
    case IQ_KERR_NOTINDB: bad(IQ_ERR_NOTINDB, "a forced pick is not in the pattern database");
    case IQ_KERR_ITER: bad(IQ_ERR_ITER, "min-cut push-relabel iteration guard reached");
    case IQ_KERR_SORTMAX: bad(IQ_ERR_UNSUPPORTED, "pattern database with more than " + std::to_string(IQ_SORT_MAX) + " entries and several distinct distances");
    case IQ_KERR_EMPTYBAND: bad(IQ_ERR_HIP, "internal error: empty pattern band");
    default: bad(IQ_ERR_HIP, "unknown kernel error " + std::to_string(code));
  }
}

// ---------------------------------------------------------------------------
// overlap masks: unions of prev/next slabs decomposed into disjoint boxes
// ---------------------------------------------------------------------------
struct HostMask {
  std::vector<uint8_t> bits;
  std::vector<IqBox> boxes;
  int V = 0;
};

std::vector<IqBox> decompose(const std::vector<uint8_t>& bits, const int t[3], const std::vector<int> brk[3]) {
  // cells of the breakpoint grid are entirely in or out of the mask
  std::vector<IqBox> cells;
  for (size_t kz = 0; kz + 1 < brk[2].size(); kz++)
    for (size_t ky = 0; ky + 1 < brk[1].size(); ky++) {
      IqBox run{};
      bool open = false;
      for (size_t kx = 0; kx + 1 < brk[0].size(); kx++) {
        int x0 = brk[0][kx], y0 = brk[1][ky], z0 = brk[2][kz];
        bool in = bits[x0 + t[0] * (y0 + t[1] * z0)] != 0;
        if (in && open) {
          run.wx = brk[0][kx + 1] - run.x0;
        } else if (in) {
          run = IqBox{x0, y0, z0, brk[0][kx + 1] - x0, brk[1][ky + 1] - y0, brk[2][kz + 1] - z0, 0, 0};
          open = true;
        } else if (open) {
          cells.push_back(run);
          open = false;
        }
      }
      if (open) cells.push_back(run);
    }
  // merge along y, then along z
  for (int axis = 1; axis <= 2; axis++) {
    bool merged = true;
    while (merged) {
      merged = false;
      for (size_t a = 0; a < cells.size() && !merged; a++)
        for (size_t b = 0; b < cells.size() && !merged; b++) {
          if (a == b) continue;
          IqBox& A = cells[a];
          IqBox& B = cells[b];
          if (A.x0 != B.x0 || A.wx != B.wx) continue;
          if (axis == 1 && A.z0 == B.z0 && A.wz == B.wz && A.y0 + A.wy == B.y0) {
            A.wy += B.wy;
            cells.erase(cells.begin() + b);
            merged = true;
          } else if (axis == 2 && A.y0 == B.y0 && A.wy == B.wy && A.z0 + A.wz == B.z0) {
            A.wz += B.wz;
            cells.erase(cells.begin() + b);
            merged = true;
          }
        }
    }
  }
  return cells;
}

// FP32 scan configurations.  kind 1: k_scan_f32 (R x-candidates per
// work-item); kind 2: k_scan2_f32 (R x RY block, lanes over x and z, four
// wavefronts over y).  A work-group covers (R*LX) x (RY*LY) x LZ candidates.
struct ScanCfg {
  int kind, R, RY, LX, LY, LZ;
};
const ScanCfg kScanCfg[] = {{1, 4, 1, 8, 8, 4}, {1, 4, 1, 16, 4, 4}, {1, 2, 1, 16, 8, 2}, {1, 4, 1, 8, 16, 2},
                            {2, 4, 4, 8, 4, 8}, {2, 4, 4, 16, 4, 4}};
const int kNumScanCfg = 6;
const size_t kScanLdsMax = 100 * 1024;

size_t scan_lds_bytes(const ScanCfg& c, const std::vector<IqBox>& boxes, bool perz) {
  size_t m = 0;
  for (const IqBox& b : boxes) {
    int wxp = (b.wx + 3) / 4 * 4;
    int S = c.LX + wxp / c.R;
    size_t n;
    if (c.kind == 1) {
      n = (size_t)c.R * S * (c.LY + b.wy - 1) * (c.LZ + b.wz - 1);
    } else {
      size_t rowst = (size_t)c.R * S;
      size_t ryt = (size_t)c.LY * c.RY + b.wy - 1;
      size_t plane = ((ryt * rowst + 55) / 64) * 64 + 8;
      n = plane * (perz ? c.LZ : c.LZ + b.wz - 1);
    }
    m = std::max(m, n * sizeof(float));
  }
  return m;
}

bool bf16_exact(const float* v, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    uint32_t u;
    std::memcpy(&u, &v[i], 4);
    if ((u & 0xffffu) != 0u) return false;
  }
  return true;
}

// K table of the MFMA scan over the union of the overlap masks: per box,
// 8-voxel chunks along x or y (whichever pads less), an even number per box.
bool build_ktab(const std::vector<uint8_t>& ubits, const int t3[3], const std::vector<int> brk[3], IqKTab& kt,
                std::vector<IqChunk>& chunks) {
  std::vector<IqBox> boxes = decompose(ubits, t3, brk);
  if ((int)boxes.size() > IQ_MAX_KBOX) return false;
  kt.nbox = 0;
  int step = 0;
  for (const IqBox& b : boxes) {
    auto r8 = [](int w) { return (w + 7) / 8 * 8; };
    const double ex = (double)b.wx / r8(b.wx), ey = (double)b.wy / r8(b.wy);
    const int dir = (ey > ex) ? 1 : 0;
    IqKBox kb{};
    kb.x0 = b.x0; kb.y0 = b.y0; kb.z0 = b.z0; kb.wz = b.wz;
    kb.wxr = dir == 0 ? r8(b.wx) : b.wx;
    kb.wyr = dir == 1 ? r8(b.wy) : b.wy;
    const size_t c0 = chunks.size();
    for (int z = 0; z < b.wz; z++) {
      if (dir == 0) {
        for (int y = 0; y < b.wy; y++)
          for (int x = 0; x < b.wx; x += 8)
            chunks.push_back(IqChunk{(int16_t)(b.x0 + x), (int16_t)(b.y0 + y), (int16_t)(b.z0 + z), 0,
                                     (int8_t)std::min(8, b.wx - x)});
      } else {
        for (int x = 0; x < b.wx; x++)
          for (int y = 0; y < b.wy; y += 8)
            chunks.push_back(IqChunk{(int16_t)(b.x0 + x), (int16_t)(b.y0 + y), (int16_t)(b.z0 + z), 1,
                                     (int8_t)std::min(8, b.wy - y)});
      }
    }
    if ((chunks.size() - c0) % 2) {  // zero chunk: same base, no voxels
      IqChunk z = chunks.back();
      z.n = 0;
      chunks.push_back(z);
    }
    kb.s0 = step;
    step += (int)(chunks.size() - c0) / 2;
    kb.s1 = step;
    kt.box[kt.nbox++] = kb;
  }
  kt.nstep = step;
  return true;
}

// Values the FP32 / split-bf16 scans take: FP32-exact, |x| <= 2^50, and zero
// or |x| >= 2^-100 (so every split part and product stays a normal number and
// the relative error bounds of scan_gerr hold).
bool fits_f32(const double* v, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    double x = v[i];
    if ((double)(float)x != x || std::fabs(x) > 0x1p50 || (x != 0.0 && std::fabs(x) < 0x1p-100)) return false;
  }
  return true;
}
bool fits_range(const float* v, int64_t n) {
  for (int64_t i = 0; i < n; i++)
    if (!(std::fabs(v[i]) <= 0x1p50f) || (v[i] != 0.0f && std::fabs(v[i]) < 0x1p-100f)) return false;
  return true;
}

// Relative error bound of the MFMA scan's correlation: |AB32 - AB| <=
// scan_gerr * sum |a||b| over a mask of nstep k-steps (16 voxels each).
//  * Accumulation: every bf16 x bf16 product is exact in FP32; the FP32 sums
//    of kterms = 16 nstep (3 or 1) terms, in any order, err by at most
//    gamma(kterms - 1) times the sum of the terms' magnitudes, which is at
//    most (1 + 2^-8)^2 (1 + 2^-7) sum |a||b| <= 1.02 sum |a||b|; gamma of
//    twice the count is used.
//  * Split (FP32 data): t = hi + lo + r with hi = bf16(t), lo = bf16(t - hi)
//    (t - hi is exact in FP32).  For 2^e <= |t|: |t - hi| <= 2^(e-8), so
//    |lo| <= 2^-8 |t| and |r| <= half a bf16 ulp of |t - hi| < 2^(e-8), i.e.
//    |r| <= 2^(e-17) <= 2^-17 |t|; |hi + lo| <= (1 + 2^-17) |t|.  The three
//    products hi hi, hi lo, lo hi drop ab - (...) = (ha + la) rb + la lb + ra b,
//    at most (2^-16 + 2^-17 (1 + 2^-17) + 2^-17) |a||b|
//    = 2^-15 (1 + 2^-19) |a||b|.  Values are zero or at least 2^-100 in
//    magnitude (fits_f32 / fits_range), so no part is subnormal.
double scan_gerr(int64_t nstep, bool split) {
  const double kterms = 16.0 * (double)nstep * (split ? 3.0 : 1.0);
  const double ga = 2.0 * kterms * 0x1p-24 / (1.0 - 2.0 * kterms * 0x1p-24);
  return 1.02 * ga + (split ? 0x1p-15 * (1.0 + 0x1p-19) : 0.0);
}

void geo_init(IqGeo& g, const int64_t ti[3], const int64_t tile[3]) {
  for (int d = 0; d < 3; d++) {
    g.ti[d] = (int32_t)ti[d];
    g.tile[d] = (int32_t)tile[d];
    g.dist[d] = (int32_t)(ti[d] - tile[d] + 1);
  }
  g.P = (int64_t)g.dist[0] * g.dist[1] * g.dist[2];
  g.tilevox = g.tile[0] * g.tile[1] * g.tile[2];
}

// One simulation grid of a plan (iqsim.jl:99-115) and its realizations.
struct HostGrid {
  int64_t ntiles[3], pad[3], padvox = 0, npath = 0, nreal = 0;
  int64_t step0 = 0;  // first entry of its steps
  int64_t real0 = 0;  // first realization
  std::vector<std::vector<int64_t>> waves;  // path positions per wave (dependency order)
};

}  // namespace

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
struct iq_plan {
  iq_ctx* ctx = nullptr;
  int dtype = IQ_F64;
  IqGeo geo{};
  int64_t ovl[3], spacing[3];
  int64_t nreal = 0, nen = 0;  // realizations over all grids; enabled candidates
  int64_t nlog = 0, nvox = 0;  // pick-log entries and grid voxels over all realizations
  std::vector<HostGrid> grids;
  std::vector<IqReal> reals;
  size_t nwaves = 0;           // lockstep waves (the longest grid's)
  double tol = 0.1;
  int naux = 0;
  bool fast = false;
  bool hard = false;
  bool any_multi = false;
  std::vector<IqStep> steps;
  std::vector<IqMask> masks;
  int scan_cfg = -1;
  size_t scan_lds = 0;
  bool scan_perz = false;
  bool mfma = false;        // multi-query MFMA scan
  bool mfma_split = true;   // FP32 data: three bf16 products
  int mfma_nw = 4;          // wavefronts per scan group (4 candidate rows each), forced by IQHIP_SCAN_NW
  bool mfma_nw_forced = false;
  int nw_tiles[3] = {4, 4, 4};  // scan group size for launches of 1 / 2+ query tiles (tuned at plan creation)
  float tune_us[3][5] = {};     // measured scan microseconds [tiles][wavefronts]
  const int32_t* kofs_nw[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // chunk offsets per group size
  size_t lds_nw[5] = {0, 0, 0, 0, 0};
  size_t mfma_lds = 0;
  double mfma_gerr = 0.0;
  int64_t mfma_ntile_max = 0;
  int nbz = 1;
  int64_t nslots = 0;
  DMem mem;
  IqDev hd{};            // host copy of the device descriptor
  IqDev* dd = nullptr;   // device descriptor
  IqQuery* dq = nullptr; // query table (all batches of a run)
  int64_t dq_cap = 0;
  int4* dt = nullptr;    // shortlist tiles (all batches of a run)
  int64_t dt_cap = 0;
  size_t cut_lds_max = 0;
  size_t multi_lds = 0;    // LDS of the tiled multi-source maps (0: gathering kernel)
  int multi_bx = 16;       // candidate block width of the tiled maps (32 or 16; 16 rows)
  bool cut_big = false;    // largest slab needs the 512/1024-lane register-resident cut
  int cut_team = 1024;     // its work-items per slab (IQHIP_CUT_TEAM=512: 512)
  int max_nslab = 1;       // most min cuts of any step: the cut launches' grid width
  bool cut_wave = false;   // every slab fits the one-wavefront cut
  std::vector<int32_t> enabled_host;
  // learned proposal
  bool lrn_ready = false, lrn_on = false, lrn_fell_back = false;
  int lrn_check = 8;
  int lrn_check_q = 0;  // queries of a check batch the full scan decides (0: all)
  double lrn_min_recall = 0.99;
  int64_t lrn_stats[4] = {0, 0, 0, 0};
  std::vector<int64_t> lrn_wofs, lrn_bofs;
  int lrn_vmax = 0;
  // convolutional encoder of the learned proposal
  bool cv_ready = false;        // banks built from the trained network's output map
  float* cv_theta = nullptr;    // CV_NW convolution parameters
  double cv_stats[8] = {};      // see iq_plan_conv_stats
  double cv_val[4] = {};        // held-out loss and shortlist recall at the start and of the kept parameters
  float last_ms = 0.0f;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // a run's launch sequence captured once and replayed (runs that take no
  // host decision between batches and record no profiling events); [0] full
  // scan, [1] learned shortlist without checks
  bool graphs = false;  // IQHIP_GRAPH=1: capture and replay (measured no faster: the runs are device-bound)
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  const IqQuery* gdq[2] = {nullptr, nullptr};
  const int4* gdt[2] = {nullptr, nullptr};
  int64_t gnq[2] = {-1, -1};
  int64_t graph_launches = 0;
  // scan profiling
  bool prof = false;
  std::vector<hipEvent_t> prof_ev;
  int64_t prof_used = 0;
  std::vector<hipEvent_t> cut_ev;  // min-cut launches (profiling): pairs of events
  std::vector<hipEvent_t> bank_ev;  // shortlist bank-scan launches (profiling): pairs of events
  int64_t bank_used = 0, bank_launches = 0;
  double bank_sec = 0.0, bank_bytes = 0.0, bank_bytes_run = 0.0;
  int64_t cut_used = 0, cut_launches = 0;
  double cut_sec = 0.0;
  double prof_flop = 0.0, prof_sec = 0.0, prof_mflop = 0.0;
  int64_t prof_launches = 0;
  std::vector<double> batch_flop;
  int64_t cut_stats[6] = {0, 0, 0, 0, 0, 0};  // push rounds, relabel BFS levels, cuts; plain queries, superset and band sizes
};

namespace {

// Waves: a path position depends on every earlier path position whose tile is
// within one tile along each axis (only those overlap when the overlap is at
// most half a tile); it runs one wave after the latest of them.  Tiles of one
// wave never overlap, every neighbour pasted before a tile in path order is
// pasted in an earlier wave and none after it is, so masks, windows and cuts
// are exactly those of the sequential loop.
std::vector<std::vector<int64_t>> build_waves(const HostGrid& G, const IqStep* steps, const int64_t* path,
                                              bool sequential) {
  std::vector<std::vector<int64_t>> waves;
  const int64_t* nt3 = G.ntiles;
  const int64_t nt = nt3[0] * nt3[1] * nt3[2];
  std::vector<int64_t> wave_of((size_t)nt, -1);
  int64_t last = -1;
  for (int64_t i = 0; i < G.npath; i++) {
    if (!steps[i].active) continue;
    int64_t w;
    if (sequential) {
      w = ++last;
    } else {
      const int64_t ind = path[i];
      int64_t c[3] = {ind % nt3[0], (ind / nt3[0]) % nt3[1], ind / (nt3[0] * nt3[1])};
      w = 0;
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            int64_t nb[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
            if (nb[0] < 0 || nb[1] < 0 || nb[2] < 0 || nb[0] >= nt3[0] || nb[1] >= nt3[1] || nb[2] >= nt3[2])
              continue;
            int64_t wi = wave_of[nb[0] + nt3[0] * (nb[1] + nt3[1] * nb[2])];
            if (wi >= 0) w = std::max(w, wi + 1);
          }
      wave_of[ind] = w;
    }
    if ((int64_t)waves.size() <= w) waves.resize(w + 1);
    waves[w].push_back(i);
  }
  return waves;
}

// Scan autotuner: the MFMA scan's group size (2 or 4 wavefronts) is measured
// on the plan itself.  A batch of the plan's own masked steps (32 queries,
// then up to 64 as the slots allow) is prepared once; each group size runs one
// warm-up and three timed launches; the faster is kept per tile count.  The
// choice changes speed only: both sizes compute the same bounds.
template <typename T>
void tune_scan(iq_plan* pl) {
  const int nwa[3] = {2, 3, 4};
  for (int t = 1; t <= 2; t++) pl->nw_tiles[t] = pl->lds_nw[4] ? 4 : 2;
  if (!pl->mfma || pl->mfma_nw_forced || !pl->lds_nw[2] || !pl->lds_nw[3] || !pl->lds_nw[4]) return;
  const IqGeo& g = pl->geo;
  hipStream_t s = pl->ctx->stream;
  std::vector<IqQuery> hq;
  const int64_t want = std::min<int64_t>(pl->nslots, 64);
  for (const HostGrid& G : pl->grids) {
    for (int64_t i = 0; i < G.npath && (int64_t)hq.size() < want; i++) {
      const IqStep& st = pl->steps[G.step0 + i];
      if (st.mask < 0 || st.hard || pl->naux > 0) continue;
      IqQuery q{};
      q.step = (int32_t)(G.step0 + i);
      q.real = (int32_t)G.real0;
      q.slot = (int32_t)hq.size();
      q.multi = 0;
      q.u = 0.5;
      q.forced = -1;
      hq.push_back(q);
    }
  }
  if (hq.size() < 32) return;
  const IqDev* D = pl->dd;
  const IqQuery* dq = pl->mem.upload(hq.data(), hq.size(), s);
  hipLaunchKernelGGL((k_fill<T>), dim3(1024), dim3(256), 0, s, (T*)pl->hd.sims, pl->nvox, (T)0);
  HCL();
  hipEvent_t a, b;
  HC(hipEventCreate(&a));
  HC(hipEventCreate(&b));
  for (int ntile = 1; ntile <= 2; ntile++) {
    const int nq = ntile == 1 ? 32 : (int)hq.size();  // two tiles: every prepared query (33 to 64)
    if (ntile == 2 && nq <= 32) break;
    hipLaunchKernelGGL((k_query_prep<T>), dim3(nq, 1), dim3(256), 0, s, D, dq);
    HCL();
    hipLaunchKernelGGL(k_mfma_bprep, dim3(pl->hd.ktab.nstep, ntile), dim3(64), 0, s, D, dq, nq);
    HCL();
    float best = 1e30f;
    for (int nw : nwa) {
      dim3 grid(nblk(g.dist[0], 32), nblk(g.dist[1], 4 * nw), (unsigned)(g.dist[2] * ntile));
      float tot = 0.0f;
      for (int rep = 0; rep < 4; rep++) {
        if (rep == 1) HC(hipEventRecord(a, s));
        if (pl->mfma_split)
          hipLaunchKernelGGL((k_scan_mfma<true>), grid, dim3(64 * nw), pl->lds_nw[nw], s, D, dq, nq, pl->mfma_gerr,
                             pl->kofs_nw[nw]);
        else
          hipLaunchKernelGGL((k_scan_mfma<false>), grid, dim3(64 * nw), pl->lds_nw[nw], s, D, dq, nq, pl->mfma_gerr,
                             pl->kofs_nw[nw]);
        HCL();
      }
      HC(hipEventRecord(b, s));
      HC(hipEventSynchronize(b));
      HC(hipEventElapsedTime(&tot, a, b));
      const float us = tot * 1e3f / 3.0f;
      pl->tune_us[ntile][nw] = us;
      if (us < best) { best = us; pl->nw_tiles[ntile] = nw; }
    }
  }
  if ((int)hq.size() <= 32) pl->nw_tiles[2] = pl->nw_tiles[1];
  HC(hipEventDestroy(a));
  HC(hipEventDestroy(b));
}

// Argument checks of a plan description (the @asserts of iqsim.jl:69-89 as
// they reach the ABI, and the geometry identities of iqsim.jl:92-115); host
// only, no device work.
void validate_descs(const iq_plan_desc* descs, int ngrid) {
  if (!descs || ngrid <= 0) bad(IQ_ERR_ARG, "no simulation grid");
  const iq_plan_desc* d = &descs[0];
  if (!d->ti || !d->disabled) bad(IQ_ERR_ARG, "null plan input");
  if (d->ti_dtype != IQ_F32 && d->ti_dtype != IQ_F64) bad(IQ_ERR_ARG, "ti_dtype must be IQ_F32 or IQ_F64");
  if (!(d->tol > 0.0 && d->tol <= 1.0)) bad(IQ_ERR_ARG, "tolerance must be in range (0,1]");
  if (d->naux < 0 || d->naux > IQ_MAX_AUX) bad(IQ_ERR_ARG, "naux must be in 0..8");
  for (int k = 0; k < 3; k++) {
    if (d->tilesize[k] <= 0 || d->tilesize[k] > d->tisize[k]) bad(IQ_ERR_ARG, "invalid tile size");
    if (d->ovlsize[k] < 0 || d->ovlsize[k] > d->tilesize[k]) bad(IQ_ERR_ARG, "invalid overlap size");
    if (d->spacing[k] != d->tilesize[k] - d->ovlsize[k]) bad(IQ_ERR_ARG, "spacing must be tilesize - ovlsize");
  }
  if (d->naux > 0) {
    if (!d->auxti) bad(IQ_ERR_ARG, "null soft data");
    for (int a = 0; a < d->naux; a++)
      if (!d->auxti[a]) bad(IQ_ERR_ARG, "null soft data");
  }
  for (int gi = 0; gi < ngrid; gi++) {
    const iq_plan_desc* e = &descs[gi];
    if (!e->path || !e->skipped) bad(IQ_ERR_ARG, "null plan input");
    if (e->nreal <= 0) bad(IQ_ERR_ARG, "invalid number of realizations");
    if (e->ti_dtype != d->ti_dtype || e->naux != d->naux || e->tol != d->tol)
      bad(IQ_ERR_ARG, "grids of one plan share the training image, tol and soft pairs");
    for (int k = 0; k < 3; k++) {
      if (e->tisize[k] != d->tisize[k] || e->tilesize[k] != d->tilesize[k] || e->ovlsize[k] != d->ovlsize[k] ||
          e->spacing[k] != d->spacing[k])
        bad(IQ_ERR_ARG, "grids of one plan share the training image and tile geometry");
      if (e->ntiles[k] <= 0) bad(IQ_ERR_ARG, "invalid ntiles");
      if (e->padsize[k] != e->ntiles[k] * e->spacing[k] + e->ovlsize[k]) bad(IQ_ERR_ARG, "padsize must be ntiles*(tilesize-ovlsize)+ovlsize");
    }
    const int64_t nt = e->ntiles[0] * e->ntiles[1] * e->ntiles[2];
    std::vector<uint8_t> seen((size_t)nt, 0);
    for (int64_t i = 0; i < e->npath; i++) {
      if (e->path[i] < 0 || e->path[i] >= nt) bad(IQ_ERR_ARG, "path entry outside the tile grid");
      if (seen[e->path[i]]++) bad(IQ_ERR_ARG, "path visits a tile twice");
    }
    if (e->naux > 0) {
      if (!e->aux) bad(IQ_ERR_ARG, "null soft data");
      for (int a = 0; a < e->naux; a++)
        if (!e->aux[a]) bad(IQ_ERR_ARG, "null soft data");
    }
  }
}

void plan_build(iq_ctx* ctx, const iq_plan_desc* descs, int ngrid, iq_plan* pl) {
  validate_descs(descs, ngrid);
  const iq_plan_desc* d = &descs[0];  // training image, tile geometry, tol and soft pairs of every grid
  hipStream_t s = ctx->stream;
  pl->ctx = ctx;
  pl->dtype = d->ti_dtype;
  pl->tol = d->tol;
  pl->naux = d->naux;
  bool half = true;
  for (int k = 0; k < 3; k++) {
    if (d->tilesize[k] <= 0 || d->tilesize[k] > d->tisize[k]) bad(IQ_ERR_ARG, "invalid tile size");
    if (d->ovlsize[k] < 0 || d->ovlsize[k] > d->tilesize[k]) bad(IQ_ERR_ARG, "invalid overlap size");
    if (d->spacing[k] != d->tilesize[k] - d->ovlsize[k]) bad(IQ_ERR_ARG, "spacing must be tilesize - ovlsize");
    pl->ovl[k] = d->ovlsize[k];
    pl->spacing[k] = d->spacing[k];
    if (2 * d->ovlsize[k] > d->tilesize[k] || d->spacing[k] < 1) half = false;
  }
  geo_init(pl->geo, d->tisize, d->tilesize);
  IqGeo& g = pl->geo;
  if (g.P > (int64_t)S2MAX * IQ_SB * IQ_SB)
    bad(IQ_ERR_UNSUPPORTED, "more than " + std::to_string((int64_t)S2MAX * IQ_SB * IQ_SB) + " candidate patterns");
  const int64_t tivox = (int64_t)g.ti[0] * g.ti[1] * g.ti[2];
  const int t3[3] = {g.tile[0], g.tile[1], g.tile[2]};
  for (int gi = 0; gi < ngrid; gi++) {
    const iq_plan_desc* e = &descs[gi];
    if (!e->path || !e->skipped) bad(IQ_ERR_ARG, "null plan input");
    if (e->nreal <= 0) bad(IQ_ERR_ARG, "invalid number of realizations");
    if (e->ti_dtype != d->ti_dtype || e->naux != d->naux || e->tol != d->tol)
      bad(IQ_ERR_ARG, "grids of one plan share the training image, tol and soft pairs");
    HostGrid G;
    for (int k = 0; k < 3; k++) {
      if (e->tisize[k] != d->tisize[k] || e->tilesize[k] != d->tilesize[k] || e->ovlsize[k] != d->ovlsize[k] ||
          e->spacing[k] != d->spacing[k])
        bad(IQ_ERR_ARG, "grids of one plan share the training image and tile geometry");
      if (e->ntiles[k] <= 0) bad(IQ_ERR_ARG, "invalid ntiles");
      if (e->padsize[k] != e->ntiles[k] * e->spacing[k] + e->ovlsize[k]) bad(IQ_ERR_ARG, "padsize must be ntiles*(tilesize-ovlsize)+ovlsize");
      G.ntiles[k] = e->ntiles[k];
      G.pad[k] = e->padsize[k];
    }
    G.padvox = G.pad[0] * G.pad[1] * G.pad[2];
    if (G.pad[0] >= (1ll << 31) || G.pad[1] >= (1ll << 31) || G.pad[2] >= (1ll << 31) || G.padvox >= (1ll << 40))
      bad(IQ_ERR_UNSUPPORTED, "simulation grid too large");
    G.npath = e->npath;
    G.nreal = e->nreal;
    const int64_t nt = G.ntiles[0] * G.ntiles[1] * G.ntiles[2];
    for (int64_t i = 0; i < e->npath; i++)
      if (e->path[i] < 0 || e->path[i] >= nt) bad(IQ_ERR_ARG, "path entry outside the tile grid");
    G.real0 = pl->nreal;
    pl->nreal += G.nreal;
    pl->grids.push_back(G);
  }

  // ---- steps and masks (iqsim.jl:172-205, 252-275): per grid, shared by its
  // realizations; mask shapes shared by all grids
  std::map<std::vector<uint8_t>, int> mask_index;
  std::vector<HostMask> hmasks;
  std::vector<int32_t> htvox;
  std::vector<int64_t> mask_uses;
  pl->hard = false;
  for (int gi = 0; gi < ngrid; gi++) {
  const iq_plan_desc* e = &descs[gi];
  HostGrid& G = pl->grids[gi];
  G.step0 = (int64_t)pl->steps.size();
  if (e->hardgrid) pl->hard = true;
  const int64_t nt = G.ntiles[0] * G.ntiles[1] * G.ntiles[2];
  std::vector<uint8_t> pasted((size_t)nt, 0);
  for (int64_t i = 0; i < e->npath; i++) {
    IqStep st{};
    int64_t ind = e->path[i];
    if (pasted[ind]) bad(IQ_ERR_ARG, "path visits a tile twice");
    if (e->skipped[ind]) { st.active = 0; st.mask = -1; pasted[ind] = 0; pl->steps.push_back(st); continue; }
    st.active = 1;
    int64_t ti3[3] = {ind % G.ntiles[0], (ind / G.ntiles[0]) % G.ntiles[1], ind / (G.ntiles[0] * G.ntiles[1])};
    for (int k = 0; k < 3; k++) st.origin[k] = (int32_t)(ti3[k] * pl->spacing[k]);
    std::vector<uint8_t> bits((size_t)g.tilevox, 0);
    int any = 0;
    for (int k = 0; k < 3; k++) {
      if (pl->ovl[k] <= 1) continue;
      for (int side = 0; side < 2; side++) {
        int64_t nb[3] = {ti3[0], ti3[1], ti3[2]};
        nb[k] += side == 0 ? -1 : 1;
        if (nb[k] < 0 || nb[k] >= G.ntiles[k]) continue;
        if (!pasted[nb[0] + G.ntiles[0] * (nb[1] + G.ntiles[1] * nb[2])]) continue;
        IqSlab sl{};
        sl.dim = k;
        sl.side = side;
        for (int q = 0; q < 3; q++) { sl.lo[q] = 0; sl.sz[q] = t3[q]; }
        if (side == 0) sl.sz[k] = (int32_t)pl->ovl[k];
        else { sl.lo[k] = (int32_t)pl->spacing[k]; sl.sz[k] = t3[k] - (int32_t)pl->spacing[k]; }
        st.slab[st.nslab++] = sl;
        for (int z = sl.lo[2]; z < sl.lo[2] + sl.sz[2]; z++)
          for (int y = sl.lo[1]; y < sl.lo[1] + sl.sz[1]; y++)
            for (int x = sl.lo[0]; x < sl.lo[0] + sl.sz[0]; x++) bits[x + t3[0] * (y + t3[1] * z)] = 1;
        any = 1;
      }
    }
    // slabs two slices thick (no interior: their cut is fixed) go last
    std::stable_partition(st.slab, st.slab + st.nslab, [](const IqSlab& x) { return x.sz[x.dim] > 2; });
    st.ncut = 0;
    while (st.ncut < st.nslab && st.slab[st.ncut].sz[st.slab[st.ncut].dim] > 2) st.ncut++;
    st.mask = -1;
    if (any) {
      auto it = mask_index.find(bits);
      if (it == mask_index.end()) {
        HostMask hm;
        hm.bits = bits;
        std::vector<int> brk[3];
        for (int k = 0; k < 3; k++) {
          brk[k] = {0, t3[k]};
          if (pl->ovl[k] > 1) { brk[k].push_back((int)pl->ovl[k]); brk[k].push_back((int)pl->spacing[k]); }
          std::sort(brk[k].begin(), brk[k].end());
          brk[k].erase(std::unique(brk[k].begin(), brk[k].end()), brk[k].end());
          brk[k].erase(std::remove_if(brk[k].begin(), brk[k].end(), [&](int v) { return v < 0 || v > t3[k]; }), brk[k].end());
        }
        hm.boxes = decompose(bits, t3, brk);
        for (uint8_t b : bits) hm.V += b;
        int vol = 0;
        for (const IqBox& b : hm.boxes) vol += b.wx * b.wy * b.wz;
        if (vol != hm.V || (int)hm.boxes.size() > IQ_MAX_BOX) bad(IQ_ERR_ARG, "internal: overlap mask decomposition failed");
        mask_index[bits] = (int)hmasks.size();
        st.mask = (int)hmasks.size();
        hmasks.push_back(hm);
        mask_uses.push_back(0);
      } else {
        st.mask = it->second;
      }
      mask_uses[st.mask]++;
    }
    // hard data in the tile (iqsim.jl:210-219, indicator!/event! utils.jl:18-36)
    if (e->hardgrid) {
      st.hoff = (int32_t)htvox.size();
      for (int z = 0; z < t3[2]; z++)
        for (int y = 0; y < t3[1]; y++)
          for (int x = 0; x < t3[0]; x++) {
            double v = e->hardgrid[st.origin[0] + x + G.pad[0] * (st.origin[1] + y + G.pad[1] * (int64_t)(st.origin[2] + z))];
            if (!std::isnan(v)) htvox.push_back(x + t3[0] * (y + t3[1] * z));
          }
      st.hcnt = (int32_t)htvox.size() - st.hoff;
      st.hard = st.hcnt > 0;
    }
    if (st.hard || pl->naux > 0) pl->any_multi = true;
    pasted[ind] = 1;
    pl->steps.push_back(st);
  }
  G.waves = build_waves(G, pl->steps.data() + G.step0, e->path, !half);
  pl->nwaves = std::max(pl->nwaves, G.waves.size());
  }

  // ---- offset tables: per mask, tile-linear and TI-linear offsets (ascending column-major)
  std::vector<int32_t> tvox, toff;
  auto ti_off = [&](int32_t t) {
    int x = t % t3[0], y = (t / t3[0]) % t3[1], z = t / (t3[0] * t3[1]);
    return (int32_t)(x + (int64_t)g.ti[0] * (y + (int64_t)g.ti[1] * z));
  };
  for (HostMask& hm : hmasks) {
    IqMask m{};
    m.nbox = (int32_t)hm.boxes.size();
    int32_t kofs = 0;
    for (int b = 0; b < m.nbox; b++) {
      IqBox bx = hm.boxes[b];
      bx.wxp = (bx.wx + 3) / 4 * 4;
      bx.kofs = kofs;
      kofs += bx.wxp * bx.wy * bx.wz;
      m.box[b] = bx;
      hm.boxes[b] = bx;
    }
    m.kbox = kofs;
    m.V = hm.V;
    m.off = (int32_t)tvox.size();
    for (int32_t t = 0; t < g.tilevox; t++)
      if (hm.bits[t]) { tvox.push_back(t); toff.push_back(ti_off(t)); }
    m.a2 = -1;
    pl->masks.push_back(m);
  }
  const int32_t full_off = (int32_t)tvox.size();
  for (int32_t t = 0; t < g.tilevox; t++) { tvox.push_back(t); toff.push_back(ti_off(t)); }
  std::vector<int32_t> htoff(htvox.size());
  for (size_t i = 0; i < htvox.size(); i++) htoff[i] = ti_off(htvox[i]);

  // ---- training image
  IqDev& H = pl->hd;
  H.g = g;
  const double* tid = (const double*)d->ti;
  const float* tif = (const float*)d->ti;
  if (d->ti_dtype == IQ_F32) {
    H.ti = pl->mem.upload(tif, tivox, s);
    pl->fast = fits_range(tif, tivox);
    if (pl->fast) H.ti32 = (const float*)H.ti;
  } else {
    H.ti = pl->mem.upload(tid, tivox, s);
    pl->fast = fits_f32(tid, tivox);
    if (pl->fast) {
      std::vector<float> t32((size_t)tivox);
      for (int64_t i = 0; i < tivox; i++) t32[i] = (float)tid[i];
      H.ti32 = pl->mem.upload(t32.data(), tivox, s);
      HC(hipStreamSynchronize(s));
    }
  }
  H.ti_f32 = d->ti_dtype == IQ_F32;
  H.naux = pl->naux;
  H.tol = pl->tol;
  H.disabled = pl->mem.upload(d->disabled, (size_t)g.P, s);
  std::vector<int32_t> en;
  en.reserve((size_t)g.P);
  for (int64_t c = 0; c < g.P; c++)
    if (!d->disabled[c]) en.push_back((int32_t)c);
  pl->nen = (int64_t)en.size();
  // every window holds an inactive voxel: the reference's band would take
  // every (infinitely distant) pattern and its relaxation would not end
  // (iqsim.jl:237, relaxation.jl:7-36); the plan is refused instead
  if (pl->nen == 0) bad(IQ_ERR_NOCAND, "every candidate pattern is disabled (all training-image windows hold inactive voxels)");
  H.enabled = pl->mem.upload(en.data(), en.size(), s);
  pl->enabled_host = en;
  H.nen = pl->nen;
  H.steps = pl->mem.upload(pl->steps.data(), pl->steps.size(), s);
  H.tvox = pl->mem.upload(tvox.data(), tvox.size(), s);
  H.toff = pl->mem.upload(toff.data(), toff.size(), s);
  H.htvox = pl->mem.upload(htvox.data(), htvox.size(), s);
  H.htoff = pl->mem.upload(htoff.data(), htoff.size(), s);
  H.full_off = full_off;

  // ---- exact A2 maps of the overlap masks (A2 chains over the TI)
  std::vector<const double*> a2ptr;
  std::vector<const float*> a2fptr;
  for (size_t mi = 0; mi < pl->masks.size(); mi++) {
    IqMask& m = pl->masks[mi];
    double* a2 = pl->mem.alloc<double>(g.P);
    if (pl->dtype == IQ_F32)
      hipLaunchKernelGGL((k_exact_map<float, 2>), dim3(nblk(g.P, 256)), dim3(256), 0, s, (const float*)H.ti, g,
                         H.toff + m.off, (const double*)nullptr, m.V, (const double*)nullptr, a2);
    else
      hipLaunchKernelGGL((k_exact_map<double, 2>), dim3(nblk(g.P, 256)), dim3(256), 0, s, (const double*)H.ti, g,
                         H.toff + m.off, (const double*)nullptr, m.V, (const double*)nullptr, a2);
    HCL();
    // padded to whole 32-candidate chunks of the bank scan, the pad NaN
    const int64_t P32 = (g.P + 31) / 32 * 32;
    float* a2f = pl->mem.alloc<float>(P32);
    if (P32 > g.P) HC(hipMemsetAsync(a2f + g.P, 0xff, (P32 - g.P) * sizeof(float), s));
    hipLaunchKernelGGL(k_a2_f32, dim3(std::min<int64_t>(nblk(g.P, 256), 4096)), dim3(256), 0, s, (const double*)a2,
                       (const uint8_t*)H.disabled, g.P, a2f);
    HCL();
    m.a2 = (int32_t)a2ptr.size();
    a2ptr.push_back(a2);
    a2fptr.push_back(a2f);
  }
  H.masks = pl->mem.upload(pl->masks.data(), pl->masks.size(), s);
  H.a2maps = (const double* const*)pl->mem.upload(a2ptr.data(), a2ptr.size(), s);
  H.a2f = (const float* const*)pl->mem.upload(a2fptr.data(), a2fptr.size(), s);
  // one scan configuration for the plan: every mask must fit its LDS, the
  // most used mask decides the tiling; 3-D grids prefer the 2-D register block
  if (pl->fast && !hmasks.empty()) {
    size_t best_mi = std::max_element(mask_uses.begin(), mask_uses.end()) - mask_uses.begin();
    double bestcost = 1e300;
    for (int c = 0; c < kNumScanCfg; c++) {
      const ScanCfg& sc = kScanCfg[c];
      if (sc.kind == 2 && g.dist[2] < sc.LZ / 2) continue;
      bool perz = false;
      size_t lmax = 0;
      for (const HostMask& hm : hmasks) lmax = std::max(lmax, scan_lds_bytes(sc, hm.boxes, false));
      if (sc.kind == 2 && lmax > 80 * 1024) {
        perz = true;
        lmax = 0;
        for (const HostMask& hm : hmasks) lmax = std::max(lmax, scan_lds_bytes(sc, hm.boxes, true));
      }
      if (lmax > (sc.kind == 1 ? (size_t)64 * 1024 : kScanLdsMax)) continue;
      double padded = (double)nblk(g.dist[0], sc.R * sc.LX) * sc.R * sc.LX * nblk(g.dist[1], sc.LY * sc.RY) * sc.LY *
                      sc.RY * nblk(g.dist[2], sc.LZ) * sc.LZ;
      size_t l = scan_lds_bytes(sc, hmasks[best_mi].boxes, perz);
      double cost = padded * (1.0 + 0.02 * (double)l / 16384.0) * (sc.kind == 2 ? 0.4 : 1.0);
      if (cost < bestcost) { bestcost = cost; pl->scan_cfg = c; pl->scan_lds = lmax; pl->scan_perz = perz; }
    }
    if (pl->scan_cfg < 0) pl->fast = false;
    else pl->nbz = (int)nblk(g.dist[2], kScanCfg[pl->scan_cfg].LZ);
    if (pl->fast && pl->scan_lds > 64 * 1024) {
      HC(hipFuncSetAttribute((const void*)k_scan2_f32<4, 4, 8, 8>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kScanLdsMax));
      HC(hipFuncSetAttribute((const void*)k_scan2_f32<4, 4, 16, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kScanLdsMax));
    }
  }
  if (!pl->fast) H.ti32 = nullptr;
  // MFMA scan over the union of the masks
  {
    const char* env = std::getenv("IQHIP_SCAN");
    const bool want = !(env && std::string(env) == "valu");
    std::vector<uint8_t> ubits((size_t)g.tilevox, 0);
    for (const HostMask& hm : hmasks)
      for (int32_t t = 0; t < g.tilevox; t++) ubits[t] |= hm.bits[t];
    std::vector<int> brk[3];
    for (int k = 0; k < 3; k++) {
      brk[k] = {0, t3[k]};
      if (pl->ovl[k] > 1) { brk[k].push_back((int)pl->ovl[k]); brk[k].push_back((int)pl->spacing[k]); }
      std::sort(brk[k].begin(), brk[k].end());
      brk[k].erase(std::unique(brk[k].begin(), brk[k].end()), brk[k].end());
      brk[k].erase(std::remove_if(brk[k].begin(), brk[k].end(), [&](int v) { return v < 0 || v > t3[k]; }), brk[k].end());
    }
    std::vector<IqChunk> chunks;
    if (want && pl->fast && !hmasks.empty() && build_ktab(ubits, t3, brk, H.ktab, chunks)) {
      // Group size: a group of nw wavefronts covers 4 nw candidate rows.
      // Launches of at most about four 4-wavefront groups per compute unit
      // (few queries, a small candidate grid) run faster with 2-wavefront
      // groups, which spread the same rows over twice as many groups; larger
      // launches keep 4 wavefronts, whose staged block serves 16 rows.  The
      // choice is made per launch (run_batch); both layouts are prepared.
      auto lds_of = [&](int nw) {
        size_t l = 0;
        for (int b = 0; b < H.ktab.nbox; b++) {
          const IqKBox& kb = H.ktab.box[b];
          l = std::max(l, (size_t)(32 + kb.wxr - 1) * (4 * nw + kb.wyr - 1) * kb.wz * 4);
        }
        return l;
      };
      if (const char* envnw = std::getenv("IQHIP_SCAN_NW")) {
        const int e = std::atoi(envnw);
        pl->mfma_nw = e >= 4 ? 4 : (e == 3 ? 3 : 2);
        pl->mfma_nw_forced = true;
      }
      size_t lmax = 0;
      for (int nw : {2, 3, 4}) {
        if (pl->mfma_nw_forced && nw != pl->mfma_nw) continue;
        if (lds_of(nw) > 150 * 1024) continue;
        pl->lds_nw[nw] = lds_of(nw);
        lmax = std::max(lmax, pl->lds_nw[nw]);
      }
      if (pl->lds_nw[2] || pl->lds_nw[4]) {
        pl->mfma = true;
        pl->mfma_lds = lmax;
        pl->mfma_split = !bf16_exact(H.ti32, tivox);
        pl->mfma_gerr = scan_gerr(H.ktab.nstep, pl->mfma_split);
        H.chunks = pl->mem.upload(chunks.data(), chunks.size(), s);
        for (int nw : {2, 3, 4}) {
          if (!pl->lds_nw[nw]) continue;
          std::vector<int32_t> kofs(chunks.size());
          for (int b = 0; b < H.ktab.nbox; b++) {
            const IqKBox& kb = H.ktab.box[b];
            const int RX = 32 + kb.wxr - 1, RY = 4 * nw + kb.wyr - 1;
            for (int k = 2 * kb.s0; k < 2 * kb.s1; k++) {
              const IqChunk& c = chunks[k];
              // x-runs: block layout x + RX (y + RY z); y-runs: y + RY (x + RX z)
              kofs[k] = c.dir ? (((c.jz - kb.z0) * RX + (c.jx - kb.x0)) * RY + (c.jy - kb.y0)) | (1 << 30)
                              : ((c.jz - kb.z0) * RY + (c.jy - kb.y0)) * RX + (c.jx - kb.x0);
            }
          }
          pl->kofs_nw[nw] = pl->mem.upload(kofs.data(), kofs.size(), s);
        }
        std::vector<uint8_t> mb((size_t)hmasks.size() * g.tilevox);
        for (size_t m = 0; m < hmasks.size(); m++) std::memcpy(&mb[m * g.tilevox], hmasks[m].bits.data(), g.tilevox);
        H.maskbits = pl->mem.upload(mb.data(), mb.size(), s);
        if (lmax > 64 * 1024) {
          HC(hipFuncSetAttribute((const void*)k_scan_mfma<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lmax));
          HC(hipFuncSetAttribute((const void*)k_scan_mfma<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lmax));
        }
      }
    }
  }

  // ---- hard and soft data: grids per simulation grid, soft training images shared
  std::vector<const double*> auxtip, a2sp;
  if (pl->naux > 0) {
    if (!d->auxti) bad(IQ_ERR_ARG, "null soft data");
    for (int a = 0; a < pl->naux; a++) {
      if (!d->auxti[a]) bad(IQ_ERR_ARG, "null soft data");
      double* at = pl->mem.upload(d->auxti[a], (size_t)tivox, s);
      auxtip.push_back(at);
      double* a2 = pl->mem.alloc<double>(g.P);
      hipLaunchKernelGGL((k_exact_map<double, 2>), dim3(nblk(g.P, 256)), dim3(256), 0, s, (const double*)at, g,
                         H.toff + full_off, (const double*)nullptr, g.tilevox, (const double*)nullptr, a2);
      HCL();
      a2sp.push_back(a2);
    }
    H.auxti = (const double* const*)pl->mem.upload(auxtip.data(), auxtip.size(), s);
    H.a2soft = (const double* const*)pl->mem.upload(a2sp.data(), a2sp.size(), s);
  }
  std::vector<std::vector<const double*>> gaux((size_t)ngrid);  // alive until the final synchronize
  int64_t sim_off = 0, log_off = 0;
  for (int gi = 0; gi < ngrid; gi++) {
    const iq_plan_desc* e = &descs[gi];
    const HostGrid& G = pl->grids[gi];
    IqReal R{};
    for (int k = 0; k < 3; k++) R.pad[k] = (int32_t)G.pad[k];
    R.step0 = (int32_t)G.step0;
    if (e->hardgrid) R.hard = pl->mem.upload(e->hardgrid, (size_t)G.padvox, s);
    if (pl->naux > 0) {
      if (!e->aux) bad(IQ_ERR_ARG, "null soft data");
      for (int a = 0; a < pl->naux; a++) {
        if (!e->aux[a]) bad(IQ_ERR_ARG, "null soft data");
        gaux[gi].push_back(pl->mem.upload(e->aux[a], (size_t)G.padvox, s));
      }
      R.aux = (const double* const*)pl->mem.upload(gaux[gi].data(), gaux[gi].size(), s);
    }
    for (int64_t r = 0; r < G.nreal; r++) {
      R.sim_off = sim_off;
      R.log_off = log_off;
      pl->reals.push_back(R);
      sim_off += G.padvox;
      log_off += G.npath;
    }
  }
  pl->nvox = sim_off;
  pl->nlog = log_off;
  H.reals = pl->mem.upload(pl->reals.data(), pl->reals.size(), s);

  // ---- per-slot scratch: as many slots as the widest batch, within a memory budget
  const int64_t P = g.P;
  const int64_t nw64 = (P + 63) / 64, nw32 = (P + 31) / 32;
  const int nmap = 2 + pl->naux;
  int kmax = 4;
  for (const IqMask& m : pl->masks) kmax = std::max(kmax, m.kbox);
  int64_t maxn = 0;
  for (const IqStep& st : pl->steps)
    for (int q = 0; q < st.nslab; q++) {
      const IqSlab& sl = st.slab[q];
      maxn = std::max(maxn, (int64_t)sl.sz[0] * sl.sz[1] * sl.sz[2] / sl.sz[sl.dim] * std::max(0, sl.sz[sl.dim] - 2));
    }
  pl->cut_big = maxn > 4 * CB && maxn <= CUT_BIG_MAX;
  for (const IqStep& st : pl->steps) pl->max_nslab = std::max(pl->max_nslab, st.ncut);
  for (const IqStep& st : pl->steps)
    for (int q = 0; q < st.nslab; q++) {
      const IqSlab& sl = st.slab[q];
      const int64_t n = (int64_t)sl.sz[0] * sl.sz[1] * sl.sz[2] / sl.sz[sl.dim] * std::max(0, sl.sz[sl.dim] - 2);
      pl->cut_lds_max = std::max(pl->cut_lds_max, cut_lds_bytes(n, pl->cut_big));
    }
  pl->cut_wave = maxn <= CUT_WAVE_MAX;
  const int64_t cutper = (maxn > IQ_PR_NLDS && !pl->cut_big) ? maxn * CUT_NODE_BYTES : 0;
  size_t per_slot = (size_t)nmap * g.tilevox * 8 + (size_t)kmax * 4 + 64 + (size_t)P * (4 + 8 + 4 + 4 + 16) + nw64 * 8 +
                    (size_t)IQ_SORT_MAX * 12 + (size_t)(nw64 + 1) * 8 + (size_t)IQ_MAX_SLAB * g.tilevox +
                    (size_t)IQ_MAX_SLAB * cutper;
  if (pl->any_multi) per_slot += (size_t)nmap * P * 8 + (size_t)nmap * nw32 * 4 + (size_t)P * 12;  // maps, bitmaps, events, values
  int64_t widest = 1;
  for (size_t k = 0; k < pl->nwaves; k++) {
    int64_t width = 0;  // lockstep wave k over every grid and realization
    for (const HostGrid& G : pl->grids)
      if (k < G.waves.size()) width += (int64_t)G.waves[k].size() * G.nreal;
    widest = std::max(widest, width);
  }
  const double budget = 24.0 * (1ull << 30);
  int64_t cap = std::max<int64_t>(1, (int64_t)(budget / (double)per_slot));
  cap = std::min<int64_t>(cap, 65535 / std::max(1, pl->nbz));
  if (pl->mfma) cap = std::min<int64_t>(cap, (int64_t)(65535 / std::max(1, g.dist[2])) * 32);
  cap = std::min<int64_t>(cap, 4096);
  pl->nslots = std::min(widest, cap);
  const int64_t NS = pl->nslots;
  H.kv = pl->mem.alloc<double>((size_t)NS * nmap * g.tilevox);
  H.kv_stride = (int64_t)nmap * g.tilevox;
  H.kv32 = pl->mem.alloc<float>((size_t)NS * kmax);
  H.kv32_stride = kmax;
  H.kv32_on = pl->fast && !pl->mfma ? 1 : 0;
  H.b2 = pl->mem.alloc<double>((size_t)NS * nmap);
  H.gmin_hi = pl->mem.alloc<uint32_t>(NS);
  H.gmin64 = pl->mem.alloc<unsigned long long>(NS);
  H.lo32 = pl->mem.alloc<float>((size_t)NS * P);
  H.dexact = pl->mem.alloc<double>((size_t)NS * P);
  H.bitmap = pl->mem.alloc<unsigned long long>((size_t)NS * nw64);
  H.nw64 = nw64;
  H.band = pl->mem.alloc<int32_t>((size_t)NS * P);
  H.sup = pl->mem.alloc<int32_t>((size_t)NS * P);
  H.rkey = pl->mem.alloc<unsigned long long>((size_t)NS * 2 * P);
  H.nsup = pl->mem.alloc<int32_t>(NS);
  H.sd = pl->mem.alloc<double>((size_t)NS * IQ_SORT_MAX);
  H.si = pl->mem.alloc<int32_t>((size_t)NS * IQ_SORT_MAX);
  H.s1 = pl->mem.alloc<double>((size_t)NS * (nw64 + 1));
  H.s1_stride = nw64 + 1;
  H.pick = pl->mem.alloc<int64_t>(NS);
  H.nband = pl->mem.alloc<int64_t>(NS);
  H.keep = pl->mem.alloc<uint8_t>((size_t)NS * IQ_MAX_SLAB * g.tilevox);
  if (cutper) {
    H.cutscratch = pl->mem.alloc<char>((size_t)NS * IQ_MAX_SLAB * cutper);
    H.cutscratch_per = cutper;
  }
  H.nmap = nmap;
  H.nw32 = nw32;
  if (pl->any_multi) {
    // tiled exact maps when the candidate block's training-image region, the
    // query values and their region offsets fit in LDS
    // (voxel lists as offsets into the staged region; read as scalar loads)
    // (candidate blocks of 32 x 16 where the region fits, else 16 x 16;
    // IQHIP_MAPS_BX=16 forces the narrow blocks)
    const char* ebx = std::getenv("IQHIP_MAPS_BX");
    for (int bx : {32, 16}) {
      if (bx == 32 && ebx && std::atoi(ebx) == 16) continue;
      const int RX = bx + g.tile[0] - 1, RY = 16 + g.tile[1] - 1;
      const size_t need = (size_t)RX * RY * g.tile[2] * 8;
      if (need > 150 * 1024) continue;
      auto region = [&](const std::vector<int32_t>& tv) {
        std::vector<int32_t> r(tv.size());
        for (size_t i = 0; i < tv.size(); i++) {
          const int t = tv[i], x = t % g.tile[0], y = (t / g.tile[0]) % g.tile[1], z = t / (g.tile[0] * g.tile[1]);
          r[i] = x + RX * (y + RY * z);
        }
        return r;
      };
      const std::vector<int32_t> rt = region(tvox), rh = region(htvox);
      H.rtvox = pl->mem.upload(rt.data(), rt.size(), s);
      H.rhtvox = rh.empty() ? nullptr : pl->mem.upload(rh.data(), rh.size(), s);
      pl->multi_lds = need;
      pl->multi_bx = bx;
      const void* fn = bx == 32 ? (const void*)k_multi_maps_tiled<32, 16> : (const void*)k_multi_maps_tiled<16, 16>;
      if (need > 64 * 1024) HC(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
      break;
    }
    H.maps = pl->mem.alloc<double>((size_t)NS * nmap * P);
    H.bits = pl->mem.alloc<uint32_t>((size_t)NS * nmap * nw32);
    H.ev = pl->mem.alloc<int32_t>((size_t)NS * P);
    H.vals = pl->mem.alloc<double>((size_t)NS * P);
  }
  if (pl->mfma) {
    pl->mfma_ntile_max = (NS + 31) / 32;
    const size_t nb = (size_t)pl->mfma_ntile_max * H.ktab.nstep * 64;
    H.bh = pl->mem.alloc<uint4>(nb);
    H.bl = pl->mem.alloc<uint4>(nb);
  }
  H.err = pl->mem.alloc<int32_t>(1);
  // superset size from which the exact recheck gives each member a work-item
  // instead of a wavefront (IQHIP_RC_SMALL; speed only, same values)
  H.rc_small = RC_SMALL;
  if (const char* erc = std::getenv("IQHIP_RC_SMALL")) H.rc_small = std::max(0, std::atoi(erc));
  if (const char* eg = std::getenv("IQHIP_GRAPH")) pl->graphs = std::atoi(eg) != 0;
  if (const char* et = std::getenv("IQHIP_CUT_TEAM")) pl->cut_team = std::atoi(et) == 512 ? 512 : 1024;
  // the exact recheck runs a lane per superset member (k_band_recheck_lanes;
  // measured faster than a wavefront per member at 720- and 4,736-voxel
  // masks); IQHIP_RECHECK=groups selects the wavefront-per-member kernel
  H.rc_lanes = 0;
  if (const char* er = std::getenv("IQHIP_RECHECK")) {
    if (std::string(er) == "groups") H.rc_lanes = 1 << 30;
    if (std::string(er) == "lanes") H.rc_lanes = 0;
  }
  H.rounds = pl->mem.alloc<int64_t>(6);  // push rounds, BFS levels, cuts; plain queries, superset, band members

  // ---- results
  const size_t esz = pl->dtype == IQ_F32 ? 4 : 8;
  H.sims = pl->mem.alloc<char>((size_t)pl->nvox * esz);
  H.cuts = pl->mem.alloc<uint8_t>((size_t)pl->nvox);
  H.log = pl->mem.alloc<int64_t>((size_t)std::max<int64_t>(pl->nlog, 1));
  pl->dd = pl->mem.upload(&pl->hd, 1, s);
  HC(hipEventCreate(&pl->e0));
  HC(hipEventCreate(&pl->e1));
  if (pl->cut_lds_max > 64 * 1024) {
    HC(hipFuncSetAttribute((const void*)k_graphcut<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl->cut_lds_max));
    HC(hipFuncSetAttribute((const void*)k_graphcut<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl->cut_lds_max));
    HC(hipFuncSetAttribute((const void*)k_graphcut_big<float, 512>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl->cut_lds_max));
    HC(hipFuncSetAttribute((const void*)k_graphcut_big<double, 512>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl->cut_lds_max));
    HC(hipFuncSetAttribute((const void*)k_graphcut_big<float, 1024>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl->cut_lds_max));
    HC(hipFuncSetAttribute((const void*)k_graphcut_big<double, 1024>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl->cut_lds_max));
  }
  if (pl->dtype == IQ_F32) tune_scan<float>(pl);
  else tune_scan<double>(pl);
  HC(hipStreamSynchronize(s));
}

// bank scan: persistent blocks, as many as are resident at once on every
// CU (each owns a staging region), none beyond the chunks
void launch_mscan(iq_plan* pl, const IqQuery* q, const int4* tiles, int ntiles, hipStream_t s) {
  static thread_local int per_cu = 0;
  if (per_cu <= 0) {
    HC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_learn_mscan, 64 * LMS_WAVES, 0));
    per_cu = std::min(std::max(per_cu, 1), LMS_BLOCKS_PER_CU);
  }
  const int64_t nchunk = (pl->geo.P + 31) / 32;
  const int64_t nb = std::min<int64_t>(nblk(nchunk, LMS_WAVES), (int64_t)per_cu * pl->ctx->ncu);
  hipLaunchKernelGGL(k_learn_mscan, dim3((unsigned)nb), dim3(64 * LMS_WAVES), 0, s, (const IqDev*)pl->dd, q, tiles, ntiles);
  HCL();
}

// exact FP64 recheck of the band supersets of a batch: with a lane per
// member, or (IQHIP_RECHECK=groups) a wavefront per member
void launch_recheck(iq_plan* pl, const IqQuery* q, int nq, hipStream_t s) {
  if (pl->hd.rc_lanes == 0)
    hipLaunchKernelGGL(k_band_recheck_lanes, dim3(nq, RCL_GROUPS), dim3(64), 0, s, (const IqDev*)pl->dd, q);
  else
    hipLaunchKernelGGL(k_band_recheck, dim3(nq, RC_GROUPS), dim3(256), 0, s, (const IqDev*)pl->dd, q);
  HCL();
}

void launch_scan_batch(iq_plan* pl, const IqQuery* q, int nq, hipStream_t s) {
  const IqGeo& g = pl->geo;
  const ScanCfg& c = kScanCfg[pl->scan_cfg];
  dim3 grid(nblk(g.dist[0], c.R * c.LX), nblk(g.dist[1], c.LY * c.RY), (unsigned)(pl->nbz * nq));
  const size_t lds = pl->scan_lds;
  const IqDev* D = pl->dd;
  const int perz = pl->scan_perz ? 1 : 0;
  switch (pl->scan_cfg) {
    case 0: hipLaunchKernelGGL((k_scan_f32<4, 8, 8, 4>), grid, dim3(256), lds, s, D, q, pl->nbz); break;
    case 1: hipLaunchKernelGGL((k_scan_f32<4, 16, 4, 4>), grid, dim3(256), lds, s, D, q, pl->nbz); break;
    case 2: hipLaunchKernelGGL((k_scan_f32<2, 16, 8, 2>), grid, dim3(256), lds, s, D, q, pl->nbz); break;
    case 3: hipLaunchKernelGGL((k_scan_f32<4, 8, 16, 2>), grid, dim3(256), lds, s, D, q, pl->nbz); break;
    case 4: hipLaunchKernelGGL((k_scan2_f32<4, 4, 8, 8>), grid, dim3(256), lds, s, D, q, pl->nbz, perz); break;
    default: hipLaunchKernelGGL((k_scan2_f32<4, 4, 16, 4>), grid, dim3(256), lds, s, D, q, pl->nbz, perz); break;
  }
  HCL();
}

// One batch of independent queries: distances, selection, cut and paste.
// Selection of the plain queries: 0 full scan; 1 learned shortlist; 2 both,
// the full scan deciding and the shortlist's recall measured on its band.
enum { SEL_SCAN = 0, SEL_LEARNED = 1, SEL_CHECK = 2 };

template <typename T>
void run_batch(iq_plan* pl, const IqQuery* hq, const IqQuery* q, int nq, bool has_plain, bool has_multi, int sel,
               const int4* tiles, const int4* htiles, int ntiles) {
  const IqGeo& g = pl->geo;
  hipStream_t s = pl->ctx->stream;
  const IqDev* D = pl->dd;
  const int64_t P = g.P;
  const int nsrc_prep = has_multi ? 2 + pl->naux : 1;
  // plain batches of the MFMA scan: query preparation and B fragments in one launch
  const bool fused_prep = !has_multi && has_plain && sel == SEL_SCAN && pl->fast && pl->mfma;
  if (fused_prep) {
    const int nitems = pl->hd.ktab.nstep * ((nq + 31) / 32);
    hipLaunchKernelGGL((k_prep_bprep<T>), dim3(nq + (nitems + 3) / 4), dim3(256), 0, s, D, q, nq, nitems);
  } else {
    hipLaunchKernelGGL((k_query_prep<T>), dim3(nq, nsrc_prep), dim3(256), 0, s, D, q);
  }
  HCL();
  if (has_plain && sel != SEL_SCAN && ntiles > 0) {
    // shortlist: query encodings -> bracket from a sample -> MFMA scan of the
    // bank keeping what falls in the bracket -> the K best
    hipLaunchKernelGGL(k_learn_embed, dim3(nq), dim3(1024), 0, s, D, q);
    HCL();
    hipLaunchKernelGGL(k_learn_bfrag, dim3(ntiles), dim3(64), 0, s, D, q, tiles);
    HCL();
    hipLaunchKernelGGL(k_learn_bracket, dim3(nq), dim3(1024), 0, s, D, q);
    HCL();
    if (pl->prof) {
      if (pl->bank_used + 2 > (int64_t)pl->bank_ev.size()) {
        hipEvent_t a, b;
        HC(hipEventCreate(&a));
        HC(hipEventCreate(&b));
        pl->bank_ev.push_back(a);
        pl->bank_ev.push_back(b);
      }
      HC(hipEventRecord(pl->bank_ev[pl->bank_used], s));
      // algorithmic bytes: each mask shape's bank rows (64 bf16) and its FP32
      // A2 map once per launch (tiles of one mask are adjacent)
      int nmk = 0;
      for (int t = 0; t < ntiles; t++) nmk += (t == 0 || htiles[t].x != htiles[t - 1].x) ? 1 : 0;
      pl->bank_bytes_run += (double)nmk * (double)P * (IQ_EMB * 2 + 4);
    }
    launch_mscan(pl, q, tiles, ntiles, s);
    HCL();
    if (pl->prof) {
      HC(hipEventRecord(pl->bank_ev[pl->bank_used + 1], s));
      pl->bank_used += 2;
    }
    hipLaunchKernelGGL(k_learn_pick, dim3(nq), dim3(1024), 0, s, D, q);
    HCL();
  }
  // a check batch may hand only its first lrn_check_q queries to the full
  // scan (which decides them and measures recall on them); the shortlist
  // decides the rest as in a learned batch
  const int nfull = sel == SEL_LEARNED ? 0 : (sel == SEL_CHECK && pl->lrn_check_q > 0 ? std::min(nq, pl->lrn_check_q) : nq);
  if (has_plain && nfull < nq) {
    // shortlist -> FP32 bounds of its members -> band superset -> exact recheck
    const IqQuery* ql = q + nfull;
    const int nl = nq - nfull;
    hipLaunchKernelGGL(k_short_bound, dim3(32, nl), dim3(512), (size_t)pl->lrn_vmax * 8, s, D, ql);
    HCL();
    hipLaunchKernelGGL(k_short_select, dim3(16, nl), dim3(256), 0, s, D, ql);
    HCL();
    launch_recheck(pl, ql, nl, s);
    HCL();
    hipLaunchKernelGGL(k_band_finalize, dim3(nl), dim3(FB), 0, s, D, ql, 1);
    HCL();
  }
  if (has_plain && nfull > 0) {
    // the full scan over the batch's first nf queries
    const int nf = nfull;
    if (pl->fast) {
      if (pl->prof && pl->prof_used + 2 > (int64_t)pl->prof_ev.size()) {
        hipEvent_t a, b;
        HC(hipEventCreate(&a));
        HC(hipEventCreate(&b));
        pl->prof_ev.push_back(a);
        pl->prof_ev.push_back(b);
      }
      if (pl->mfma) {
        const int ntile = (nf + 31) / 32;
        if (!fused_prep) {
          hipLaunchKernelGGL(k_mfma_bprep, dim3(pl->hd.ktab.nstep, ntile), dim3(64), 0, s, D, q, nf);
          HCL();
        }
        const int nw = pl->mfma_nw_forced ? pl->mfma_nw : pl->nw_tiles[ntile > 1 ? 2 : 1];
        const int YB = 4 * nw;
        dim3 grid(nblk(g.dist[0], 32), nblk(g.dist[1], YB), (unsigned)(g.dist[2] * ntile));
        if (pl->prof) {
          HC(hipEventRecord(pl->prof_ev[pl->prof_used], s));
          // four 32-candidate row blocks per wavefront with a row inside the
          // grid, per x block and z x query tile: k-steps x (3 or 1) MFMAs
          double rows = 0.0;
          for (int y0 = 0; y0 < g.dist[1]; y0 += YB) rows += 4.0 * ((std::min(YB, g.dist[1] - y0) + 3) / 4);
          pl->prof_mflop += (double)grid.x * rows * grid.z * pl->hd.ktab.nstep * (pl->mfma_split ? 3.0 : 1.0) *
                            (2.0 * 32 * 32 * 16);
        }
        if (pl->mfma_split)
          hipLaunchKernelGGL((k_scan_mfma<true>), grid, dim3(64 * nw), pl->lds_nw[nw], s, D, q, nf, pl->mfma_gerr,
                             pl->kofs_nw[nw]);
        else
          hipLaunchKernelGGL((k_scan_mfma<false>), grid, dim3(64 * nw), pl->lds_nw[nw], s, D, q, nf, pl->mfma_gerr,
                             pl->kofs_nw[nw]);
        HCL();
      } else {
        if (pl->prof) HC(hipEventRecord(pl->prof_ev[pl->prof_used], s));
        launch_scan_batch(pl, q, nf, s);
      }
      if (pl->prof) {
        HC(hipEventRecord(pl->prof_ev[pl->prof_used + 1], s));
        pl->prof_used += 2;
        double f = 0.0;
        for (int i = 0; i < nf; i++) {
          const IqStep& st = pl->steps[hq[i].step];
          if (!hq[i].multi && st.mask >= 0) f += 2.0 * (double)P * (double)pl->masks[st.mask].V;
        }
        pl->batch_flop.push_back(f);
        pl->prof_flop += f;
      }
      hipLaunchKernelGGL(k_band_select, dim3(nblk(P, 256 * SEL_PT), nf), dim3(256), 0, s, D, q);
      HCL();
      launch_recheck(pl, q, nf, s);
      HCL();
    } else {
      hipLaunchKernelGGL(k_band_exact_all, dim3(nblk(P, 256), nf), dim3(256), 0, s, D, q);
      HCL();
    }
    hipLaunchKernelGGL(k_band_finalize, dim3(nf), dim3(FB), 0, s, D, q, pl->fast ? 1 : 0);
    HCL();
    if (sel == SEL_CHECK) {
      hipLaunchKernelGGL(k_learn_recall, dim3(nf), dim3(256), 0, s, D, q);
      HCL();
    }
  }
  if (has_multi) {
    if (pl->multi_lds > 0) {
      const int bx = pl->multi_bx;
      const unsigned nb = nblk(g.dist[0], bx) * nblk(g.dist[1], 16) * (unsigned)g.dist[2];
      if (bx == 32)
        hipLaunchKernelGGL((k_multi_maps_tiled<32, 16>), dim3(nb, nq, 2 + pl->naux), dim3(512), pl->multi_lds, s, D, q);
      else
        hipLaunchKernelGGL((k_multi_maps_tiled<16, 16>), dim3(nb, nq, 2 + pl->naux), dim3(256), pl->multi_lds, s, D, q);
    } else {
      hipLaunchKernelGGL(k_multi_maps, dim3(nblk(P, 256), nq, 2 + pl->naux), dim3(256), 0, s, D, q);
    }
    HCL();
    hipLaunchKernelGGL(k_relax_finalize, dim3(nq), dim3(FB), 0, s, D, q);
    HCL();
  }
  if (pl->prof) {
    if (pl->cut_used + 2 > (int64_t)pl->cut_ev.size()) {
      hipEvent_t a, b;
      HC(hipEventCreate(&a));
      HC(hipEventCreate(&b));
      pl->cut_ev.push_back(a);
      pl->cut_ev.push_back(b);
    }
    HC(hipEventRecord(pl->cut_ev[pl->cut_used], s));
  }
  if (pl->cut_wave)
    hipLaunchKernelGGL((k_graphcut_wave<T>), dim3(pl->max_nslab, nq), dim3(64), pl->cut_lds_max, s, D, q);
  else if (pl->cut_big && pl->cut_team == 1024)
    hipLaunchKernelGGL((k_graphcut_big<T, 1024>), dim3(pl->max_nslab, nq), dim3(1024), pl->cut_lds_max, s, D, q);
  else if (pl->cut_big)
    hipLaunchKernelGGL((k_graphcut_big<T, CB_BIG>), dim3(pl->max_nslab, nq), dim3(CB_BIG), pl->cut_lds_max, s, D, q);
  else
    hipLaunchKernelGGL((k_graphcut<T>), dim3(pl->max_nslab, nq), dim3(CB), pl->cut_lds_max, s, D, q);
  HCL();
  if (pl->prof) {
    HC(hipEventRecord(pl->cut_ev[pl->cut_used + 1], s));
    pl->cut_used += 2;
  }
  hipLaunchKernelGGL((k_paste<T>), dim3(nblk(g.tilevox, 256), nq), dim3(256), 0, s, D, q);
  HCL();
}

// One batch of a run: queries [off, off + n) of the run's query table and
// shortlist tiles [toff, toff + nt).
struct RunBatch {
  int64_t off;
  int n;
  bool plain, multi;
  int64_t toff;
  int nt;
};
template <typename T>
void plan_run_launches(iq_plan* pl, const std::vector<IqQuery>& qs, const std::vector<RunBatch>& batches,
                       const std::vector<int4>& tiles, bool learned);
void plan_run_finish(iq_plan* pl);

// The query table of a run: wave k of every grid, all realizations, in
// batches of at most nslots queries (each query's slot is its batch
// position); with the learned proposal, queries of one mask side by side and
// the shortlist tiles of each batch.
void plan_tables(iq_plan* pl, const double* uniforms, const int64_t* forced, bool learned, std::vector<IqQuery>& qs,
                 std::vector<int4>& tiles, std::vector<RunBatch>& batches) {
  // sampled checks: the batches plan_run will check (every lrn_check-th
  // batch with plain queries) put an evenly strided sample of their queries
  // first, so the sample spans the batch's mask shapes and realizations
  const bool sampled = learned && pl->lrn_check > 0 && pl->lrn_check_q > 0;
  int64_t nplain = 0;
  for (size_t k = 0; k < pl->nwaves; k++) {
    // lockstep: wave k of every grid, all realizations
    std::vector<IqQuery> wq;
    for (const HostGrid& G : pl->grids) {
      if (k >= G.waves.size()) continue;
      for (int64_t r = G.real0; r < G.real0 + G.nreal; r++)
        for (int64_t i : G.waves[k]) {
          const int64_t st = G.step0 + i, lo = pl->reals[r].log_off + i;
          IqQuery q{};
          q.step = (int32_t)st;
          q.real = (int32_t)r;
          q.multi = (pl->steps[st].hard || pl->naux > 0) ? 1 : 0;
          q.u = uniforms ? uniforms[lo] : 0.5;
          q.forced = forced ? forced[lo] : -1;
          wq.push_back(q);
        }
    }
    if (learned)  // queries of one mask side by side: one bank row load serves them
      std::stable_sort(wq.begin(), wq.end(), [&](const IqQuery& a, const IqQuery& b) {
        return pl->steps[a.step].mask < pl->steps[b.step].mask;
      });
    for (size_t b0 = 0; b0 < wq.size(); b0 += (size_t)pl->nslots) {
      RunBatch bt{(int64_t)qs.size(), 0, false, false, 0, 0};
      const size_t b1 = std::min(wq.size(), b0 + (size_t)pl->nslots);
      std::vector<size_t> order(b1 - b0);
      for (size_t k = b0; k < b1; k++) order[k - b0] = k;
      bool plain = false;
      for (size_t k = b0; k < b1; k++) plain |= !wq[k].multi;
      if (plain && sampled && nplain++ % pl->lrn_check == 0 && (size_t)pl->lrn_check_q < order.size()) {
        const size_t n = order.size(), ns = (size_t)pl->lrn_check_q;
        std::vector<char> taken(n, 0);
        std::vector<size_t> re;
        for (size_t j = 0; j < ns; j++) {
          const size_t i = j * n / ns;
          taken[i] = 1;
          re.push_back(b0 + i);
        }
        for (size_t i = 0; i < n; i++)
          if (!taken[i]) re.push_back(b0 + i);
        order = re;
      } else if (plain && !sampled) {
        nplain++;
      }
      for (size_t k : order) {
        IqQuery q = wq[k];
        q.slot = (int32_t)(qs.size() - bt.off);
        bt.plain |= !q.multi;
        bt.multi |= q.multi != 0;
        qs.push_back(q);
        bt.n++;
      }
      if (learned) {
        // tiles of up to 32 consecutive plain queries of one overlap mask
        bt.toff = (int64_t)tiles.size();
        for (int i = 0; i < bt.n; i++) {
          const IqQuery& q = qs[bt.off + i];
          const int mk = pl->steps[q.step].mask;
          if (q.multi || mk < 0) continue;
          if (bt.nt > 0 && tiles.back().x == mk && tiles.back().z < 32 && tiles.back().y + tiles.back().z == i) {
            tiles.back().z++;
          } else {
            tiles.push_back(make_int4(mk, i, 1, 0));
            bt.nt++;
          }
        }
      }
      batches.push_back(bt);
    }
  }
}

template <typename T>
void plan_run(iq_plan* pl, const double* uniforms, const int64_t* forced) {
  hipStream_t s = pl->ctx->stream;
  const IqGeo& g = pl->geo;
  IqDev& H = pl->hd;
  pl->prof_used = 0;
  pl->prof_flop = 0.0;
  pl->prof_mflop = 0.0;
  pl->prof_sec = 0.0;
  pl->prof_launches = 0;
  pl->cut_used = 0;
  pl->cut_sec = 0.0;
  pl->cut_launches = 0;
  pl->bank_used = 0;
  pl->bank_sec = 0.0;
  pl->bank_launches = 0;
  pl->bank_bytes_run = 0.0;
  pl->batch_flop.clear();
  // query table of every batch
  std::vector<IqQuery> qs;
  std::vector<int4> tiles;
  std::vector<RunBatch> batches;
  plan_tables(pl, uniforms, forced, pl->lrn_ready && pl->lrn_on, qs, tiles, batches);
  if ((int64_t)qs.size() > pl->dq_cap) {
    pl->dq = pl->mem.alloc<IqQuery>(qs.size());
    pl->dq_cap = (int64_t)qs.size();
  }
  if (!qs.empty()) HC(hipMemcpyAsync(pl->dq, qs.data(), qs.size() * sizeof(IqQuery), hipMemcpyHostToDevice, s));
  if ((int64_t)tiles.size() > pl->dt_cap) {
    pl->dt = pl->mem.alloc<int4>(tiles.size());
    pl->dt_cap = (int64_t)tiles.size();
  }
  if (!tiles.empty()) HC(hipMemcpyAsync(pl->dt, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice, s));
  bool learned = pl->lrn_ready && pl->lrn_on;
  // replay: the launch sequence depends on the plan, the mode and the query
  // table's place only (uniforms and forced picks are read from the table)
  const int gk = learned ? 1 : 0;
  const bool graphable = pl->graphs && !pl->prof && !(learned && pl->lrn_check > 0);
  HC(hipEventRecord(pl->e0, s));
  if (graphable && pl->gexec[gk] &&
      (pl->gdq[gk] != pl->dq || pl->gdt[gk] != pl->dt || pl->gnq[gk] != (int64_t)qs.size())) {
    HC(hipGraphExecDestroy(pl->gexec[gk]));
    pl->gexec[gk] = nullptr;
  }
  if (graphable && pl->gexec[gk]) {
    HC(hipGraphLaunch(pl->gexec[gk], s));
    pl->graph_launches++;
    for (int k = 0; k < 4; k++) pl->lrn_stats[k] = 0;
    pl->lrn_fell_back = false;
  } else {
    if (graphable) HC(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      plan_run_launches<T>(pl, qs, batches, tiles, learned);
    } catch (...) {
      if (graphable) {
        hipGraph_t gr = nullptr;
        (void)hipStreamEndCapture(s, &gr);
        if (gr) (void)hipGraphDestroy(gr);
      }
      throw;
    }
    if (graphable) {
      hipGraph_t gr = nullptr;
      HC(hipStreamEndCapture(s, &gr));
      hipGraphExec_t ex = nullptr;
      const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      HC(e);
      pl->gexec[gk] = ex;
      pl->gdq[gk] = pl->dq;
      pl->gdt[gk] = pl->dt;
      pl->gnq[gk] = (int64_t)qs.size();
      HC(hipGraphLaunch(ex, s));
      pl->graph_launches++;
    }
  }
  HC(hipEventRecord(pl->e1, s));
  HC(hipStreamSynchronize(s));
  plan_run_finish(pl);
}

// State a run starts from: error word, counters, pick log, empty grids.
template <typename T>
void plan_resets(iq_plan* pl) {
  hipStream_t s = pl->ctx->stream;
  IqDev& H = pl->hd;
  hipLaunchKernelGGL((k_fill<int32_t>), dim3(1), dim3(64), 0, s, H.err, (int64_t)1, 0);
  hipLaunchKernelGGL((k_fill<int64_t>), dim3(1), dim3(64), 0, s, H.rounds, (int64_t)6, (int64_t)0);
  hipLaunchKernelGGL((k_fill<int64_t>), dim3(256), dim3(256), 0, s, H.log, pl->nlog, (int64_t)-1);
  hipLaunchKernelGGL((k_fill<T>), dim3(1024), dim3(256), 0, s, (T*)H.sims, pl->nvox, (T)0);
  hipLaunchKernelGGL((k_fill<uint8_t>), dim3(1024), dim3(256), 0, s, H.cuts, pl->nvox, (uint8_t)0);
  hipLaunchKernelGGL((k_fill<int64_t>), dim3(16), dim3(256), 0, s, H.pick, pl->nslots, (int64_t)-1);
  HCL();
}

// The recall checks' fallback rule: a recall below the threshold ends the
// shortlist once this many queries have been checked, or at once when it
// misses by more than LRN_GROSS times the allowed loss 1 - min_recall.
#define LRN_DECIDE_QUERIES 256
#define LRN_GROSS 4.0

// Every launch of one run, in order: resets, then the batches.  Host decisions
// between batches (the recall checks' fallback) read back device counters.
template <typename T>
void plan_run_launches(iq_plan* pl, const std::vector<IqQuery>& qs, const std::vector<RunBatch>& batches,
                       const std::vector<int4>& tiles, bool learned) {
  hipStream_t s = pl->ctx->stream;
  IqDev& H = pl->hd;
  plan_resets<T>(pl);
  pl->lrn_fell_back = false;
  for (int k = 0; k < 4; k++) pl->lrn_stats[k] = 0;
  if (learned) {
    hipLaunchKernelGGL((k_fill<unsigned long long>), dim3(1), dim3(64), 0, s, H.lrn.stats, (int64_t)4, 0ull);
    HCL();
  }
  int64_t nplain = 0;
  for (const RunBatch& b : batches) {
    int sel = SEL_SCAN;
    if (learned && b.plain) sel = (pl->lrn_check > 0 && nplain++ % pl->lrn_check == 0) ? SEL_CHECK : SEL_LEARNED;
    run_batch<T>(pl, qs.data() + b.off, pl->dq + b.off, b.n, b.plain, b.multi, sel, pl->dt + b.toff,
                 tiles.data() + b.toff, b.nt);
    if (sel == SEL_CHECK) {
      // the full scan decided this batch: stop using the shortlist once its
      // cumulative recall on the exact bands falls below the threshold.  The
      // recall pools band members, and a few queries with wide bands weigh
      // heavily in it, so a small sample is noisy: a shortfall counts once
      // LRN_DECIDE_QUERIES queries have been checked, or at once when the
      // recall misses by more than LRN_GROSS times the allowed loss
      unsigned long long st[4];
      HC(hipMemcpyAsync(st, H.lrn.stats, sizeof(st), hipMemcpyDeviceToHost, s));
      HC(hipStreamSynchronize(s));
      for (int k = 0; k < 4; k++) pl->lrn_stats[k] = (int64_t)st[k];
      const double rec = st[0] > 0 ? (double)st[1] / (double)st[0] : 1.0;
      const double loss = 1.0 - pl->lrn_min_recall;
      if (rec < pl->lrn_min_recall && (st[2] >= LRN_DECIDE_QUERIES || rec < 1.0 - LRN_GROSS * loss)) {
        learned = false;
        pl->lrn_fell_back = true;
      }
    }
  }
}

// After a run: device time, profiling events, counters, error word.
void plan_run_finish(iq_plan* pl) {
  IqDev& H = pl->hd;
  HC(hipEventElapsedTime(&pl->last_ms, pl->e0, pl->e1));
  for (int64_t k = 0; k + 1 < pl->prof_used; k += 2) {
    float ms = 0.0f;
    HC(hipEventElapsedTime(&ms, pl->prof_ev[k], pl->prof_ev[k + 1]));
    pl->prof_sec += ms * 1e-3;
    pl->prof_launches++;
  }
  for (int64_t k = 0; k + 1 < pl->cut_used; k += 2) {
    float ms = 0.0f;
    HC(hipEventElapsedTime(&ms, pl->cut_ev[k], pl->cut_ev[k + 1]));
    pl->cut_sec += ms * 1e-3;
    pl->cut_launches++;
  }
  for (int64_t k = 0; k + 1 < pl->bank_used; k += 2) {
    float ms = 0.0f;
    HC(hipEventElapsedTime(&ms, pl->bank_ev[k], pl->bank_ev[k + 1]));
    pl->bank_sec += ms * 1e-3;
    pl->bank_launches++;
  }
  pl->bank_bytes = pl->bank_bytes_run;
  int32_t code = 0;
  HC(hipMemcpy(&code, H.err, sizeof(int32_t), hipMemcpyDeviceToHost));
  HC(hipMemcpy(pl->cut_stats, H.rounds, 6 * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (pl->lrn_ready && pl->lrn_on) {
    unsigned long long st[4];
    HC(hipMemcpy(st, H.lrn.stats, sizeof(st), hipMemcpyDeviceToHost));
    for (int k = 0; k < 4; k++) pl->lrn_stats[k] = (int64_t)st[k];
  }
  check_kernel_error(code);
}

void sgemm(hipStream_t s, int M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbn,
           int64_t sbk, float* C, int64_t ldc) {
  hipLaunchKernelGGL(k_sgemm, dim3(nblk(N, LG_T), nblk(M, LG_T)), dim3(256), 0, s, M, N, K, A, sam, sak, B, sbn, sbk, C,
                     ldc);
  HCL();
}

// Transposed heads of every mask (read by the query encodings).
void refresh_wt(iq_plan* pl) {
  hipStream_t s = pl->ctx->stream;
  for (int m = 0; m < (int)pl->masks.size(); m++) {
    const int V = pl->masks[m].V;
    hipLaunchKernelGGL(k_learn_transpose, dim3(V), dim3(IQ_EMB), 0, s, pl->hd.lrn.W + pl->lrn_wofs[m], V,
                       (float*)pl->hd.lrn.WT + pl->lrn_wofs[m]);
    HCL();
  }
}

// Rows of Y (64 x V) made orthonormal into W: W = L^-1 Y with Y Y^T = L L^T
// (Cholesky in FP64 on the host), applied twice.
void orthonormalize(hipStream_t s, DMem& tmp, const float* Y, float* W, int V) {
  float* G = tmp.alloc<float>(IQ_EMB * IQ_EMB);
  float* Li = tmp.alloc<float>(IQ_EMB * IQ_EMB);
  float* Y1 = tmp.alloc<float>((size_t)IQ_EMB * V);
  const float* src = Y;
  for (int pass = 0; pass < 2; pass++) {
    float* dst = pass == 0 ? Y1 : W;
    sgemm(s, IQ_EMB, IQ_EMB, V, src, V, 1, src, V, 1, G, IQ_EMB);
    std::vector<float> gh(IQ_EMB * IQ_EMB);
    HC(hipMemcpyAsync(gh.data(), G, gh.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
    std::vector<double> L(IQ_EMB * IQ_EMB, 0.0);
    double tr = 0.0;
    for (int i = 0; i < IQ_EMB; i++) tr += gh[i * IQ_EMB + i];
    const double ridge = 1e-9 * tr / IQ_EMB + 1e-30;
    for (int i = 0; i < IQ_EMB; i++)
      for (int j = 0; j <= i; j++) {
        double v = 0.5 * ((double)gh[i * IQ_EMB + j] + (double)gh[j * IQ_EMB + i]) + (i == j ? ridge : 0.0);
        for (int k = 0; k < j; k++) v -= L[i * IQ_EMB + k] * L[j * IQ_EMB + k];
        if (i == j) {
          if (!(v > 0.0)) bad(IQ_ERR_UNSUPPORTED, "learned proposal: training windows span fewer than 64 directions");
          L[i * IQ_EMB + i] = std::sqrt(v);
        } else {
          L[i * IQ_EMB + j] = v / L[j * IQ_EMB + j];
        }
      }
    std::vector<float> li(IQ_EMB * IQ_EMB, 0.0f);  // inverse of the lower-triangular L
    for (int c = 0; c < IQ_EMB; c++) {
      std::vector<double> x(IQ_EMB, 0.0);
      for (int i = 0; i < IQ_EMB; i++) {
        double v = i == c ? 1.0 : 0.0;
        for (int k = 0; k < i; k++) v -= L[i * IQ_EMB + k] * x[k];
        x[i] = v / L[i * IQ_EMB + i];
      }
      for (int i = 0; i < IQ_EMB; i++) li[i * IQ_EMB + c] = (float)x[i];
    }
    HC(hipMemcpyAsync(Li, li.data(), li.size() * sizeof(float), hipMemcpyHostToDevice, s));
    sgemm(s, IQ_EMB, V, IQ_EMB, Li, IQ_EMB, 1, src, 1, V, dst, V);
    HC(hipStreamSynchronize(s));
    src = dst;
  }
}

// Encoders of every overlap mask from nsamp training-image windows (iters
// rounds of block subspace iteration Y = X^T X W^T, rows orthonormalized),
// then the bank of every candidate's encoding.
void plan_learn(iq_plan* pl, int64_t nsamp, int iters, uint64_t seed) {
  hipStream_t s = pl->ctx->stream;
  IqDev& H = pl->hd;
  const IqGeo& g = pl->geo;
  const int nm = (int)pl->masks.size();
  if (nm == 0) bad(IQ_ERR_ARG, "learned proposal: the plan has no overlap mask");
  if (nsamp < IQ_EMB || iters < 0) bad(IQ_ERR_ARG, "learned proposal: need at least 64 samples");
  if (pl->enabled_host.empty()) bad(IQ_ERR_NOCAND, "every candidate pattern is disabled");
  if (!pl->fast) bad(IQ_ERR_UNSUPPORTED, "learned proposal: training image values must be exact in FP32");
  if (!pl->lrn_ready) {
    const int64_t P32 = (g.P + 31) / 32 * 32;
    std::vector<int64_t> wofs(nm), bofs(nm);
    size_t wtot = 0;
    for (int m = 0; m < nm; m++) {
      wofs[m] = (int64_t)wtot;
      wtot += (size_t)IQ_EMB * pl->masks[m].V;
      bofs[m] = (int64_t)m * IQ_EMB * P32;
    }
    H.lrn.W = pl->mem.alloc<float>(wtot);
    H.lrn.WT = pl->mem.alloc<float>(wtot);
    // each mask's rows padded to whole 32-candidate chunks (the pad rows zero)
    H.lrn.bank = pl->mem.alloc<uint16_t>((size_t)nm * IQ_EMB * P32);
    HC(hipMemsetAsync((void*)H.lrn.bank, 0, (size_t)nm * IQ_EMB * P32 * sizeof(uint16_t), s));
    H.lrn.wofs = pl->mem.upload(wofs.data(), wofs.size(), s);
    H.lrn.bofs = pl->mem.upload(bofs.data(), bofs.size(), s);
    H.lrn.eq = pl->mem.alloc<float>((size_t)pl->nslots * IQ_EMB);
    H.lrn.bq = pl->mem.alloc<uint4>((size_t)pl->nslots * (IQ_EMB / 16) * 64);
    H.lrn.brk = pl->mem.alloc<uint32_t>((size_t)pl->nslots * 4);
    H.lrn.binfo = pl->mem.alloc<uint4>((size_t)pl->nslots);
    H.lrn.stage_cap = LMS_STAGE;
    H.lrn.stage = pl->mem.alloc<int4>((size_t)LMS_BLOCKS_PER_CU * pl->ctx->ncu * LMS_STAGE);
    H.lrn.lcnt = pl->mem.alloc<int32_t>(pl->nslots);
    H.lrn.sbits = pl->mem.alloc<unsigned long long>((size_t)pl->nslots * H.nw64);
    H.lrn.stats = pl->mem.alloc<unsigned long long>(4);
    H.lrn.slist = pl->mem.alloc<int32_t>((size_t)pl->nslots * g.P);
    H.lrn.nslist = pl->mem.alloc<int32_t>(pl->nslots);
    for (int m = 0; m < nm; m++) pl->lrn_vmax = std::max(pl->lrn_vmax, pl->masks[m].V);
    const size_t bound_lds = (size_t)pl->lrn_vmax * 8;  // query values and window offsets (+ 2 KB of unit starts)
    if (bound_lds > 156 * 1024) bad(IQ_ERR_UNSUPPORTED, "learned proposal: overlap masks of more than 19,968 voxels");
    if (bound_lds > 64 * 1024)
      HC(hipFuncSetAttribute((const void*)k_short_bound, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bound_lds));
    pl->lrn_wofs = wofs;
    pl->lrn_bofs = bofs;
  }
  std::mt19937_64 rng(seed);
  const int64_t n = nsamp;
  std::vector<int32_t> cands((size_t)n);
  std::uniform_int_distribution<size_t> pick(0, pl->enabled_host.size() - 1);
  for (int64_t i = 0; i < n; i++) cands[i] = pl->enabled_host[pick(rng)];
  DMem tmp;
  int32_t* dc = tmp.upload(cands.data(), cands.size(), s);
  std::normal_distribution<float> nrm(0.0f, 1.0f);
  for (int m = 0; m < nm; m++) {
    const int V = pl->masks[m].V;
    DMem it;
    float* X = it.alloc<float>((size_t)n * V);
    float* Z = it.alloc<float>((size_t)n * IQ_EMB);
    float* Y = it.alloc<float>((size_t)IQ_EMB * V);
    float* Wm = (float*)H.lrn.W + pl->lrn_wofs[m];
    hipLaunchKernelGGL(k_learn_gather, dim3((unsigned)n), dim3(256), 0, s, (const IqDev*)pl->dd, m, (const int32_t*)dc, X);
    HCL();
    std::vector<float> w0((size_t)IQ_EMB * V);
    for (float& v : w0) v = nrm(rng);
    HC(hipMemcpyAsync(Y, w0.data(), w0.size() * sizeof(float), hipMemcpyHostToDevice, s));
    orthonormalize(s, it, Y, Wm, V);
    for (int r = 0; r < iters; r++) {
      sgemm(s, (int)n, IQ_EMB, V, X, V, 1, Wm, V, 1, Z, IQ_EMB);      // Z = X W^T
      sgemm(s, IQ_EMB, V, (int)n, Z, 1, IQ_EMB, X, 1, V, Y, V);       // Y = Z^T X
      orthonormalize(s, it, Y, Wm, V);
    }
    hipLaunchKernelGGL(k_learn_bank, dim3(nblk(g.P, 256)), dim3(256), 0, s, (const IqDev*)pl->dd, m, (const float*)Wm,
                       (const float*)nullptr, (uint16_t*)H.lrn.bank + pl->lrn_bofs[m]);
    HCL();
    HC(hipStreamSynchronize(s));
  }
  refresh_wt(pl);
  pl->lrn_ready = true;
  pl->cv_ready = false;  // banks now hold the linear encodings of the training image
  HC(hipMemcpyAsync(pl->dd, &pl->hd, sizeof(IqDev), hipMemcpyHostToDevice, s));
  HC(hipStreamSynchronize(s));
}


// ---------------------------------------------------------------------------
// convolutional encoder: training on (query, exact distance) pairs of the full
// scan, banks from the trained network's output map
// ---------------------------------------------------------------------------
struct ConvBuf {
  DMem mem;
  CvDims d{};
  const float* x = nullptr;  // training image (FP32)
  float *h1 = nullptr, *h2 = nullptr, *g1 = nullptr, *g2 = nullptr, *f = nullptr, *gf = nullptr;
  float *wr = nullptr, *part = nullptr, *grad = nullptr;
  unsigned long long* gfix = nullptr;
  unsigned* amax = nullptr;
  int nparts = 0;
  void init(iq_plan* pl) {
    const IqGeo& g = pl->geo;
    d.X = g.ti[0];
    d.Y = g.ti[1];
    d.Z = g.ti[2];
    d.N = (int64_t)d.X * d.Y * d.Z;
    x = pl->hd.ti32;
    if (!x) bad(IQ_ERR_UNSUPPORTED, "convolutional encoder: training image values must be exact in FP32");
    h1 = mem.alloc<float>((size_t)CV_C * d.N);
    h2 = mem.alloc<float>((size_t)CV_C * d.N);
    g1 = mem.alloc<float>((size_t)CV_C * d.N);
    g2 = mem.alloc<float>((size_t)CV_C * d.N);
    f = mem.alloc<float>(d.N);
    gf = mem.alloc<float>(d.N);
    gfix = mem.alloc<unsigned long long>(d.N);
    amax = mem.alloc<unsigned>(1);
    wr = mem.alloc<float>(CV_C * CV_C * 27);
    const int64_t ntiles = (int64_t)((d.X + CVW_T - 1) / CVW_T) * ((d.Y + CVW_T - 1) / CVW_T) * ((d.Z + 1) / 2);
    nparts = (int)std::min<int64_t>(1024, ntiles);
    part = mem.alloc<float>((size_t)nparts * (CV_C * CV_C * 27 + CV_C));
    grad = mem.alloc<float>(CV_NW);
  }
};

unsigned cv_mid_blocks(const CvDims& d) {
  return (unsigned)(((d.X + CVM_TX - 1) / CVM_TX) * ((d.Y + CVM_TY - 1) / CVM_TY) * ((d.Z + 1) / 2));
}

// f = x + conv3(relu(conv2(relu(conv1(x))))) for parameters th (device)
void conv_forward(hipStream_t s, ConvBuf& B, const float* th) {
  const CvDims& d = B.d;
  hipLaunchKernelGGL((k_cv_in<false>), dim3(nblk(d.N, 256)), dim3(256), 0, s, d, B.x, th + CV_W1, th + CV_B1,
                     (const float*)nullptr, B.h1);
  hipLaunchKernelGGL((k_cv_wprep<false>), dim3(nblk(CV_C * CV_C * 27, 256)), dim3(256), 0, s, th + CV_W2, B.wr);
  hipLaunchKernelGGL((k_cv_mid<false>), dim3(cv_mid_blocks(d)), dim3(256), 0, s, d, (const float*)B.h1,
                     (const float*)B.wr, th + CV_B2, (const float*)nullptr, B.h2);
  hipLaunchKernelGGL(k_cv_out, dim3(nblk(d.N, 256)), dim3(256), 0, s, d, (const float*)B.h2, th + CV_W3, th + CV_B3,
                     B.x, B.f);
  HCL();
}

template <int NGA, int NGB, int AG, int VG>
void conv_wgrad(hipStream_t s, ConvBuf& B, const float* G, const float* A, float* out) {
  hipLaunchKernelGGL((k_cv_wgrad<NGA, NGB, AG, VG>), dim3(B.nparts), dim3(256), 0, s, B.d, G, A, B.part);
  const int n = NGA * NGB * 27 + NGA;
  hipLaunchKernelGGL(k_cv_wreduce, dim3(nblk(n, 256)), dim3(256), 0, s, (const float*)B.part, B.nparts, n, out);
  HCL();
}

// gradient of the convolution parameters from dL/df (B.gf) into B.grad;
// each layer's weights and biases are adjacent in the parameter vector, as
// in the partial rows of k_cv_wgrad
void conv_backward(hipStream_t s, ConvBuf& B, const float* th) {
  const CvDims& d = B.d;
  conv_wgrad<1, CV_C, 1, 8>(s, B, B.gf, B.h2, B.grad + CV_W3);
  hipLaunchKernelGGL((k_cv_in<true>), dim3(nblk(d.N, 256)), dim3(256), 0, s, d, (const float*)B.gf, th + CV_W3,
                     (const float*)nullptr, (const float*)B.h2, B.g2);
  conv_wgrad<CV_C, CV_C, 4, 1>(s, B, B.g2, B.h1, B.grad + CV_W2);
  hipLaunchKernelGGL((k_cv_wprep<true>), dim3(nblk(CV_C * CV_C * 27, 256)), dim3(256), 0, s, th + CV_W2, B.wr);
  hipLaunchKernelGGL((k_cv_mid<true>), dim3(cv_mid_blocks(d)), dim3(256), 0, s, d, (const float*)B.g2,
                     (const float*)B.wr, (const float*)nullptr, (const float*)B.h1, B.g1);
  conv_wgrad<CV_C, 1, 1, 8>(s, B, B.g1, B.x, B.grad + CV_W1);
  HCL();
}

// Loss of nq queries of mask m over ns candidates each (host arrays), with
// the network th and head W (device); with gW != nullptr also dL/dW (device,
// 64 x V) and the convolution gradient (B.grad).
double conv_loss_grad(iq_plan* pl, ConvBuf& B, const float* th, const float* W, int m, int nq, int ns, const float* hQ,
                      const int32_t* hc, const double* hD, const double* htau, float* gW, bool forward = true) {
  hipStream_t s = pl->ctx->stream;
  const int V = pl->masks[m].V;
  if (ns < 1 || ns > CVL_MAXS || nq < 1) bad(IQ_ERR_ARG, "convolutional encoder: 1..4096 candidates per query");
  const int64_t nw = (int64_t)nq * ns, rows = nw + nq;
  if (rows > 65535LL * 64) bad(IQ_ERR_UNSUPPORTED, "convolutional encoder: too many training pairs in one batch");
  DMem t;
  float* dQ = t.upload(hQ, (size_t)nq * V, s);
  int32_t* dc = t.upload(hc, (size_t)nw, s);
  double* dD = t.upload(hD, (size_t)nw, s);
  double* dtau = t.upload(htau, (size_t)nq, s);
  double* A2c = t.alloc<double>(nw);
  hipLaunchKernelGGL(k_cv_a2, dim3(nblk(nw, 256)), dim3(256), 0, s, (const IqDev*)pl->dd, m, (const int32_t*)dc, nw, A2c);
  if (forward) conv_forward(s, B, th);  // else B.f already holds the network's output for th
  float* F = t.alloc<float>((size_t)rows * V);
  hipLaunchKernelGGL(k_cv_gather, dim3((unsigned)rows), dim3(256), 0, s, (const IqDev*)pl->dd, m, (const float*)B.f,
                     (const int32_t*)dc, (int)nw, (const float*)dQ, F);
  HCL();
  float* E = t.alloc<float>((size_t)rows * IQ_EMB);
  sgemm(s, (int)rows, IQ_EMB, V, F, V, 1, W, V, 1, E, IQ_EMB);  // E = F W^T
  float* G = t.alloc<float>((size_t)rows * IQ_EMB);
  double* lq = t.alloc<double>(nq);
  double* ls = t.alloc<double>(1);
  hipLaunchKernelGGL(k_cv_rankloss, dim3(nq), dim3(256), 0, s, nq, ns, (const float*)E, (const double*)A2c,
                     (const double*)dD, (const double*)dtau, 1.0 / nq, G, lq);
  hipLaunchKernelGGL(k_cv_losssum, dim3(1), dim3(64), 0, s, (const double*)lq, nq, ls);
  HCL();
  if (gW) {
    sgemm(s, IQ_EMB, V, (int)rows, G, 1, IQ_EMB, F, 1, V, gW, V);  // dL/dW = G^T F
    float* U = t.alloc<float>((size_t)nw * V);
    sgemm(s, (int)nw, V, IQ_EMB, G, IQ_EMB, 1, W, 1, V, U, V);      // dL/dF = G W (window rows)
    hipLaunchKernelGGL((k_fill<int64_t>), dim3(1024), dim3(256), 0, s, (int64_t*)B.gfix, B.d.N, (int64_t)0);
    hipLaunchKernelGGL((k_fill<int32_t>), dim3(1), dim3(64), 0, s, (int32_t*)B.amax, (int64_t)1, 0);
    hipLaunchKernelGGL(k_cv_absmax, dim3(1024), dim3(256), 0, s, (const float*)U, nw * V, B.amax);
    hipLaunchKernelGGL(k_cv_scatter, dim3((unsigned)nw), dim3(256), 0, s, (const IqDev*)pl->dd, m, (const int32_t*)dc,
                       (const float*)U, (int)nw, (const unsigned*)B.amax, B.gfix);
    hipLaunchKernelGGL(k_cv_unfix, dim3(1024), dim3(256), 0, s, (const unsigned long long*)B.gfix, B.d.N,
                       (const unsigned*)B.amax, (int)nw, B.gf);
    HCL();
    conv_backward(s, B, th);
  }
  double loss = 0.0;
  HC(hipMemcpyAsync(&loss, ls, sizeof(double), hipMemcpyDeviceToHost, s));
  HC(hipStreamSynchronize(s));
  return loss;
}

void adam(hipStream_t s, int64_t n, float* p, const float* g, float* m, float* v, float lr, int step) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - (float)std::pow(0.9, step), c2 = 1.0f - (float)std::pow(0.999, step);
  hipLaunchKernelGGL(k_cv_adam, dim3(nblk(n, 256)), dim3(256), 0, s, n, p, g, m, v, lr, b1, b2, c1, c2);
  HCL();
}

// Training pairs: per query, its values, its `topn` nearest candidates by
// exact FP64 distance (preselected from the full scan's lower bounds) and
// `nrand` random enabled candidates, with exact distances, and a
// temperature (the spread of its eight nearest distances).
struct ConvSet {
  int ns = 0;
  std::vector<int> mask;
  std::vector<int32_t> step, real;  // the tile step and realization of each query
  std::vector<int64_t> qoff;  // each query's values start at kv[qoff]
  std::vector<double> kv, b2;  // query values (FP64, V of its mask) and B2
  std::vector<int32_t> cands;
  std::vector<double> D, tau;
};

// The learned shortlist of the queries hq (device copy dq; slots 0..nq-1,
// query values and B2 already in their slots, queries of one mask adjacent),
// launched as iq_plan_run launches it: encodings, bracket, bank scan, pick.
// dt: room for nq tiles.
void launch_shortlist(iq_plan* pl, const std::vector<IqQuery>& hq, const IqQuery* dq, int4* dt) {
  hipStream_t s = pl->ctx->stream;
  const IqDev* D = pl->dd;
  const int nq = (int)hq.size();
  std::vector<int4> tiles;
  for (int i = 0; i < nq; i++) {
    const int mk = pl->steps[hq[i].step].mask;
    if (!tiles.empty() && tiles.back().x == mk && tiles.back().z < 32 && tiles.back().y + tiles.back().z == i)
      tiles.back().z++;
    else
      tiles.push_back(make_int4(mk, i, 1, 0));
  }
  const int nt = (int)tiles.size();
  HC(hipMemcpyAsync(dt, tiles.data(), nt * sizeof(int4), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_learn_embed, dim3(nq), dim3(1024), 0, s, D, dq);
  hipLaunchKernelGGL(k_learn_bfrag, dim3(nt), dim3(64), 0, s, D, dq, (const int4*)dt);
  hipLaunchKernelGGL(k_learn_bracket, dim3(nq), dim3(1024), 0, s, D, dq);
  launch_mscan(pl, dq, (const int4*)dt, nt, s);
  hipLaunchKernelGGL(k_learn_pick, dim3(nq), dim3(1024), 0, s, D, dq);
  HCL();
}

// Device scratch of the training-pair harvest (one batch of up to nb queries).
struct HarvestWs {
  DMem t;
  int nb = 0, cap = 0, nrand = 0, nhard = 0;
  IqQuery* dq = nullptr;
  int4* dt = nullptr;
  int32_t *hist = nullptr, *list = nullptr, *cnt = nullptr, *dslot = nullptr, *dcand = nullptr;
  uint32_t* sel = nullptr;
  double *dex = nullptr, *dkv = nullptr;
  void init(iq_plan* pl, int topn, int nrand_, int nhard_) {
    nb = (int)std::min<int64_t>(pl->nslots, 256);
    cap = 4 * topn;
    nrand = nrand_;
    nhard = nhard_;
    dq = t.alloc<IqQuery>(nb);
    dt = t.alloc<int4>(nb);
    hist = t.alloc<int32_t>((size_t)nb * CVT_BINS);
    sel = t.alloc<uint32_t>((size_t)nb * 4);
    list = t.alloc<int32_t>((size_t)nb * cap);
    cnt = t.alloc<int32_t>(nb);
    dslot = t.alloc<int32_t>(nb);
    dcand = t.alloc<int32_t>((size_t)nb * (cap + nhard + nrand));
    dex = t.alloc<double>((size_t)nb * (cap + nhard + nrand));
  }
};

// Training pairs of the plain tile steps `pool` (step, realization), read from
// the grids' current state (so a query is what its tile step sees when it
// runs): the query values (FP64) and B2, the `topn` nearest candidates by
// exact FP64 distance among those the full scan's lower bounds put first,
// `nhard` members of the current encoder's shortlist that are not among the
// listed nearest (the competitors a shortlist must rank below the band) and
// `nrand` random enabled candidates, all with their exact distances.
template <typename T>
void harvest_queries(iq_plan* pl, std::vector<std::pair<int32_t, int32_t>> pool, int topn, int nrand,
                     std::mt19937_64& rng, HarvestWs& ws, ConvSet& set) {
  const int nhard = ws.nhard;
  hipStream_t s = pl->ctx->stream;
  const IqGeo& g = pl->geo;
  const IqDev* D = pl->dd;
  std::stable_sort(pool.begin(), pool.end(), [&](const std::pair<int32_t, int32_t>& a, const std::pair<int32_t, int32_t>& b) {
    return pl->steps[a.first].mask < pl->steps[b.first].mask;
  });
  const int cap = ws.cap, nb = ws.nb;
  set.ns = topn + nhard + nrand;
  std::uniform_int_distribution<size_t> pick(0, pl->enabled_host.size() - 1);
  IqQuery* dq = ws.dq;
  int32_t *hist = ws.hist, *list = ws.list, *cnt = ws.cnt, *dslot = ws.dslot, *dcand = ws.dcand;
  uint32_t* sel = ws.sel;
  double* dex = ws.dex;
  for (size_t b0 = 0, b1 = 0; b0 < pool.size(); b0 = b1) {
    // queries of one mask per batch (exact distances and values per mask)
    const int m = pl->steps[pool[b0].first].mask;
    b1 = b0;
    while (b1 < pool.size() && b1 < b0 + (size_t)nb && pl->steps[pool[b1].first].mask == m) b1++;
    const int nq = (int)(b1 - b0);
    const int V = pl->masks[m].V;
    std::vector<IqQuery> hq((size_t)nq);
    std::vector<int32_t> hs((size_t)nq);
    for (int i = 0; i < nq; i++) {
      IqQuery q{};
      q.step = pool[b0 + i].first;
      q.real = pool[b0 + i].second;
      q.slot = i;
      q.multi = 0;
      q.u = 0.5;
      q.forced = -1;
      hq[i] = q;
      hs[i] = i;
    }
    HC(hipMemcpyAsync(dq, hq.data(), nq * sizeof(IqQuery), hipMemcpyHostToDevice, s));
    HC(hipMemcpyAsync(dslot, hs.data(), nq * sizeof(int32_t), hipMemcpyHostToDevice, s));
    // the full scan's lower bounds of every candidate, per slot
    hipLaunchKernelGGL((k_query_prep<T>), dim3(nq, 1), dim3(256), 0, s, D, (const IqQuery*)dq);
    HCL();
    if (pl->mfma) {
      const int ntile = (nq + 31) / 32;
      hipLaunchKernelGGL(k_mfma_bprep, dim3(pl->hd.ktab.nstep, ntile), dim3(64), 0, s, D, (const IqQuery*)dq, nq);
      const int nw = pl->mfma_nw_forced ? pl->mfma_nw : pl->nw_tiles[ntile > 1 ? 2 : 1];
      dim3 grid(nblk(g.dist[0], 32), nblk(g.dist[1], 4 * nw), (unsigned)(g.dist[2] * ntile));
      if (pl->mfma_split)
        hipLaunchKernelGGL((k_scan_mfma<true>), grid, dim3(64 * nw), pl->lds_nw[nw], s, D, (const IqQuery*)dq, nq,
                           pl->mfma_gerr, pl->kofs_nw[nw]);
      else
        hipLaunchKernelGGL((k_scan_mfma<false>), grid, dim3(64 * nw), pl->lds_nw[nw], s, D, (const IqQuery*)dq, nq,
                           pl->mfma_gerr, pl->kofs_nw[nw]);
    } else {
      launch_scan_batch(pl, dq, nq, s);
    }
    HCL();
    // the topn smallest lower bounds per slot: two histogram passes, then the list
    const unsigned hb = std::min<unsigned>(nblk(g.P, 256 * 16), 512);
    hipLaunchKernelGGL((k_fill<int32_t>), dim3(256), dim3(256), 0, s, hist, (int64_t)nb * CVT_BINS, 0);
    hipLaunchKernelGGL((k_cv_tophist<0>), dim3(hb, nq), dim3(256), 0, s, D, (const IqQuery*)dq, hist, (const uint32_t*)sel);
    hipLaunchKernelGGL((k_cv_topsel<0>), dim3(nq), dim3(64), 0, s, (const IqQuery*)dq, (const int32_t*)hist, topn, sel);
    hipLaunchKernelGGL((k_fill<int32_t>), dim3(256), dim3(256), 0, s, hist, (int64_t)nb * CVT_BINS, 0);
    hipLaunchKernelGGL((k_cv_tophist<1>), dim3(hb, nq), dim3(256), 0, s, D, (const IqQuery*)dq, hist, (const uint32_t*)sel);
    hipLaunchKernelGGL((k_cv_topsel<1>), dim3(nq), dim3(64), 0, s, (const IqQuery*)dq, (const int32_t*)hist, topn, sel);
    hipLaunchKernelGGL((k_fill<int32_t>), dim3(1), dim3(256), 0, s, cnt, (int64_t)nb, 0);
    hipLaunchKernelGGL(k_cv_topcollect, dim3(hb, nq), dim3(256), 0, s, D, (const IqQuery*)dq, (const uint32_t*)sel, list,
                       cnt, cap);
    HCL();
    std::vector<int32_t> hcnt((size_t)nq), hlist((size_t)nq * cap);
    HC(hipMemcpyAsync(hcnt.data(), cnt, nq * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HC(hipMemcpyAsync(hlist.data(), list, (size_t)nq * cap * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    // the current encoder's shortlists of the same queries
    std::vector<int32_t> hns((size_t)nq, 0);
    std::vector<std::vector<int32_t>> hsl((size_t)nq);
    if (nhard > 0) {
      launch_shortlist(pl, hq, dq, ws.dt);
      HC(hipMemcpyAsync(hns.data(), pl->hd.lrn.nslist, nq * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      HC(hipStreamSynchronize(s));
      for (int i = 0; i < nq; i++) {
        hsl[i].resize((size_t)std::max(hns[i], 0));
        if (hns[i] > 0)
          HC(hipMemcpyAsync(hsl[i].data(), pl->hd.lrn.slist + (int64_t)i * g.P, hns[i] * sizeof(int32_t),
                            hipMemcpyDeviceToHost, s));
      }
    }
    HC(hipStreamSynchronize(s));
    // listed, hard and random candidates -> exact distances
    const int w = cap + nhard + nrand;
    std::vector<int32_t> hc((size_t)nq * w);
    std::vector<int> nl((size_t)nq);
    for (int i = 0; i < nq; i++) {
      nl[i] = std::min(hcnt[i], cap);
      std::vector<int32_t> li(hlist.begin() + (size_t)i * cap, hlist.begin() + (size_t)i * cap + nl[i]);
      std::sort(li.begin(), li.end());
      for (int j = 0; j < w; j++) hc[(size_t)i * w + j] = j < nl[i] ? li[j] : pl->enabled_host[0];  // padding
      // hard candidates: shortlist members outside the list (fewer than
      // nhard: the rest random)
      std::vector<int32_t> hard;
      for (int32_t c : hsl[i])
        if (!std::binary_search(li.begin(), li.end(), c)) hard.push_back(c);
      std::shuffle(hard.begin(), hard.end(), rng);
      // random candidates outside the list (repeats allowed when the list
      // holds most of the enabled candidates)
      const bool distinct = pl->enabled_host.size() >= 2 * (size_t)(nl[i] + nhard + nrand);
      for (int j = nl[i]; j < nl[i] + nhard + nrand;) {
        const int k = j - nl[i];
        int32_t c;
        if (k < nhard && k < (int)hard.size()) {
          c = hard[k];
        } else {
          c = pl->enabled_host[pick(rng)];
          if (distinct && std::binary_search(li.begin(), li.end(), c)) continue;
        }
        hc[(size_t)i * w + j++] = c;
      }
    }
    HC(hipMemcpyAsync(dcand, hc.data(), hc.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_cv_exact, dim3(w, nq), dim3(64), 0, s, D, m, (const int32_t*)dslot, (const int32_t*)dcand, w, dex);
    HCL();
    std::vector<double> hd((size_t)nq * w);
    std::vector<double> hkv((size_t)nq * V), hb2((size_t)nq);
    HC(hipMemcpyAsync(hd.data(), dex, hd.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    for (int i = 0; i < nq; i++) {
      HC(hipMemcpyAsync(hkv.data() + (size_t)i * V, pl->hd.kv + (int64_t)i * pl->hd.kv_stride, V * sizeof(double),
                        hipMemcpyDeviceToHost, s));
      HC(hipMemcpyAsync(hb2.data() + i, pl->hd.b2 + (int64_t)i * (2 + pl->naux), sizeof(double), hipMemcpyDeviceToHost, s));
    }
    HC(hipStreamSynchronize(s));
    for (int i = 0; i < nq; i++) {
      if (nl[i] < topn) continue;  // fewer enabled candidates than topn
      std::vector<int> ord((size_t)nl[i]);
      for (int j = 0; j < nl[i]; j++) ord[j] = j;
      const double* di = hd.data() + (size_t)i * w;
      const int32_t* ci = hc.data() + (size_t)i * w;
      std::sort(ord.begin(), ord.end(), [&](int a, int b) { return di[a] < di[b] || (di[a] == di[b] && ci[a] < ci[b]); });
      set.mask.push_back(m);
      set.step.push_back(hq[i].step);
      set.real.push_back(hq[i].real);
      set.qoff.push_back((int64_t)set.kv.size());
      set.kv.insert(set.kv.end(), hkv.begin() + (size_t)i * V, hkv.begin() + (size_t)(i + 1) * V);
      set.b2.push_back(hb2[i]);
      for (int j = 0; j < topn; j++) {
        set.cands.push_back(ci[ord[j]]);
        set.D.push_back(di[ord[j]]);
      }
      for (int j = 0; j < nhard + nrand; j++) {
        set.cands.push_back(ci[nl[i] + j]);
        set.D.push_back(di[nl[i] + j]);
      }
      // temperature: the band's width tol * min D (iqsim.jl:237), at least the
      // spread of the eight nearest, so the target spreads over the band
      const double d0 = di[ord[0]], d8 = di[ord[std::min(7, topn - 1)]];
      set.tau.push_back(std::max({pl->tol * d0, d8 - d0, 1e-3 * d8, 1e-30}));
    }
  }
}


// Training pairs from the plain method at work: a full-scan run of every
// realization (uniforms from rng; the grids then hold its result) in which
// about nq_total plain tile steps, spread over up to 48 batches, are
// harvested right before their batch runs.
template <typename T>
void conv_harvest(iq_plan* pl, int64_t nq_total, int topn, int nhard, int nrand, std::mt19937_64& rng, ConvSet& set) {
  hipStream_t s = pl->ctx->stream;
  std::vector<double> U((size_t)pl->nlog);
  std::uniform_real_distribution<double> uni(0.0, 1.0);
  for (double& u : U) u = uni(rng);
  std::vector<IqQuery> qs;
  std::vector<int4> tiles;
  std::vector<RunBatch> batches;
  plan_tables(pl, U.data(), nullptr, false, qs, tiles, batches);
  auto eligible = [&](const IqQuery& q) {
    const IqStep& st = pl->steps[q.step];
    return !q.multi && st.active && st.mask >= 0;
  };
  std::vector<int> hb;  // batches with plain queries that have an overlap
  for (int b = 0; b < (int)batches.size(); b++)
    for (int i = 0; i < batches[b].n; i++)
      if (eligible(qs[batches[b].off + i])) { hb.push_back(b); break; }
  if (hb.empty()) bad(IQ_ERR_ARG, "convolutional encoder: the plan has no plain tile step with an overlap");
  const int npts = (int)std::min<size_t>(48, hb.size());
  std::vector<int> pts;  // evenly spread over the run
  for (int k = 0; k < npts; k++) pts.push_back(hb[(size_t)k * hb.size() / npts]);
  const int64_t per = (nq_total + npts - 1) / npts;
  if ((int64_t)qs.size() > pl->dq_cap) {
    pl->dq = pl->mem.alloc<IqQuery>(qs.size());
    pl->dq_cap = (int64_t)qs.size();
  }
  HC(hipMemcpyAsync(pl->dq, qs.data(), qs.size() * sizeof(IqQuery), hipMemcpyHostToDevice, s));
  HarvestWs ws;
  ws.init(pl, topn, nrand, nhard);
  plan_resets<T>(pl);
  size_t next = 0;
  int64_t left = nq_total;
  for (int b = 0; b < (int)batches.size(); b++) {
    const RunBatch& bt = batches[b];
    if (next < pts.size() && pts[next] == b) {
      next++;
      std::vector<std::pair<int32_t, int32_t>> pool;
      for (int i = 0; i < bt.n; i++)
        if (eligible(qs[bt.off + i])) pool.push_back({qs[bt.off + i].step, qs[bt.off + i].real});
      std::shuffle(pool.begin(), pool.end(), rng);
      const int64_t take = std::min<int64_t>({(int64_t)pool.size(), per, left, (int64_t)ws.nb});
      pool.resize((size_t)take);
      left -= take;
      if (!pool.empty()) harvest_queries<T>(pl, pool, topn, nrand, rng, ws, set);
    }
    run_batch<T>(pl, qs.data() + bt.off, pl->dq + bt.off, bt.n, bt.plain, bt.multi, SEL_SCAN, nullptr, nullptr, 0);
  }
  HC(hipStreamSynchronize(s));
  int32_t code = 0;
  HC(hipMemcpy(&code, pl->hd.err, sizeof(int32_t), hipMemcpyDeviceToHost));
  check_kernel_error(code);
  if (set.mask.empty()) bad(IQ_ERR_ARG, "convolutional encoder: no training query has enough enabled candidates");
}

// Banks of every mask from the network's output map B.f and the current heads.
void conv_banks(iq_plan* pl, ConvBuf& B) {
  hipStream_t s = pl->ctx->stream;
  IqDev& H = pl->hd;
  for (int m = 0; m < (int)pl->masks.size(); m++) {
    hipLaunchKernelGGL(k_learn_bank, dim3(nblk(pl->geo.P, 256)), dim3(256), 0, s, (const IqDev*)pl->dd, m,
                       (const float*)H.lrn.W + pl->lrn_wofs[m], (const float*)B.f, (uint16_t*)H.lrn.bank + pl->lrn_bofs[m]);
    HCL();
  }
}

// Recall of the shortlist of K (the deployed bank scan on the current banks)
// on the exact bands of held-out queries: each query's harvested values (what
// its tile step saw) are shortlisted exactly as iq_plan_run does; its band is
// the members of its harvested nearest list within (1 + tol) of the nearest
// (iqsim.jl:237).  Returns (band members in the shortlist) / (band members).
template <typename T>
double conv_val_recall(iq_plan* pl, const ConvSet& set, const std::vector<int>& val, int topn) {
  hipStream_t s = pl->ctx->stream;
  const IqDev* D = pl->dd;
  int64_t hits = 0, total = 0;
  const int nb = (int)std::min<int64_t>(pl->nslots, 256);
  DMem t;
  IqQuery* dq = t.alloc<IqQuery>(nb);
  int4* dt = t.alloc<int4>(nb);
  int32_t* dband = t.alloc<int32_t>((size_t)nb * topn);
  int32_t* dnb = t.alloc<int32_t>(nb);
  unsigned long long* dhits = t.alloc<unsigned long long>(1);
  for (size_t b0 = 0; b0 < val.size(); b0 += (size_t)nb) {
    const int nq = (int)std::min<size_t>(nb, val.size() - b0);
    std::vector<IqQuery> hq((size_t)nq);
    std::vector<int32_t> hband((size_t)nq * topn, 0), hnb((size_t)nq, 0);
    for (int i = 0; i < nq; i++) {
      const int qi = val[b0 + i];
      IqQuery q{};
      q.step = set.step[qi];
      q.real = set.real[qi];
      q.slot = i;
      q.u = 0.5;
      q.forced = -1;
      hq[i] = q;
      const double* Di = set.D.data() + (size_t)qi * set.ns;
      const int32_t* ci = set.cands.data() + (size_t)qi * set.ns;
      const double thr = (1.0 + pl->tol) * Di[0];  // the topn list is ascending in D
      int k = 0;
      while (k < topn && Di[k] <= thr) {
        hband[(size_t)i * topn + k] = ci[k];
        k++;
      }
      hnb[i] = k;
      total += k;
    }
    HC(hipMemcpyAsync(dq, hq.data(), nq * sizeof(IqQuery), hipMemcpyHostToDevice, s));
    HC(hipMemcpyAsync(dband, hband.data(), hband.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    HC(hipMemcpyAsync(dnb, hnb.data(), nq * sizeof(int32_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((k_fill<int64_t>), dim3(1), dim3(64), 0, s, (int64_t*)dhits, (int64_t)1, (int64_t)0);
    for (int i = 0; i < nq; i++) {  // the query values its tile step saw, and B2
      const int qi = val[b0 + i];
      const int V = pl->masks[set.mask[qi]].V;
      HC(hipMemcpyAsync(pl->hd.kv + (int64_t)i * pl->hd.kv_stride, set.kv.data() + set.qoff[qi], V * sizeof(double),
                        hipMemcpyHostToDevice, s));
      HC(hipMemcpyAsync(pl->hd.b2 + (int64_t)i * (2 + pl->naux), set.b2.data() + qi, sizeof(double),
                        hipMemcpyHostToDevice, s));
    }
    launch_shortlist(pl, hq, dq, dt);
    hipLaunchKernelGGL(k_cv_bandhits, dim3(nq), dim3(256), 0, s, D, (const IqQuery*)dq, (const int32_t*)dband,
                       (const int32_t*)dnb, topn, dhits);
    HCL();
    unsigned long long h = 0;
    HC(hipMemcpyAsync(&h, dhits, sizeof(h), hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
    hits += (int64_t)h;
  }
  return total ? (double)hits / (double)total : 1.0;
}

// Harvest, train (Adam), then rebuild every mask's bank from the network's
// output map with the trained heads.
template <typename T>
void conv_learn(iq_plan* pl, const iq_conv_opts& o) {
  hipStream_t s = pl->ctx->stream;
  IqDev& H = pl->hd;
  const IqGeo& g = pl->geo;
  if (!pl->lrn_ready) bad(IQ_ERR_ARG, "convolutional encoder: learn the linear encoders first (iq_plan_learn)");
  if (o.nqueries < 1 || o.topn < 8 || o.nrandom < 0 || o.nhard < 0 || o.topn + o.nhard + o.nrandom > CVL_MAXS ||
      o.steps < 0 || o.batch < 1 ||
      !(o.lr >= 0.0f) || !(o.lr_head >= 0.0f))
    bad(IQ_ERR_ARG, "convolutional encoder: invalid options");
  const auto t0 = std::chrono::steady_clock::now();
  std::mt19937_64 rng(o.seed);
  ConvBuf B;
  B.init(pl);
  // shortlist size of the held-out recall check, and room for its lists
  if (o.val_k > 0) H.lrn.K = (int32_t)std::min<int64_t>(o.val_k, g.P);
  else if (H.lrn.K <= 0) H.lrn.K = (int32_t)std::min<int64_t>(2048, g.P);
  const int64_t lcap = std::min<int64_t>(g.P, std::max<int64_t>(16 * (int64_t)H.lrn.K, 65536));
  if (lcap > H.lrn.lcap) {
    H.lrn.list = pl->mem.alloc<int2>((size_t)pl->nslots * lcap);
    H.lrn.lcap = lcap;
  }
  HC(hipMemcpyAsync(pl->dd, &pl->hd, sizeof(IqDev), hipMemcpyHostToDevice, s));
  // network: He-normal hidden layers, zero output layer (f = x at the start)
  std::vector<float> th(CV_NW, 0.0f);
  std::normal_distribution<float> nrm(0.0f, 1.0f);
  for (int i = 0; i < CV_C * 27; i++) th[CV_W1 + i] = nrm(rng) * std::sqrt(2.0f / 27.0f);
  for (int i = 0; i < CV_C * CV_C * 27; i++) th[CV_W2 + i] = nrm(rng) * std::sqrt(2.0f / (27.0f * CV_C));
  if (!pl->cv_theta) pl->cv_theta = pl->mem.alloc<float>(CV_NW);
  HC(hipMemcpyAsync(pl->cv_theta, th.data(), CV_NW * sizeof(float), hipMemcpyHostToDevice, s));
  ConvSet set;
  conv_harvest<T>(pl, o.nqueries, o.topn, o.nhard, o.nrandom, rng, set);
  const auto t1 = std::chrono::steady_clock::now();
  // every fifth query is held out; batches of one mask each, visited in a
  // shuffled order every epoch
  std::map<int, std::vector<int>> bymask, valmask;
  for (int i = 0; i < (int)set.mask.size(); i++) (i % 5 == 4 ? valmask : bymask)[set.mask[i]].push_back(i);
  std::vector<std::vector<int>> batches, vbatches;
  for (auto& kv : bymask)
    for (size_t a = 0; a < kv.second.size(); a += (size_t)o.batch)
      batches.emplace_back(kv.second.begin() + a, kv.second.begin() + std::min(kv.second.size(), a + (size_t)o.batch));
  for (auto& kv : valmask)
    for (size_t a = 0; a < kv.second.size(); a += 64)
      vbatches.emplace_back(kv.second.begin() + a, kv.second.begin() + std::min(kv.second.size(), a + (size_t)64));
  if (batches.empty()) bad(IQ_ERR_ARG, "convolutional encoder: too few training queries");
  size_t wtot = 0;
  for (const IqMask& m : pl->masks) wtot += (size_t)IQ_EMB * m.V;
  DMem adm;
  float* cm = adm.alloc<float>(CV_NW);
  float* cv = adm.alloc<float>(CV_NW);
  float* hm = adm.alloc<float>(wtot);
  float* hv = adm.alloc<float>(wtot);
  float* gW = adm.alloc<float>((size_t)IQ_EMB * pl->lrn_vmax);
  hipLaunchKernelGGL((k_fill<float>), dim3(64), dim3(256), 0, s, cm, (int64_t)CV_NW, 0.0f);
  hipLaunchKernelGGL((k_fill<float>), dim3(64), dim3(256), 0, s, cv, (int64_t)CV_NW, 0.0f);
  hipLaunchKernelGGL((k_fill<float>), dim3(1024), dim3(256), 0, s, hm, (int64_t)wtot, 0.0f);
  hipLaunchKernelGGL((k_fill<float>), dim3(1024), dim3(256), 0, s, hv, (int64_t)wtot, 0.0f);
  HCL();
  std::vector<int> hstep(pl->masks.size(), 0);
  std::vector<double> losses;
  std::vector<size_t> order(batches.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = i;
  const int ns = set.ns;
  // loss of the queries `bt` (one mask) for the current parameters
  auto batch_loss = [&](const std::vector<int>& bt, float* grad_head, bool forward) {
    const int m = set.mask[bt[0]];
    const int V = pl->masks[m].V;
    const int nq = (int)bt.size();
    std::vector<float> hQ((size_t)nq * V);
    std::vector<int32_t> hc((size_t)nq * ns);
    std::vector<double> hD((size_t)nq * ns), htau((size_t)nq);
    for (int i = 0; i < nq; i++) {
      const int qi = bt[i];
      for (int v = 0; v < V; v++) hQ[(size_t)i * V + v] = (float)set.kv[set.qoff[qi] + v];
      std::copy(set.cands.begin() + (size_t)qi * ns, set.cands.begin() + (size_t)(qi + 1) * ns, hc.begin() + (size_t)i * ns);
      std::copy(set.D.begin() + (size_t)qi * ns, set.D.begin() + (size_t)(qi + 1) * ns, hD.begin() + (size_t)i * ns);
      htau[i] = set.tau[qi];
    }
    const float* Wm = (const float*)H.lrn.W + pl->lrn_wofs[m];
    return conv_loss_grad(pl, B, pl->cv_theta, Wm, m, nq, ns, hQ.data(), hc.data(), hD.data(), htau.data(), grad_head,
                          forward);
  };
  // held-out queries (all masks, in mask order) for the recall check
  std::vector<int> vq;
  for (auto& kv : valmask) vq.insert(vq.end(), kv.second.begin(), kv.second.end());
  // held-out loss (mean over held-out queries)
  auto val_loss = [&]() {
    double tot = 0.0;
    int64_t n = 0;
    for (size_t k = 0; k < vbatches.size(); k++) {
      tot += batch_loss(vbatches[k], nullptr, k == 0) * (double)vbatches[k].size();
      n += (int64_t)vbatches[k].size();
    }
    return n ? tot / (double)n : 0.0;
  };
  float* best_th = adm.alloc<float>(CV_NW);
  float* best_W = adm.alloc<float>(wtot);
  auto keep_best = [&]() {
    HC(hipMemcpyAsync(best_th, pl->cv_theta, CV_NW * sizeof(float), hipMemcpyDeviceToDevice, s));
    HC(hipMemcpyAsync(best_W, H.lrn.W, wtot * sizeof(float), hipMemcpyDeviceToDevice, s));
  };
  // the kept parameters: highest held-out recall of the deployed shortlist
  // (banks rebuilt from the network's output), lower held-out loss on ties;
  // the starting point (network = identity) is a candidate
  const int eval_every = std::max(1, std::min(25, o.steps / 4 + 1));
  auto evaluate = [&](double& rec, double& vl) {
    vl = vbatches.empty() ? 0.0 : val_loss();  // leaves B.f = the network's output
    rec = 0.0;
    if (!vq.empty()) {
      refresh_wt(pl);
      conv_banks(pl, B);
      rec = conv_val_recall<T>(pl, set, vq, o.topn);
    }
  };
  double best_rec = 0.0, best = 0.0;
  if (vbatches.empty()) conv_forward(s, B, pl->cv_theta);
  evaluate(best_rec, best);
  const double val0 = best, rec0 = best_rec;
  int best_step = 0;
  keep_best();
  for (int step = 0; step < o.steps; step++) {
    if (step % (int)batches.size() == 0) std::shuffle(order.begin(), order.end(), rng);
    const std::vector<int>& bt = batches[order[step % batches.size()]];
    const int m = set.mask[bt[0]];
    const int V = pl->masks[m].V;
    float* Wm = (float*)H.lrn.W + pl->lrn_wofs[m];
    losses.push_back(batch_loss(bt, gW, true));
    adam(s, CV_NW, pl->cv_theta, B.grad, cm, cv, o.lr, step + 1);
    if (o.lr_head > 0.0f)
      adam(s, (int64_t)IQ_EMB * V, Wm, gW, hm + pl->lrn_wofs[m], hv + pl->lrn_wofs[m], o.lr_head, ++hstep[m]);
    if (!vbatches.empty() && ((step + 1) % eval_every == 0 || step + 1 == o.steps)) {
      double r, v;
      evaluate(r, v);
      if (r > best_rec || (r == best_rec && v < best)) {
        best_rec = r;
        best = v;
        best_step = step + 1;
        keep_best();
      }
    }
  }
  HC(hipMemcpyAsync(pl->cv_theta, best_th, CV_NW * sizeof(float), hipMemcpyDeviceToDevice, s));
  HC(hipMemcpyAsync((float*)H.lrn.W, best_W, wtot * sizeof(float), hipMemcpyDeviceToDevice, s));
  // banks from the trained network's output map
  refresh_wt(pl);
  conv_forward(s, B, pl->cv_theta);
  for (int m = 0; m < (int)pl->masks.size(); m++) {
    hipLaunchKernelGGL(k_learn_bank, dim3(nblk(g.P, 256)), dim3(256), 0, s, (const IqDev*)pl->dd, m,
                       (const float*)H.lrn.W + pl->lrn_wofs[m], (const float*)B.f, (uint16_t*)H.lrn.bank + pl->lrn_bofs[m]);
    HCL();
  }
  HC(hipStreamSynchronize(s));
  const auto t2 = std::chrono::steady_clock::now();
  pl->cv_ready = true;
  auto mean = [&](size_t a, size_t b) {
    double x = 0.0;
    for (size_t i = a; i < b; i++) x += losses[i];
    return b > a ? x / (double)(b - a) : 0.0;
  };
  const size_t k = std::min<size_t>(10, losses.size());
  pl->cv_stats[0] = std::chrono::duration<double>(t1 - t0).count();
  pl->cv_stats[1] = std::chrono::duration<double>(t2 - t1).count();
  pl->cv_stats[2] = mean(0, k);
  pl->cv_stats[3] = mean(losses.size() - k, losses.size());
  pl->cv_stats[4] = (double)set.mask.size() * ns;
  pl->cv_stats[5] = (double)o.steps;
  pl->cv_stats[6] = (double)set.mask.size();
  pl->cv_stats[7] = (double)best_step;
  pl->cv_val[0] = val0;
  pl->cv_val[1] = best;
  pl->cv_val[2] = rec0;
  pl->cv_val[3] = best_rec;
}
}  // namespace

// ---------------------------------------------------------------------------
// allocation-free entry point: its scratch is this static per-device array
// (status word, then min-cut node state), held until the process ends.  The
// array's name is not a device address in host code: its address on each
// device is looked up once (iq_ctx_create, that device current) and only
// that address is used.  Every call on one device shares the array.
// ---------------------------------------------------------------------------
#define IQ_NOALLOC_HEAD 256
static __device__ __align__(256) char scratch[IQ_NOALLOC_HEAD + (size_t)IQ_NOALLOC_MAX_NODES * CUT_NODE_BYTES];

namespace {
constexpr int IQ_MAX_DEVICES = 64;
char* g_scratch[IQ_MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// with `device` current
void scratch_setup(int device) {
  if (device >= IQ_MAX_DEVICES) return;  // iq_graphcut_noalloc then reports the device as unsupported
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return;
  void* p = nullptr;
  HC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch[device] = (char*)p;
}

char* scratch_ptr(int device) {
  if (device >= IQ_MAX_DEVICES) bad(IQ_ERR_UNSUPPORTED, "allocation-free entry points support devices 0-63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) bad(IQ_ERR_HIP, "device scratch of the allocation-free entry points not set up");
  return g_scratch[device];
}

// ---------------------------------------------------------------------------
// standalone min cut: argument checks and the launch shared by iq_graphcut
// and iq_graphcut_noalloc
// ---------------------------------------------------------------------------
// interior nodes of the slab (the first and last slices along `dim` are the
// terminals, graphcut.jl:57-70)
int64_t cut_check(int dtype, const int64_t dims[3], int dim) {
  if (dtype != IQ_F32 && dtype != IQ_F64) bad(IQ_ERR_ARG, "dtype must be IQ_F32 or IQ_F64");
  if (dim < 0 || dim > 2) bad(IQ_ERR_ARG, "cut axis must be 0, 1 or 2");
  for (int d = 0; d < 3; d++)
    if (dims[d] <= 0 || dims[d] > (1 << 20)) bad(IQ_ERR_ARG, "invalid slab extents");
  if (dims[dim] < 2) bad(IQ_ERR_ARG, "cut axis needs at least two slices");
  const int64_t nvox = dims[0] * dims[1] * dims[2];
  if (nvox >= (1ll << 31)) bad(IQ_ERR_UNSUPPORTED, "slab of 2^31 or more voxels");
  const int L = (int)dims[dim];
  return nvox / L * std::max(0, L - 2);
}

// A, B, keep: device slabs of extents dims; gscr: node state for slabs of
// more than IQ_PR_NLDS interior nodes (n * CUT_NODE_BYTES), else unused;
// err: device status word, zero before the launch.
void cut_launch(hipStream_t s, const void* dA, const void* dB, int dtype, const int64_t dims[3], int dim, int64_t n,
                uint8_t* keep, char* gscr, int32_t* err) {
  const size_t lds = cut_lds_bytes(n, false);
  if (lds > 64 * 1024) {
    HC(hipFuncSetAttribute((const void*)k_graphcut_host<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    HC(hipFuncSetAttribute((const void*)k_graphcut_host<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  const int s0 = (int)dims[0], s1 = (int)dims[1], s2 = (int)dims[2];
  if (n <= CUT_WAVE_MAX && dtype == IQ_F32)
    hipLaunchKernelGGL((k_graphcut_host_wave<float>), dim3(1), dim3(64), lds, s, (const float*)dA, (const float*)dB,
                       s0, s1, s2, dim, keep, err);
  else if (n <= CUT_WAVE_MAX)
    hipLaunchKernelGGL((k_graphcut_host_wave<double>), dim3(1), dim3(64), lds, s, (const double*)dA,
                       (const double*)dB, s0, s1, s2, dim, keep, err);
  else if (dtype == IQ_F32)
    hipLaunchKernelGGL((k_graphcut_host<float>), dim3(1), dim3(CB), lds, s, (const float*)dA, (const float*)dB,
                       s0, s1, s2, dim, keep, gscr, err);
  else
    hipLaunchKernelGGL((k_graphcut_host<double>), dim3(1), dim3(CB), lds, s, (const double*)dA, (const double*)dB,
                       s0, s1, s2, dim, keep, gscr, err);
  HCL();
}
}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int iq_device_count(int* count) {
  return guard(nullptr, [&] {
    if (!count) bad(IQ_ERR_ARG, "null count");
    int n = 0;
    HC(hipGetDeviceCount(&n));
    *count = n;
  });
}

int iq_ctx_create(int device, iq_ctx** out) {
  if (!out) return fail(nullptr, IQ_ERR_ARG, "null output handle");
  iq_ctx* c = new (std::nothrow) iq_ctx();
  if (!c) return fail(nullptr, IQ_ERR_NOMEM, "host allocation failed");
  int rc = guard(nullptr, [&] {
    int n = 0;
    HC(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) bad(IQ_ERR_ARG, "device " + std::to_string(device) + " not present (" + std::to_string(n) + " visible)");
    c->device = device;
    HC(hipSetDevice(device));
    HC(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HC(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device));
    scratch_setup(device);
  });
  if (rc != IQ_OK) { delete c; return rc; }
  *out = c;
  return IQ_OK;
}

int iq_ctx_destroy(iq_ctx* ctx) {
  if (!ctx) return IQ_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return IQ_OK;
}

const char* iq_last_error(const iq_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  return g_last_error.c_str();
}

int iq_ctx_synchronize(iq_ctx* ctx) {
  if (!ctx) return fail(nullptr, IQ_ERR_ARG, "null context");
  return guard(ctx, [&] { HC(hipStreamSynchronize(ctx->stream)); });
}

int iq_host_alloc(iq_ctx* ctx, size_t bytes, void** out) {
  if (!ctx || !out) return fail(ctx, IQ_ERR_ARG, "null context or output");
  return guard(ctx, [&] {
    *out = nullptr;
    HC(hipSetDevice(ctx->device));
    if (hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
      *out = nullptr;
      bad(IQ_ERR_NOMEM, "page-locked host allocation failed");
    }
  });
}

int iq_host_free(iq_ctx* ctx, void* p) {
  if (!ctx) return fail(nullptr, IQ_ERR_ARG, "null context");
  return guard(ctx, [&] {
    if (p) HC(hipHostFree(p));
  });
}

int iq_imfilter(iq_ctx* ctx, const double* img, const int64_t idims[3], const double* krn, const int64_t kdims[3],
                double* out) {
  if (!ctx) return fail(nullptr, IQ_ERR_ARG, "null context");
  return guard(ctx, [&] {
    if (!img || !krn || !out || !idims || !kdims) bad(IQ_ERR_ARG, "null argument");
    int64_t od[3];
    for (int d = 0; d < 3; d++) {
      od[d] = idims[d] - kdims[d] + 1;
      if (kdims[d] <= 0 || od[d] <= 0) bad(IQ_ERR_ARG, "kernel larger than image");
    }
    hipStream_t s = ctx->stream;
    DMem m;
    int64_t ni = idims[0] * idims[1] * idims[2], nk = kdims[0] * kdims[1] * kdims[2], no = od[0] * od[1] * od[2];
    double* di = m.upload(img, ni, s);
    double* dk = m.upload(krn, nk, s);
    double* dout = m.alloc<double>(no);
    hipLaunchKernelGGL(k_imfilter, dim3(nblk(no, 256)), dim3(256), 0, s, (const double*)di, (int)idims[0],
                       (int)idims[1], (int)idims[2], (const double*)dk, (int)kdims[0], (int)kdims[1], (int)kdims[2],
                       dout, (int)od[0], (int)od[1], (int)od[2]);
    HCL();
    HC(hipMemcpyAsync(out, dout, no * sizeof(double), hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
  });
}

int iq_fastdistance(iq_ctx* ctx, const void* img, int img_dtype, const int64_t idims[3], const double* kern,
                    const uint8_t* weights, const int64_t kdims[3], double* out) {
  if (!ctx) return fail(nullptr, IQ_ERR_ARG, "null context");
  return guard(ctx, [&] {
    if (!img || !kern || !out || !idims || !kdims) bad(IQ_ERR_ARG, "null argument");
    if (img_dtype != IQ_F32 && img_dtype != IQ_F64) bad(IQ_ERR_ARG, "img_dtype must be IQ_F32 or IQ_F64");
    IqGeo g{};
    geo_init(g, idims, kdims);
    for (int d = 0; d < 3; d++)
      if (kdims[d] <= 0 || g.dist[d] <= 0) bad(IQ_ERR_ARG, "kernel larger than image");
    hipStream_t s = ctx->stream;
    DMem m;
    std::vector<int32_t> tv, to;
    for (int32_t t = 0; t < g.tilevox; t++)
      if (!weights || weights[t]) {
        int x = t % g.tile[0], y = (t / g.tile[0]) % g.tile[1], z = t / (g.tile[0] * g.tile[1]);
        tv.push_back(t);
        to.push_back((int32_t)(x + (int64_t)g.ti[0] * (y + (int64_t)g.ti[1] * z)));
      }
    int V = (int)tv.size();
    int64_t ni = idims[0] * idims[1] * idims[2];
    const size_t esz = img_dtype == IQ_F32 ? 4 : 8;
    char* di = m.alloc<char>(ni * esz);
    HC(hipMemcpyAsync(di, img, ni * esz, hipMemcpyHostToDevice, s));
    double* dk = m.upload(kern, (size_t)g.tilevox, s);
    int32_t* dtv = m.upload(tv.data(), tv.size(), s);
    int32_t* dto = m.upload(to.data(), to.size(), s);
    double* kv = m.alloc<double>(std::max(V, 1));
    double* b2 = m.alloc<double>(1);
    double* dout = m.alloc<double>(g.P);
    hipLaunchKernelGGL(k_gather_b2, dim3(1), dim3(256), 0, s, (const double*)dk, (const int32_t*)dtv, V, kv, b2);
    HCL();
    if (img_dtype == IQ_F32)
      hipLaunchKernelGGL((k_exact_map<float, 0>), dim3(nblk(g.P, 256)), dim3(256), 0, s, (const float*)di, g,
                         (const int32_t*)dto, (const double*)kv, V, (const double*)b2, dout);
    else
      hipLaunchKernelGGL((k_exact_map<double, 0>), dim3(nblk(g.P, 256)), dim3(256), 0, s, (const double*)di, g,
                         (const int32_t*)dto, (const double*)kv, V, (const double*)b2, dout);
    HCL();
    HC(hipMemcpyAsync(out, dout, g.P * sizeof(double), hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
  });
}

int iq_band_pick(iq_ctx* ctx, const double* D, int64_t n, double tol, int64_t* out_idx, int64_t* out_count) {
  if (!ctx) return fail(nullptr, IQ_ERR_ARG, "null context");
  return guard(ctx, [&] {
    if (!D || !out_idx || !out_count || n <= 0) bad(IQ_ERR_ARG, "invalid argument");
    if (n >= (1ll << 31)) bad(IQ_ERR_UNSUPPORTED, "more than 2^31 candidates");
    hipStream_t s = ctx->stream;
    DMem m;
    double* dD = m.upload(D, n, s);
    int32_t* dout = m.alloc<int32_t>(n);
    int64_t* dcnt = m.alloc<int64_t>(1);
    hipLaunchKernelGGL(k_band_pick_host, dim3(1), dim3(FB), 0, s, (const double*)dD, n, tol, dout, dcnt);
    HCL();
    int64_t cnt = 0;
    HC(hipMemcpyAsync(&cnt, dcnt, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
    std::vector<int32_t> h((size_t)cnt);
    if (cnt) HC(hipMemcpy(h.data(), dout, cnt * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < cnt; i++) out_idx[i] = h[i];
    *out_count = cnt;
  });
}

int iq_relaxation(iq_ctx* ctx, const double* D, const double* const* Ds, int naux, int64_t n, double tol,
                  int64_t* out_idx, int64_t* out_count) {
  if (!ctx) return fail(nullptr, IQ_ERR_ARG, "null context");
  return guard(ctx, [&] {
    if (!D || !out_idx || !out_count || n <= 0 || naux < 1 || naux > IQ_MAX_AUX + 1 || !Ds) bad(IQ_ERR_ARG, "invalid argument");
    if (n >= (1ll << 31)) bad(IQ_ERR_UNSUPPORTED, "more than 2^31 candidates");
    hipStream_t s = ctx->stream;
    DMem m;
    int nsrc = naux + 1;
    std::vector<const double*> maps;
    for (int j = 0; j < nsrc; j++) {
      const double* src = j == 0 ? D : Ds[j - 1];
      if (!src) bad(IQ_ERR_ARG, "null auxiliary map");
      maps.push_back(m.upload(src, n, s));
    }
    const double** dmaps = (const double**)m.upload(maps.data(), maps.size(), s);
    uint32_t* bits = m.alloc<uint32_t>((size_t)((n + 31) / 32) * nsrc);
    int32_t* ev = m.alloc<int32_t>(n);
    int64_t* cnt = m.alloc<int64_t>(1);
    int32_t* err = m.alloc<int32_t>(1);
    HC(hipMemsetAsync(err, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_relax_host, dim3(1), dim3(FB), 0, s, n, nsrc, dmaps, tol, bits, ev, cnt, err);
    HCL();
    int32_t code = 0;
    int64_t c = 0;
    HC(hipMemcpyAsync(&code, err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HC(hipMemcpyAsync(&c, cnt, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
    check_kernel_error(code);
    std::vector<int32_t> h((size_t)c);
    if (c) HC(hipMemcpy(h.data(), ev, c * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < c; i++) out_idx[i] = h[i];
    *out_count = c;
  });
}

int iq_taumodel(iq_ctx* ctx, const int64_t* events, int64_t nev, const double* D, const double* const* Ds, int naux,
                int64_t n, double* out_p) {
  if (!ctx) return fail(nullptr, IQ_ERR_ARG, "null context");
  return guard(ctx, [&] {
    if (!events || nev <= 0 || !D || !out_p || n <= 0 || naux < 0 || (naux > 0 && !Ds)) bad(IQ_ERR_ARG, "invalid argument");
    hipStream_t s = ctx->stream;
    DMem m;
    std::vector<int32_t> ev32((size_t)nev);
    for (int64_t i = 0; i < nev; i++) {
      if (events[i] < 0 || events[i] >= n) bad(IQ_ERR_ARG, "event index out of range");
      ev32[i] = (int32_t)events[i];
    }
    int32_t* dev = m.upload(ev32.data(), ev32.size(), s);
    std::vector<const double*> maps;
    maps.push_back(m.upload(D, n, s));
    for (int j = 0; j < naux; j++) maps.push_back(m.upload(Ds[j], n, s));
    const double** dmaps = (const double**)m.upload(maps.data(), maps.size(), s);
    double* vals = m.alloc<double>(nev);
    int32_t* rank = m.alloc<int32_t>(nev);
    double* p = m.alloc<double>(nev);
    unsigned long long* keys = m.alloc<unsigned long long>(2 * nev);
    int32_t* err = m.alloc<int32_t>(1);
    HC(hipMemsetAsync(err, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_taumodel_host, dim3(1), dim3(FB), 0, s, (const int32_t*)dev, nev, naux + 1, dmaps, vals, rank,
                       p, keys, err);
    HCL();
    int32_t code = 0;
    HC(hipMemcpyAsync(&code, err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HC(hipMemcpyAsync(out_p, p, nev * sizeof(double), hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
    check_kernel_error(code);
  });
}

int iq_sample(iq_ctx* ctx, const double* p, int64_t n, double u, int64_t* out_index) {
  if (!ctx) return fail(nullptr, IQ_ERR_ARG, "null context");
  return guard(ctx, [&] {
    if (!p || n <= 0 || !out_index || !(u >= 0.0 && u < 1.0)) bad(IQ_ERR_ARG, "invalid argument");
    if ((n + IQ_SB * IQ_SB - 1) / (IQ_SB * IQ_SB) > S2MAX) bad(IQ_ERR_UNSUPPORTED, "too many weights");
    hipStream_t s = ctx->stream;
    DMem m;
    double* dp = m.upload(p, n, s);
    double* s1 = m.alloc<double>((n + IQ_SB - 1) / IQ_SB + 1);
    int64_t* o = m.alloc<int64_t>(1);
    hipLaunchKernelGGL(k_sample_host, dim3(1), dim3(FB), 0, s, (const double*)dp, n, u, s1, o);
    HCL();
    HC(hipMemcpyAsync(out_index, o, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
  });
}

int iq_graphcut(iq_ctx* ctx, const void* A, const void* B, int dtype, const int64_t dims[3], int dim,
                uint8_t* out_mask) {
  if (!ctx) return fail(nullptr, IQ_ERR_ARG, "null context");
  return guard(ctx, [&] {
    if (!A || !B || !dims || !out_mask) bad(IQ_ERR_ARG, "null argument");
    const int64_t n = cut_check(dtype, dims, dim);
    const int64_t nvox = dims[0] * dims[1] * dims[2];
    hipStream_t s = ctx->stream;
    DMem m;
    const size_t esz = dtype == IQ_F32 ? 4 : 8;
    char* dA = m.alloc<char>(nvox * esz);
    char* dB = m.alloc<char>(nvox * esz);
    HC(hipMemcpyAsync(dA, A, nvox * esz, hipMemcpyHostToDevice, s));
    HC(hipMemcpyAsync(dB, B, nvox * esz, hipMemcpyHostToDevice, s));
    uint8_t* keep = m.alloc<uint8_t>(nvox);
    int32_t* err = m.alloc<int32_t>(1);
    HC(hipMemsetAsync(err, 0, sizeof(int32_t), s));
    char* gscr = n > IQ_PR_NLDS ? m.alloc<char>((size_t)n * CUT_NODE_BYTES) : nullptr;
    cut_launch(s, dA, dB, dtype, dims, dim, n, keep, gscr, err);
    int32_t code = 0;
    HC(hipMemcpyAsync(&code, err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HC(hipMemcpyAsync(out_mask, keep, nvox, hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
    check_kernel_error(code);
  });
}

int iq_graphcut_noalloc(iq_ctx* ctx, const void* A, const void* B, int dtype, const int64_t dims[3], int dim,
                        uint8_t* out_mask) {
  if (!ctx) return fail(nullptr, IQ_ERR_ARG, "null context");
  return guard(ctx, [&] {
    if (!A || !B || !dims || !out_mask) bad(IQ_ERR_ARG, "null argument");
    const int64_t n = cut_check(dtype, dims, dim);
    if (n > IQ_NOALLOC_MAX_NODES)
      bad(IQ_ERR_UNSUPPORTED, "slab with " + std::to_string(n) + " interior nodes: the allocation-free cut holds at most " +
                                  std::to_string(IQ_NOALLOC_MAX_NODES));
    char* base = scratch_ptr(ctx->device);
    int32_t* err = (int32_t*)base;
    char* gscr = n > IQ_PR_NLDS ? base + IQ_NOALLOC_HEAD : nullptr;
    hipStream_t s = ctx->stream;
    // the status word keeps the previous call's value: cleared by a kernel
    hipLaunchKernelGGL((k_fill<int32_t>), dim3(1), dim3(64), 0, s, err, (int64_t)1, (int32_t)0);
    HCL();
    cut_launch(s, A, B, dtype, dims, dim, n, out_mask, gscr, err);
    int32_t code = 0;
    HC(hipMemcpyAsync(&code, err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
    check_kernel_error(code);
  });
}

int iq_scan_error_bound(int64_t nstep, int split, double* gerr) {
  return guard(nullptr, [&] {
    if (nstep <= 0 || !gerr) bad(IQ_ERR_ARG, "nstep must be positive");
    *gerr = scan_gerr(nstep, split != 0);
  });
}

int iq_plan_validate(const iq_plan_desc* descs, int ngrid) { return guard(nullptr, [&] { validate_descs(descs, ngrid); }); }

int iq_plan_create(iq_ctx* ctx, const iq_plan_desc* desc, iq_plan** out) {
  if (!ctx) return fail(nullptr, IQ_ERR_ARG, "null context");
  if (!out) return fail(ctx, IQ_ERR_ARG, "null output handle");
  iq_plan* pl = new (std::nothrow) iq_plan();
  if (!pl) return fail(ctx, IQ_ERR_NOMEM, "host allocation failed");
  int rc = guard(ctx, [&] { plan_build(ctx, desc, 1, pl); });
  if (rc != IQ_OK) { iq_plan_destroy(pl); return rc; }
  *out = pl;
  return IQ_OK;
}

int iq_plan_create_multi(iq_ctx* ctx, const iq_plan_desc* descs, int ngrid, iq_plan** out) {
  if (!ctx) return fail(nullptr, IQ_ERR_ARG, "null context");
  if (!out) return fail(ctx, IQ_ERR_ARG, "null output handle");
  iq_plan* pl = new (std::nothrow) iq_plan();
  if (!pl) return fail(ctx, IQ_ERR_NOMEM, "host allocation failed");
  int rc = guard(ctx, [&] { plan_build(ctx, descs, ngrid, pl); });
  if (rc != IQ_OK) { iq_plan_destroy(pl); return rc; }
  *out = pl;
  return IQ_OK;
}

int iq_plan_destroy(iq_plan* plan) {
  if (!plan) return IQ_OK;
  if (plan->ctx) (void)hipSetDevice(plan->ctx->device);
  for (hipEvent_t e : plan->prof_ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : plan->cut_ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : plan->bank_ev) (void)hipEventDestroy(e);
  for (hipGraphExec_t g : plan->gexec)
    if (g) (void)hipGraphExecDestroy(g);
  if (plan->e0) (void)hipEventDestroy(plan->e0);
  if (plan->e1) (void)hipEventDestroy(plan->e1);
  plan->mem.release();
  delete plan;
  return IQ_OK;
}

int iq_plan_run(iq_ctx* ctx, iq_plan* plan, const double* uniforms, const int64_t* forced) {
  if (!ctx || !plan) return fail(ctx, IQ_ERR_ARG, "null context or plan");
  return guard(ctx, [&] {
    if (plan->ctx != ctx) bad(IQ_ERR_ARG, "plan belongs to another context");
    if (uniforms)
      for (int64_t i = 0; i < plan->nlog; i++)
        if (!(uniforms[i] >= 0.0 && uniforms[i] < 1.0)) bad(IQ_ERR_ARG, "uniforms must lie in [0,1)");
    if (plan->dtype == IQ_F32) plan_run<float>(plan, uniforms, forced);
    else plan_run<double>(plan, uniforms, forced);
  });
}

int iq_plan_results(iq_ctx* ctx, iq_plan* plan, void* out, int64_t* chosen, double* cutgrid) {
  if (!ctx || !plan) return fail(ctx, IQ_ERR_ARG, "null context or plan");
  return guard(ctx, [&] {
    hipStream_t s = ctx->stream;
    const size_t esz = plan->dtype == IQ_F32 ? 4 : 8;
    if (out) HC(hipMemcpyAsync(out, plan->hd.sims, (size_t)plan->nvox * esz, hipMemcpyDeviceToHost, s));
    if (chosen && plan->nlog)
      HC(hipMemcpyAsync(chosen, plan->hd.log, (size_t)plan->nlog * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    std::vector<uint8_t> c;
    if (cutgrid) {
      c.resize((size_t)plan->nvox);
      HC(hipMemcpyAsync(c.data(), plan->hd.cuts, c.size(), hipMemcpyDeviceToHost, s));
    }
    HC(hipStreamSynchronize(s));
    for (size_t i = 0; i < c.size(); i++) cutgrid[i] = (double)c[i];
  });
}

int iq_plan_last_time(iq_plan* plan, double* seconds) {
  if (!plan || !seconds) return fail(nullptr, IQ_ERR_ARG, "null argument");
  *seconds = plan->last_ms * 1e-3;
  return IQ_OK;
}

int iq_plan_profile(iq_plan* plan, int enable) {
  if (!plan) return fail(nullptr, IQ_ERR_ARG, "null plan");
  plan->prof = enable != 0;
  return IQ_OK;
}

int iq_plan_scan_tuning(iq_plan* plan, int nw[2], float us[4]) {
  if (!plan || !nw || !us) return fail(nullptr, IQ_ERR_ARG, "null argument");
  nw[0] = plan->mfma ? (plan->mfma_nw_forced ? plan->mfma_nw : plan->nw_tiles[1]) : 0;
  nw[1] = plan->mfma ? (plan->mfma_nw_forced ? plan->mfma_nw : plan->nw_tiles[2]) : 0;
  us[0] = plan->tune_us[1][2];
  us[1] = plan->tune_us[1][4];
  us[2] = plan->tune_us[2][2];
  us[3] = plan->tune_us[2][4];
  return IQ_OK;
}

int iq_plan_cut_time(iq_plan* plan, int64_t* launches, double* seconds) {
  if (!plan) return fail(nullptr, IQ_ERR_ARG, "null plan");
  if (launches) *launches = plan->cut_launches;
  if (seconds) *seconds = plan->cut_sec;
  return IQ_OK;
}

int iq_plan_scan_stats(iq_plan* plan, int64_t* launches, double* seconds, double* flop, double* mfma_flop) {
  if (!plan) return fail(nullptr, IQ_ERR_ARG, "null plan");
  if (launches) *launches = plan->prof_launches;
  if (seconds) *seconds = plan->prof_sec;
  if (flop) *flop = plan->prof_flop;
  if (mfma_flop) *mfma_flop = plan->prof_mflop;
  return IQ_OK;
}

int iq_plan_bank_stats(iq_plan* plan, int64_t* launches, double* seconds, double* bytes) {
  if (!plan) return fail(nullptr, IQ_ERR_ARG, "null plan");
  if (launches) *launches = plan->bank_launches;
  if (seconds) *seconds = plan->bank_sec;
  if (bytes) *bytes = plan->bank_bytes;
  return IQ_OK;
}

int iq_plan_learn(iq_ctx* ctx, iq_plan* plan, int64_t nsamples, int iters, uint64_t seed) {
  if (!ctx || !plan) return fail(ctx, IQ_ERR_ARG, "null context or plan");
  return guard(ctx, [&] {
    if (plan->ctx != ctx) bad(IQ_ERR_ARG, "plan belongs to another context");
    plan_learn(plan, nsamples, iters, seed);
  });
}

int iq_plan_learn_conv(iq_ctx* ctx, iq_plan* plan, const iq_conv_opts* opts) {
  if (!ctx || !plan || !opts) return fail(ctx, IQ_ERR_ARG, "null context, plan or options");
  return guard(ctx, [&] {
    if (plan->ctx != ctx) bad(IQ_ERR_ARG, "plan belongs to another context");
    if (plan->dtype == IQ_F32) conv_learn<float>(plan, *opts);
    else conv_learn<double>(plan, *opts);
  });
}

int iq_plan_conv_stats(iq_plan* plan, double* out) {
  if (!plan || !out) return fail(nullptr, IQ_ERR_ARG, "null plan or output");
  return guard(plan->ctx, [&] {
    if (!plan->cv_ready) bad(IQ_ERR_ARG, "call iq_plan_learn_conv first");
    for (int k = 0; k < 8; k++) out[k] = plan->cv_stats[k];
    out[8] = plan->cv_val[0];
    out[9] = plan->cv_val[1];
    out[10] = plan->cv_val[2];
    out[11] = plan->cv_val[3];
  });
}

int iq_plan_conv_params(iq_plan* plan, float* theta) {
  if (!plan || !theta) return fail(nullptr, IQ_ERR_ARG, "null plan or output");
  return guard(plan->ctx, [&] {
    if (!plan->cv_theta) bad(IQ_ERR_ARG, "call iq_plan_learn_conv first");
    HC(hipMemcpy(theta, plan->cv_theta, CV_NW * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int iq_plan_conv_grad(iq_ctx* ctx, iq_plan* plan, int mask, const float* theta, int64_t nq, int64_t ns, const float* Q,
                      const int32_t* cands, const double* D, const double* tau, double* loss, float* grad) {
  if (!ctx || !plan) return fail(ctx, IQ_ERR_ARG, "null context or plan");
  return guard(ctx, [&] {
    if (plan->ctx != ctx) bad(IQ_ERR_ARG, "plan belongs to another context");
    if (!theta || !Q || !cands || !D || !tau || !loss) bad(IQ_ERR_ARG, "null argument");
    if (mask < 0 || mask >= (int)plan->masks.size()) bad(IQ_ERR_ARG, "mask index out of range");
    if (plan->masks[mask].a2 < 0) bad(IQ_ERR_ARG, "mask without an A2 map");
    if (nq < 1 || nq > 65535 || ns < 1 || ns > CVL_MAXS) bad(IQ_ERR_ARG, "1..65535 queries of 1..4096 candidates");
    for (int64_t i = 0; i < nq * ns; i++)
      if (cands[i] < 0 || cands[i] >= plan->geo.P) bad(IQ_ERR_ARG, "candidate index out of range");
    for (int64_t i = 0; i < nq; i++)
      if (!(tau[i] > 0.0)) bad(IQ_ERR_ARG, "temperatures must be positive");
    const int V = plan->masks[mask].V;
    hipStream_t s = ctx->stream;
    ConvBuf B;
    B.init(plan);
    DMem t;
    float* th = t.upload(theta, CV_NW, s);
    float* W = t.upload(theta + CV_NW, (size_t)IQ_EMB * V, s);
    float* gW = grad ? t.alloc<float>((size_t)IQ_EMB * V) : nullptr;
    *loss = conv_loss_grad(plan, B, th, W, mask, (int)nq, (int)ns, Q, cands, D, tau, gW);
    if (grad) {
      HC(hipMemcpyAsync(grad, B.grad, CV_NW * sizeof(float), hipMemcpyDeviceToHost, s));
      HC(hipMemcpyAsync(grad + CV_NW, gW, (size_t)IQ_EMB * V * sizeof(float), hipMemcpyDeviceToHost, s));
      HC(hipStreamSynchronize(s));
    }
  });
}

int iq_plan_bank_rows(iq_ctx* ctx, iq_plan* plan, int mask, const int32_t* cands, int64_t n, float* out) {
  if (!ctx || !plan) return fail(ctx, IQ_ERR_ARG, "null context or plan");
  return guard(ctx, [&] {
    if (plan->ctx != ctx) bad(IQ_ERR_ARG, "plan belongs to another context");
    if (!plan->lrn_ready) bad(IQ_ERR_ARG, "call iq_plan_learn first");
    if (mask < 0 || mask >= (int)plan->masks.size()) bad(IQ_ERR_ARG, "mask index out of range");
    if (!cands || !out || n < 0) bad(IQ_ERR_ARG, "invalid argument");
    std::vector<uint16_t> row(IQ_EMB);
    for (int64_t i = 0; i < n; i++) {
      if (cands[i] < 0 || cands[i] >= plan->geo.P) bad(IQ_ERR_ARG, "candidate index out of range");
      HC(hipMemcpy(row.data(), plan->hd.lrn.bank + plan->lrn_bofs[mask] + (int64_t)cands[i] * IQ_EMB,
                   IQ_EMB * sizeof(uint16_t), hipMemcpyDeviceToHost));
      for (int j = 0; j < IQ_EMB; j++) {
        const uint32_t b = (uint32_t)row[j] << 16;
        float f;
        std::memcpy(&f, &b, 4);
        out[i * IQ_EMB + j] = f;
      }
    }
  });
}

int iq_plan_set_check_sample(iq_plan* plan, int queries) {
  if (!plan) return fail(nullptr, IQ_ERR_ARG, "null plan");
  if (queries < 0) return fail(plan->ctx, IQ_ERR_ARG, "negative check sample");
  plan->lrn_check_q = queries;
  return IQ_OK;
}

int iq_plan_set_proposal(iq_plan* plan, int enable, int64_t K, int check_every, double min_recall) {
  if (!plan) return fail(nullptr, IQ_ERR_ARG, "null plan");
  return guard(plan->ctx, [&] {
    if (enable && !plan->lrn_ready) bad(IQ_ERR_ARG, "call iq_plan_learn first");
    if (K < 1 || check_every < 0 || !(min_recall >= 0.0 && min_recall <= 1.0)) bad(IQ_ERR_ARG, "invalid proposal settings");
    plan->lrn_on = enable != 0;
    plan->lrn_check = check_every;
    plan->lrn_min_recall = min_recall;
    plan->hd.lrn.K = (int32_t)std::min<int64_t>(K, plan->geo.P);
    // candidate lists inside the bracket: room for several times K
    const int64_t cap = std::min<int64_t>(plan->geo.P, std::max<int64_t>(16 * (int64_t)plan->hd.lrn.K, 65536));
    if (enable && cap > plan->hd.lrn.lcap) {
      plan->hd.lrn.list = plan->mem.alloc<int2>((size_t)plan->nslots * cap);
      plan->hd.lrn.lcap = cap;
    }
    hipStream_t s = plan->ctx->stream;
    HC(hipMemcpyAsync(plan->dd, &plan->hd, sizeof(IqDev), hipMemcpyHostToDevice, s));
    HC(hipStreamSynchronize(s));
  });
}

int iq_plan_encoder(iq_plan* plan, int mask, float* W, int32_t* tvox, int64_t* V, int* nmasks) {
  if (!plan) return fail(nullptr, IQ_ERR_ARG, "null plan");
  return guard(plan->ctx, [&] {
    if (nmasks) *nmasks = (int)plan->masks.size();
    if (mask < 0 || mask >= (int)plan->masks.size()) bad(IQ_ERR_ARG, "mask index out of range");
    if (V) *V = plan->masks[mask].V;
    if (tvox) HC(hipMemcpy(tvox, plan->hd.tvox + plan->masks[mask].off, (size_t)plan->masks[mask].V * sizeof(int32_t),
                           hipMemcpyDeviceToHost));
    if (W) {
      if (!plan->lrn_ready) bad(IQ_ERR_ARG, "call iq_plan_learn first");
      HC(hipMemcpy(W, plan->hd.lrn.W + plan->lrn_wofs[mask], (size_t)IQ_EMB * plan->masks[mask].V * sizeof(float),
                   hipMemcpyDeviceToHost));
    }
  });
}

int iq_plan_proposal_stats(iq_plan* plan, int64_t* out) {
  if (!plan || !out) return fail(nullptr, IQ_ERR_ARG, "null argument");
  for (int k = 0; k < 4; k++) out[k] = plan->lrn_stats[k];
  out[4] = plan->lrn_fell_back ? 1 : 0;
  return IQ_OK;
}

int iq_plan_select_stats(iq_plan* plan, int64_t* queries, int64_t* superset, int64_t* band) {
  if (!plan) return fail(nullptr, IQ_ERR_ARG, "null plan");
  if (queries) *queries = plan->cut_stats[3];
  if (superset) *superset = plan->cut_stats[4];
  if (band) *band = plan->cut_stats[5];
  return IQ_OK;
}

int iq_plan_cut_stats(iq_plan* plan, int64_t* rounds, int64_t* levels, int64_t* cuts) {
  if (!plan) return fail(nullptr, IQ_ERR_ARG, "null plan");
  if (rounds) *rounds = plan->cut_stats[0];
  if (levels) *levels = plan->cut_stats[1];
  if (cuts) *cuts = plan->cut_stats[2];
  return IQ_OK;
}

int iq_simulate(iq_ctx* ctx, const iq_plan_desc* desc, const double* uniforms, const int64_t* forced, void* out,
                int64_t* chosen, double* cutgrid) {
  iq_plan* pl = nullptr;
  int rc = iq_plan_create(ctx, desc, &pl);
  if (rc != IQ_OK) return rc;
  rc = iq_plan_run(ctx, pl, uniforms, forced);
  if (rc == IQ_OK) rc = iq_plan_results(ctx, pl, out, chosen, cutgrid);
  iq_plan_destroy(pl);
  return rc;
}

}  // extern "C"

#include "iq_prep.hip"
