// iq_api.hip -- host driver and C ABI (include/iqhip.h) of libiqhip.
//
// One translation unit holds the host code and the kernels (k_*.hip are
// included below); the radix sort lives in iq_sort.hip.
#include "../../include/iqhip.h"
#include "iq_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "k_distance.hip"
#include "k_select.hip"
#include "k_cut.hip"

size_t iq_sort_temp_bytes(int n);
hipError_t iq_sort_pairs(void* tmp, size_t tmp_bytes, const double* keys_in, double* keys_out,
                         const int32_t* vals_in, int32_t* vals_out, int n, hipStream_t s);

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
  hipStream_t stream = nullptr;
  std::string err;
};

namespace {
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

// FP32 scan configurations: R, LX, LY, LZ
struct ScanCfg {
  int R, LX, LY, LZ;
};
const ScanCfg kScanCfg[] = {{4, 8, 8, 4}, {4, 16, 4, 4}, {2, 16, 8, 2}, {4, 8, 16, 2}};
const int kNumScanCfg = 4;

size_t scan_lds_bytes(const ScanCfg& c, const std::vector<IqBox>& boxes) {
  size_t m = 0;
  for (const IqBox& b : boxes) {
    int S = c.LX + b.wxp / c.R;
    size_t n = (size_t)c.R * S * (c.LY + b.wy - 1) * (c.LZ + b.wz - 1);
    m = std::max(m, n * sizeof(float));
  }
  return m;
}

template <int R, int LX, int LY, int LZ>
void launch_scan_cfg(dim3 grid, size_t lds, hipStream_t s, const float* ti, IqGeo g, const IqMask* mask,
                     const float* kv32, const double* b2, const double* a2, const uint8_t* dis, float* lo,
                     uint32_t* gmin) {
  hipLaunchKernelGGL((k_scan_f32<R, LX, LY, LZ>), grid, dim3(LX * LY * LZ), lds, s, ti, g, mask, kv32, b2, a2,
                     dis, lo, gmin);
}

void launch_scan(int cfg, const IqGeo& g, size_t lds, hipStream_t s, const float* ti, const IqMask* mask,
                 const float* kv32, const double* b2, const double* a2, const uint8_t* dis, float* lo,
                 uint32_t* gmin) {
  const ScanCfg& c = kScanCfg[cfg];
  dim3 grid(nblk(g.dist[0], c.R * c.LX), nblk(g.dist[1], c.LY), nblk(g.dist[2], c.LZ));
  switch (cfg) {
    case 0: launch_scan_cfg<4, 8, 8, 4>(grid, lds, s, ti, g, mask, kv32, b2, a2, dis, lo, gmin); break;
    case 1: launch_scan_cfg<4, 16, 4, 4>(grid, lds, s, ti, g, mask, kv32, b2, a2, dis, lo, gmin); break;
    case 2: launch_scan_cfg<2, 16, 8, 2>(grid, lds, s, ti, g, mask, kv32, b2, a2, dis, lo, gmin); break;
    default: launch_scan_cfg<4, 8, 16, 2>(grid, lds, s, ti, g, mask, kv32, b2, a2, dis, lo, gmin); break;
  }
  HCL();
}

int pick_scan_cfg(const IqGeo& g, const std::vector<IqBox>& boxes, size_t* lds) {
  int best = -1;
  double bestcost = 1e300;
  for (int i = 0; i < kNumScanCfg; i++) {
    const ScanCfg& c = kScanCfg[i];
    size_t l = scan_lds_bytes(c, boxes);
    if (l > 64 * 1024) continue;
    double padded = (double)nblk(g.dist[0], c.R * c.LX) * c.R * c.LX * nblk(g.dist[1], c.LY) * c.LY *
                    nblk(g.dist[2], c.LZ) * c.LZ;
    double cost = padded * (1.0 + 0.02 * (double)l / 16384.0);
    if (cost < bestcost) { bestcost = cost; best = i; *lds = l; }
  }
  return best;
}

bool fits_f32(const double* v, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    double x = v[i];
    if ((double)(float)x != x || std::fabs(x) > 0x1p50) return false;
  }
  return true;
}
bool fits_range(const float* v, int64_t n) {
  for (int64_t i = 0; i < n; i++)
    if (!(std::fabs(v[i]) <= 0x1p50f)) return false;
  return true;
}

void geo_init(IqGeo& g, const int64_t ti[3], const int64_t tile[3], const int64_t pad[3]) {
  for (int d = 0; d < 3; d++) {
    g.ti[d] = (int32_t)ti[d];
    g.tile[d] = (int32_t)tile[d];
    g.pad[d] = (int32_t)pad[d];
    g.dist[d] = (int32_t)(ti[d] - tile[d] + 1);
  }
  g.P = (int64_t)g.dist[0] * g.dist[1] * g.dist[2];
  g.padvox = (int64_t)g.pad[0] * g.pad[1] * g.pad[2];
  g.tilevox = g.tile[0] * g.tile[1] * g.tile[2];
}

}  // namespace

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
struct iq_plan {
  iq_ctx* ctx = nullptr;
  int dtype = IQ_F64;
  IqGeo geo{};
  int64_t ovl[3], spacing[3], ntiles[3];
  int64_t npath = 0, nreal = 0, nen = 0;
  double tol = 0.1;
  int naux = 0;
  bool fast = false;
  bool hard = false;
  std::vector<IqStep> steps;
  std::vector<IqMask> masks;
  std::vector<int> mask_cfg;
  std::vector<size_t> mask_lds;
  DMem mem;
  void* ti = nullptr;
  float* ti32 = nullptr;
  uint8_t* disabled = nullptr;
  int32_t* enabled = nullptr;
  IqStep* dsteps = nullptr;
  IqMask* dmasks = nullptr;
  int32_t *dtvox = nullptr, *dtoff = nullptr, *dhtvox = nullptr, *dhtoff = nullptr;
  int32_t full_off = 0;
  std::vector<double*> a2maps;
  double* hardgrid = nullptr;
  std::vector<double*> aux, auxti, a2soft;
  // per-query scratch
  double* kv = nullptr;
  float* kv32 = nullptr;
  double* b2 = nullptr;
  uint32_t* gmin_hi = nullptr;
  unsigned long long* gmin64 = nullptr;
  float* lo32 = nullptr;
  double* dexact = nullptr;
  unsigned long long* bitmap = nullptr;
  int32_t* band = nullptr;
  double* bandd = nullptr;
  int32_t* rank = nullptr;
  double* p = nullptr;
  double* s1 = nullptr;
  int64_t* pick = nullptr;
  int32_t* err = nullptr;
  int64_t* nband = nullptr;
  int64_t* rounds = nullptr;
  uint8_t* keep = nullptr;
  char* cutscratch = nullptr;
  int64_t cutscratch_per = 0;
  // multi-source selection
  std::vector<double*> maps;      // 0 overlap, 1 hard, 2+s soft
  std::vector<int32_t*> sorted;   // stable ascending orders of maps
  double* keys_out = nullptr;
  int32_t* iota = nullptr;
  void* sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  const double** dmaps_hard = nullptr;
  const double** dmaps_soft = nullptr;
  const int32_t** dsorted_hard = nullptr;
  const int32_t** dsorted_soft = nullptr;
  uint32_t* bits = nullptr;
  int32_t* ev = nullptr;
  double* vals = nullptr;
  double* kvh = nullptr;
  double* b2h = nullptr;
  double* kvs = nullptr;
  double* b2s = nullptr;
  // results
  void* sims = nullptr;
  uint8_t* cuts = nullptr;
  int64_t* log = nullptr;
  float last_ms = 0.0f;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  size_t cut_lds_max = 0;
  // scan profiling
  bool prof = false;
  std::vector<hipEvent_t> prof_ev;
  int64_t prof_used = 0;
  double prof_flop = 0.0, prof_sec = 0.0;
  int64_t prof_launches = 0;
};

namespace {

template <typename T>
T* sim_of(iq_plan* pl, int64_t r) {
  return (T*)pl->sims + r * pl->geo.padvox;
}

void plan_build(iq_ctx* ctx, const iq_plan_desc* d, iq_plan* pl) {
  if (!d || !d->ti || !d->disabled || !d->path || !d->skipped) bad(IQ_ERR_ARG, "null plan input");
  if (d->ti_dtype != IQ_F32 && d->ti_dtype != IQ_F64) bad(IQ_ERR_ARG, "ti_dtype must be IQ_F32 or IQ_F64");
  if (d->nreal <= 0) bad(IQ_ERR_ARG, "invalid number of realizations");
  if (!(d->tol > 0.0 && d->tol <= 1.0)) bad(IQ_ERR_ARG, "tolerance must be in range (0,1]");
  if (d->naux < 0 || d->naux > 8) bad(IQ_ERR_ARG, "naux must be in 0..8");
  hipStream_t s = ctx->stream;
  pl->ctx = ctx;
  pl->dtype = d->ti_dtype;
  pl->npath = d->npath;
  pl->nreal = d->nreal;
  pl->tol = d->tol;
  pl->naux = d->naux;
  for (int k = 0; k < 3; k++) {
    if (d->tilesize[k] <= 0 || d->tilesize[k] > d->tisize[k]) bad(IQ_ERR_ARG, "invalid tile size");
    if (d->ovlsize[k] < 0 || d->ovlsize[k] > d->tilesize[k]) bad(IQ_ERR_ARG, "invalid overlap size");
    if (d->spacing[k] != d->tilesize[k] - d->ovlsize[k]) bad(IQ_ERR_ARG, "spacing must be tilesize - ovlsize");
    if (d->ntiles[k] <= 0) bad(IQ_ERR_ARG, "invalid ntiles");
    if (d->padsize[k] != d->ntiles[k] * d->spacing[k] + d->ovlsize[k]) bad(IQ_ERR_ARG, "padsize must be ntiles*(tilesize-ovlsize)+ovlsize");
    pl->ovl[k] = d->ovlsize[k];
    pl->spacing[k] = d->spacing[k];
    pl->ntiles[k] = d->ntiles[k];
  }
  geo_init(pl->geo, d->tisize, d->tilesize, d->padsize);
  IqGeo& g = pl->geo;
  if (g.P > (int64_t)S2MAX * IQ_SB * IQ_SB || g.padvox >= (1ll << 40))
    bad(IQ_ERR_UNSUPPORTED, "more than " + std::to_string((int64_t)S2MAX * IQ_SB * IQ_SB) + " candidate patterns");
  const int64_t tivox = (int64_t)g.ti[0] * g.ti[1] * g.ti[2];
  const int64_t nt = pl->ntiles[0] * pl->ntiles[1] * pl->ntiles[2];
  for (int64_t i = 0; i < d->npath; i++)
    if (d->path[i] < 0 || d->path[i] >= nt) bad(IQ_ERR_ARG, "path entry outside the tile grid");
  const int t3[3] = {g.tile[0], g.tile[1], g.tile[2]};

  // ---- steps and masks (iqsim.jl:172-205, 252-275), shared by all realizations
  std::vector<uint8_t> pasted((size_t)nt, 0);
  std::map<std::vector<uint8_t>, int> mask_index;
  std::vector<HostMask> hmasks;
  std::vector<int32_t> htvox;
  pl->hard = d->hardgrid != nullptr;
  for (int64_t i = 0; i < d->npath; i++) {
    IqStep st{};
    int64_t ind = d->path[i];
    if (d->skipped[ind]) { st.active = 0; st.mask = -1; pl->steps.push_back(st); continue; }
    st.active = 1;
    int64_t ti3[3] = {ind % pl->ntiles[0], (ind / pl->ntiles[0]) % pl->ntiles[1], ind / (pl->ntiles[0] * pl->ntiles[1])};
    for (int k = 0; k < 3; k++) st.origin[k] = (int32_t)(ti3[k] * pl->spacing[k]);
    std::vector<uint8_t> bits((size_t)g.tilevox, 0);
    int any = 0;
    for (int k = 0; k < 3; k++) {
      if (pl->ovl[k] <= 1) continue;
      for (int side = 0; side < 2; side++) {
        int64_t nb[3] = {ti3[0], ti3[1], ti3[2]};
        nb[k] += side == 0 ? -1 : 1;
        if (nb[k] < 0 || nb[k] >= pl->ntiles[k]) continue;
        if (!pasted[nb[0] + pl->ntiles[0] * (nb[1] + pl->ntiles[1] * nb[2])]) continue;
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
      } else {
        st.mask = it->second;
      }
    }
    // hard data in the tile (iqsim.jl:210-219, indicator!/event! utils.jl:18-36)
    if (d->hardgrid) {
      st.hoff = (int32_t)htvox.size();
      for (int z = 0; z < t3[2]; z++)
        for (int y = 0; y < t3[1]; y++)
          for (int x = 0; x < t3[0]; x++) {
            double v = d->hardgrid[iq_lin3(g.pad, st.origin[0] + x, st.origin[1] + y, st.origin[2] + z)];
            if (!std::isnan(v)) htvox.push_back(x + t3[0] * (y + t3[1] * z));
          }
      st.hcnt = (int32_t)htvox.size() - st.hoff;
      st.hard = st.hcnt > 0;
    }
    pasted[ind] = 1;
    pl->steps.push_back(st);
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
  pl->full_off = (int32_t)tvox.size();
  for (int32_t t = 0; t < g.tilevox; t++) { tvox.push_back(t); toff.push_back(ti_off(t)); }
  std::vector<int32_t> htoff(htvox.size());
  for (size_t i = 0; i < htvox.size(); i++) htoff[i] = ti_off(htvox[i]);

  // ---- training image
  const double* tid = (const double*)d->ti;
  const float* tif = (const float*)d->ti;
  if (d->ti_dtype == IQ_F32) {
    pl->ti = pl->mem.upload(tif, tivox, s);
    pl->fast = fits_range(tif, tivox);
    if (pl->fast) pl->ti32 = (float*)pl->ti;
  } else {
    pl->ti = pl->mem.upload(tid, tivox, s);
    pl->fast = fits_f32(tid, tivox);
    if (pl->fast) {
      std::vector<float> t32((size_t)tivox);
      for (int64_t i = 0; i < tivox; i++) t32[i] = (float)tid[i];
      pl->ti32 = pl->mem.upload(t32.data(), tivox, s);
      HC(hipStreamSynchronize(s));
    }
  }
  pl->disabled = pl->mem.upload(d->disabled, (size_t)g.P, s);
  std::vector<int32_t> en;
  en.reserve((size_t)g.P);
  for (int64_t c = 0; c < g.P; c++)
    if (!d->disabled[c]) en.push_back((int32_t)c);
  pl->nen = (int64_t)en.size();
  pl->enabled = pl->mem.upload(en.data(), en.size(), s);
  pl->dtvox = pl->mem.upload(tvox.data(), tvox.size(), s);
  pl->dtoff = pl->mem.upload(toff.data(), toff.size(), s);
  pl->dhtvox = pl->mem.upload(htvox.data(), htvox.size(), s);
  pl->dhtoff = pl->mem.upload(htoff.data(), htoff.size(), s);
  pl->dsteps = pl->mem.upload(pl->steps.data(), pl->steps.size(), s);

  // ---- per-query scratch
  const int64_t P = g.P;
  const int64_t nw64 = (P + 63) / 64;
  int kmax = 4;
  for (const IqMask& m : pl->masks) kmax = std::max(kmax, m.kbox);
  pl->kv = pl->mem.alloc<double>(g.tilevox);
  pl->kv32 = pl->mem.alloc<float>(kmax);
  pl->b2 = pl->mem.alloc<double>(1);
  pl->gmin_hi = pl->mem.alloc<uint32_t>(1);
  pl->gmin64 = pl->mem.alloc<unsigned long long>(1);
  pl->lo32 = pl->mem.alloc<float>(P);
  pl->dexact = pl->mem.alloc<double>(P);
  pl->bitmap = pl->mem.alloc<unsigned long long>(nw64);
  pl->band = pl->mem.alloc<int32_t>(P);
  pl->bandd = pl->mem.alloc<double>(P);
  pl->rank = pl->mem.alloc<int32_t>(P);
  pl->p = pl->mem.alloc<double>(P);
  pl->s1 = pl->mem.alloc<double>(nw64 + 1);
  pl->pick = pl->mem.alloc<int64_t>(1);
  pl->err = pl->mem.alloc<int32_t>(1);
  pl->nband = pl->mem.alloc<int64_t>(1);
  pl->rounds = pl->mem.alloc<int64_t>(1);
  pl->keep = pl->mem.alloc<uint8_t>((size_t)IQ_MAX_SLAB * g.tilevox);
  int64_t maxn = 0;
  for (const IqStep& st : pl->steps)
    for (int q = 0; q < st.nslab; q++) {
      const IqSlab& sl = st.slab[q];
      int64_t n = (int64_t)sl.sz[0] * sl.sz[1] * sl.sz[2] / sl.sz[sl.dim] * std::max(0, sl.sz[sl.dim] - 2);
      if (n <= IQ_PR_NLDS) pl->cut_lds_max = std::max(pl->cut_lds_max, (size_t)n * CUT_NODE_BYTES);
      maxn = std::max(maxn, n);
    }
  if (maxn > IQ_PR_NLDS) {
    pl->cutscratch_per = maxn * CUT_NODE_BYTES;
    pl->cutscratch = pl->mem.alloc<char>((size_t)pl->cutscratch_per * IQ_MAX_SLAB);
  }

  // ---- exact A2 maps of the overlap masks (MODE 2 chains)
  IqMask* dm = pl->mem.upload(pl->masks.data(), pl->masks.size(), s);
  pl->dmasks = dm;
  for (size_t mi = 0; mi < pl->masks.size(); mi++) {
    IqMask& m = pl->masks[mi];
    double* a2 = pl->mem.alloc<double>(P);
    if (pl->dtype == IQ_F32)
      hipLaunchKernelGGL((k_exact_map<float, 2>), dim3(nblk(P, 256)), dim3(256), 0, s, (const float*)pl->ti, g,
                         pl->dtoff + m.off, (const double*)nullptr, m.V, (const double*)nullptr,
                         (const double*)nullptr, (const uint8_t*)nullptr, a2);
    else
      hipLaunchKernelGGL((k_exact_map<double, 2>), dim3(nblk(P, 256)), dim3(256), 0, s, (const double*)pl->ti, g,
                         pl->dtoff + m.off, (const double*)nullptr, m.V, (const double*)nullptr,
                         (const double*)nullptr, (const uint8_t*)nullptr, a2);
    HCL();
    m.a2 = (int32_t)pl->a2maps.size();
    pl->a2maps.push_back(a2);
    size_t lds = 0;
    int cfg = pl->fast ? pick_scan_cfg(g, hmasks[mi].boxes, &lds) : -1;
    if (cfg < 0) pl->fast = false;
    pl->mask_cfg.push_back(cfg);
    pl->mask_lds.push_back(lds);
  }
  HC(hipMemcpyAsync(dm, pl->masks.data(), pl->masks.size() * sizeof(IqMask), hipMemcpyHostToDevice, s));

  // ---- hard and soft data, multi-source selection buffers
  const bool multi = pl->hard || pl->naux > 0;
  if (pl->hard) {
    pl->hardgrid = pl->mem.upload(d->hardgrid, (size_t)g.padvox, s);
    pl->kvh = pl->mem.alloc<double>(g.tilevox);
    pl->b2h = pl->mem.alloc<double>(1);
  }
  if (pl->naux > 0) {
    if (!d->aux || !d->auxti) bad(IQ_ERR_ARG, "null soft data");
    pl->kvs = pl->mem.alloc<double>((size_t)g.tilevox * pl->naux);
    pl->b2s = pl->mem.alloc<double>(pl->naux);
    for (int a = 0; a < pl->naux; a++) {
      if (!d->aux[a] || !d->auxti[a]) bad(IQ_ERR_ARG, "null soft data");
      pl->aux.push_back(pl->mem.upload(d->aux[a], (size_t)g.padvox, s));
      double* at = pl->mem.upload(d->auxti[a], (size_t)tivox, s);
      pl->auxti.push_back(at);
      double* a2 = pl->mem.alloc<double>(P);
      hipLaunchKernelGGL((k_exact_map<double, 2>), dim3(nblk(P, 256)), dim3(256), 0, s, (const double*)at, g,
                         pl->dtoff + pl->full_off, (const double*)nullptr, g.tilevox, (const double*)nullptr,
                         (const double*)nullptr, (const uint8_t*)nullptr, a2);
      HCL();
      pl->a2soft.push_back(a2);
    }
  }
  if (multi) {
    int nm = 2 + pl->naux;
    for (int k = 0; k < nm; k++) {
      pl->maps.push_back(pl->mem.alloc<double>(P));
      pl->sorted.push_back(pl->mem.alloc<int32_t>(P));
    }
    pl->keys_out = pl->mem.alloc<double>(P);
    pl->iota = pl->mem.alloc<int32_t>(P);
    hipLaunchKernelGGL(k_iota32, dim3(nblk(P, 256)), dim3(256), 0, s, pl->iota, P);
    HCL();
    pl->sort_tmp_bytes = iq_sort_temp_bytes((int)P);
    pl->sort_tmp = pl->mem.alloc<char>(pl->sort_tmp_bytes);
    std::vector<const double*> mh, ms;
    std::vector<const int32_t*> sh_, ss;
    mh.push_back(pl->maps[1]); sh_.push_back(pl->sorted[1]);
    mh.push_back(pl->maps[0]); sh_.push_back(pl->sorted[0]);
    ms.push_back(pl->maps[0]); ss.push_back(pl->sorted[0]);
    for (int a = 0; a < pl->naux; a++) {
      mh.push_back(pl->maps[2 + a]); sh_.push_back(pl->sorted[2 + a]);
      ms.push_back(pl->maps[2 + a]); ss.push_back(pl->sorted[2 + a]);
    }
    pl->dmaps_hard = (const double**)pl->mem.upload(mh.data(), mh.size(), s);
    pl->dmaps_soft = (const double**)pl->mem.upload(ms.data(), ms.size(), s);
    pl->dsorted_hard = (const int32_t**)pl->mem.upload(sh_.data(), sh_.size(), s);
    pl->dsorted_soft = (const int32_t**)pl->mem.upload(ss.data(), ss.size(), s);
    pl->bits = pl->mem.alloc<uint32_t>((size_t)((P + 31) / 32) * nm);
    pl->ev = pl->mem.alloc<int32_t>(P);
    pl->vals = pl->mem.alloc<double>(P);
  }

  // ---- results
  const size_t esz = pl->dtype == IQ_F32 ? 4 : 8;
  pl->sims = pl->mem.alloc<char>((size_t)pl->nreal * g.padvox * esz);
  pl->cuts = pl->mem.alloc<uint8_t>((size_t)pl->nreal * g.padvox);
  pl->log = pl->mem.alloc<int64_t>((size_t)pl->nreal * std::max<int64_t>(pl->npath, 1));
  HC(hipEventCreate(&pl->e0));
  HC(hipEventCreate(&pl->e1));
  if (pl->cut_lds_max > 64 * 1024) {
    HC(hipFuncSetAttribute((const void*)k_graphcut<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl->cut_lds_max));
    HC(hipFuncSetAttribute((const void*)k_graphcut<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl->cut_lds_max));
  }
  HC(hipStreamSynchronize(s));
}

template <typename T>
void run_step(iq_plan* pl, int64_t r, int64_t i, double u, int64_t forced) {
  const IqGeo& g = pl->geo;
  hipStream_t s = pl->ctx->stream;
  const IqStep& st = pl->steps[i];
  const IqStep* dst = pl->dsteps + i;
  T* sim = sim_of<T>(pl, r);
  const T* ti = (const T*)pl->ti;
  const int64_t P = g.P;
  SelOut o{pl->pick, pl->log + r * pl->npath + i, pl->err, pl->nband};
  const bool multi = st.hard || pl->naux > 0;
  const int32_t* origin = &dst->origin[0];
  if (!multi) {
    if (st.mask < 0) {
      hipLaunchKernelGGL(k_band_enabled, dim3(1), dim3(FB), 0, s, (const int32_t*)pl->enabled, pl->nen, u, forced,
                         pl->p, pl->s1, o);
      HCL();
    } else {
      const IqMask& m = pl->masks[st.mask];
      const IqMask* dm = pl->dmasks + st.mask;
      const double* a2 = pl->a2maps[m.a2];
      hipLaunchKernelGGL((k_query_prep<T>), dim3(1), dim3(256), 0, s, (const T*)sim, g, origin, pl->dtvox + m.off,
                         m.V, pl->fast ? dm : (const IqMask*)nullptr, pl->kv, pl->kv32, pl->b2, pl->gmin_hi,
                         pl->gmin64);
      HCL();
      if (pl->fast) {
        if (pl->prof) {
          if (pl->prof_used + 2 > (int64_t)pl->prof_ev.size()) {
            hipEvent_t a, b;
            HC(hipEventCreate(&a));
            HC(hipEventCreate(&b));
            pl->prof_ev.push_back(a);
            pl->prof_ev.push_back(b);
          }
          HC(hipEventRecord(pl->prof_ev[pl->prof_used], s));
        }
        launch_scan(pl->mask_cfg[st.mask], g, pl->mask_lds[st.mask], s, pl->ti32, dm, pl->kv32, pl->b2, a2,
                    pl->disabled, pl->lo32, pl->gmin_hi);
        if (pl->prof) {
          HC(hipEventRecord(pl->prof_ev[pl->prof_used + 1], s));
          pl->prof_used += 2;
          pl->prof_flop += 2.0 * (double)P * (double)m.V;
        }
        hipLaunchKernelGGL((k_band_superset<T>), dim3(nblk(P, 256)), dim3(256), 0, s, ti, g, (const float*)pl->lo32,
                           (const uint32_t*)pl->gmin_hi, pl->tol, pl->dtoff + m.off, (const double*)pl->kv, m.V,
                           (const double*)pl->b2, a2, pl->dexact, pl->bitmap, pl->gmin64);
        HCL();
      } else {
        hipLaunchKernelGGL((k_exact_map<T, 1>), dim3(nblk(P, 256)), dim3(256), 0, s, ti, g, pl->dtoff + m.off,
                           (const double*)pl->kv, m.V, (const double*)pl->b2, a2, (const uint8_t*)pl->disabled,
                           pl->dexact);
        HCL();
        hipLaunchKernelGGL(k_band_all, dim3(nblk(P, 256)), dim3(256), 0, s, (const double*)pl->dexact, g, pl->bitmap,
                           pl->gmin64);
        HCL();
      }
      hipLaunchKernelGGL(k_band_finalize, dim3(1), dim3(FB), 0, s, P, (const unsigned long long*)pl->bitmap,
                         (const double*)pl->dexact, (const unsigned long long*)pl->gmin64, pl->tol, u, forced,
                         pl->band, pl->bandd, pl->rank, pl->p, pl->s1, o);
      HCL();
    }
  } else {
    // overlap map (iqsim.jl:206-207)
    if (st.mask >= 0) {
      const IqMask& m = pl->masks[st.mask];
      hipLaunchKernelGGL((k_query_prep<T>), dim3(1), dim3(256), 0, s, (const T*)sim, g, origin, pl->dtvox + m.off,
                         m.V, (const IqMask*)nullptr, pl->kv, (float*)nullptr, pl->b2, (uint32_t*)nullptr,
                         (unsigned long long*)nullptr);
      HCL();
      hipLaunchKernelGGL((k_exact_map<T, 1>), dim3(nblk(P, 256)), dim3(256), 0, s, ti, g, pl->dtoff + m.off,
                         (const double*)pl->kv, m.V, (const double*)pl->b2, (const double*)pl->a2maps[m.a2],
                         (const uint8_t*)pl->disabled, pl->maps[0]);
      HCL();
    } else {
      hipLaunchKernelGGL(k_zero_map, dim3(nblk(P, 256)), dim3(256), 0, s, pl->maps[0], (const uint8_t*)pl->disabled, P);
      HCL();
    }
    // hard map (iqsim.jl:210-219)
    if (st.hard) {
      hipLaunchKernelGGL((k_query_prep<double>), dim3(1), dim3(256), 0, s, (const double*)pl->hardgrid, g, origin,
                         pl->dhtvox + st.hoff, st.hcnt, (const IqMask*)nullptr, pl->kvh, (float*)nullptr, pl->b2h,
                         (uint32_t*)nullptr, (unsigned long long*)nullptr);
      HCL();
      hipLaunchKernelGGL((k_exact_map<T, 0>), dim3(nblk(P, 256)), dim3(256), 0, s, ti, g, pl->dhtoff + st.hoff,
                         (const double*)pl->kvh, st.hcnt, (const double*)pl->b2h, (const double*)nullptr,
                         (const uint8_t*)pl->disabled, pl->maps[1]);
      HCL();
    }
    // soft maps (iqsim.jl:222-227)
    for (int a = 0; a < pl->naux; a++) {
      hipLaunchKernelGGL((k_query_prep<double>), dim3(1), dim3(256), 0, s, (const double*)pl->aux[a], g, origin,
                         pl->dtvox + pl->full_off, g.tilevox, (const IqMask*)nullptr, pl->kvs + (int64_t)a * g.tilevox,
                         (float*)nullptr, pl->b2s + a, (uint32_t*)nullptr, (unsigned long long*)nullptr);
      HCL();
      hipLaunchKernelGGL((k_exact_map<double, 1>), dim3(nblk(P, 256)), dim3(256), 0, s, (const double*)pl->auxti[a],
                         g, pl->dtoff + pl->full_off, (const double*)(pl->kvs + (int64_t)a * g.tilevox), g.tilevox,
                         (const double*)(pl->b2s + a), (const double*)pl->a2soft[a], (const uint8_t*)pl->disabled,
                         pl->maps[2 + a]);
      HCL();
    }
    // stable orders (partialsortperm, relaxation.jl:12,27)
    std::vector<int> srcs;
    if (st.hard) srcs.push_back(1);
    srcs.push_back(0);
    for (int a = 0; a < pl->naux; a++) srcs.push_back(2 + a);
    for (int k : srcs)
      HC(iq_sort_pairs(pl->sort_tmp, pl->sort_tmp_bytes, pl->maps[k], pl->keys_out, pl->iota, pl->sorted[k], (int)P, s));
    const int nsrc = (int)srcs.size();
    hipLaunchKernelGGL(k_relax_finalize, dim3(1), dim3(FB), 0, s, P, nsrc,
                       st.hard ? pl->dmaps_hard : pl->dmaps_soft, st.hard ? pl->dsorted_hard : pl->dsorted_soft,
                       pl->tol, u, forced, pl->bits, pl->ev, pl->vals, pl->rank, pl->p, pl->s1, o);
    HCL();
  }
  // boundary cut (iqsim.jl:252-275) and paste (iqsim.jl:278-281)
  if (st.nslab > 0) {
    hipLaunchKernelGGL((k_graphcut<T>), dim3(st.nslab), dim3(CB), pl->cut_lds_max, s, (const T*)sim, ti, g, dst,
                       (const int64_t*)pl->pick, pl->keep, pl->cutscratch, pl->cutscratch_per, pl->err, pl->rounds);
    HCL();
  }
  hipLaunchKernelGGL((k_paste<T>), dim3(nblk(g.tilevox, 256)), dim3(256), 0, s, sim, ti, g, dst,
                     (const int64_t*)pl->pick, (const uint8_t*)pl->keep, pl->cuts + r * g.padvox);
  HCL();
}

template <typename T>
void plan_run(iq_plan* pl, const double* uniforms, const int64_t* forced) {
  hipStream_t s = pl->ctx->stream;
  const IqGeo& g = pl->geo;
  pl->prof_used = 0;
  pl->prof_flop = 0.0;
  pl->prof_sec = 0.0;
  pl->prof_launches = 0;
  HC(hipEventRecord(pl->e0, s));
  hipLaunchKernelGGL((k_fill<int32_t>), dim3(1), dim3(64), 0, s, pl->err, (int64_t)1, 0);
  hipLaunchKernelGGL((k_fill<int64_t>), dim3(1), dim3(64), 0, s, pl->rounds, (int64_t)1, (int64_t)0);
  hipLaunchKernelGGL((k_fill<int64_t>), dim3(256), dim3(256), 0, s, pl->log, pl->nreal * pl->npath, (int64_t)-1);
  HCL();
  for (int64_t r = 0; r < pl->nreal; r++) {
    hipLaunchKernelGGL((k_fill<T>), dim3(1024), dim3(256), 0, s, sim_of<T>(pl, r), g.padvox, (T)0);
    hipLaunchKernelGGL((k_fill<uint8_t>), dim3(1024), dim3(256), 0, s, pl->cuts + r * g.padvox, g.padvox, (uint8_t)0);
    hipLaunchKernelGGL((k_fill<int64_t>), dim3(1), dim3(64), 0, s, pl->pick, (int64_t)1, (int64_t)-1);
    HCL();
    for (int64_t i = 0; i < pl->npath; i++) {
      if (!pl->steps[i].active) continue;
      const double u = uniforms ? uniforms[r * pl->npath + i] : 0.5;
      const int64_t f = forced ? forced[r * pl->npath + i] : -1;
      run_step<T>(pl, r, i, u, f);
    }
  }
  HC(hipEventRecord(pl->e1, s));
  HC(hipStreamSynchronize(s));
  HC(hipEventElapsedTime(&pl->last_ms, pl->e0, pl->e1));
  for (int64_t k = 0; k + 1 < pl->prof_used; k += 2) {
    float ms = 0.0f;
    HC(hipEventElapsedTime(&ms, pl->prof_ev[k], pl->prof_ev[k + 1]));
    pl->prof_sec += ms * 1e-3;
    pl->prof_launches++;
  }
  int32_t code = 0;
  HC(hipMemcpy(&code, pl->err, sizeof(int32_t), hipMemcpyDeviceToHost));
  check_kernel_error(code);
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
    int64_t one[3] = {kdims[0], kdims[1], kdims[2]};
    geo_init(g, idims, kdims, one);  // query grid = the kernel array itself
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
    int32_t zero3[3] = {0, 0, 0};
    int32_t* dorg = m.upload(zero3, 3, s);
    hipLaunchKernelGGL((k_query_prep<double>), dim3(1), dim3(256), 0, s, (const double*)dk, g, (const int32_t*)dorg,
                       (const int32_t*)dtv, V, (const IqMask*)nullptr, kv, (float*)nullptr, b2, (uint32_t*)nullptr,
                       (unsigned long long*)nullptr);
    HCL();
    if (img_dtype == IQ_F32)
      hipLaunchKernelGGL((k_exact_map<float, 0>), dim3(nblk(g.P, 256)), dim3(256), 0, s, (const float*)di, g,
                         (const int32_t*)dto, (const double*)kv, V, (const double*)b2, (const double*)nullptr,
                         (const uint8_t*)nullptr, dout);
    else
      hipLaunchKernelGGL((k_exact_map<double, 0>), dim3(nblk(g.P, 256)), dim3(256), 0, s, (const double*)di, g,
                         (const int32_t*)dto, (const double*)kv, V, (const double*)b2, (const double*)nullptr,
                         (const uint8_t*)nullptr, dout);
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
    if (!D || !out_idx || !out_count || n <= 0 || naux < 1 || naux > 15 || !Ds) bad(IQ_ERR_ARG, "invalid argument");
    if (n >= (1ll << 31)) bad(IQ_ERR_UNSUPPORTED, "more than 2^31 candidates");
    hipStream_t s = ctx->stream;
    DMem m;
    int nsrc = naux + 1;
    std::vector<const double*> maps;
    std::vector<const int32_t*> sorted;
    double* keys = m.alloc<double>(n);
    int32_t* iota = m.alloc<int32_t>(n);
    hipLaunchKernelGGL(k_iota32, dim3(nblk(n, 256)), dim3(256), 0, s, iota, n);
    HCL();
    size_t tb = iq_sort_temp_bytes((int)n);
    void* tmp = m.alloc<char>(tb);
    for (int j = 0; j < nsrc; j++) {
      const double* src = j == 0 ? D : Ds[j - 1];
      if (!src) bad(IQ_ERR_ARG, "null auxiliary map");
      double* dm = m.upload(src, n, s);
      int32_t* so = m.alloc<int32_t>(n);
      HC(iq_sort_pairs(tmp, tb, dm, keys, iota, so, (int)n, s));
      maps.push_back(dm);
      sorted.push_back(so);
    }
    const double** dmaps = (const double**)m.upload(maps.data(), maps.size(), s);
    const int32_t** dsorted = (const int32_t**)m.upload(sorted.data(), sorted.size(), s);
    uint32_t* bits = m.alloc<uint32_t>((size_t)((n + 31) / 32) * nsrc);
    int32_t* ev = m.alloc<int32_t>(n);
    int64_t* cnt = m.alloc<int64_t>(1);
    int32_t* err = m.alloc<int32_t>(1);
    HC(hipMemsetAsync(err, 0, sizeof(int32_t), s));
    SelOut o{nullptr, nullptr, err, cnt};
    hipLaunchKernelGGL(k_relax_finalize, dim3(1), dim3(FB), 0, s, n, nsrc, dmaps, dsorted, tol, 0.0, (int64_t)-1,
                       bits, ev, (double*)nullptr, (int32_t*)nullptr, (double*)nullptr, (double*)nullptr, o);
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
    int32_t* err = m.alloc<int32_t>(1);
    HC(hipMemsetAsync(err, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_taumodel_host, dim3(1), dim3(FB), 0, s, (const int32_t*)dev, nev, naux + 1, dmaps, vals, rank,
                       p, err);
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
    if (dtype != IQ_F32 && dtype != IQ_F64) bad(IQ_ERR_ARG, "dtype must be IQ_F32 or IQ_F64");
    if (dim < 0 || dim > 2 || dims[dim] < 2) bad(IQ_ERR_ARG, "cut axis needs at least two slices");
    for (int d = 0; d < 3; d++)
      if (dims[d] <= 0) bad(IQ_ERR_ARG, "invalid slab extents");
    IqGeo g{};
    geo_init(g, dims, dims, dims);  // A is the grid, B is the training image, one candidate at 0
    IqStep st{};
    st.active = 1;
    st.nslab = 1;
    st.slab[0].dim = dim;
    st.slab[0].side = 0;
    for (int d = 0; d < 3; d++) st.slab[0].sz[d] = (int32_t)dims[d];
    int L = (int)dims[dim];
    int64_t n = g.tilevox / L * std::max(0, L - 2);
    hipStream_t s = ctx->stream;
    DMem m;
    const size_t esz = dtype == IQ_F32 ? 4 : 8;
    char* dA = m.alloc<char>(g.tilevox * esz);
    char* dB = m.alloc<char>(g.tilevox * esz);
    HC(hipMemcpyAsync(dA, A, g.tilevox * esz, hipMemcpyHostToDevice, s));
    HC(hipMemcpyAsync(dB, B, g.tilevox * esz, hipMemcpyHostToDevice, s));
    IqStep* dst = m.upload(&st, 1, s);
    int64_t zero = 0;
    int64_t* pick = m.upload(&zero, 1, s);
    uint8_t* keep = m.alloc<uint8_t>(g.tilevox);
    int32_t* err = m.alloc<int32_t>(1);
    HC(hipMemsetAsync(err, 0, sizeof(int32_t), s));
    char* scratch = nullptr;
    size_t lds = 0;
    if (n > IQ_PR_NLDS) scratch = m.alloc<char>((size_t)n * CUT_NODE_BYTES);
    else lds = (size_t)n * CUT_NODE_BYTES;
    if (lds > 64 * 1024) {
      HC(hipFuncSetAttribute((const void*)k_graphcut<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      HC(hipFuncSetAttribute((const void*)k_graphcut<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    if (dtype == IQ_F32)
      hipLaunchKernelGGL((k_graphcut<float>), dim3(1), dim3(CB), lds, s, (const float*)dA, (const float*)dB, g,
                         (const IqStep*)dst, (const int64_t*)pick, keep, scratch, (int64_t)(n * CUT_NODE_BYTES), err,
                         (int64_t*)nullptr);
    else
      hipLaunchKernelGGL((k_graphcut<double>), dim3(1), dim3(CB), lds, s, (const double*)dA, (const double*)dB, g,
                         (const IqStep*)dst, (const int64_t*)pick, keep, scratch, (int64_t)(n * CUT_NODE_BYTES), err,
                         (int64_t*)nullptr);
    HCL();
    int32_t code = 0;
    HC(hipMemcpyAsync(&code, err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HC(hipMemcpyAsync(out_mask, keep, g.tilevox, hipMemcpyDeviceToHost, s));
    HC(hipStreamSynchronize(s));
    check_kernel_error(code);
  });
}

int iq_plan_create(iq_ctx* ctx, const iq_plan_desc* desc, iq_plan** out) {
  if (!ctx) return fail(nullptr, IQ_ERR_ARG, "null context");
  if (!out) return fail(ctx, IQ_ERR_ARG, "null output handle");
  iq_plan* pl = new (std::nothrow) iq_plan();
  if (!pl) return fail(ctx, IQ_ERR_NOMEM, "host allocation failed");
  int rc = guard(ctx, [&] { plan_build(ctx, desc, pl); });
  if (rc != IQ_OK) { iq_plan_destroy(pl); return rc; }
  *out = pl;
  return IQ_OK;
}

int iq_plan_destroy(iq_plan* plan) {
  if (!plan) return IQ_OK;
  if (plan->ctx) (void)hipSetDevice(plan->ctx->device);
  for (hipEvent_t e : plan->prof_ev) (void)hipEventDestroy(e);
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
      for (int64_t i = 0; i < plan->nreal * plan->npath; i++)
        if (!(uniforms[i] >= 0.0 && uniforms[i] < 1.0)) bad(IQ_ERR_ARG, "uniforms must lie in [0,1)");
    if (plan->dtype == IQ_F32) plan_run<float>(plan, uniforms, forced);
    else plan_run<double>(plan, uniforms, forced);
  });
}

int iq_plan_results(iq_ctx* ctx, iq_plan* plan, void* out, int64_t* chosen, double* cutgrid) {
  if (!ctx || !plan) return fail(ctx, IQ_ERR_ARG, "null context or plan");
  return guard(ctx, [&] {
    hipStream_t s = ctx->stream;
    const IqGeo& g = plan->geo;
    const size_t esz = plan->dtype == IQ_F32 ? 4 : 8;
    if (out) HC(hipMemcpyAsync(out, plan->sims, (size_t)plan->nreal * g.padvox * esz, hipMemcpyDeviceToHost, s));
    if (chosen && plan->npath)
      HC(hipMemcpyAsync(chosen, plan->log, (size_t)plan->nreal * plan->npath * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    std::vector<uint8_t> c;
    if (cutgrid) {
      c.resize((size_t)plan->nreal * g.padvox);
      HC(hipMemcpyAsync(c.data(), plan->cuts, c.size(), hipMemcpyDeviceToHost, s));
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

int iq_plan_scan_stats(iq_plan* plan, int64_t* launches, double* seconds, double* flop) {
  if (!plan) return fail(nullptr, IQ_ERR_ARG, "null plan");
  if (launches) *launches = plan->prof_launches;
  if (seconds) *seconds = plan->prof_sec;
  if (flop) *flop = plan->prof_flop;
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
