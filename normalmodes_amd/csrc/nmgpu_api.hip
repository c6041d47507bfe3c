// C-ABI of the nmgpu library and the device-resident driver of the
// filtered-Lanczos path (LanBounds, FindPol, ChebLanNr, results, residuals).
//
// The call sequence mirrors what NormalModes does through pEVSL:
//   setupmatvec  (src/mod_matvec.f90:24-250):   nmg_parcsr_create, nmg_csr_diagscale,
//                nmg_start/nmg_set_amv/nmg_lanbounds on B~, nmg_setup_chebiter
//   pnm_apply_pevsl (src/mod_pevsl.f90:54-162): nmg_start, nmg_setprobsizes,
//                nmg_set_bmv, nmg_set_bsol_chebiter, nmg_set_amv_dad, nmg_set_geneig,
//                nmg_lanbounds, nmg_findpol, nmg_cheblannr, nmg_get_nev,
//                nmg_copy_result, nmg_residuals, nmg_freepol, nmg_finish
// Everything between upload and copy-out stays in HBM; the only host work per
// Lanczos step is the periodic tridiagonal eigenvalue check.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <unordered_set>
#include <vector>

#include "nmgpu_internal.h"

namespace nmg {

static thread_local std::string g_err;
void set_error(int code, const std::string& msg) {
  (void)code;
  g_err = msg;
}

// ---------------------------------------------------------------------------
// handle registry
enum Kind { K_SOLVER = 1, K_MAT = 2, K_CHEB = 3, K_POL = 4 };
static std::mutex g_mu;
static std::unordered_set<const void*> g_live[5];

template <class T>
static nmg_handle reg(T* p, int kind) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_live[kind].insert(p);
  return (nmg_handle)(intptr_t)p;
}
template <class T>
static T* get(nmg_handle h, int kind) {
  std::lock_guard<std::mutex> lk(g_mu);
  const void* p = (const void*)(intptr_t)h;
  if (!p || !g_live[kind].count(p)) throw Error{NMG_EINVAL, "invalid handle"};
  return (T*)p;
}
static void unreg(const void* p, int kind) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_live[kind].erase(p);
}

template <class F>
static int guarded(F&& f) {
  // library calls (rocBLAS/rocSOLVER set-up) can leave a sticky HIP error
  // behind; start every entry from a clean state so launch checks are exact
  (void)hipGetLastError();
  try {
    f();
    return NMG_OK;
  } catch (const Error& e) {
    set_error(e.code, e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error(NMG_ENOMEM, "host allocation failed");
    return NMG_ENOMEM;
  } catch (...) {
    set_error(NMG_EINVAL, "unknown failure");
    return NMG_EINVAL;
  }
}

template <class T>
static T* dmalloc(size_t count) {
  void* p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&p, sizeof(T) * count);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error{NMG_ENOMEM, "hipMalloc of " + std::to_string(sizeof(T) * count) + " bytes failed"};
  }
  return (T*)p;
}
template <class T>
static void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

static hipStream_t default_stream() {
  static hipStream_t s = nullptr;
  static std::once_flag once;
  std::call_once(once, [] {
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
  });
  NMG_CHECK(s != nullptr, NMG_EHIP, "cannot create HIP stream");
  return s;
}

// ---------------------------------------------------------------------------
// matrices
Tiling::~Tiling() {
  dfree(trp);
  dfree(tcp);
  dfree(tcols);
  dfree(lcol);
}

RowGroups::~RowGroups() {
  dfree(grow);
  dfree(gk);
}

void DevMat::release() {
  for (auto& t : tiling) {
    delete t;
    t = nullptr;
  }
  delete rgroups;
  rgroups = nullptr;
  dfree(rowptr);
  dfree(col);
  dfree(val);
  dfree(val32);
  dfree(browptr);
  dfree(bcol);
  dfree(bval);
  dfree(bval32);
}
DevMat::~DevMat() { release(); }

double DevMat::matrix_bytes(bool fp32) const {
  const double sv = fp32 ? 4.0 : 8.0;
  if (layout == L_CSR) return (double)nnz * (sv + 4.0) + (double)(nrow + 1) * 8.0;
  const double vals = layout == L_BSR3 ? 9.0 : 1.0;
  return (double)nbnz * (vals * sv + 4.0) + (double)(nb + 1) * 8.0;
}

double DevMat::algorithmic_bytes(bool fp32) const {
  return matrix_bytes(fp32) + (double)ncol * 8.0 + (double)nrow * 8.0;
}

// Layout detection on the host copy.
static int detect_layout(int64_t nrow, int64_t ncol, const int64_t* rp, const int32_t* ci,
                         const double* va) {
  if (nrow % 3 != 0 || ncol % 3 != 0 || nrow == 0) return L_CSR;
  const int64_t nb = nrow / 3;
  bool bsr = true, mass = true;
  for (int64_t v = 0; v < nb && (bsr || mass); ++v) {
    const int64_t r0 = rp[3 * v], l0 = rp[3 * v + 1] - r0;
    for (int p = 1; p < 3; ++p)
      if (rp[3 * v + p + 1] - rp[3 * v + p] != l0) bsr = mass = false;
    if (!bsr && !mass) break;
    if (bsr) {
      if (l0 % 3 != 0) bsr = false;
      for (int p = 0; p < 3 && bsr; ++p) {
        const int64_t rs = rp[3 * v + p];
        for (int64_t j = 0; j < l0 && bsr; j += 3) {
          const int32_t c0 = ci[rs + j];
          if (c0 % 3 != 0 || ci[rs + j + 1] != c0 + 1 || ci[rs + j + 2] != c0 + 2) bsr = false;
          if (ci[rs + j] != ci[r0 + j]) bsr = false;
          if (j > 0 && ci[rs + j] <= ci[rs + j - 1]) bsr = false;
        }
      }
    }
    if (mass) {
      for (int p = 0; p < 3 && mass; ++p) {
        const int64_t rs = rp[3 * v + p];
        for (int64_t j = 0; j < l0 && mass; ++j) {
          if ((ci[rs + j] % 3) != p) mass = false;
          else if (ci[rs + j] / 3 != ci[r0 + j] / 3) mass = false;
          else if (va[rs + j] != va[r0 + j]) mass = false;
        }
      }
    }
  }
  if (bsr) return L_BSR3;
  if (mass) return L_MASS3;
  return L_CSR;
}

DevMat* make_devmat(int64_t nrow, int64_t ncol, const int64_t* rp, const int32_t* ci,
                    const double* va, bool allow_blocked) {
  NMG_CHECK(nrow > 0 && ncol > 0, NMG_EINVAL, "empty matrix");
  NMG_CHECK(rp[0] == 0, NMG_EINVAL, "row offsets must start at 0");
  for (int64_t i = 0; i < nrow; ++i)
    NMG_CHECK(rp[i + 1] >= rp[i], NMG_EINVAL, "row offsets must be non-decreasing");
  const int64_t nnz = rp[nrow];
  for (int64_t k = 0; k < nnz; ++k)
    NMG_CHECK(ci[k] >= 0 && ci[k] < ncol, NMG_EINVAL, "column index out of range");
  std::unique_ptr<DevMat> M(new DevMat());
  M->nrow = nrow;
  M->ncol = ncol;
  M->nnz = nnz;
  M->layout = allow_blocked ? detect_layout(nrow, ncol, rp, ci, va) : L_CSR;
  if (M->layout == L_CSR) {
    M->rowptr = dmalloc<int64_t>(nrow + 1);
    M->col = dmalloc<int32_t>(nnz);
    M->val = dmalloc<double>(nnz);
    NMG_HIP(hipMemcpy(M->rowptr, rp, sizeof(int64_t) * (nrow + 1), hipMemcpyHostToDevice));
    NMG_HIP(hipMemcpy(M->col, ci, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
    NMG_HIP(hipMemcpy(M->val, va, sizeof(double) * nnz, hipMemcpyHostToDevice));
    return M.release();
  }
  const int64_t nb = nrow / 3;
  const bool bsr = M->layout == L_BSR3;
  std::vector<int64_t> brp(nb + 1, 0);
  for (int64_t v = 0; v < nb; ++v) {
    const int64_t l0 = rp[3 * v + 1] - rp[3 * v];
    brp[v + 1] = brp[v] + (bsr ? l0 / 3 : l0);
  }
  const int64_t nbnz = brp[nb];
  const int planes = bsr ? 9 : 1;
  std::vector<int32_t> bc(nbnz);
  std::vector<double> bv((size_t)planes * nbnz);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < nb; ++v) {
    for (int64_t j = 0; j < brp[v + 1] - brp[v]; ++j) {
      const int64_t k = brp[v] + j;
      if (bsr) {
        bc[k] = ci[rp[3 * v] + 3 * j] / 3;
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) bv[(size_t)(3 * r + c) * nbnz + k] = va[rp[3 * v + r] + 3 * j + c];
      } else {
        bc[k] = ci[rp[3 * v] + j] / 3;
        bv[k] = va[rp[3 * v] + j];
      }
    }
  }
  M->nb = nb;
  M->nbnz = nbnz;
  M->h_browptr = brp;
  M->h_bcol = bc;
  M->browptr = dmalloc<int64_t>(nb + 1);
  M->bcol = dmalloc<int32_t>(nbnz);
  M->bval = dmalloc<double>((size_t)planes * nbnz);
  NMG_HIP(hipMemcpy(M->browptr, brp.data(), sizeof(int64_t) * (nb + 1), hipMemcpyHostToDevice));
  NMG_HIP(hipMemcpy(M->bcol, bc.data(), sizeof(int32_t) * nbnz, hipMemcpyHostToDevice));
  NMG_HIP(hipMemcpy(M->bval, bv.data(), sizeof(double) * planes * nbnz, hipMemcpyHostToDevice));
  return M.release();
}

// Greedy row tiling on the host: consecutive node rows are added to the
// current tile while the number of distinct referenced nodes stays within
// max_nodes.  Outputs tile row offsets, node-list offsets, the sorted node
// lists and each entry's index within its tile's list.
struct HostTiling {
  std::vector<int64_t> trp{0}, tcp{0};
  std::vector<int32_t> tcols;
  std::vector<uint16_t> lcol;
  int max_nodes = 0;
};
static void compute_tiling(int64_t nb, const int64_t* rp, const int32_t* ci, int max_nodes, HostTiling& H) {
  NMG_CHECK(max_nodes > 0 && max_nodes <= 65535, NMG_EINVAL, "tile node list must fit 16-bit local indices");
  const int64_t nnz = rp[nb];
  std::vector<int64_t> stamp(nb, -1);
  std::vector<int32_t> local(nb, 0), cur;
  H.lcol.assign((size_t)nnz, 0);
  int64_t t = 0, start = 0, row = 0;
  auto close_tile = [&]() {
    std::sort(cur.begin(), cur.end());
    for (size_t i = 0; i < cur.size(); ++i) local[cur[i]] = (int32_t)i;
    for (int64_t r = start; r < row; ++r)
      for (int64_t k = rp[r]; k < rp[r + 1]; ++k) H.lcol[k] = (uint16_t)local[ci[k]];
    H.tcols.insert(H.tcols.end(), cur.begin(), cur.end());
    H.max_nodes = std::max(H.max_nodes, (int)cur.size());
    H.trp.push_back(row);
    H.tcp.push_back((int64_t)H.tcols.size());
    cur.clear();
    start = row;
    t++;
  };
  while (row < nb) {
    int add = 0;
    for (int64_t k = rp[row]; k < rp[row + 1]; ++k) {
      NMG_CHECK(ci[k] >= 0 && ci[k] < nb, NMG_EINVAL, "column out of range");
      if (stamp[ci[k]] != t) add++;
    }
    if ((int)cur.size() + add > max_nodes && row > start) {
      close_tile();
      continue;
    }
    NMG_CHECK(add <= max_nodes, NMG_EINVAL, "a row references more nodes than one tile holds");
    for (int64_t k = rp[row]; k < rp[row + 1]; ++k)
      if (stamp[ci[k]] != t) {
        stamp[ci[k]] = t;
        cur.push_back(ci[k]);
      }
    row++;
  }
  if (row > start) close_tile();
}

void build_row_groups(DevMat& M, int per_wave) {
  M.touch();
  delete M.rgroups;
  M.rgroups = nullptr;
  if (per_wave <= 0 || M.layout == L_CSR) return;
  NMG_CHECK(per_wave == 1 || per_wave == 2 || per_wave == 4 || per_wave == 8, NMG_EINVAL,
            "groups per wavefront must be 1, 2, 4 or 8");
  NMG_CHECK((int64_t)M.h_browptr.size() == M.nb + 1, NMG_ESTATE, "no host pattern for row groups");
  const int64_t* rp = M.h_browptr.data();
  std::vector<int64_t> gr, gk;
  int64_t r = 0;
  while (r < M.nb) {
    gr.push_back(r);
    gk.push_back(rp[r]);
    if (rp[r + 1] - rp[r] > 64) {  // a long row is a group of its own
      ++r;
      continue;
    }
    const int64_t start = r;
    while (r < M.nb && r - start < 64 && rp[r + 1] - rp[start] <= 64) ++r;
  }
  gr.push_back(M.nb);
  gk.push_back(rp[M.nb]);
  std::unique_ptr<RowGroups> G(new RowGroups());
  G->ngroups = (int64_t)gr.size() - 1;
  G->per_wave = per_wave;
  G->grow = dmalloc<int64_t>(gr.size());
  G->gk = dmalloc<int64_t>(gk.size());
  NMG_HIP(hipMemcpy(G->grow, gr.data(), sizeof(int64_t) * gr.size(), hipMemcpyHostToDevice));
  NMG_HIP(hipMemcpy(G->gk, gk.data(), sizeof(int64_t) * gk.size(), hipMemcpyHostToDevice));
  M.rgroups = G.release();
}

void build_tiling(DevMat& M, int K, int max_nodes, int rows_per_wave, int block, int wide) {
  M.touch();
  int idx = -1;
  for (int i = 0; i < 7; ++i)
    if ((1 << i) == K) idx = i;
  NMG_CHECK(idx >= 0, NMG_EINVAL, "K must be 1, 2, 4, 8, 16, 32 or 64");
  delete M.tiling[idx];
  M.tiling[idx] = nullptr;
  NMG_CHECK(max_nodes <= 0 || K <= 16, NMG_EINVAL, "LDS tiles are built for K <= 16");
  M.tiling[idx] = nullptr;
  if (max_nodes <= 0 || M.layout == L_CSR) return;
  NMG_CHECK((int64_t)M.h_browptr.size() == M.nb + 1, NMG_ESTATE, "no host pattern for tiling");
  HostTiling H;
  compute_tiling(M.nb, M.h_browptr.data(), M.h_bcol.data(), max_nodes, H);
  std::unique_ptr<Tiling> T(new Tiling());
  T->ntiles = (int64_t)H.trp.size() - 1;
  T->max_nodes = H.max_nodes;
  T->cap = max_nodes;
  T->rows_per_wave = rows_per_wave > 0 ? rows_per_wave : 1;
  NMG_CHECK(block == 256 || block == 1024, NMG_EINVAL, "tile workgroups have 256 or 1024 threads");
  T->block = block;
  NMG_CHECK(wide == 0 || ((wide == 4 || wide == 8 || wide == 16) && K >= 2 && M.layout == L_MASS3 && block == 1024),
            NMG_EINVAL, "tiled wide rows: 4, 8 or 16 lanes per row, K >= 2, scalar mass, 1024 threads");
  T->wide = wide;
  T->staged_nodes = (double)H.tcols.size();
  T->trp = dmalloc<int64_t>(H.trp.size());
  T->tcp = dmalloc<int64_t>(H.tcp.size());
  T->tcols = dmalloc<int32_t>(H.tcols.size());
  T->lcol = dmalloc<uint16_t>(H.lcol.size());
  NMG_HIP(hipMemcpy(T->trp, H.trp.data(), sizeof(int64_t) * H.trp.size(), hipMemcpyHostToDevice));
  NMG_HIP(hipMemcpy(T->tcp, H.tcp.data(), sizeof(int64_t) * H.tcp.size(), hipMemcpyHostToDevice));
  NMG_HIP(hipMemcpy(T->tcols, H.tcols.data(), sizeof(int32_t) * H.tcols.size(), hipMemcpyHostToDevice));
  NMG_HIP(hipMemcpy(T->lcol, H.lcol.data(), sizeof(uint16_t) * H.lcol.size(), hipMemcpyHostToDevice));
  M.tiling[idx] = T.release();
}

DevMat* clone_devmat(const DevMat& S, hipStream_t s) {
  std::unique_ptr<DevMat> M(new DevMat());
  M->nrow = S.nrow;
  M->ncol = S.ncol;
  M->nnz = S.nnz;
  M->layout = S.layout;
  M->nb = S.nb;
  M->nbnz = S.nbnz;
  M->h_browptr = S.h_browptr;
  M->h_bcol = S.h_bcol;
  M->group = S.group;
  for (int i = 0; i < 7; ++i) M->block_map[i] = S.block_map[i];
  for (int i = 0; i < 7; ++i) M->wide[i] = S.wide[i];
  auto cp = [&](auto* src, auto*& dst, size_t count) {
    using T = std::remove_pointer_t<std::remove_reference_t<decltype(dst)>>;
    if (!src) return;
    dst = dmalloc<T>(count);
    NMG_HIP(hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyDeviceToDevice, s));
  };
  if (S.layout == L_CSR) {
    cp(S.rowptr, M->rowptr, S.nrow + 1);
    cp(S.col, M->col, S.nnz);
    cp(S.val, M->val, S.nnz);
  } else {
    cp(S.browptr, M->browptr, S.nb + 1);
    cp(S.bcol, M->bcol, S.nbnz);
    cp(S.bval, M->bval, S.value_count());
  }
  M->single_blocked = S.single_blocked;
  M->pipe = S.pipe;
  M->xcd = S.xcd;
  M->nt = S.nt;
  M->tiled_vals = S.tiled_vals;
  for (int i = 0; i < 5; ++i)
    if (S.tiling[i])
      build_tiling(*M, 1 << i, S.tiling[i]->cap, S.tiling[i]->rows_per_wave, S.tiling[i]->block, S.tiling[i]->wide);
  if (S.rgroups) build_row_groups(*M, S.rgroups->per_wave);
  return M.release();
}

// ---------------------------------------------------------------------------
struct DVec {
  double* p = nullptr;
  int64_t n = 0;
  void ensure(int64_t m) {
    if (m <= n) return;
    dfree(p);
    p = dmalloc<double>(m);
    n = m;
  }
  ~DVec() { dfree(p); }
};
struct FVec {
  float* p = nullptr;
  int64_t n = 0;
  void ensure(int64_t m) {
    if (m <= n) return;
    dfree(p);
    p = dmalloc<float>(m);
    n = m;
  }
  ~FVec() { dfree(p); }
};

struct Cheb {
  DevMat* mat = nullptr;
  double lmin = 0, lmax = 0;
  int deg = 0;
  int64_t ncalls = 0, nmatvec = 0;
  double seconds = 0;
  // 0 (default): one fused launch per step; 1: one persistent launch per
  // solve where planned (scalar mass x I3 whose rows fit the LDS budget).
  // Measured on MI355X the grid barrier between steps costs 19 us (69
  // workgroups) to 62 us (256 workgroups), against ~5.6 us for a whole
  // fused step launch at 52k DoF, so the per-step launches are the default.
  int mode = 0;
  // mixed-precision refinement (0: off): `refine` passes, each an inner
  // Chebyshev solve of degree inner_deg with FP32 values and FP32 iterates
  // on the FP64 residual b - B~x of the sum so far, added in FP64
  int refine = 0, inner_deg = 0;
  FVec own32;           // inner work (3 n K floats) when the caller brings none
  PersistPlan plan[2];  // FP64 / FP32 values
  bool planned[2] = {false, false};
  const PersistPlan* persistent(bool fp32) {
    if (mode != 1) return nullptr;
    const int i = fp32 ? 1 : 0;
    if (!planned[i]) {
      plan_cheb_persistent(*mat, fp32, plan[i]);
      planned[i] = true;
    }
    return plan[i].ok ? &plan[i] : nullptr;
  }
};

// Sampled kernel timing on the solver's stream: every `every`-th launch of a
// kernel class is bracketed by two HIP events; the mean of the sampled
// durations is reported with the algorithmic bytes of one launch.
struct KernelProbe {
  int every = 0;
  int64_t count = 0;
  double bytes = 0.0;  // algorithmic bytes of one launch
  int launches = 1;    // launches inside one sampled bracket
  std::vector<hipEvent_t> e0, e1;
  size_t used = 0;
  bool armed(void) { return every > 0 && (count++ % every) == 0 && used < 4096; }
  void before(hipStream_t s) {
    if (used >= e0.size()) {
      hipEvent_t a, b;
      NMG_HIP(hipEventCreate(&a));
      NMG_HIP(hipEventCreate(&b));
      e0.push_back(a);
      e1.push_back(b);
    }
    NMG_HIP(hipEventRecord(e0[used], s));
  }
  void after(hipStream_t s) {
    NMG_HIP(hipEventRecord(e1[used], s));
    used++;
  }
  void reset() {
    count = 0;
    used = 0;
  }
  double mean_seconds(int64_t* n) {
    double tot = 0.0;
    int64_t k = 0;
    for (size_t i = 0; i < used; ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, e0[i], e1[i]) == hipSuccess) {
        tot += ms * 1e-3;
        k++;
      }
    }
    if (n) *n = k;
    return k ? tot / k : 0.0;
  }
  ~KernelProbe() {
    for (auto e : e0) (void)hipEventDestroy(e);
    for (auto e : e1) (void)hipEventDestroy(e);
  }
};


// Fluid-solid operator of sparsefsAV (src/mod_matvec.f90:498-520):
//   w = D [Ad (D v) + E (Dp C_p (Dp ET (D v)))],  C_p ~ Ap~^-1 by Chebyshev
// iteration, Ap~ = Dp (-Ap) Dp (Apdiagscaling, src/mod_matvec.f90:345-441);
// i.e. D (Ad - E Ap^-1 ET) D, the displacement Schur complement.
struct FsWork {
  DVec t[3], p[3];  // n*K and np*K
  DVec aw;          // operator output of a split filter chain (n*K)
  FVec f32;         // inner FP32 work of a refined Ap~ solve (3 np*K)
};
struct FsOp {
  DevMat* Ad = nullptr;  // n x n
  DevMat* E = nullptr;   // n x np
  DevMat* ET = nullptr;  // np x n
  Cheb* cp = nullptr;    // Chebyshev solve with Ap~ (np x np)
  int64_t n = 0, np = 0;
  DVec d, dp;
  FsWork part[4];  // per concurrent chain (part 0: the solver's stream)
};

struct Solver {
  int device = 0;
  hipStream_t s = nullptr;
  rocblas_handle blas = nullptr;
  int64_t n = 0;
  DevMat* amat = nullptr;            // plain operator (nmg_set_amv)
  std::unique_ptr<DevMat> adad;      // owned D A D (nmg_set_amv_dad)
  DevMat* bmat = nullptr;
  Cheb* bsol = nullptr;
  bool geneig = false;
  int precision = NMG_FP64;
  uint64_t seed = 7;
  int ntest = 30, cycle = 20;
  double rtol = 1e-9;
  DVec dscale;  // D of the D A D operator (for physical eigenvectors)
  double resid_target = 5e-11;  // stop when every extracted pair meets this (0 = trace test only)
  int block = 1;     // Lanczos block width (1 = the reference's single-vector ChebLanNr)
  DVec bw[6];        // interleaved n*block work blocks
  // work vectors
  DVec w[8];
  DVec red;  // reductions
  DVec ab;   // alpha / beta device scalars
  double* V = nullptr;
  double* Z = nullptr;
  int cap = 0;
  // results
  int nev = 0;
  std::vector<double> lam;
  double* X = nullptr;
  // stats
  KernelProbe probe[2];  // 0: fused Chebyshev B-solve step, 1: fused filter step
  int lan_steps = 0;
  int64_t filter_applies = 0;
  int ncand = 0;
  double seconds = 0;
  int target_met = 1;       // last cheblannr: every pair met resid_target (1), or it stopped on stagnation (0)
  double resid_best = 0.0;  // max ||A~x - lam B~x||/|lam| of the returned pairs at the residual stop (0: not checked)
  DVec mvx, mvy;            // staging of the MVFunc-shaped callbacks (nmg_amv / nmg_bmv / nmg_bsol)
  // dense products: 2 (default) MFMA for A B and for A^T B with K >= 16 (the
  // FMA split reduction for narrower A^T B, where a 16x16 MFMA tile would be
  // mostly empty); 1 MFMA for all; 0 rocBLAS DGEMM + the FMA split reduction
  int mfma = 2;

  std::shared_ptr<FsOp> fs;  // fluid-solid Schur operator in place of A (nmg_set_amv_fs)
  DVec aw;                   // product A~ in of the unfused filter step (n*K)
  DVec gw;                   // partial sums of the tall-skinny products
  bool borrowed = false;  // stream and rocBLAS handle belong to a parent solver
  // k-column B-solves split into `split` parts on as many streams (default
  // priority): the handle's stream and split-1 more
  int split = 0;
  bool split_filter = true;  // with split: whole filter chains per part (nmg_set_split_streams mode)
  hipStream_t sx[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  DVec hb[12];
  FVec hf[4];  // FP32 work of the k-column B-solve parts (3 blocks each)
  DVec pf[4][7];  // per-part filter work of the split block filter
  std::vector<int> slice_nev;  // pairs per slice after nmg_solve_slices
  int verbose = 0;  // progress line on stderr every `verbose` Lanczos steps (0: silent)
  // B-solves replayed from captured HIP graphs (nmg_set_graphs): one
  // executable graph per distinct set of buffers, built on its second use
  bool graphs = false;
  struct GraphKey {
    const void* p[24];
    uint64_t gen;  // layout generation of the matrix
    int K, parts, fp32, deg, mode;
    double lmin, lmax;
  };
  struct GraphEntry {
    GraphKey key;
    int uses;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> gcache;
  void drop_graphs() {
    // a launched graph may still run on the stream: let it finish first
    if (s && !gcache.empty()) (void)hipStreamSynchronize(s);
    for (auto& e : gcache)
      if (e.exec) (void)hipGraphExecDestroy(e.exec);
    gcache.clear();
  }

  ~Solver() {
    free_basis();
    dfree(X);
    drop_graphs();
    if (borrowed) return;
    for (auto x : sx)
      if (x) (void)hipStreamDestroy(x);
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
    if (blas) rocblas_destroy_handle(blas);
    if (s) (void)hipStreamDestroy(s);
  }
  void free_basis() {
    // Z aliases V for a standard problem: free it only when it is separate
    if (Z && Z != V) dfree(Z);
    Z = nullptr;
    dfree(V);
    cap = 0;
  }
  const DevMat& A() const {
    if (adad) return *adad;
    NMG_CHECK(amat != nullptr, NMG_ESTATE, "no A operator set");
    return *amat;
  }
  void check_A() const {
    if (!fs) (void)A();
  }
  bool fp32() const { return precision != NMG_FP64; }                // FP32 values in the B-solve
  bool fp32v() const { return precision == NMG_FP32_BSOLVE; }        // and FP32 vectors (k-column solves)
  bool fp32a() const { return precision == NMG_FP32_FILTER_A; }      // and FP32 stiffness values in the filter
};

static void apply_A_block(Solver& S, int K, const double* x, double* y);
static void apply_fs(FsOp& F, int K, const double* x, double* y, hipStream_t s, FsWork& W, bool f32);
static void apply_A(Solver& S, const double* x, double* y);

// grow the Lanczos basis to hold `need` columns (copying the first `keep`)
static void ensure_basis(Solver& S, int need, int keep, int maxcap) {
  if (need <= S.cap) return;
  int nc = std::max(need, std::min(maxcap, std::max(64, S.cap * 2)));
  double* V = dmalloc<double>((size_t)S.n * nc);
  double* Z = S.geneig ? dmalloc<double>((size_t)S.n * nc) : V;
  if (keep > 0) {
    NMG_HIP(hipMemcpyAsync(V, S.V, sizeof(double) * S.n * keep, hipMemcpyDeviceToDevice, S.s));
    if (S.geneig)
      NMG_HIP(hipMemcpyAsync(Z, S.Z, sizeof(double) * S.n * keep, hipMemcpyDeviceToDevice, S.s));
    NMG_HIP(hipStreamSynchronize(S.s));
  }
  S.free_basis();
  S.V = V;
  S.Z = Z;
  S.cap = nc;
}

// Runs `body` (launches on S.s, possibly forked to S.sx and joined back) as
// a captured HIP graph when graphs are on: the first use of a key runs
// eagerly (so lazily created device data exists before any capture), the
// second captures, instantiates and launches, later ones only launch.  The
// body issues kernels and event waits only: no allocation, no memset, no
// synchronisation.
template <class F>
static void run_graphed(Solver& S, const Solver::GraphKey& key, F&& body) {
  if (!S.graphs) {
    body();
    return;
  }
  Solver::GraphEntry* hit = nullptr;
  for (auto& e : S.gcache)
    if (std::memcmp(&e.key, &key, sizeof(key)) == 0) {
      hit = &e;
      break;
    }
  if (hit && hit->exec) {
    NMG_HIP(hipGraphLaunch(hit->exec, S.s));
    return;
  }
  if (!hit) {
    if (S.gcache.size() >= 16) {
      if (S.gcache.front().exec) {
        NMG_HIP(hipStreamSynchronize(S.s));  // its last launch may still run
        (void)hipGraphExecDestroy(S.gcache.front().exec);
      }
      S.gcache.erase(S.gcache.begin());
    }
    S.gcache.push_back(Solver::GraphEntry{key, 1, nullptr});
    body();
    return;
  }
  hipGraph_t g = nullptr;
  NMG_HIP(hipStreamBeginCapture(S.s, hipStreamCaptureModeThreadLocal));
  try {
    body();
  } catch (...) {
    hipGraph_t junk = nullptr;
    (void)hipStreamEndCapture(S.s, &junk);
    if (junk) (void)hipGraphDestroy(junk);
    throw;
  }
  NMG_HIP(hipStreamEndCapture(S.s, &g));
  hipGraphExec_t ex = nullptr;
  const hipError_t rc = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  NMG_HIP(rc);
  hit->exec = ex;
  hit->uses++;
  NMG_HIP(hipGraphLaunch(ex, S.s));
}

static Solver::GraphKey graph_key(const Solver& S, const Cheb& C, int K, int parts, const void* b, const void* out,
                                  const void* tmp) {
  Solver::GraphKey k;
  std::memset(&k, 0, sizeof(k));
  k.p[0] = b;
  k.p[1] = out;
  k.p[2] = tmp;
  k.p[3] = &C;
  k.p[4] = C.mat;
  k.gen = C.mat->gen;
  k.mode = C.mode;
  for (int i = 0; i < 3 * parts && i < 12; ++i) k.p[5 + i] = S.hb[i].p;
  for (int i = 0; i < 4; ++i) k.p[17 + i] = S.hf[i].p;
  k.K = K;
  k.parts = parts;
  k.fp32 = S.fp32() ? 1 : 0;
  k.deg = C.deg;
  k.lmin = C.lmin;
  k.lmax = C.lmax;
  return k;
}

// x ~ B~^{-1} b by the Chebyshev semi-iteration (x_0 = 0, deg products):
//   x_1 = b/th,  x_{k+1} = x_{k-1} + w_{k+1}(x_k - x_{k-1} + (b - B x_k)/th),
//   w_1 = 2, w_{k+1} = 1/(1 - w_k/(4 s^2)), th = (lmax+lmin)/2, s = th/de.
// The result lands in `out`; `tmp` is a second buffer.
// The Chebyshev semi-iteration as one fused launch per step: x_1 = b/th,
// x_0 = 0 (cleared here), x_{k+1} from the fused product + update; the
// iterates alternate between out and tmp so that the last lands in out.
// With `probe`, the middle step's launch is bracketed by its events.
static void cheb_steps(const Cheb& C, const double* b, double* out, double* tmp, hipStream_t s, bool fp32,
                       KernelProbe* probe) {
  const int64_t n = C.mat->nrow;
  const double th = 0.5 * (C.lmax + C.lmin), de = 0.5 * (C.lmax - C.lmin), sg = th / de;
  double* xk = (C.deg % 2 == 0) ? out : tmp;
  double* xm = (C.deg % 2 == 0) ? tmp : out;
  k_scale_copy(n, 1.0 / th, b, xk, s);
  k_fill(n, 0.0, xm, s);
  const int mid = (C.deg + 1) / 2;
  double w = 2.0;
  for (int k = 1; k <= C.deg; ++k) {
    w = 1.0 / (1.0 - w / (4.0 * sg * sg));
    if (probe && k == mid) probe->before(s);
    k_cheb_step(*C.mat, b, xk, xm, w, 1.0 / th, s, fp32);
    if (probe && k == mid) probe->after(s);
    std::swap(xk, xm);
  }
}

// x ~ B~^-1 b on K interleaved columns by mixed-precision refinement:
//   x_1 = S(b),  x_p = x_{p-1} + S(b - B~ x_{p-1})
// S an inner Chebyshev solve of degree inner_deg on FP32 values and FP32
// iterates (FP64 row accumulation), the residual and the sum in FP64.  Each
// pass divides the error by max(2 rho^inner_deg, the FP32 floor ~1e-6).
static void refine_solve(Cheb& C, int K, const double* b, double* out, double* tmp, hipStream_t s,
                         KernelProbe* probe, float* w32) {
  const int64_t nk = C.mat->nrow * K;
  if (!w32) {
    C.own32.ensure(3 * nk);
    w32 = C.own32.p;
  }
  float* b32 = w32;
  float* fa = w32 + nk;
  float* fb = w32 + 2 * nk;
  const double th = 0.5 * (C.lmax + C.lmin), de = 0.5 * (C.lmax - C.lmin), sg = th / de;
  const bool pr = probe && probe->armed();
  for (int p = 0; p < C.refine; ++p) {
    const double* rhs = b;
    if (p > 0) {
      k_block_spmv(K, *C.mat, out, tmp, s, false);  // FP64 values
      k_sub_from(nk, b, tmp, s);
      rhs = tmp;
    }
    float* fk = fa;
    float* fm = fb;
    k_cheb32_start(nk, 1.0 / th, rhs, b32, fk, fm, s);
    const int mid = (C.inner_deg + 1) / 2;
    double w = 2.0;
    for (int k = 1; k <= C.inner_deg; ++k) {
      w = 1.0 / (1.0 - w / (4.0 * sg * sg));
      if (pr && p == 0 && k == mid) probe->before(s);
      k_block_cheb_step32(K, *C.mat, b32, fk, fm, w, 1.0 / th, s, true);
      if (pr && p == 0 && k == mid) probe->after(s);
      std::swap(fk, fm);
    }
    if (p == 0) k_f32_to_f64(nk, fk, out, s);
    else k_add_f32(nk, fk, out, s);
  }
  if (probe) {
    probe->bytes = C.mat->matrix_bytes(true) + 4.0 * 4.0 * (double)nk;
    probe->launches = 1;
  }
}

static void cheb_solve(Cheb& C, const double* b, double* out, double* tmp, hipStream_t s,
                       bool fp32, KernelProbe* probe = nullptr) {
  if (C.refine > 0) return refine_solve(C, 1, b, out, tmp, s, probe, nullptr);
  const int64_t n = C.mat->nrow;
  const double step_bytes = C.mat->matrix_bytes(fp32) + 4.0 * 8.0 * (double)n;
  if (const PersistPlan* P = C.persistent(fp32)) {
    const bool pr = probe && probe->armed();
    if (pr) probe->before(s);
    k_cheb_persistent(1, *C.mat, *P, b, out, tmp, C.deg, C.lmin, C.lmax, s);
    if (pr) probe->after(s);
    if (probe) {
      probe->bytes = step_bytes * C.deg;
      probe->launches = 1;
    }
    return;
  }
  // in a sampled solve one step launch (the middle one) is bracketed
  const bool pr = probe && probe->armed();
  cheb_steps(C, b, out, tmp, s, fp32, pr ? probe : nullptr);
  if (probe) {
    probe->bytes = step_bytes;
    probe->launches = 1;
  }
}

static void apply_B(Solver& S, const double* x, double* y) {
  if (S.bmat) k_spmv(*S.bmat, x, y, S.s);
  else k_copy(S.n, x, y, S.s);
}

static void bsolve(Solver& S, const double* b, double* out, double* tmp) {
  if (!S.geneig) {
    if (out != b) k_copy(S.n, b, out, S.s);
    return;
  }
  NMG_CHECK(S.bsol != nullptr, NMG_ESTATE, "generalized problem without a B-solve");
  Cheb& C = *S.bsol;
  if (S.graphs && !C.persistent(S.fp32())) {
    // the whole solve is one graph; a sampled solve is bracketed around it
    KernelProbe& pb = S.probe[0];
    const bool pr = pb.armed();
    if (pr) pb.before(S.s);
    run_graphed(S, graph_key(S, C, 1, 0, b, out, tmp), [&] { cheb_solve(C, b, out, tmp, S.s, S.fp32(), nullptr); });
    if (pr) pb.after(S.s);
    pb.bytes = C.mat->matrix_bytes(S.fp32()) + 4.0 * 8.0 * (double)C.mat->nrow;
    pb.launches = C.deg;
  } else {
    cheb_solve(C, b, out, tmp, S.s, S.fp32(), &S.probe[0]);
  }
  C.ncalls++;
  C.nmatvec += C.deg;
}

// y = p((A B^-1 - cc)/dd) z  (z-space; p((A - cc)/dd) z for a standard problem)
static void filter_apply(Solver& S, const PolyFilter& P, const double* z, double* y) {
  const int64_t n = S.n;
  for (int i = 0; i < 4; ++i) S.w[i].ensure(n);
  double* vk = S.w[0].p;
  double* vkm = S.w[1].p;
  double* b1 = S.w[2].p;
  double* b2 = S.w[3].p;
  k_copy(n, z, vk, S.s);
  k_scale_copy(n, P.mu[0], vk, y, S.s);
  if (S.fs) S.w[7].ensure(n);
  for (int k = 1; k <= P.deg; ++k) {
    const double t = (k == 1 ? 1.0 : 2.0) / P.dd;
    const double* in = vk;
    if (S.geneig) {
      bsolve(S, vk, b1, b2);
      in = b1;
    }
    if (S.fs) {
      apply_fs(*S.fs, 1, in, S.w[7].p, S.s, S.fs->part[0], S.fp32a());
      k_filter_update(n, 1, S.w[7].p, vk, vkm, y, nullptr, P.mu[k], t, P.cc, k == 1, S.s);
    } else {
      const bool pr = S.probe[1].armed();
      if (pr) S.probe[1].before(S.s);
      k_filter_step(S.A(), in, vk, vkm, y, t, P.cc, P.mu[k], k == 1, S.s, S.fp32a());
      if (pr) S.probe[1].after(S.s);
    }
    std::swap(vk, vkm);
  }
  S.filter_applies++;
}

static void setup_work(Solver& S) {
  for (int i = 0; i < 8; ++i) S.w[i].ensure(S.n);
  S.red.ensure((int64_t)std::max(dot_work_size(), (size_t)16));
}

// Lanczos step k, split so that several independent recurrences can share
// their sparse products: (1) op(.) into z_{k+1}; (2) three-term recurrence and
// CGS2 against (Z, V)(:, 0:k); (3) v_{k+1} = B^-1 z_{k+1}; (4) B-normalise.
static void lanczos_orth(Solver& S, int k, double* hwork) {
  const int64_t n = S.n;
  double* v = S.V + (size_t)k * n;
  double* z = S.Z + (size_t)k * n;
  double* zn = S.Z + (size_t)(k + 1) * n;
  double* alpha = S.ab.p + k;
  double* beta = S.ab.p + S.ab.n / 2 + k;
  if (k > 0) k_axpy_dev(n, beta - 1, S.Z + (size_t)(k - 1) * n, zn, S.s);
  k_dot(n, v, zn, alpha, false, S.red.p, S.s);
  k_axpy_dev(n, alpha, z, zn, S.s);
  // classical Gram-Schmidt, twice
  double* h = hwork;
  double* gw = hwork + (k + 2);
  for (int pass = 0; pass < 2; ++pass) {
    k_gemv_t(n, k + 1, S.V, n, zn, h, gw, S.s);
    k_gemv_n_sub(n, k + 1, S.Z, n, h, zn, S.s);
  }
}

static void lanczos_normalise(Solver& S, int k) {
  const int64_t n = S.n;
  double* vn = S.V + (size_t)(k + 1) * n;
  double* zn = S.Z + (size_t)(k + 1) * n;
  double* beta = S.ab.p + S.ab.n / 2 + k;
  if (S.geneig) {
    k_dot(n, vn, zn, beta, true, S.red.p, S.s);
    k_invscale_dev(n, beta, vn, zn, S.s);
  } else {
    k_dot(n, zn, zn, beta, true, S.red.p, S.s);
    k_invscale_dev(n, beta, zn, nullptr, S.s);
  }
}

// op: 0 = A v_k (bounds), 1 = filter of z_k
static void lanczos_step(Solver& S, int k, int op, const PolyFilter* P, double* hwork) {
  const int64_t n = S.n;
  double* zn = S.Z + (size_t)(k + 1) * n;
  if (op == 0) apply_A(S, S.V + (size_t)k * n, zn);
  else filter_apply(S, *P, S.Z + (size_t)k * n, zn);
  lanczos_orth(S, k, hwork);
  if (S.geneig) bsolve(S, zn, S.V + (size_t)(k + 1) * n, S.w[4].p);
  lanczos_normalise(S, k);
}

static void lanczos_start(Solver& S) {
  const int64_t n = S.n;
  std::vector<double> v0(n);
  start_vector(n, S.seed, v0.data());
  NMG_HIP(hipMemcpyAsync(S.V, v0.data(), sizeof(double) * n, hipMemcpyHostToDevice, S.s));
  double* nrm = S.ab.p + S.ab.n - 1;
  if (S.geneig) {
    apply_B(S, S.V, S.Z);
    k_dot(n, S.V, S.Z, nrm, true, S.red.p, S.s);
    k_invscale_dev(n, nrm, S.V, S.Z, S.s);
  } else {
    k_dot(n, S.V, S.V, nrm, true, S.red.p, S.s);
    k_invscale_dev(n, nrm, S.V, nullptr, S.s);
  }
  NMG_HIP(hipStreamSynchronize(S.s));
}

static void fetch_tridiag(Solver& S, int kdim, std::vector<double>& al, std::vector<double>& be) {
  al.resize(kdim);
  be.resize(kdim);
  NMG_HIP(hipMemcpyAsync(al.data(), S.ab.p, sizeof(double) * kdim, hipMemcpyDeviceToHost, S.s));
  NMG_HIP(hipMemcpyAsync(be.data(), S.ab.p + S.ab.n / 2, sizeof(double) * kdim, hipMemcpyDeviceToHost, S.s));
  NMG_HIP(hipStreamSynchronize(S.s));
  for (int i = 0; i < kdim; ++i)
    NMG_CHECK(std::isfinite(al[i]) && std::isfinite(be[i]), NMG_ENUM,
              "Lanczos breakdown: non-finite recurrence coefficient");
}

// rocBLAS / rocSOLVER calls: check the status and clear any HIP error state
// the library left behind, so later kernel-launch checks see only their own
#define NMG_LIB(call, what)                                                            \
  do {                                                                                 \
    rocblas_status st_ = (call);                                                       \
    (void)hipGetLastError();                                                           \
    NMG_CHECK(st_ == rocblas_status_success, NMG_EHIP,                                 \
              std::string(what) + " failed: " + rocblas_status_to_string(st_));        \
  } while (0)

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// progress line of a long solve (nmg_set_verbose)
static void progress(const Solver& S, const char* what, int step, double t0) {
  if (S.verbose <= 0 || step % S.verbose != 0) return;
  std::fprintf(stderr, "[nmg] %s: step %d, %.1f s\n", what, step, now() - t0);
  std::fflush(stderr);
}

static void do_lanbounds(Solver& S, int mlan, int maxit, double tol, double* lmin, double* lmax) {
  NMG_CHECK(S.n > 0, NMG_ESTATE, "problem size not set");
  const double t0 = now();
  int m = std::min<int64_t>(std::min(mlan, maxit), S.n);
  NMG_CHECK(m >= 2, NMG_EINVAL, "mlan/maxit too small");
  setup_work(S);
  S.ab.ensure(2 * (int64_t)(m + 2));
  S.free_basis();
  ensure_basis(S, std::min(m + 1, 64), 0, m + 1);
  std::vector<double> hbuf;
  DVec hw;
  lanczos_start(S);
  std::vector<double> al, be, w, sv;
  int k = 0;
  bool done = false;
  for (k = 0; k < m && !done; ++k) {
    ensure_basis(S, k + 2, k + 1, m + 1);
    hw.ensure((int64_t)gemv_work_size(S.n, k + 2) + k + 4);
    lanczos_step(S, k, 0, nullptr, hw.p);
    const int kd = k + 1;
    progress(S, "lanbounds", kd, t0);
    if (kd % 10 == 0 || kd == m) {
      fetch_tridiag(S, kd, al, be);
      if (tridiag_eigvals(kd, al.data(), be.data(), w) != NMG_OK)
        throw Error{NMG_ENUM, "tridiagonal eigenvalues did not converge"};
      tridiag_eigvec(kd, al.data(), be.data(), w[0], sv);
      const double rmin = std::fabs(be[kd - 1] * sv[kd - 1]);
      tridiag_eigvec(kd, al.data(), be.data(), w[kd - 1], sv);
      const double rmax = std::fabs(be[kd - 1] * sv[kd - 1]);
      *lmin = w[0] - rmin;
      *lmax = w[kd - 1] + rmax;
      const double spread = std::fabs(w[kd - 1]) + std::fabs(w[0]);
      if (be[kd - 1] == 0.0 || (rmin <= tol * spread && rmax <= tol * spread)) done = true;
    }
  }
  S.lan_steps = k;
  S.free_basis();
  S.seconds = now() - t0;
}

// C (m x kc) = A^T B for any kc, as tall-skinny products on column groups
// of B of 16, 8, 4, 2 and 1
static void gemm_tn_any(Solver& S, int64_t n, int m, int kc, const double* A, int64_t lda, const double* B,
                        int64_t ldb, double* C, int64_t ldc) {
  for (int c0 = 0; c0 < kc;) {
    int w = 16;
    while (w > kc - c0) w >>= 1;
    S.gw.ensure((int64_t)gemm_tn_work_size(n, m, w));
    k_gemm_tn(n, m, w, A, lda, B + (int64_t)c0 * ldb, ldb, C + (int64_t)c0 * ldc, ldc, S.gw.p, S.s);
    c0 += w;
  }
}

// Dense products of the block recurrence and the extraction, on the device:
// C (n x K) = alpha A (n x m) B (m x K) + beta C and C (m x K) = A^T B.
static void gemm(Solver& S, rocblas_operation ta, rocblas_operation tb, int64_t m, int64_t n, int64_t k,
                 double alpha, const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                 double* C, int64_t ldc, const char* what);
static void dense_nn(Solver& S, int64_t n, int K, int m, double alpha, const double* A, int64_t lda,
                     const double* B, int64_t ldb, double beta, double* C, int64_t ldc, const char* what) {
  if (S.mfma) k_gemm_nn_mfma(n, K, m, alpha, A, lda, B, ldb, beta, C, ldc, S.s);
  else gemm(S, rocblas_operation_none, rocblas_operation_none, n, K, m, alpha, A, lda, B, ldb, beta, C, ldc, what);
}
static void dense_tn(Solver& S, int64_t n, int m, int K, const double* A, int64_t lda, const double* B,
                     int64_t ldb, double* C, int64_t ldc) {
  if (S.mfma == 0 || (S.mfma == 2 && K < 16)) return gemm_tn_any(S, n, m, K, A, lda, B, ldb, C, ldc);
  for (int c0 = 0; c0 < K; c0 += 64) {
    const int w = std::min(64, K - c0);
    S.gw.ensure((int64_t)gemm_tn_mfma_work_size(n, m, w));
    k_gemm_tn_mfma(n, m, w, A, lda, B + (int64_t)c0 * ldb, ldb, C + (int64_t)c0 * ldc, ldc, S.gw.p, S.s);
  }
}

// Ritz vectors of the candidates (theta >= bar) from the Lanczos basis V(:, 0:kdim)
// and the eigenvectors dC (kdim x kdim, device) of the projected matrix, then
// Rayleigh-Ritz of (A~, B~) on their span; keeps the pairs inside xintv[0..1].
static void extract_ritz(Solver& S, int kdim, const double* dC, const std::vector<double>& theta, double bar,
                         const double* xintv) {
  const int64_t n = S.n;
  DVec dinfo_v;
  dinfo_v.ensure(1);
  rocblas_int* dinfo = (rocblas_int*)dinfo_v.p;
  rocblas_int info = 0;
  std::vector<int> sel;
  for (int i = 0; i < kdim; ++i)
    if (theta[i] >= bar) sel.push_back(i);
  const int nc = (int)sel.size();
  S.ncand = nc;
  S.nev = 0;
  S.lam.clear();
  dfree(S.X);
  if (nc > 0) {
    DVec Ssel, Y, AY, BY, GA, GB, mu, ee;
    Ssel.ensure((int64_t)kdim * nc);
    for (int j = 0; j < nc; ++j)
      NMG_HIP(hipMemcpyAsync(Ssel.p + (size_t)j * kdim, dC + (size_t)sel[j] * kdim,
                             sizeof(double) * kdim, hipMemcpyDeviceToDevice, S.s));
    Y.ensure(n * nc);
    dense_nn(S, n, nc, kdim, 1.0, S.V, n, Ssel.p, kdim, 0.0, Y.p, n, "rocblas_dgemm (Ritz vectors) failed");
    // Rayleigh-Ritz of (A~, B~) on span(Y)
    AY.ensure(n * nc);
    BY.ensure(n * nc);
    for (int j = 0; j < nc; ++j) {
      apply_A(S, Y.p + (size_t)j * n, AY.p + (size_t)j * n);
      apply_B(S, Y.p + (size_t)j * n, BY.p + (size_t)j * n);
    }
    GA.ensure((int64_t)nc * nc);
    GB.ensure((int64_t)nc * nc);
    dense_tn(S, n, nc, nc, Y.p, n, AY.p, n, GA.p, nc);  // Y^T A Y
    dense_tn(S, n, nc, nc, Y.p, n, BY.p, n, GB.p, nc);  // Y^T B Y
    // symmetrise (the products agree to rounding)
    {
      std::vector<double> ga((size_t)nc * nc), gb((size_t)nc * nc);
      NMG_HIP(hipMemcpyAsync(ga.data(), GA.p, sizeof(double) * nc * nc, hipMemcpyDeviceToHost, S.s));
      NMG_HIP(hipMemcpyAsync(gb.data(), GB.p, sizeof(double) * nc * nc, hipMemcpyDeviceToHost, S.s));
      NMG_HIP(hipStreamSynchronize(S.s));
      for (int i = 0; i < nc; ++i)
        for (int j = i + 1; j < nc; ++j) {
          double a = 0.5 * (ga[(size_t)i * nc + j] + ga[(size_t)j * nc + i]);
          double b = 0.5 * (gb[(size_t)i * nc + j] + gb[(size_t)j * nc + i]);
          ga[(size_t)i * nc + j] = ga[(size_t)j * nc + i] = a;
          gb[(size_t)i * nc + j] = gb[(size_t)j * nc + i] = b;
        }
      NMG_HIP(hipMemcpyAsync(GA.p, ga.data(), sizeof(double) * nc * nc, hipMemcpyHostToDevice, S.s));
      NMG_HIP(hipMemcpyAsync(GB.p, gb.data(), sizeof(double) * nc * nc, hipMemcpyHostToDevice, S.s));
    }
    mu.ensure(nc);
    ee.ensure(nc);
    NMG_LIB(rocsolver_dsygvd(S.blas, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, nc,
                               GA.p, nc, GB.p, nc, mu.p, ee.p, dinfo), "rocsolver_dsygvd failed");
    std::vector<double> rmu(nc);
    NMG_HIP(hipMemcpyAsync(rmu.data(), mu.p, sizeof(double) * nc, hipMemcpyDeviceToHost, S.s));
    NMG_HIP(hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, S.s));
    NMG_HIP(hipStreamSynchronize(S.s));
    NMG_CHECK(info == 0, NMG_ENUM, "Rayleigh-Ritz eigensolver failed (B~ Gram matrix not SPD)");
    std::vector<int> keep;
    for (int j = 0; j < nc; ++j)
      if (rmu[j] >= xintv[0] && rmu[j] <= xintv[1]) keep.push_back(j);
    const int nev = (int)keep.size();
    if (nev > 0) {
      DVec Wsel;
      Wsel.ensure((int64_t)nc * nev);
      for (int j = 0; j < nev; ++j)
        NMG_HIP(hipMemcpyAsync(Wsel.p + (size_t)j * nc, GA.p + (size_t)keep[j] * nc, sizeof(double) * nc,
                               hipMemcpyDeviceToDevice, S.s));
      S.X = dmalloc<double>((size_t)n * nev);
      dense_nn(S, n, nev, nc, 1.0, Y.p, n, Wsel.p, nc, 0.0, S.X, n, "rocblas_dgemm (eigenvectors) failed");
      for (int j = 0; j < nev; ++j) S.lam.push_back(rmu[keep[j]]);
      S.nev = nev;
    }
    NMG_HIP(hipStreamSynchronize(S.s));
  }
}


// eigenpairs of the Lanczos tridiagonal on the device (rocSOLVER divide & conquer)
static void tridiag_eig_device(Solver& S, int kdim, const std::vector<double>& al, const std::vector<double>& be,
                               DVec& dC, std::vector<double>& theta) {
  DVec dD, dE, dinfo_v;
  dD.ensure(kdim);
  dE.ensure(kdim);
  dC.ensure((int64_t)kdim * kdim);
  dinfo_v.ensure(1);
  NMG_HIP(hipMemcpyAsync(dD.p, al.data(), sizeof(double) * kdim, hipMemcpyHostToDevice, S.s));
  NMG_HIP(hipMemcpyAsync(dE.p, be.data(), sizeof(double) * kdim, hipMemcpyHostToDevice, S.s));
  NMG_LIB(rocsolver_dstedc(S.blas, rocblas_evect_tridiagonal, kdim, dD.p, dE.p, dC.p, kdim,
                           (rocblas_int*)dinfo_v.p), "rocsolver_dstedc failed");
  theta.resize(kdim);
  rocblas_int info = 0;
  NMG_HIP(hipMemcpyAsync(theta.data(), dD.p, sizeof(double) * kdim, hipMemcpyDeviceToHost, S.s));
  NMG_HIP(hipMemcpyAsync(&info, dinfo_v.p, sizeof(info), hipMemcpyDeviceToHost, S.s));
  NMG_HIP(hipStreamSynchronize(S.s));
  NMG_CHECK(info == 0, NMG_ENUM, "tridiagonal eigensolver failed");
}

// Residual-based stop: after the trace test has passed, extract the pairs and
// accept when max ||A~x - lam B~x||/|lam| <= resid_target.  The residual has
// a floor set by the B-solve accuracy; when it has not improved by 10% over
// the last max(60, kdim/4) Lanczos vectors the iteration stops, and the pairs
// are extracted again at the Lanczos dimension that gave the best residual
// (`reextract`).  When the converged candidates yield no pair inside the
// interval (they are neighbours whose filter value reaches bar), the
// iteration stops with no pairs, as ChebLanNr does.  S.target_met and
// S.resid_best record the outcome.  Returns true when the iteration should
// stop.
static void do_residuals(Solver& S, double* res_ref, double* res2);
struct ResidualTracker {
  double best = 1e300;
  int best_k = 0;
};
static bool residual_stop(Solver& S, int kdim, const double* dC, const std::vector<double>& theta, double bar,
                          const double* xintv, ResidualTracker& rt,
                          const std::function<void(int)>& reextract) {
  if (S.resid_target <= 0.0) return true;
  extract_ritz(S, kdim, dC, theta, bar, xintv);
  if (S.nev == 0) {
    S.target_met = 1;
    S.resid_best = 0.0;
    return true;
  }
  std::vector<double> r2(S.nev);
  do_residuals(S, nullptr, r2.data());
  double mx = 0.0;
  for (double v : r2) mx = std::max(mx, std::isfinite(v) ? v : 1e300);
  if (mx <= S.resid_target) {
    S.target_met = 1;
    S.resid_best = mx;
    return true;
  }
  if (mx < 0.9 * rt.best) {
    rt.best = mx;
    rt.best_k = kdim;
    return false;
  }
  if (kdim - rt.best_k < std::max(60, kdim / 4)) return false;
  // stagnated: the pairs of the best dimension, unless the latest are better
  if (mx > rt.best && rt.best_k > 0 && rt.best_k < kdim) {
    reextract(rt.best_k);
    S.resid_best = rt.best;
  } else {
    S.resid_best = mx;
  }
  S.target_met = 0;
  return true;
}

// ChebLanNr's test at Lanczos dimension kdim: on the reference cadence
// (ntest, then every cycle steps) the trace of the Ritz values >= bar must
// settle (|tr1 - tr0| < tol |tr1|), then every candidate's Ritz residual
// |beta_k s_k| must be <= rtol, then (with a residual target) the extracted
// pairs must meet it.  With no candidate, the stop comes when the largest
// Ritz value has converged below bar (an empty interval).  Returns 0 to go
// on, 1 to stop and extract, 2 when the stop already left the extracted
// pairs in S.
struct LanczosTest {
  double tr0 = 0.0;
  ResidualTracker rt;
  std::vector<double> al, be, w, sv;
};
static int lanczos_test(Solver& S, int kdim, const PolyFilter& P, const double* xintv, double tol, LanczosTest& L) {
  if (kdim < S.ntest || ((kdim - S.ntest) % S.cycle) != 0) return 0;
  fetch_tridiag(S, kdim, L.al, L.be);
  if (L.be[kdim - 1] == 0.0) return 1;
  if (tridiag_eigvals(kdim, L.al.data(), L.be.data(), L.w) != NMG_OK)
    throw Error{NMG_ENUM, "tridiagonal eigenvalues did not converge"};
  double tr1 = 0.0;
  int cnt = 0;
  for (int i = 0; i < kdim; ++i)
    if (L.w[i] >= P.bar) {
      tr1 += L.w[i];
      cnt++;
    }
  if (cnt == 0) {
    // no filtered Ritz value reaches bar: once the largest one has converged
    // (|beta_k s_k| <= rtol) the top of p's spectrum lies below bar and the
    // interval holds no eigenvalue
    L.tr0 = 0.0;
    tridiag_eigvec(kdim, L.al.data(), L.be.data(), L.w[kdim - 1], L.sv);
    return std::fabs(L.be[kdim - 1] * L.sv[kdim - 1]) <= S.rtol ? 1 : 0;
  }
  const bool conv = std::fabs(tr1 - L.tr0) < tol * std::fabs(tr1);
  L.tr0 = tr1;
  if (!conv) return 0;
  double rmax = 0.0;
  for (int i = 0; i < kdim; ++i)
    if (L.w[i] >= P.bar) {
      tridiag_eigvec(kdim, L.al.data(), L.be.data(), L.w[i], L.sv);
      rmax = std::max(rmax, std::fabs(L.be[kdim - 1] * L.sv[kdim - 1]));
    }
  if (rmax > S.rtol) return 0;
  if (S.resid_target <= 0.0) return 1;
  std::vector<double> th;
  DVec dC;
  tridiag_eig_device(S, kdim, L.al, L.be, dC, th);
  auto reextract = [&](int k) {
    std::vector<double> al, be, th2;
    fetch_tridiag(S, k, al, be);
    DVec dC2;
    tridiag_eig_device(S, k, al, be, dC2, th2);
    extract_ritz(S, k, dC2.p, th2, P.bar, xintv);
  };
  return residual_stop(S, kdim, dC.p, th, P.bar, xintv, L.rt, reextract) ? 2 : 0;
}

// extraction after the loop: eigenvectors of T on the device (rocSOLVER
// divide & conquer), Ritz vectors, Rayleigh-Ritz
static void lanczos_extract(Solver& S, int kdim, const PolyFilter& P, const double* xintv) {
  std::vector<double> al, be, theta;
  fetch_tridiag(S, kdim, al, be);
  DVec dC;
  tridiag_eig_device(S, kdim, al, be, dC, theta);
  extract_ritz(S, kdim, dC.p, theta, P.bar, xintv);
}

static void do_cheblannr(Solver& S, const double* xintv, int maxit, double tol, const PolyFilter& P) {
  NMG_CHECK(S.n > 0, NMG_ESTATE, "problem size not set");
  NMG_CHECK(xintv[1] > xintv[0], NMG_EINVAL, "empty interval");
  const double t0 = now();
  const int64_t n = S.n;
  maxit = (int)std::min<int64_t>(maxit, n);
  setup_work(S);
  S.ab.ensure(2 * (int64_t)(maxit + 2));
  S.free_basis();
  ensure_basis(S, std::min(maxit + 1, 64), 0, maxit + 1);
  S.filter_applies = 0;
  S.target_met = 1;
  S.resid_best = 0.0;
  DVec hw;
  lanczos_start(S);
  LanczosTest L;
  int kdim = 0, stop = 0;
  for (int k = 0; k < maxit && stop == 0; ++k) {
    ensure_basis(S, k + 2, k + 1, maxit + 1);
    hw.ensure((int64_t)gemv_work_size(n, k + 2) + k + 4);
    lanczos_step(S, k, 1, &P, hw.p);
    kdim = k + 1;
    progress(S, "cheblannr", kdim, t0);
    stop = lanczos_test(S, kdim, P, xintv, tol, L);
  }
  S.lan_steps = kdim;
  if (stop != 2) lanczos_extract(S, kdim, P, xintv);
  S.free_basis();
  S.seconds = now() - t0;
}

// ---------------------------------------------------------------------------
// Block filtered Lanczos (block width K > 1): the same filter and extraction
// as ChebLanNr, applied to K vectors at once, so that every sparse product of
// the filter and of the Chebyshev B-solve reads the matrices once for K
// columns.  Recurrence in the z-space with B-inner products:
//   W = p(A B^-1) Z_j - Z_{j-1} R_j^T - Z_j A_j,  A_j = V_j^T p(A B^-1) Z_j,
//   full re-orthogonalisation W -= Z (V^T W) twice (tall-skinny GEMMs),
//   V' = B^-1 W,  V'^T W = R^T R (Cholesky),  V_{j+1} = V' R^-1, Z_{j+1} = W R^-1,
// giving the block tridiagonal T with diagonal blocks A_j and sub-diagonal
// blocks R_{j+1}.  Candidates are the Ritz values of T >= bar; Ritz residual
// of candidate s: ||R_{j+1} s_last||.

// K interleaved columns of x ~ M^{-1} b by the Chebyshev semi-iteration of
// cheb_solve, one K-column fused step per degree
// With w32 (3 n K floats), the iterates and b are held in FP32 (the product
// and the update still formed in FP64 per entry): half the vector bytes.
static void block_cheb(Cheb& C, int K, const double* b, double* out, double* tmp, hipStream_t s, bool fp32,
                       KernelProbe* probe, float* w32 = nullptr) {
  if (C.refine > 0) return refine_solve(C, K, b, out, tmp, s, probe, w32);
  if (K == 1 && !w32) return cheb_solve(C, b, out, tmp, s, fp32, probe);
  const int64_t nk = C.mat->nrow * K;
  const double step_bytes = C.mat->matrix_bytes(fp32) + 4.0 * 8.0 * (double)nk;
  // the persistent form is built for K <= 16
  if (const PersistPlan* P = K <= 16 ? C.persistent(fp32) : nullptr) {
    const bool pr = probe && probe->armed();
    if (pr) probe->before(s);
    k_cheb_persistent(K, *C.mat, *P, b, out, tmp, C.deg, C.lmin, C.lmax, s);
    if (pr) probe->after(s);
    if (probe) {
      probe->bytes = step_bytes * C.deg;
      probe->launches = 1;
    }
    return;
  }
  const double th = 0.5 * (C.lmax + C.lmin), de = 0.5 * (C.lmax - C.lmin), sg = th / de;
  if (w32) {
    float* b32 = w32;
    float* fk = w32 + nk;
    float* fm = w32 + 2 * nk;
    k_cheb32_start(nk, 1.0 / th, b, b32, fk, fm, s);
    const bool pr = probe && probe->armed();
    const int mid = (C.deg + 1) / 2;
    double w = 2.0;
    for (int k = 1; k <= C.deg; ++k) {
      w = 1.0 / (1.0 - w / (4.0 * sg * sg));
      if (pr && k == mid) probe->before(s);
      k_block_cheb_step32(K, *C.mat, b32, fk, fm, w, 1.0 / th, s, fp32);
      if (pr && k == mid) probe->after(s);
      std::swap(fk, fm);
    }
    k_f32_to_f64(nk, fk, out, s);
    if (probe) {
      probe->bytes = C.mat->matrix_bytes(fp32) + 4.0 * 4.0 * (double)nk;
      probe->launches = 1;
    }
    return;
  }
  double* xk = (C.deg % 2 == 0) ? out : tmp;
  double* xm = (C.deg % 2 == 0) ? tmp : out;
  k_scale_copy(nk, 1.0 / th, b, xk, s);
  k_fill(nk, 0.0, xm, s);
  const bool pr = probe && probe->armed();
  const int mid = (C.deg + 1) / 2;
  double w = 2.0;
  for (int k = 1; k <= C.deg; ++k) {
    w = 1.0 / (1.0 - w / (4.0 * sg * sg));
    if (pr && k == mid) probe->before(s);
    k_block_cheb_step(K, *C.mat, b, xk, xm, w, 1.0 / th, s, fp32);
    if (pr && k == mid) probe->after(s);
    std::swap(xk, xm);
  }
  if (probe) {
    probe->bytes = step_bytes;
    probe->launches = 1;
  }
}

// y (interleaved n*K) ~ B~^{-1} b, all K columns at once
static void block_bsolve(Solver& S, int K, const double* b, double* out, double* tmp) {
  NMG_CHECK(S.bsol != nullptr, NMG_ESTATE, "generalized problem without a B-solve");
  Cheb& C = *S.bsol;
  const int P = (S.split > 1 && K % S.split == 0 && K / S.split >= 2 && !S.borrowed) ? S.split : 1;
  const int H = K / P;
  if (P > 1) {
    // P solves of K/P columns at once on P streams: independent columns, the
    // same matrix read by all
    if (!S.sx[0]) {
      for (int q = 0; q < 3; ++q) NMG_HIP(hipStreamCreateWithFlags(&S.sx[q], hipStreamNonBlocking));
      for (int q = 0; q < 4; ++q) NMG_HIP(hipEventCreateWithFlags(&S.ev[q], hipEventDisableTiming));
    }
    for (int i = 0; i < 3 * P; ++i) S.hb[i].ensure(S.n * H);
  }
  // FP32 work per part: FP32 iterates, or the inner passes of a refined
  // solve (each concurrent part needs its own)
  const bool v32 = (S.fp32v() && !C.persistent(S.fp32())) || C.refine > 0;
  if (v32)
    for (int q = 0; q < P; ++q) S.hf[q].ensure(3 * S.n * H);
  const bool graphed = S.graphs && !C.persistent(S.fp32());
  KernelProbe* probe = &S.probe[0];
  auto body = [&](KernelProbe* pb) {
    if (P == 1) {
      block_cheb(C, K, b, out, tmp, S.s, S.fp32(), pb, v32 ? S.hf[0].p : nullptr);
      return;
    }
    double* in[4];
    double* res[4];
    for (int q = 0; q < P; ++q) {
      in[q] = S.hb[3 * q].p;
      res[q] = S.hb[3 * q + 1].p;
    }
    k_split_parts(S.n, K, P, b, in, S.s);
    NMG_HIP(hipEventRecord(S.ev[0], S.s));
    for (int q = 1; q < P; ++q) NMG_HIP(hipStreamWaitEvent(S.sx[q - 1], S.ev[0], 0));
    for (int q = 0; q < P; ++q)
      block_cheb(C, H, in[q], res[q], S.hb[3 * q + 2].p, q == 0 ? S.s : S.sx[q - 1], S.fp32(),
                 q == 0 ? pb : nullptr, v32 ? S.hf[q].p : nullptr);
    for (int q = 1; q < P; ++q) {
      NMG_HIP(hipEventRecord(S.ev[q], S.sx[q - 1]));
      NMG_HIP(hipStreamWaitEvent(S.s, S.ev[q], 0));
    }
    k_merge_parts(S.n, K, P, res, out, S.s);
  };
  if (graphed) {
    // the whole solve (all parts) is one graph; a sampled solve is
    // bracketed around it on the solver's stream
    const bool pr = probe->armed();
    if (pr) probe->before(S.s);
    run_graphed(S, graph_key(S, C, K, P, b, out, tmp), [&] { body(nullptr); });
    if (pr) probe->after(S.s);
    probe->bytes = C.mat->matrix_bytes(S.fp32()) + 4.0 * (v32 ? 4.0 : 8.0) * (double)C.mat->nrow * H;
    probe->launches = C.deg;
  } else {
    body(probe);
  }
  C.ncalls += K;
  C.nmatvec += (int64_t)C.deg * K;
}

// A~ applied to K interleaved columns (K = 1: one vector): the sparse product,
// or the fluid-solid Schur operator
// the fluid-solid Schur operator on K interleaved columns, on stream s with
// the work buffers W (one set per concurrent chain)
static void apply_fs(FsOp& F, int K, const double* x, double* y, hipStream_t s, FsWork& W, bool f32) {
  for (int i = 0; i < 3; ++i) {
    W.t[i].ensure(F.n * K);
    W.p[i].ensure(F.np * K);
  }
  double *v0 = W.t[0].p, *w0 = W.t[1].p, *w1 = W.t[2].p;
  double *x0 = W.p[0].p, *y0 = W.p[1].p, *tp = W.p[2].p;
  k_rowmul(F.n, K, F.d.p, x, v0, s);        // v0 = D v
  k_block_spmv(K, *F.Ad, v0, w0, s, f32);   // w0 = Ad v0
  k_block_spmv(K, *F.ET, v0, x0, s, f32);   // x0 = ET v0
  k_rowmul(F.np, K, F.dp.p, x0, tp, s);     // x1 = Dp x0
  float* w32 = nullptr;
  if (F.cp->refine > 0) {
    W.f32.ensure(3 * F.np * K);
    w32 = W.f32.p;
  }
  block_cheb(*F.cp, K, tp, y0, x0, s, f32, nullptr, w32);  // y0 ~ Ap~^-1 x1
  k_rowmul(F.np, K, F.dp.p, y0, tp, s);     // y1 = Dp y0
  k_block_spmv(K, *F.E, tp, w1, s, f32);    // w1 = E y1
  k_add_rowmul(F.n, K, F.d.p, w0, w1, y, s);  // w = D (w0 + w1)
  F.cp->ncalls += K;
  F.cp->nmatvec += (int64_t)F.cp->deg * K;
}

static void apply_A_block(Solver& S, int K, const double* x, double* y) {
  if (!S.fs) {
    if (K == 1) k_spmv(S.A(), x, y, S.s);
    else k_block_spmv(K, S.A(), x, y, S.s);
    return;
  }
  apply_fs(*S.fs, K, x, y, S.s, S.fs->part[0], false);
}
static void apply_A(Solver& S, const double* x, double* y) { apply_A_block(S, 1, x, y); }

// one Chebyshev filter step on K interleaved columns: vkm <- t (A~ in - cc vk)
// - vkm, y += mu vkm (fused into the sparse product's epilogue, or after the
// fluid-solid operator)
static void block_filter_step(Solver& S, int K, const double* in, const double* vk, double* vkm, double* y,
                              const double* mu_k, double t, double cc, bool first) {
  if (S.fs) {
    S.aw.ensure(S.n * K);
    apply_fs(*S.fs, K, in, S.aw.p, S.s, S.fs->part[0], S.fp32a());
    k_filter_update(S.n, K, S.aw.p, vk, vkm, y, mu_k, 0.0, t, cc, first, S.s);
    return;
  }
  const bool pr = S.probe[1].armed();
  if (pr) S.probe[1].before(S.s);
  k_block_filter_step(K, S.A(), in, vk, vkm, y, mu_k, t, cc, first, S.s, S.fp32a());
  if (pr) S.probe[1].after(S.s);
}

// the parts of a split filter: the columns of K split into P groups of
// K/P, each filtered as its own chain (B-solve, product, recurrence) on its
// own stream; 1 when the filter is not split
static int filter_parts(const Solver& S, int K) {
  const int P = S.split;
  // fluid-solid chains may be single columns: the latency-bound Ap~ solve of
  // one chain then runs beside the bandwidth-bound B~ solve of another
  return (P > 1 && K % P == 0 && (K / P >= 2 || S.fs) && !S.borrowed && S.geneig) ? P : 1;
}

// block_filter with the columns split into P independent chains: the chains
// share only the matrices, so each part's stiffness product overlaps the
// other parts' B-solves and no per-degree split, merge or join is needed
static void block_filter_split(Solver& S, int K, int P, const PolyFilter& Pf, const double* mu_dev,
                               const double* z, double* y) {
  const int H = K / P;
  const int64_t nh = S.n * H;
  Cheb& C = *S.bsol;
  if (!S.sx[0]) {
    for (int q = 0; q < 3; ++q) NMG_HIP(hipStreamCreateWithFlags(&S.sx[q], hipStreamNonBlocking));
    for (int q = 0; q < 4; ++q) NMG_HIP(hipEventCreateWithFlags(&S.ev[q], hipEventDisableTiming));
  }
  const bool v32 = (S.fp32v() && !C.persistent(S.fp32())) || C.refine > 0;  // per-chain FP32 work
  double* zq[4];
  double* yq[4];
  double* mq[4];
  for (int q = 0; q < P; ++q) {
    for (int i = 0; i < 6; ++i) S.pf[q][i].ensure(nh);
    S.pf[q][6].ensure((int64_t)(Pf.deg + 1) * H);
    zq[q] = S.pf[q][0].p;
    yq[q] = S.pf[q][1].p;
    mq[q] = S.pf[q][6].p;
    if (v32) S.hf[q].ensure(3 * nh);
  }
  k_split_parts(S.n, K, P, z, zq, S.s);
  k_split_parts(Pf.deg + 1, K, P, mu_dev, mq, S.s);  // per-part filter coefficients
  NMG_HIP(hipEventRecord(S.ev[0], S.s));
  for (int q = 1; q < P; ++q) NMG_HIP(hipStreamWaitEvent(S.sx[q - 1], S.ev[0], 0));
  // the host issues the chains degree by degree, alternating between the
  // streams, so that every stream always has queued work
  hipStream_t sq[4];
  double* vk[4];
  double* vkm[4];
  for (int q = 0; q < P; ++q) {
    sq[q] = q == 0 ? S.s : S.sx[q - 1];
    vk[q] = S.pf[q][2].p;
    vkm[q] = S.pf[q][3].p;
    k_copy(nh, zq[q], vk[q], sq[q]);
    k_block_colscale(S.n, H, mq[q], vk[q], yq[q], sq[q]);
  }
  for (int k = 1; k <= Pf.deg; ++k) {
    const double t = (k == 1 ? 1.0 : 2.0) / Pf.dd;
    for (int q = 0; q < P; ++q) {
      double* b1 = S.pf[q][4].p;
      double* b2 = S.pf[q][5].p;
      block_cheb(C, H, vk[q], b1, b2, sq[q], S.fp32(), q == 0 ? &S.probe[0] : nullptr,
                 v32 ? S.hf[q].p : nullptr);
      if (S.fs) {
        FsWork& W = S.fs->part[q];
        W.aw.ensure(nh);
        apply_fs(*S.fs, H, b1, W.aw.p, sq[q], W, S.fp32a());
        k_filter_update(S.n, H, W.aw.p, vk[q], vkm[q], yq[q], mq[q] + (size_t)k * H, 0.0, t, Pf.cc, k == 1,
                        sq[q]);
        std::swap(vk[q], vkm[q]);
        continue;
      }
      const bool pr = q == 0 && S.probe[1].armed();
      if (pr) S.probe[1].before(sq[q]);
      k_block_filter_step(H, S.A(), b1, vk[q], vkm[q], yq[q], mq[q] + (size_t)k * H, t, Pf.cc, k == 1, sq[q],
                          S.fp32a());
      if (pr) S.probe[1].after(sq[q]);
      std::swap(vk[q], vkm[q]);
    }
  }
  for (int q = 1; q < P; ++q) {
    NMG_HIP(hipEventRecord(S.ev[q], S.sx[q - 1]));
    NMG_HIP(hipStreamWaitEvent(S.s, S.ev[q], 0));
  }
  k_merge_parts(S.n, K, P, yq, y, S.s);
  C.ncalls += (int64_t)K * Pf.deg;
  C.nmatvec += (int64_t)C.deg * K * Pf.deg;
  S.filter_applies += K;
}

// y = p((A B^-1 - cc)/dd) z for K interleaved columns; mu_dev: (deg+1)*K
static void block_filter(Solver& S, int K, const PolyFilter& P, const double* mu_dev, const double* z,
                         double* y) {
  if (S.split_filter) {
    const int parts = filter_parts(S, K);
    if (parts > 1) return block_filter_split(S, K, parts, P, mu_dev, z, y);
  }
  const int64_t nk = S.n * K;
  for (int i = 0; i < 4; ++i) S.bw[i].ensure(nk);
  double* vk = S.bw[0].p;
  double* vkm = S.bw[1].p;
  double* b1 = S.bw[2].p;
  double* b2 = S.bw[3].p;
  k_copy(nk, z, vk, S.s);
  k_block_colscale(S.n, K, mu_dev, vk, y, S.s);
  for (int k = 1; k <= P.deg; ++k) {
    const double t = (k == 1 ? 1.0 : 2.0) / P.dd;
    const double* in = vk;
    if (S.geneig) {
      block_bsolve(S, K, vk, b1, b2);
      in = b1;
    }
    block_filter_step(S, K, in, vk, vkm, y, mu_dev + (size_t)k * K, t, P.cc, k == 1);
    std::swap(vk, vkm);
  }
  S.filter_applies += K;
}

// G = R^T R for a small SPD G (row-major K x K); returns R^-1 (row-major, upper)
static void chol_inv_upper(int K, std::vector<double>& G, std::vector<double>& R, std::vector<double>& Ri) {
  R.assign((size_t)K * K, 0.0);
  for (int j = 0; j < K; ++j) {
    double s = G[(size_t)j * K + j];
    for (int k = 0; k < j; ++k) s -= R[(size_t)k * K + j] * R[(size_t)k * K + j];
    NMG_CHECK(s > 0.0 && std::isfinite(s), NMG_ENUM,
              "block Lanczos: the new block lost rank (B-Gram matrix not positive definite)");
    const double rjj = std::sqrt(s);
    R[(size_t)j * K + j] = rjj;
    for (int i = j + 1; i < K; ++i) {
      double t = G[(size_t)j * K + i];
      for (int k = 0; k < j; ++k) t -= R[(size_t)k * K + j] * R[(size_t)k * K + i];
      R[(size_t)j * K + i] = t / rjj;
    }
  }
  Ri.assign((size_t)K * K, 0.0);
  for (int c = 0; c < K; ++c) {  // solve R x = e_c (back substitution)
    for (int i = K - 1; i >= 0; --i) {
      double t = (i == c) ? 1.0 : 0.0;
      for (int k = i + 1; k < K; ++k) t -= R[(size_t)i * K + k] * Ri[(size_t)k * K + c];
      Ri[(size_t)i * K + c] = t / R[(size_t)i * K + i];
    }
  }
}

// small K x K products on the device: C = op(A) op(B) with rocBLAS (column-major)
static void gemm(Solver& S, rocblas_operation ta, rocblas_operation tb, int64_t m, int64_t n, int64_t k,
                 double alpha, const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                 double* C, int64_t ldc, const char* what) {
  NMG_LIB(rocblas_dgemm(S.blas, ta, tb, (rocblas_int)m, (rocblas_int)n, (rocblas_int)k, &alpha, A,
                        (rocblas_int)lda, B, (rocblas_int)ldb, &beta, C, (rocblas_int)ldc),
          what);
}

// B-orthonormalise the column block V (n x K col-major) with Z = B V given in Zc:
// V <- V R^-1, Z <- Z R^-1; returns R (row-major upper) on the host
static void block_borth(Solver& S, int K, double* V, double* Z, std::vector<double>& R) {
  const int64_t n = S.n;
  DVec G, Rinv, T;
  G.ensure((int64_t)K * K);
  Rinv.ensure((int64_t)K * K);
  T.ensure(n * K);
  S.gw.ensure((int64_t)gemm_tn_work_size(n, K, K));
  dense_tn(S, n, K, K, V, n, Z, n, G.p, K);  // block Gram V^T Z
  std::vector<double> g((size_t)K * K), Ri;
  NMG_HIP(hipMemcpyAsync(g.data(), G.p, sizeof(double) * K * K, hipMemcpyDeviceToHost, S.s));
  NMG_HIP(hipStreamSynchronize(S.s));
  for (int i = 0; i < K; ++i)
    for (int j = i + 1; j < K; ++j) g[(size_t)i * K + j] = g[(size_t)j * K + i] = 0.5 * (g[(size_t)i * K + j] + g[(size_t)j * K + i]);
  chol_inv_upper(K, g, R, Ri);
  // column-major copy of R^-1 (row-major Ri -> transpose)
  std::vector<double> ric((size_t)K * K);
  for (int i = 0; i < K; ++i)
    for (int j = 0; j < K; ++j) ric[(size_t)j * K + i] = Ri[(size_t)i * K + j];
  NMG_HIP(hipMemcpyAsync(Rinv.p, ric.data(), sizeof(double) * K * K, hipMemcpyHostToDevice, S.s));
  dense_nn(S, n, K, K, 1.0, V, n, Rinv.p, K, 0.0, T.p, n, "V R^-1");
  NMG_HIP(hipMemcpyAsync(V, T.p, sizeof(double) * n * K, hipMemcpyDeviceToDevice, S.s));
  if (Z != V) {
    dense_nn(S, n, K, K, 1.0, Z, n, Rinv.p, K, 0.0, T.p, n, "Z R^-1");
    NMG_HIP(hipMemcpyAsync(Z, T.p, sizeof(double) * n * K, hipMemcpyDeviceToDevice, S.s));
  }
  NMG_HIP(hipStreamSynchronize(S.s));
}

// eigenvalues/vectors of the dense symmetric T (m x m, column-major) on the device
static void dense_eig(Solver& S, int m, const std::vector<double>& T, DVec& dT, std::vector<double>& w) {
  dT.ensure((int64_t)m * m);
  DVec dw, de, dinfo_v;
  dw.ensure(m);
  de.ensure(m);
  dinfo_v.ensure(1);
  NMG_HIP(hipMemcpyAsync(dT.p, T.data(), sizeof(double) * m * m, hipMemcpyHostToDevice, S.s));
  NMG_LIB(rocsolver_dsyevd(S.blas, rocblas_evect_original, rocblas_fill_upper, m, dT.p, m, dw.p, de.p,
                           (rocblas_int*)dinfo_v.p),
          "rocsolver_dsyevd failed");
  w.resize(m);
  rocblas_int info = 0;
  NMG_HIP(hipMemcpyAsync(w.data(), dw.p, sizeof(double) * m, hipMemcpyDeviceToHost, S.s));
  NMG_HIP(hipMemcpyAsync(&info, dinfo_v.p, sizeof(info), hipMemcpyDeviceToHost, S.s));
  NMG_HIP(hipStreamSynchronize(S.s));
  NMG_CHECK(info == 0, NMG_ENUM, "dense eigensolver failed");
}

static void do_cheblannr_block(Solver& S, const double* xintv, int maxit, double tol, const PolyFilter& P) {
  NMG_CHECK(S.n > 0, NMG_ESTATE, "problem size not set");
  NMG_CHECK(xintv[1] > xintv[0], NMG_EINVAL, "empty interval");
  const int K = S.block;
  const double t0 = now();
  const int64_t n = S.n;
  maxit = (int)std::min<int64_t>(maxit, n);
  const int maxblk = std::max(2, maxit / K);
  const int maxcols = (maxblk + 1) * K;
  setup_work(S);
  for (int i = 0; i < 6; ++i) S.bw[i].ensure(n * K);
  S.free_basis();
  ensure_basis(S, std::min(maxcols, std::max(2 * K, 64)), 0, maxcols);
  S.filter_applies = 0;
  S.target_met = 1;
  S.resid_best = 0.0;
  // per-column filter coefficients (all columns share the interval here)
  DVec mu_dev;
  mu_dev.ensure((int64_t)(P.deg + 1) * K);
  {
    std::vector<double> m((size_t)(P.deg + 1) * K);
    for (int k = 0; k <= P.deg; ++k)
      for (int c = 0; c < K; ++c) m[(size_t)k * K + c] = P.mu[k];
    NMG_HIP(hipMemcpyAsync(mu_dev.p, m.data(), sizeof(double) * m.size(), hipMemcpyHostToDevice, S.s));
  }
  // start block: column c is the start vector of seed S.seed over indices c*n + i
  {
    std::vector<double> v0(n * K);
    for (int c = 0; c < K; ++c) {
      std::vector<double> col(n);
      start_vector(n, S.seed + 0x9E3779B97F4A7C15ULL * (uint64_t)c, col.data());
      std::memcpy(v0.data() + (size_t)c * n, col.data(), sizeof(double) * n);
    }
    NMG_HIP(hipMemcpyAsync(S.V, v0.data(), sizeof(double) * n * K, hipMemcpyHostToDevice, S.s));
    for (int c = 0; c < K; ++c) apply_B(S, S.V + (size_t)c * n, S.Z + (size_t)c * n);
    std::vector<double> R0;
    block_borth(S, K, S.V, S.Z, R0);
  }
  std::vector<std::vector<double>> Ablk, Rblk;  // row-major K x K
  DVec Hd, Ad, Rd, dT;
  std::vector<double> Tm, w;
  double tr0 = 0.0;
  ResidualTracker rt;
  int nblk = 0;
  bool converged = false, stopped_with_result = false;
  for (int j = 0; j < maxblk && !converged; ++j) {
    progress(S, "block cheblannr", j * K, t0);
    ensure_basis(S, (j + 2) * K, (j + 1) * K, maxcols);
    double* Vj = S.V + (size_t)j * K * n;
    double* Zj = S.Z + (size_t)j * K * n;
    double* W = S.Z + (size_t)(j + 1) * K * n;
    double* Vn = S.V + (size_t)(j + 1) * K * n;
    // filter in the interleaved layout
    k_to_interleaved(n, K, Zj, n, S.bw[4].p, S.s);
    block_filter(S, K, P, mu_dev.p, S.bw[4].p, S.bw[5].p);
    k_from_interleaved(n, K, S.bw[5].p, W, n, S.s);
    // three-term part
    Ad.ensure((int64_t)K * K);
    Rd.ensure((int64_t)K * K);
    if (j > 0) {
      // column-major R_j^T: element (r, c) = R_j[c][r], and R_j is row-major
      std::vector<double> rt((size_t)K * K);
      for (int r = 0; r < K; ++r)
        for (int c = 0; c < K; ++c) rt[(size_t)c * K + r] = Rblk[j - 1][(size_t)c * K + r];
      NMG_HIP(hipMemcpyAsync(Rd.p, rt.data(), sizeof(double) * K * K, hipMemcpyHostToDevice, S.s));
      dense_nn(S, n, K, K, -1.0, S.Z + (size_t)(j - 1) * K * n, n, Rd.p, K, 1.0, W, n, "W -= Z_{j-1} R_j^T");
    }
    dense_tn(S, n, K, K, Vj, n, W, n, Ad.p, K);  // A_j = V_j^T W
    std::vector<double> a((size_t)K * K);
    NMG_HIP(hipMemcpyAsync(a.data(), Ad.p, sizeof(double) * K * K, hipMemcpyDeviceToHost, S.s));
    dense_nn(S, n, K, K, -1.0, Zj, n, Ad.p, K, 1.0, W, n, "W -= Z_j A_j");
    // full re-orthogonalisation, classical Gram-Schmidt twice, as GEMMs
    const int64_t m = (int64_t)(j + 1) * K;
    Hd.ensure(m * K);
    for (int pass = 0; pass < 2; ++pass) {
      dense_tn(S, n, (int)m, K, S.V, n, W, n, Hd.p, m);  // H = V^T W
      dense_nn(S, n, K, (int)m, -1.0, S.Z, n, Hd.p, m, 1.0, W, n, "W -= Z H");
    }
    // V_{j+1} = B^-1 W, B-orthonormalise
    if (S.geneig) {
      k_to_interleaved(n, K, W, n, S.bw[4].p, S.s);
      block_bsolve(S, K, S.bw[4].p, S.bw[5].p, S.bw[2].p);
      k_from_interleaved(n, K, S.bw[5].p, Vn, n, S.s);
    }
    NMG_HIP(hipStreamSynchronize(S.s));
    for (int r = 0; r < K; ++r)  // column-major -> row-major, symmetrised
      for (int c = r; c < K; ++c) {
        const double v = 0.5 * (a[(size_t)c * K + r] + a[(size_t)r * K + c]);
        a[(size_t)c * K + r] = a[(size_t)r * K + c] = v;
      }
    for (double v : a) NMG_CHECK(std::isfinite(v), NMG_ENUM, "block Lanczos breakdown: non-finite coefficient");
    std::vector<double> R;
    block_borth(S, K, Vn, W, R);
    Ablk.push_back(a);
    Rblk.push_back(R);
    nblk = j + 1;
    const int mm = nblk * K;
    if (mm < S.ntest) continue;
    // dense block tridiagonal T (column-major) and its eigenpairs
    Tm.assign((size_t)mm * mm, 0.0);
    for (int b = 0; b < nblk; ++b)
      for (int r = 0; r < K; ++r)
        for (int c = 0; c < K; ++c) {
          Tm[(size_t)(b * K + c) * mm + b * K + r] = Ablk[b][(size_t)r * K + c];
          if (b + 1 < nblk) {
            const double v = Rblk[b][(size_t)r * K + c];  // T[(b+1)K + r, bK + c] = R_{b+1}[r][c]
            Tm[(size_t)(b * K + c) * mm + (b + 1) * K + r] = v;
            Tm[(size_t)((b + 1) * K + r) * mm + b * K + c] = v;
          }
        }
    dense_eig(S, mm, Tm, dT, w);
    double tr1 = 0.0;
    int cnt = 0;
    for (int i = 0; i < mm; ++i)
      if (w[i] >= P.bar) {
        tr1 += w[i];
        cnt++;
      }
    // Ritz residual ||R_{j+1} s_last|| of Ritz vector i of T
    std::vector<double> last;
    auto ritz_resid = [&](int i) {
      double nrm = 0.0;
      for (int r = 0; r < K; ++r) {
        double t = 0.0;
        for (int c = r; c < K; ++c) t += R[(size_t)r * K + c] * last[(size_t)i * K + c];
        nrm += t * t;
      }
      return std::sqrt(nrm);
    };
    auto fetch_last = [&]() {
      last.resize((size_t)K * mm);
      for (int i = 0; i < mm; ++i)
        NMG_HIP(hipMemcpyAsync(last.data() + (size_t)i * K, dT.p + (size_t)i * mm + (mm - K), sizeof(double) * K,
                               hipMemcpyDeviceToHost, S.s));
      NMG_HIP(hipStreamSynchronize(S.s));
    };
    if (cnt == 0) {
      // no filtered Ritz value reaches bar: once the largest one has
      // converged, the top of p's spectrum lies below bar and the interval
      // holds no eigenvalue (the single-vector rule)
      tr0 = 0.0;
      fetch_last();
      if (ritz_resid(mm - 1) <= S.rtol) converged = true;
      continue;
    }
    const bool conv = std::fabs(tr1 - tr0) < tol * std::fabs(tr1);
    tr0 = tr1;
    if (!conv) continue;
    fetch_last();
    double rmax = 0.0;
    for (int i = 0; i < mm; ++i)
      if (w[i] >= P.bar) rmax = std::max(rmax, ritz_resid(i));
    if (rmax > S.rtol) continue;
    auto reextract = [&](int k) {
      std::vector<double> Tk((size_t)k * k), wk;
      for (int c = 0; c < k; ++c)
        for (int r = 0; r < k; ++r) Tk[(size_t)c * k + r] = Tm[(size_t)c * mm + r];
      DVec dTk;
      dense_eig(S, k, Tk, dTk, wk);
      extract_ritz(S, k, dTk.p, wk, P.bar, xintv);
    };
    if (residual_stop(S, mm, dT.p, w, P.bar, xintv, rt, reextract)) {
      converged = true;
      stopped_with_result = S.resid_target > 0.0;
    }
  }
  const int kdim = nblk * K;
  S.lan_steps = kdim;
  if (stopped_with_result) {
    S.free_basis();
    S.seconds = now() - t0;
    return;
  }
  if (w.size() != (size_t)kdim) {
    Tm.assign((size_t)kdim * kdim, 0.0);
    for (int b = 0; b < nblk; ++b)
      for (int r = 0; r < K; ++r)
        for (int c = 0; c < K; ++c) {
          Tm[(size_t)(b * K + c) * kdim + b * K + r] = Ablk[b][(size_t)r * K + c];
          if (b + 1 < nblk) {
            const double v = Rblk[b][(size_t)r * K + c];
            Tm[(size_t)(b * K + c) * kdim + (b + 1) * K + r] = v;
            Tm[(size_t)((b + 1) * K + r) * kdim + b * K + c] = v;
          }
        }
    dense_eig(S, kdim, Tm, dT, w);
  }
  extract_ritz(S, kdim, dT.p, w, P.bar, xintv);
  S.free_basis();
  S.seconds = now() - t0;
}

// ---------------------------------------------------------------------------
// Spectrum slices advanced in lockstep.  Slice s has its own interval
// [sli[s], sli[s+1]], its own FindPol filter p_s and its own single-vector
// ChebLanNr recurrence (basis, tridiagonal, tests, extraction), exactly as a
// separate nmg_cheblannr call would run it.  What the slices share is every
// sparse product: at each Lanczos step the K = (#slices rounded up to a block
// width) current vectors go through one K-column filter (per-column Chebyshev
// coefficients; a filter of lower degree has zero coefficients past its own
// degree, so its polynomial is unchanged) and one K-column B-solve, so A~ and
// B~ are read once per product for all slices.  All filters share cc and dd
// (the centre and half-width of [lmin, lmax]).
static int block_width_for(int ns) {
  for (int K : {1, 2, 4, 8, 16, 32, 64})
    if (K >= ns) return K;
  return -1;
}

static void do_solve_slices(Solver& S, int ns, const double* sli, double lmin, double lmax, double thr_int,
                            double thr_ext, int maxit, double tol) {
  NMG_CHECK(S.n > 0, NMG_ESTATE, "problem size not set");
  const int K = block_width_for(ns);
  NMG_CHECK(ns >= 1 && K > 0, NMG_EINVAL, "between 1 and 64 slices");
  for (int i = 0; i < ns; ++i) NMG_CHECK(sli[i + 1] > sli[i], NMG_EINVAL, "slice boundaries must increase");
  const double t0 = now();
  const int64_t n = S.n;
  maxit = (int)std::min<int64_t>(maxit, n);
  // filters
  std::vector<PolyFilter> pol(ns);
  std::vector<std::array<double, 4>> xf(ns), xi(ns);
  int maxdeg = 0;
  // a pair on an interior boundary can come out of Rayleigh-Ritz a rounding
  // above (or below) it, while its filtered Ritz value sits at the bar of the
  // neighbour's filter: extraction accepts pairs up to bnd_tol beyond an
  // interior boundary, and the copies found by both neighbours are merged
  auto bnd_tol = [](double x) { return 1e-9 * std::max(1.0, std::fabs(x)); };
  for (int i = 0; i < ns; ++i) {
    xf[i] = {sli[i], sli[i + 1], lmin, lmax};
    xi[i] = {i > 0 ? sli[i] - bnd_tol(sli[i]) : sli[i], i + 1 < ns ? sli[i + 1] + bnd_tol(sli[i + 1]) : sli[i + 1],
             lmin, lmax};
    const int rc = find_pol(xf[i].data(), thr_int, thr_ext, 2, 10000, 2, pol[i]);
    NMG_CHECK(rc == NMG_OK, NMG_ENUM, "no filter meets the thresholds for slice " + std::to_string(i));
    maxdeg = std::max(maxdeg, pol[i].deg);
  }
  PolyFilter Pmax = pol[0];
  Pmax.deg = maxdeg;
  DVec mu_dev;
  mu_dev.ensure((int64_t)(maxdeg + 1) * K);
  {
    std::vector<double> m((size_t)(maxdeg + 1) * K, 0.0);
    for (int k = 0; k <= maxdeg; ++k)
      for (int c = 0; c < ns; ++c) m[(size_t)k * K + c] = k <= pol[c].deg ? pol[c].mu[k] : 0.0;
    NMG_HIP(hipMemcpyAsync(mu_dev.p, m.data(), sizeof(double) * m.size(), hipMemcpyHostToDevice, S.s));
  }
  // one recurrence per slice, sharing the operators, stream and rocBLAS handle
  std::vector<std::unique_ptr<Solver>> sub(ns);
  for (int i = 0; i < ns; ++i) {
    Solver* C = new Solver();
    sub[i].reset(C);
    C->borrowed = true;
    C->device = S.device;
    C->s = S.s;
    C->blas = S.blas;
    C->n = n;
    C->fs = S.fs;
    C->amat = S.fs ? nullptr : const_cast<DevMat*>(&S.A());
    C->bmat = S.bmat;
    C->bsol = S.bsol;
    C->geneig = S.geneig;
    C->precision = S.precision;
    C->seed = S.seed + 0x9E3779B97F4A7C15ULL * (uint64_t)i;
    C->ntest = S.ntest;
    C->cycle = S.cycle;
    C->rtol = S.rtol;
    C->resid_target = S.resid_target;
    setup_work(*C);
    C->ab.ensure(2 * (int64_t)(maxit + 2));
    ensure_basis(*C, std::min(maxit + 1, 64), 0, maxit + 1);
    lanczos_start(*C);
  }
  setup_work(S);
  for (int i = 0; i < 6; ++i) S.bw[i].ensure(n * K);
  DVec cm, hw;
  cm.ensure(n * K);
  S.filter_applies = 0;
  std::vector<LanczosTest> L(ns);
  std::vector<int> stop(ns, 0), kdim(ns, 0);
  // gather column c of the K-column block from slice c's vector (zero for
  // finished or padding columns), and back
  auto gather = [&](auto vec_of) {
    for (int c = 0; c < K; ++c) {
      const double* src = (c < ns && stop[c] == 0) ? vec_of(c) : nullptr;
      if (src) NMG_HIP(hipMemcpyAsync(cm.p + (size_t)c * n, src, sizeof(double) * n, hipMemcpyDeviceToDevice, S.s));
      else k_fill(n, 0.0, cm.p + (size_t)c * n, S.s);
    }
    k_to_interleaved(n, K, cm.p, n, S.bw[4].p, S.s);
  };
  auto scatter = [&](auto vec_of) {
    k_from_interleaved(n, K, S.bw[5].p, cm.p, n, S.s);
    for (int c = 0; c < ns; ++c)
      if (stop[c] == 0)
        NMG_HIP(hipMemcpyAsync(vec_of(c), cm.p + (size_t)c * n, sizeof(double) * n, hipMemcpyDeviceToDevice, S.s));
  };
  int steps = 0;
  for (int k = 0; k < maxit; ++k) {
    int active = 0;
    for (int i = 0; i < ns; ++i) active += stop[i] == 0;
    if (active == 0) break;
    progress(S, "slices", k, t0);
    for (int i = 0; i < ns; ++i)
      if (stop[i] == 0) ensure_basis(*sub[i], k + 2, k + 1, maxit + 1);
    // z_{k+1} = p_s(A~ B~^-1) z_k for every slice, one K-column filter
    gather([&](int c) { return sub[c]->Z + (size_t)k * n; });
    block_filter(S, K, Pmax, mu_dev.p, S.bw[4].p, S.bw[5].p);
    scatter([&](int c) { return sub[c]->Z + (size_t)(k + 1) * n; });
    hw.ensure((int64_t)gemv_work_size(n, k + 2) + k + 4);
    for (int i = 0; i < ns; ++i)
      if (stop[i] == 0) lanczos_orth(*sub[i], k, hw.p);
    // v_{k+1} = B~^-1 z_{k+1}, one K-column B-solve
    if (S.geneig) {
      gather([&](int c) { return sub[c]->Z + (size_t)(k + 1) * n; });
      block_bsolve(S, K, S.bw[4].p, S.bw[5].p, S.bw[2].p);
      scatter([&](int c) { return sub[c]->V + (size_t)(k + 1) * n; });
    }
    for (int i = 0; i < ns; ++i)
      if (stop[i] == 0) lanczos_normalise(*sub[i], k);
    steps = k + 1;
    for (int i = 0; i < ns; ++i) {
      if (stop[i] != 0) continue;
      kdim[i] = k + 1;
      sub[i]->filter_applies++;
      stop[i] = lanczos_test(*sub[i], kdim[i], pol[i], xi[i].data(), tol, L[i]);
    }
  }
  // extraction and the union of the slices' pairs
  dfree(S.X);
  S.lam.clear();
  S.nev = 0;
  S.slice_nev.assign(ns, 0);
  S.ncand = 0;
  int total = 0;
  for (int i = 0; i < ns; ++i) {
    Solver& C = *sub[i];
    if (stop[i] != 2) lanczos_extract(C, kdim[i], pol[i], xi[i].data());
    C.free_basis();
    total += C.nev;
    S.ncand += C.ncand;
  }
  // a pair on a shared boundary can be found by both neighbours: drop the
  // second copy when the two are B~-parallel (|x_i^T B~ x_j| > 0.5)
  std::vector<std::vector<char>> keep(ns);
  for (int i = 0; i < ns; ++i) keep[i].assign(sub[i]->nev, 1);
  DVec bx, dd;
  bx.ensure(n);
  dd.ensure(1);
  for (int i = 0; i + 1 < ns; ++i) {
    const double b = sli[i + 1], del = 2.0 * bnd_tol(b);
    Solver &L0 = *sub[i], &L1 = *sub[i + 1];
    for (int a = 0; a < L0.nev; ++a) {
      if (std::fabs(L0.lam[a] - b) > del) continue;
      apply_B(S, L0.X + (size_t)a * n, bx.p);
      for (int c = 0; c < L1.nev; ++c) {
        if (!keep[i + 1][c] || std::fabs(L1.lam[c] - L0.lam[a]) > 2 * del) continue;
        k_dot(n, bx.p, L1.X + (size_t)c * n, dd.p, false, S.red.p, S.s);
        double g = 0.0;
        NMG_HIP(hipMemcpyAsync(&g, dd.p, sizeof(double), hipMemcpyDeviceToHost, S.s));
        NMG_HIP(hipStreamSynchronize(S.s));
        if (std::fabs(g) > 0.5) keep[i + 1][c] = 0;
      }
    }
  }
  total = 0;
  for (int i = 0; i < ns; ++i)
    for (char kf : keep[i]) total += kf;
  if (total > 0) {
    S.X = dmalloc<double>((size_t)n * total);
    int j = 0;
    for (int i = 0; i < ns; ++i)
      for (int a = 0; a < sub[i]->nev; ++a) {
        if (!keep[i][a]) continue;
        NMG_HIP(hipMemcpyAsync(S.X + (size_t)j * n, sub[i]->X + (size_t)a * n, sizeof(double) * n,
                               hipMemcpyDeviceToDevice, S.s));
        S.lam.push_back(sub[i]->lam[a]);
        S.slice_nev[i]++;
        ++j;
      }
  }
  S.nev = total;
  NMG_HIP(hipStreamSynchronize(S.s));
  S.filter_applies = 0;
  S.target_met = 1;
  S.resid_best = 0.0;
  for (int i = 0; i < ns; ++i) {
    S.filter_applies += sub[i]->filter_applies;
    S.target_met = std::min(S.target_met, sub[i]->target_met);
    S.resid_best = std::max(S.resid_best, sub[i]->resid_best);
  }
  S.lan_steps = steps;
  S.seconds = now() - t0;
}

// Trace moments tau_k ~ tr T_k((A~ B~^-1 - cc)/dd), k = 0..mdeg, by
// Hutchinson's estimator over nvec Rademacher vectors (tr M = E[v^T M v]),
// the vectors advanced 16 (or nvec) at a time through the K-column Chebyshev
// step and B-solve.  A~B~^-1 is similar to B~^-1 A~, so the trace is
// sum_i T_k(t_i) over the pencil's eigenvalues.
static void do_kpm(Solver& S, int mdeg, int nvec, double lmin, double lmax, std::vector<double>& tau) {
  NMG_CHECK(S.n > 0, NMG_ESTATE, "problem size not set");
  NMG_CHECK(mdeg >= 1 && mdeg <= 100000, NMG_EINVAL, "mdeg must be in [1, 100000]");
  NMG_CHECK(lmax > lmin, NMG_EINVAL, "lmax must exceed lmin");
  const int K = nvec >= 16 ? 16 : block_width_for(nvec);
  NMG_CHECK(nvec >= 1 && K > 0 && (nvec <= 16 ? K == nvec : nvec % 16 == 0), NMG_EINVAL,
            "nvec must be 1, 2, 4, 8 or a multiple of 16");
  const int64_t n = S.n;
  const double cc = 0.5 * (lmax + lmin), dd = 0.5 * (lmax - lmin);
  setup_work(S);
  for (int i = 0; i < 6; ++i) S.bw[i].ensure(n * K);
  DVec v0, mom, zmu, work;
  v0.ensure(n * K);
  mom.ensure((int64_t)(mdeg + 1) * K);
  zmu.ensure(K);
  work.ensure((int64_t)block_dot_work_size(K));
  k_fill(K, 0.0, zmu.p, S.s);
  tau.assign(mdeg + 1, 0.0);
  std::vector<double> h((size_t)(mdeg + 1) * K), r((size_t)n * K);
  for (int batch = 0; batch < nvec / K; ++batch) {
    start_vector(n * K, S.seed ^ (0xD1B54A32D192ED03ULL * (uint64_t)(batch + 1)), r.data());
    for (auto& x : r) x = x < 0.0 ? -1.0 : 1.0;
    NMG_HIP(hipMemcpyAsync(v0.p, r.data(), sizeof(double) * n * K, hipMemcpyHostToDevice, S.s));
    double* vk = S.bw[0].p;
    double* vkm = S.bw[1].p;
    k_copy(n * K, v0.p, vk, S.s);
    k_fill(n * K, 0.0, S.bw[5].p, S.s);
    k_block_dots(n, K, v0.p, vk, mom.p, work.p, S.s);
    for (int k = 1; k <= mdeg; ++k) {
      const double t = (k == 1 ? 1.0 : 2.0) / dd;
      const double* in = vk;
      if (S.geneig) {
        block_bsolve(S, K, vk, S.bw[2].p, S.bw[3].p);
        in = S.bw[2].p;
      }
      block_filter_step(S, K, in, vk, vkm, S.bw[5].p, zmu.p, t, cc, k == 1);
      std::swap(vk, vkm);
      k_block_dots(n, K, v0.p, vk, mom.p + (size_t)k * K, work.p, S.s);
    }
    NMG_HIP(hipMemcpyAsync(h.data(), mom.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, S.s));
    NMG_HIP(hipStreamSynchronize(S.s));
    for (int k = 0; k <= mdeg; ++k)
      for (int c = 0; c < K; ++c) tau[k] += h[(size_t)k * K + c] / nvec;
  }
}

static void do_residuals(Solver& S, double* res_ref, double* res2) {
  const int64_t n = S.n;
  setup_work(S);
  DVec out;
  out.ensure(2 * std::max(S.nev, 1));
  for (int j = 0; j < S.nev; ++j) {
    const double* x = S.X + (size_t)j * n;
    apply_A(S, x, S.w[5].p);
    apply_B(S, x, S.w[6].p);
    k_resid(n, S.w[5].p, S.w[6].p, x, S.lam[j], out.p + 2 * j, S.red.p, S.s);
  }
  std::vector<double> h(2 * std::max(S.nev, 1));
  NMG_HIP(hipMemcpyAsync(h.data(), out.p, sizeof(double) * 2 * std::max(S.nev, 1), hipMemcpyDeviceToHost, S.s));
  NMG_HIP(hipStreamSynchronize(S.s));
  for (int j = 0; j < S.nev; ++j) {
    const double al = std::fabs(S.lam[j]) > 0 ? std::fabs(S.lam[j]) : 1.0;
    if (res_ref) res_ref[j] = std::sqrt(h[2 * j] / (double)n) / al;
    if (res2) res2[j] = std::sqrt(h[2 * j] / h[2 * j + 1]) / al;
  }
}

}  // namespace nmg

using namespace nmg;

// ===========================================================================
extern "C" {

const char* nmg_last_error(void) { return g_err.c_str(); }

int nmg_version(int* major, int* minor) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  return NMG_OK;
}

int nmg_device_info(char* name, int name_len, int* cu_count, double* hbm_gbytes) {
  return guarded([&] {
    int dev = 0;
    NMG_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    NMG_HIP(hipGetDeviceProperties(&p, dev));
    if (name && name_len > 0) {
      std::snprintf(name, name_len, "%s", p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (hbm_gbytes) *hbm_gbytes = (double)p.totalGlobalMem / 1e9;
  });
}

int nmg_start(int comm, nmg_handle* h) {
  (void)comm;
  return guarded([&] {
    NMG_CHECK(h != nullptr, NMG_EINVAL, "null handle pointer");
    std::unique_ptr<Solver> S(new Solver());
    NMG_HIP(hipGetDevice(&S->device));
    NMG_HIP(hipStreamCreateWithFlags(&S->s, hipStreamNonBlocking));
    NMG_LIB(rocblas_create_handle(&S->blas), "rocblas handle");
    NMG_LIB(rocblas_set_stream(S->blas, S->s), "rocblas stream");
    (void)hipGetLastError();
    *h = reg(S.release(), K_SOLVER);
  });
}

int nmg_finish(nmg_handle h) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    (void)hipStreamSynchronize(S->s);
    unreg(S, K_SOLVER);
    delete S;
  });
}

int nmg_setprobsizes(nmg_handle h, int64_t n_glob, int64_t n_loc, int64_t nfirst) {
  (void)nfirst;
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(n_glob > 0 && n_loc == n_glob, NMG_EINVAL,
              "one device holds all rows: n_loc must equal n_glob and be positive");
    S->n = n_loc;
  });
}

int nmg_csr_create(int64_t nrow, int64_t ncol, const int64_t* rowptr, const int32_t* col,
                   const double* val, nmg_handle* mat) {
  return guarded([&] {
    NMG_CHECK(rowptr && col && val && mat, NMG_EINVAL, "null argument");
    DevMat* M = make_devmat(nrow, ncol, rowptr, col, val, true);
    *mat = reg(M, K_MAT);
  });
}

int nmg_parcsr_create(int nrow_g, int ncol_g, const int* row_starts, const int* col_starts,
                      const int* ia, const int* ja, const double* a, int comm, nmg_handle* mat) {
  (void)comm;
  return guarded([&] {
    NMG_CHECK(ia && ja && a && mat, NMG_EINVAL, "null argument");
    if (row_starts)
      NMG_CHECK(row_starts[0] == 0 && row_starts[1] == nrow_g, NMG_EINVAL,
                "row_starts must describe one owner of all rows");
    if (col_starts)
      NMG_CHECK(col_starts[0] == 0 && col_starts[1] == ncol_g, NMG_EINVAL,
                "col_starts must describe one owner of all columns");
    std::vector<int64_t> rp(nrow_g + 1);
    for (int i = 0; i <= nrow_g; ++i) rp[i] = ia[i];
    DevMat* M = make_devmat(nrow_g, ncol_g, rp.data(), (const int32_t*)ja, a, true);
    *mat = reg(M, K_MAT);
  });
}

int nmg_csr_free(nmg_handle mat) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    unreg(M, K_MAT);
    delete M;
  });
}

int nmg_csr_info(nmg_handle mat, int64_t* nrow, int64_t* ncol, int64_t* nnz, int* layout) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    if (nrow) *nrow = M->nrow;
    if (ncol) *ncol = M->ncol;
    if (nnz) *nnz = M->nnz;
    if (layout) *layout = M->layout;
  });
}

int nmg_parcsr_matvec(const double* x, double* y, nmg_handle mat) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    NMG_CHECK(x && y, NMG_EINVAL, "null vector");
    hipStream_t s = default_stream();
    DVec dx, dy;
    dx.ensure(M->ncol);
    dy.ensure(M->nrow);
    NMG_HIP(hipMemcpyAsync(dx.p, x, sizeof(double) * M->ncol, hipMemcpyHostToDevice, s));
    k_spmv(*M, dx.p, dy.p, s);
    NMG_HIP(hipMemcpyAsync(y, dy.p, sizeof(double) * M->nrow, hipMemcpyDeviceToHost, s));
    NMG_HIP(hipStreamSynchronize(s));
  });
}

int nmg_csr_diagscale(nmg_handle mat, double sign, double* d) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    NMG_CHECK(M->nrow == M->ncol, NMG_EINVAL, "diagonal scaling needs a square matrix");
    NMG_CHECK(sign == 1.0 || sign == -1.0, NMG_EINVAL, "sign must be +1 or -1");
    hipStream_t s = default_stream();
    const int64_t n = M->nrow;
    std::vector<double> hd(n);
    DVec dd;
    dd.ensure(n);
    diag_devmat(*M, dd.p, s);
    NMG_HIP(hipMemcpyAsync(hd.data(), dd.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    NMG_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < n; ++i) {
      const double v = sign * hd[i];
      NMG_CHECK(v > 0.0 && std::isfinite(v), NMG_ENUM,
                "diagonal entry " + std::to_string(i) + " is not positive after the sign: cannot scale");
      hd[i] = 1.0 / std::sqrt(v);
    }
    if (M->layout == L_MASS3)
      for (int64_t v = 0; v < n / 3; ++v)
        NMG_CHECK(hd[3 * v] == hd[3 * v + 1] && hd[3 * v] == hd[3 * v + 2], NMG_EINVAL,
                  "mass layout with component-dependent diagonal");
    NMG_HIP(hipMemcpyAsync(dd.p, hd.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    scale_devmat(*M, dd.p, dd.p, sign, s);
    M->touch();
    NMG_HIP(hipStreamSynchronize(s));
    if (d) std::memcpy(d, hd.data(), sizeof(double) * n);
  });
}

int nmg_set_amv_dad(nmg_handle h, nmg_handle matA, const double* d) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    DevMat* M = get<DevMat>(matA, K_MAT);
    NMG_CHECK(S->n > 0 && M->nrow == S->n && M->ncol == S->n, NMG_EINVAL, "A size mismatch");
    NMG_CHECK(d != nullptr, NMG_EINVAL, "null scaling vector");
    for (int64_t i = 0; i < S->n; ++i)
      NMG_CHECK(std::isfinite(d[i]) && d[i] > 0.0, NMG_EINVAL, "scaling vector must be positive");
    std::unique_ptr<DevMat> C(clone_devmat(*M, S->s));
    DVec dd;
    dd.ensure(S->n);
    NMG_HIP(hipMemcpyAsync(dd.p, d, sizeof(double) * S->n, hipMemcpyHostToDevice, S->s));
    scale_devmat(*C, dd.p, dd.p, 1.0, S->s);
    NMG_HIP(hipStreamSynchronize(S->s));
    S->adad = std::move(C);
    S->amat = nullptr;
    S->fs.reset();
    S->dscale.ensure(S->n);
    NMG_HIP(hipMemcpyAsync(S->dscale.p, d, sizeof(double) * S->n, hipMemcpyHostToDevice, S->s));
    NMG_HIP(hipStreamSynchronize(S->s));
  });
}

int nmg_set_amv_fs(nmg_handle h, nmg_handle matAd, nmg_handle matE, nmg_handle matET, nmg_handle chebAp,
                   const double* d, const double* dp) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    auto F = std::make_shared<FsOp>();
    F->Ad = get<DevMat>(matAd, K_MAT);
    F->E = get<DevMat>(matE, K_MAT);
    F->ET = get<DevMat>(matET, K_MAT);
    F->cp = get<Cheb>(chebAp, K_CHEB);
    const int64_t n = S->n, np = F->cp->mat->nrow;
    NMG_CHECK(n > 0 && F->Ad->nrow == n && F->Ad->ncol == n, NMG_EINVAL, "Ad must be n x n");
    NMG_CHECK(F->E->nrow == n && F->E->ncol == np, NMG_EINVAL, "E must be n x np (np = rows of Ap)");
    NMG_CHECK(F->ET->nrow == np && F->ET->ncol == n, NMG_EINVAL, "ET must be np x n");
    NMG_CHECK(d != nullptr && dp != nullptr, NMG_EINVAL, "null scaling vector");
    for (int64_t i = 0; i < n; ++i)
      NMG_CHECK(std::isfinite(d[i]) && d[i] > 0.0, NMG_EINVAL, "scaling vector must be positive");
    for (int64_t i = 0; i < np; ++i)
      NMG_CHECK(std::isfinite(dp[i]) && dp[i] > 0.0, NMG_EINVAL, "pressure scaling vector must be positive");
    F->n = n;
    F->np = np;
    F->d.ensure(n);
    F->dp.ensure(np);
    NMG_HIP(hipMemcpyAsync(F->d.p, d, sizeof(double) * n, hipMemcpyHostToDevice, S->s));
    NMG_HIP(hipMemcpyAsync(F->dp.p, dp, sizeof(double) * np, hipMemcpyHostToDevice, S->s));
    S->dscale.ensure(n);
    NMG_HIP(hipMemcpyAsync(S->dscale.p, d, sizeof(double) * n, hipMemcpyHostToDevice, S->s));
    NMG_HIP(hipStreamSynchronize(S->s));
    S->fs = std::move(F);
    S->adad.reset();
    S->amat = nullptr;
  });
}

int nmg_set_amv(nmg_handle h, nmg_handle mat) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    DevMat* M = get<DevMat>(mat, K_MAT);
    NMG_CHECK(S->n > 0 && M->nrow == S->n && M->ncol == S->n, NMG_EINVAL, "operator size mismatch");
    S->amat = M;
    S->adad.reset();
    S->fs.reset();
  });
}

int nmg_set_bmv(nmg_handle h, nmg_handle matB) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    DevMat* M = get<DevMat>(matB, K_MAT);
    NMG_CHECK(S->n > 0 && M->nrow == S->n && M->ncol == S->n, NMG_EINVAL, "B size mismatch");
    S->bmat = M;
  });
}

int nmg_set_bsol_chebiter(nmg_handle h, int type, nmg_handle cheb) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    Cheb* C = get<Cheb>(cheb, K_CHEB);
    NMG_CHECK(type == 1 || type == 2, NMG_EINVAL, "chebiter type must be 1 or 2");
    NMG_CHECK(S->n > 0 && C->mat->nrow == S->n, NMG_EINVAL, "B-solve size mismatch");
    S->bsol = C;
  });
}

int nmg_set_geneig(nmg_handle h) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(S->bmat && S->bsol, NMG_ESTATE, "set B and its solve before the generalized flag");
    S->geneig = true;
  });
}

int nmg_set_precision(nmg_handle h, int mode) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(mode == NMG_FP64 || mode == NMG_FP32_FILTER || mode == NMG_FP32_BSOLVE || mode == NMG_FP32_FILTER_A,
              NMG_EINVAL, "unknown precision");
    if (mode != NMG_FP64) NMG_CHECK(S->bsol != nullptr, NMG_ESTATE, "set the B-solve before the precision");
    if (mode == NMG_FP32_FILTER_A)
      NMG_CHECK(S->fs || S->adad || S->amat, NMG_ESTATE, "set the operator before the precision");
    S->precision = mode;
    S->drop_graphs();
    if (mode != NMG_FP64) {
      // B~ is well conditioned (kappa ~ 15 at P2): its FP32 values cost the
      // B-solve ~1e-7 relative.  FP32 stiffness values (NMG_FP32_FILTER_A)
      // perturb the low modes by ~eps32 lambda_max / lambda instead
      make_fp32(*S->bsol->mat, S->s);
      if (mode == NMG_FP32_FILTER_A) {
        if (S->fs) {
          for (DevMat* M : {S->fs->Ad, S->fs->E, S->fs->ET, S->fs->cp->mat})
            if (!(M->val32 || M->bval32)) make_fp32(*M, S->s);
        } else {
          DevMat& A = const_cast<DevMat&>(S->A());
          if (!(A.val32 || A.bval32)) make_fp32(A, S->s);
        }
      }
      NMG_HIP(hipStreamSynchronize(S->s));
    }
  });
}

int nmg_set_seed(nmg_handle h, uint64_t seed) {
  return guarded([&] { get<Solver>(h, K_SOLVER)->seed = seed; });
}

int nmg_set_lanczos_controls(nmg_handle h, int ntest, int cycle, double rtol) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(ntest >= 2 && cycle >= 1 && rtol > 0.0, NMG_EINVAL, "bad Lanczos controls");
    S->ntest = ntest;
    S->cycle = cycle;
    S->rtol = rtol;
  });
}

int nmg_setup_chebiter(double lmin, double lmax, int deg, nmg_handle mat, nmg_handle* cheb) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    NMG_CHECK(cheb != nullptr, NMG_EINVAL, "null handle pointer");
    NMG_CHECK(lmin > 0.0 && lmax > lmin && std::isfinite(lmax), NMG_EINVAL,
              "chebiter needs 0 < lmin < lmax (an SPD matrix)");
    NMG_CHECK(deg >= 1 && deg <= 100000, NMG_EINVAL, "chebiter degree must be >= 1");
    NMG_CHECK(M->nrow == M->ncol, NMG_EINVAL, "chebiter needs a square matrix");
    Cheb* C = new Cheb();
    C->mat = M;
    C->lmin = lmin;
    C->lmax = lmax;
    C->deg = deg;
    *cheb = reg(C, K_CHEB);
  });
}

int nmg_chebiter(int type, const double* b, double* x, nmg_handle cheb) {
  return guarded([&] {
    Cheb* C = get<Cheb>(cheb, K_CHEB);
    NMG_CHECK(type == 1 || type == 2, NMG_EINVAL, "chebiter type must be 1 or 2");
    NMG_CHECK(b && x, NMG_EINVAL, "null vector");
    hipStream_t s = default_stream();
    const int64_t n = C->mat->nrow;
    DVec db, d1, d2;
    db.ensure(n);
    d1.ensure(n);
    d2.ensure(n);
    NMG_HIP(hipMemcpyAsync(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice, s));
    const double t0 = now();
    cheb_solve(*C, db.p, d1.p, d2.p, s, false);
    NMG_HIP(hipMemcpyAsync(x, d1.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    NMG_HIP(hipStreamSynchronize(s));
    C->seconds += now() - t0;
    C->ncalls++;
    C->nmatvec += C->deg;
  });
}

// ---------------------------------------------------------------------------
// Allocation-free Chebyshev solve.  Its one scratch vector, the second iterate
// of the recurrence, is a fixed device array of kScratchDoubles doubles (32 MiB
// on every device the library runs on, held until the process ends).  Host
// code never passes the array's name to a kernel: the name is a host-side
// stand-in; its device address is looked up once per device with
// hipGetSymbolAddress and kept in scratch_ptr.
constexpr int64_t kScratchDoubles = (int64_t)1 << 22;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kScratchDoubles];

static double* scratch_on_current_device() {
  static std::mutex mu;
  static double* scratch_ptr[kMaxDevices] = {};
  int dev = -1;
  NMG_HIP(hipGetDevice(&dev));
  NMG_CHECK(dev >= 0 && dev < kMaxDevices, NMG_EHIP, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    NMG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    NMG_CHECK(p != nullptr, NMG_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
    scratch_ptr[dev] = static_cast<double*>(p);
  }
  return scratch_ptr[dev];
}

int nmg_chebiter_noalloc(int type, const double* b, double* x, nmg_handle cheb, void* stream) {
  return guarded([&] {
    Cheb* C = get<Cheb>(cheb, K_CHEB);
    NMG_CHECK(type == 1 || type == 2, NMG_EINVAL, "chebiter type must be 1 or 2");
    NMG_CHECK(b && x, NMG_EINVAL, "null vector");
    NMG_CHECK(b != x, NMG_EINVAL, "b and x must be different device vectors");
    const int64_t n = C->mat->nrow;
    NMG_CHECK(n <= kScratchDoubles, NMG_EINVAL,
              "nmg_chebiter_noalloc: the matrix has more rows than the scratch array holds (4194304)");
    double* tmp = scratch_on_current_device();
    cheb_steps(*C, b, x, tmp, static_cast<hipStream_t>(stream), false, nullptr);
    NMG_HIP(hipGetLastError());
    C->ncalls++;
    C->nmatvec += C->deg;
  });
}

int nmg_chebiter_stats(nmg_handle cheb, int64_t* ncalls, int64_t* nmatvec, double* seconds) {
  return guarded([&] {
    Cheb* C = get<Cheb>(cheb, K_CHEB);
    if (ncalls) *ncalls = C->ncalls;
    if (nmatvec) *nmatvec = C->nmatvec;
    if (seconds) *seconds = C->seconds;
  });
}

int nmg_chebiter_free(nmg_handle cheb) {
  return guarded([&] {
    Cheb* C = get<Cheb>(cheb, K_CHEB);
    unreg(C, K_CHEB);
    delete C;
  });
}

int nmg_lanbounds(nmg_handle h, int mlan, int maxit, double tol, double* lmin, double* lmax) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(lmin && lmax, NMG_EINVAL, "null output");
    NMG_CHECK(tol > 0.0, NMG_EINVAL, "tol must be positive");
    do_lanbounds(*S, mlan, maxit, tol, lmin, lmax);
  });
}

int nmg_findpol(const double* xintv, double thr_int, double thr_ext, nmg_handle* pol) {
  return guarded([&] {
    NMG_CHECK(xintv && pol, NMG_EINVAL, "null argument");
    NMG_CHECK(thr_int > 0 && thr_int < 1 && thr_ext > 0 && thr_ext < 1, NMG_EINVAL,
              "thresholds must lie in (0, 1)");
    std::unique_ptr<PolyFilter> P(new PolyFilter());
    int rc = find_pol(xintv, thr_int, thr_ext, 2, 10000, 2, *P);
    NMG_CHECK(rc != NMG_EINVAL, NMG_EINVAL, "interval [a,b] does not intersect [lmin,lmax]");
    NMG_CHECK(rc == NMG_OK, NMG_ENUM, "no filter of degree <= 10000 meets the threshold");
    *pol = reg(P.release(), K_POL);
  });
}

int nmg_pol_info(nmg_handle pol, int* deg, double* cc, double* dd, double* gam, double* bar, double* mu) {
  return guarded([&] {
    PolyFilter* P = get<PolyFilter>(pol, K_POL);
    if (deg) *deg = P->deg;
    if (cc) *cc = P->cc;
    if (dd) *dd = P->dd;
    if (gam) *gam = P->gam;
    if (bar) *bar = P->bar;
    if (mu) std::memcpy(mu, P->mu.data(), sizeof(double) * (P->deg + 1));
  });
}

int nmg_freepol(nmg_handle pol) {
  return guarded([&] {
    PolyFilter* P = get<PolyFilter>(pol, K_POL);
    unreg(P, K_POL);
    delete P;
  });
}

int nmg_cheblannr(nmg_handle h, const double* xintv, int maxit, double tol, nmg_handle pol) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    PolyFilter* P = get<PolyFilter>(pol, K_POL);
    NMG_CHECK(xintv != nullptr, NMG_EINVAL, "null interval");
    NMG_CHECK(maxit >= 2 && tol > 0.0, NMG_EINVAL, "bad maxit/tol");
    S->check_A();
    if (S->block > 1) do_cheblannr_block(*S, xintv, maxit, tol, *P);
    else do_cheblannr(*S, xintv, maxit, tol, *P);
  });
}

int nmg_get_nev(nmg_handle h, int* nev) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(nev != nullptr, NMG_EINVAL, "null output");
    *nev = S->nev;
  });
}

int nmg_copy_result(nmg_handle h, double* val, double* vec, int64_t n_loc) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(n_loc == S->n, NMG_EINVAL, "n_loc does not match the problem size");
    if (val) std::memcpy(val, S->lam.data(), sizeof(double) * S->nev);
    if (vec && S->nev > 0) {
      NMG_HIP(hipMemcpyAsync(vec, S->X, sizeof(double) * S->n * S->nev, hipMemcpyDeviceToHost, S->s));
      NMG_HIP(hipStreamSynchronize(S->s));
    }
  });
}

int nmg_residuals(nmg_handle h, double* res_ref, double* res2) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    do_residuals(*S, res_ref, res2);
  });
}

static int backward_impl(nmg_handle h, int64_t n_elem, int np, const int32_t* elems, const double* detJ,
                         const double* invJ, const double* refM, const double* refD, const double* gnode,
                         const double* gsym, double* d_rho, double* d_lam, double* d_mu) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(S->nev > 0 && S->X, NMG_ESTATE, "no eigenpairs: run nmg_cheblannr first");
    NMG_CHECK(S->dscale.n >= S->n, NMG_ESTATE, "the operator was not set with nmg_set_amv_dad");
    NMG_CHECK(!S->fs, NMG_ENOTIMPL,
              "a fluid-solid body's sensitivities come from nmg_backward_fs (its element layout and pressures)");
    NMG_CHECK(np == 4 || np == 10, NMG_EINVAL, "np must be 4 (P1) or 10 (P2)");
    NMG_CHECK(n_elem > 0 && elems && detJ && invJ && refM && refD && d_rho && d_lam && d_mu, NMG_EINVAL,
              "null argument");
    NMG_CHECK((gnode == nullptr) == (gsym == nullptr), NMG_EINVAL, "gravity and its gradients go together");
    const int64_t nn = S->n / 3;
    for (int64_t k = 0; k < n_elem * np; ++k)
      NMG_CHECK(elems[k] >= 0 && elems[k] < nn, NMG_EINVAL, "element node id out of range");
    const int nev = S->nev;
    DVec dE, dJ, dI, dM, dD, dlam, gr, gl, gm, dg, dgs;
    dE.ensure((n_elem * np + 1) / 2 + 1);
    dJ.ensure(n_elem);
    dI.ensure(9 * n_elem);
    dM.ensure(np * np);
    dD.ensure(3 * np * np);
    dlam.ensure(nev);
    gr.ensure(n_elem * nev);
    gl.ensure(n_elem * nev);
    gm.ensure(n_elem * nev);
    NMG_HIP(hipMemcpyAsync(dE.p, elems, sizeof(int32_t) * n_elem * np, hipMemcpyHostToDevice, S->s));
    NMG_HIP(hipMemcpyAsync(dJ.p, detJ, sizeof(double) * n_elem, hipMemcpyHostToDevice, S->s));
    NMG_HIP(hipMemcpyAsync(dI.p, invJ, sizeof(double) * 9 * n_elem, hipMemcpyHostToDevice, S->s));
    NMG_HIP(hipMemcpyAsync(dM.p, refM, sizeof(double) * np * np, hipMemcpyHostToDevice, S->s));
    NMG_HIP(hipMemcpyAsync(dD.p, refD, sizeof(double) * 3 * np * np, hipMemcpyHostToDevice, S->s));
    NMG_HIP(hipMemcpyAsync(dlam.p, S->lam.data(), sizeof(double) * nev, hipMemcpyHostToDevice, S->s));
    if (gnode) {
      dg.ensure(n_elem * np * 3);
      dgs.ensure(n_elem * 9);
      NMG_HIP(hipMemcpyAsync(dg.p, gnode, sizeof(double) * n_elem * np * 3, hipMemcpyHostToDevice, S->s));
      NMG_HIP(hipMemcpyAsync(dgs.p, gsym, sizeof(double) * n_elem * 9, hipMemcpyHostToDevice, S->s));
    }
    k_adjoint(np, n_elem, nev, (const int32_t*)dE.p, dJ.p, dI.p, dM.p, dD.p, S->X, S->n, S->dscale.p, dlam.p,
              gnode ? dg.p : nullptr, gnode ? dgs.p : nullptr, gr.p, gl.p, gm.p, S->s);
    NMG_HIP(hipMemcpyAsync(d_rho, gr.p, sizeof(double) * n_elem * nev, hipMemcpyDeviceToHost, S->s));
    NMG_HIP(hipMemcpyAsync(d_lam, gl.p, sizeof(double) * n_elem * nev, hipMemcpyDeviceToHost, S->s));
    NMG_HIP(hipMemcpyAsync(d_mu, gm.p, sizeof(double) * n_elem * nev, hipMemcpyDeviceToHost, S->s));
    NMG_HIP(hipStreamSynchronize(S->s));
  });
}

int nmg_backward(nmg_handle h, int64_t n_elem, int np, const int32_t* elems, const double* detJ,
                 const double* invJ, const double* refM, const double* refD, int job, double* d_rho,
                 double* d_lam, double* d_mu) {
  if (job != 1)
    return guarded([&] {
      NMG_CHECK(false, NMG_EINVAL, "job 2 (reference gravity) needs the gravity field: use nmg_backward_grav");
    });
  return backward_impl(h, n_elem, np, elems, detJ, invJ, refM, refD, nullptr, nullptr, d_rho, d_lam, d_mu);
}

int nmg_backward_grav(nmg_handle h, int64_t n_elem, int np, const int32_t* elems, const double* detJ,
                      const double* invJ, const double* refM, const double* refD, const double* gnode,
                      const double* gsym, double* d_rho, double* d_lam, double* d_mu) {
  if (!gnode || !gsym)
    return guarded([&] { NMG_CHECK(false, NMG_EINVAL, "null gravity argument"); });
  return backward_impl(h, n_elem, np, elems, detJ, invJ, refM, refD, gnode, gsym, d_rho, d_lam, d_mu);
}

int nmg_kpmdos(nmg_handle h, int mdeg, int nvec, const double* intv, double* mu, double* ecnt) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(intv && mu, NMG_EINVAL, "null argument");
    std::vector<double> tau;
    do_kpm(*S, mdeg, nvec, intv[2], intv[3], tau);
    std::memcpy(mu, tau.data(), sizeof(double) * (mdeg + 1));
    if (ecnt) *ecnt = dos_count(mdeg, tau.data(), intv[2], intv[3], intv[0], intv[1]);
  });
}

int nmg_dos_count(const double* mu, int mdeg, const double* intv, double* count) {
  return guarded([&] {
    NMG_CHECK(mu && intv && count && mdeg >= 1, NMG_EINVAL, "bad argument");
    NMG_CHECK(intv[3] > intv[2], NMG_EINVAL, "lmax must exceed lmin");
    *count = dos_count(mdeg, mu, intv[2], intv[3], intv[0], intv[1]);
  });
}

int nmg_spslicer(double* sli, const double* mu, int mdeg, const double* intv, int n_slices) {
  return guarded([&] {
    NMG_CHECK(sli && mu && intv && mdeg >= 1, NMG_EINVAL, "bad argument");
    NMG_CHECK(intv[3] > intv[2] && intv[1] > intv[0], NMG_EINVAL, "bad interval");
    NMG_CHECK(n_slices >= 1 && n_slices <= 64, NMG_EINVAL, "between 1 and 64 slices");
    const int rc = dos_slicer(mdeg, mu, intv[2], intv[3], intv[0], intv[1], n_slices, sli);
    NMG_CHECK(rc == NMG_OK, NMG_ENUM, "the density estimate gives no increasing slice boundaries");
  });
}

int nmg_solve_slices(nmg_handle h, int n_slices, const double* sli, double lmin, double lmax, double thr_int,
                     double thr_ext, int maxit, double tol) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(sli != nullptr, NMG_EINVAL, "null slice boundaries");
    NMG_CHECK(maxit >= 2 && tol > 0.0, NMG_EINVAL, "bad maxit/tol");
    NMG_CHECK(lmax > lmin, NMG_EINVAL, "lmax must exceed lmin");
    NMG_CHECK(thr_int > 0 && thr_int < 1 && thr_ext > 0 && thr_ext < 1, NMG_EINVAL,
              "thresholds must lie in (0, 1)");
    S->check_A();
    do_solve_slices(*S, n_slices, sli, lmin, lmax, thr_int, thr_ext, maxit, tol);
  });
}

int nmg_slice_counts(nmg_handle h, int* counts) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(counts != nullptr, NMG_EINVAL, "null output");
    for (size_t i = 0; i < S->slice_nev.size(); ++i) counts[i] = S->slice_nev[i];
  });
}

int nmg_solve_stats(nmg_handle h, int* lan_steps, int64_t* filter_applies, int* ncand, double* seconds) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    if (lan_steps) *lan_steps = S->lan_steps;
    if (filter_applies) *filter_applies = S->filter_applies;
    if (ncand) *ncand = S->ncand;
    if (seconds) *seconds = S->seconds;
  });
}

int nmg_solve_quality(nmg_handle h, int* target_met, double* resid_max) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    if (target_met) *target_met = S->target_met;
    if (resid_max) *resid_max = S->resid_best;
  });
}

int nmg_set_residual_target(nmg_handle h, double target) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(target >= 0.0 && std::isfinite(target), NMG_EINVAL, "target must be >= 0");
    S->resid_target = target;
  });
}

int nmg_set_split_streams(nmg_handle h, int parts) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(parts >= 0 && parts <= 4 && parts != 3, NMG_EINVAL, "parts must be 0, 1, 2 or 4");
    S->split = parts;
  });
}

int nmg_set_split_filter(nmg_handle h, int on) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(on == 0 || on == 1, NMG_EINVAL, "on must be 0 or 1");
    S->split_filter = on != 0;
  });
}

int nmg_set_verbose(nmg_handle h, int every) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(every >= 0, NMG_EINVAL, "every must be >= 0");
    S->verbose = every;
  });
}

int nmg_set_graphs(nmg_handle h, int on) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(on == 0 || on == 1, NMG_EINVAL, "on must be 0 or 1");
    S->drop_graphs();
    S->graphs = on != 0;
  });
}

int nmg_set_block_size(nmg_handle h, int block) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(block == 1 || block == 2 || block == 4 || block == 8 || block == 16 || block == 32 || block == 64,
              NMG_EINVAL, "block size must be 1, 2, 4, 8, 16, 32 or 64");
    S->block = block;
  });
}

int nmg_set_kernel_probe(nmg_handle h, int every) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(every >= 0, NMG_EINVAL, "every must be >= 0");
    for (auto& p : S->probe) {
      p.every = every;
      p.reset();
    }
  });
}

int nmg_kernel_probe(nmg_handle h, int which, double* mean_seconds, int64_t* samples,
                     double* bytes_per_launch) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(which == 0 || which == 1, NMG_EINVAL, "which must be 0 (B-solve step) or 1 (filter step)");
    NMG_HIP(hipStreamSynchronize(S->s));
    KernelProbe& p = S->probe[which];
    if (mean_seconds) *mean_seconds = p.mean_seconds(samples) / std::max(p.launches, 1);
    if (bytes_per_launch) {
      // matrix once per launch; vector streams per column: Chebyshev step
      // x, b, xm read, xm write (4); filter step x, vk, vkm r/w, y r/w (6)
      const bool f32 = S->fp32();
      const double nk = (double)S->n * (double)S->block;
      if (which == 0) {
        NMG_CHECK(S->bsol != nullptr, NMG_ESTATE, "no B-solve set");
        // the launch that was sampled: one step, or a whole persistent solve
        *bytes_per_launch = p.bytes > 0.0 ? p.bytes : S->bsol->mat->matrix_bytes(f32) + 4.0 * 8.0 * nk;
      } else {
        // the fluid-solid operator is not one fused launch: no figure
        *bytes_per_launch = S->fs ? 0.0 : S->A().matrix_bytes(false) + 6.0 * 8.0 * nk;
      }
    }
  });
}

int nmg_filter_apply(nmg_handle h, nmg_handle pol, const double* z, double* y) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    PolyFilter* P = get<PolyFilter>(pol, K_POL);
    NMG_CHECK(z && y, NMG_EINVAL, "null vector");
    setup_work(*S);
    DVec dz, dy;
    dz.ensure(S->n);
    dy.ensure(S->n);
    NMG_HIP(hipMemcpyAsync(dz.p, z, sizeof(double) * S->n, hipMemcpyHostToDevice, S->s));
    filter_apply(*S, *P, dz.p, dy.p);
    NMG_HIP(hipMemcpyAsync(y, dy.p, sizeof(double) * S->n, hipMemcpyDeviceToHost, S->s));
    NMG_HIP(hipStreamSynchronize(S->s));
  });
}

int nmg_bench_filter(nmg_handle h, nmg_handle pol, int K, int reps, double* seconds) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    PolyFilter* P = get<PolyFilter>(pol, K_POL);
    NMG_CHECK(seconds != nullptr, NMG_EINVAL, "null output");
    NMG_CHECK(K >= 1 && K <= 64 && (K & (K - 1)) == 0 && reps >= 1, NMG_EINVAL, "K must be 1..64, a power of 2");
    setup_work(*S);
    const int64_t nk = S->n * K;
    DVec dz, dy, mu;
    dz.ensure(nk);
    dy.ensure(nk);
    mu.ensure((int64_t)(P->deg + 1) * K);
    {
      std::vector<double> m((size_t)(P->deg + 1) * K), z((size_t)nk);
      for (int k = 0; k <= P->deg; ++k)
        for (int c = 0; c < K; ++c) m[(size_t)k * K + c] = P->mu[k];
      uint64_t st = 0x9E3779B97F4A7C15ULL;
      for (int64_t i = 0; i < nk; ++i) {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        z[(size_t)i] = (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
      }
      NMG_HIP(hipMemcpyAsync(mu.p, m.data(), sizeof(double) * m.size(), hipMemcpyHostToDevice, S->s));
      NMG_HIP(hipMemcpyAsync(dz.p, z.data(), sizeof(double) * z.size(), hipMemcpyHostToDevice, S->s));
      NMG_HIP(hipStreamSynchronize(S->s));
    }
    auto once = [&] {
      if (K == 1) filter_apply(*S, *P, dz.p, dy.p);
      else block_filter(*S, K, *P, mu.p, dz.p, dy.p);
    };
    once();  // warm: work buffers, streams, FP32 copies
    NMG_HIP(hipStreamSynchronize(S->s));
    const auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; ++r) once();
    NMG_HIP(hipStreamSynchronize(S->s));
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / reps;
  });
}

int nmg_apply(nmg_handle h, int which, const double* x, double* y) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(x && y, NMG_EINVAL, "null vector");
    NMG_CHECK(which >= 0 && which <= 2, NMG_EINVAL, "which must be 0 (A), 1 (B) or 2 (B-solve)");
    setup_work(*S);
    DVec dx, dy;
    dx.ensure(S->n);
    dy.ensure(S->n);
    NMG_HIP(hipMemcpyAsync(dx.p, x, sizeof(double) * S->n, hipMemcpyHostToDevice, S->s));
    if (which == 0) {
      if (S->fs) apply_A(*S, dx.p, dy.p);
      else k_spmv(S->A(), dx.p, dy.p, S->s, S->fp32());
    }
    else if (which == 1) apply_B(*S, dx.p, dy.p);
    else {
      NMG_CHECK(S->bsol != nullptr, NMG_ESTATE, "no B-solve set");
      bool g = S->geneig;
      S->geneig = true;
      bsolve(*S, dx.p, dy.p, S->w[7].p);
      S->geneig = g;
    }
    NMG_HIP(hipMemcpyAsync(y, dy.p, sizeof(double) * S->n, hipMemcpyDeviceToHost, S->s));
    NMG_HIP(hipStreamSynchronize(S->s));
  });
}

// MVFunc-shaped callbacks: host x in, host y out, through the solver's
// device operators and two staging vectors kept in the solver.  pEVSL's
// MVFunc returns nothing, so a failure leaves y filled with NaN (when the
// handle is valid) and the message in nmg_last_error().
static void mvfunc(int which, double* x, double* y, void* data) {
  Solver* S = nullptr;
  const int rc = guarded([&] {
    NMG_CHECK(data != nullptr, NMG_EINVAL, "MVFunc data must point at a solver handle");
    S = get<Solver>(*static_cast<const nmg_handle*>(data), K_SOLVER);
    NMG_CHECK(x && y, NMG_EINVAL, "null vector");
    setup_work(*S);
    S->mvx.ensure(S->n);
    S->mvy.ensure(S->n);
    NMG_HIP(hipMemcpyAsync(S->mvx.p, x, sizeof(double) * S->n, hipMemcpyHostToDevice, S->s));
    if (which == 0) {
      if (S->fs) apply_A(*S, S->mvx.p, S->mvy.p);
      else k_spmv(S->A(), S->mvx.p, S->mvy.p, S->s, S->fp32());
    } else if (which == 1) {
      apply_B(*S, S->mvx.p, S->mvy.p);
    } else {
      NMG_CHECK(S->bsol != nullptr, NMG_ESTATE, "no B-solve set (nmg_set_bsol_chebiter)");
      const bool g = S->geneig;
      S->geneig = true;
      bsolve(*S, S->mvx.p, S->mvy.p, S->w[7].p);
      S->geneig = g;
    }
    NMG_HIP(hipMemcpyAsync(y, S->mvy.p, sizeof(double) * S->n, hipMemcpyDeviceToHost, S->s));
    NMG_HIP(hipStreamSynchronize(S->s));
  });
  if (rc != NMG_OK && S && y)
    for (int64_t i = 0; i < S->n; ++i) y[i] = std::numeric_limits<double>::quiet_NaN();
}

void nmg_amv(double* x, double* y, void* data) { mvfunc(0, x, y, data); }
void nmg_bmv(double* x, double* y, void* data) { mvfunc(1, x, y, data); }
void nmg_bsol(double* x, double* y, void* data) { mvfunc(2, x, y, data); }

int nmg_apply_block(nmg_handle h, int which, int K, const double* X, double* Y) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(X && Y, NMG_EINVAL, "null block");
    NMG_CHECK(which >= 0 && which <= 2, NMG_EINVAL, "which must be 0 (A), 1 (B) or 2 (B-solve)");
    NMG_CHECK(K == 1 || K == 2 || K == 4 || K == 8 || K == 16 || K == 32 || K == 64, NMG_EINVAL,
              "K must be 1, 2, 4, 8, 16, 32 or 64");
    const int64_t n = S->n;
    setup_work(*S);
    for (int i = 0; i < 6; ++i) S->bw[i].ensure(n * K);
    DVec cm;
    cm.ensure(n * K);
    NMG_HIP(hipMemcpyAsync(cm.p, X, sizeof(double) * n * K, hipMemcpyHostToDevice, S->s));
    k_to_interleaved(n, K, cm.p, n, S->bw[4].p, S->s);
    if (which == 0) {
      if (S->fs) apply_A_block(*S, K, S->bw[4].p, S->bw[5].p);
      else k_block_spmv(K, S->A(), S->bw[4].p, S->bw[5].p, S->s, S->fp32());
    }
    else if (which == 1) {
      NMG_CHECK(S->bmat != nullptr, NMG_ESTATE, "no B set");
      k_block_spmv(K, *S->bmat, S->bw[4].p, S->bw[5].p, S->s);
    } else {
      NMG_CHECK(S->bsol != nullptr, NMG_ESTATE, "no B-solve set");
      block_bsolve(*S, K, S->bw[4].p, S->bw[5].p, S->bw[2].p);
    }
    k_from_interleaved(n, K, S->bw[5].p, cm.p, n, S->s);
    NMG_HIP(hipMemcpyAsync(Y, cm.p, sizeof(double) * n * K, hipMemcpyDeviceToHost, S->s));
    NMG_HIP(hipStreamSynchronize(S->s));
  });
}

int nmg_set_xcd_order(nmg_handle mat, int on) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    M->xcd = on ? 1 : 0;
    M->touch();
  });
}

int nmg_set_value_order(nmg_handle mat, int tiled) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    set_value_order(*M, tiled, default_stream());
  });
}

int nmg_set_stream_loads(nmg_handle mat, int on) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    M->nt = on ? 1 : 0;
    M->touch();
  });
}

int nmg_set_kernel_map(nmg_handle mat, int group, int K, int rows_per_wave) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    NMG_CHECK(group == 0 || group == 8 || group == 16 || group == 32 || group == 64, NMG_EINVAL,
              "group must be 0, 8, 16, 32 or 64");
    M->group = group;
    M->pipe = 0;
    M->touch();
    if (group > 0 || K == 1) build_row_groups(*M, 0);  // the single-column route is set explicitly
    if (K > 0) {
      int idx = -1;
      for (int i = 0; i < 7; ++i)
        if ((1 << i) == K) idx = i;
      NMG_CHECK(idx >= 0, NMG_EINVAL, "K must be 1, 2, 4, 8, 16, 32 or 64");
      NMG_CHECK(rows_per_wave == 0 || rows_per_wave == 1 || rows_per_wave == 2 || rows_per_wave == 4 ||
                    rows_per_wave == 8,
                NMG_EINVAL, "rows per wavefront must be 0, 1, 2, 4 or 8");
      NMG_CHECK(K * std::max(rows_per_wave, 1) <= 64, NMG_EINVAL, "K x rows per wavefront must be <= 64");
      M->block_map[idx] = rows_per_wave;
      M->wide[idx] = 0;
    }
  });
}

}  // extern "C"

namespace nmg {
// ---------------------------------------------------------------------------
// Autotuner of the sparse-product mapping of one matrix for K columns: times
// the candidate lane layouts on the matrix itself (untiled rows-per-wavefront
// 1..8; LDS-tiled at tile capacities on a geometric ladder x rows per
// wavefront 1..8; for K = 1 also the lanes-per-row layout) and keeps the
// fastest.
template <class F>
static double time_launches(hipStream_t s, int reps, F&& f) {
  f();  // warm-up
  hipEvent_t e0, e1;
  NMG_HIP(hipEventCreate(&e0));
  NMG_HIP(hipEventCreate(&e1));
  NMG_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) f();
  NMG_HIP(hipEventRecord(e1, s));
  NMG_HIP(hipEventSynchronize(e1));
  float ms = 0;
  NMG_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms * 1e-3 / reps;
}

struct TuneResult {
  double seconds = 1e30;
  int kind = 0;  // 0 lanes-per-row (K = 1), 1 untiled block, 2 tiled block, 3 row groups (K = 1),
                 // 4 wide (scalar mass, K >= 2), 5 tiled wide (scalar mass, K >= 2),
                 // 6 pipelined lanes-per-row (BSR3, K = 1)
  int group = 0, rows = 1, cap = 0;
  int64_t ntiles = 0;
  int block = 1024;  // tile workgroup threads
};

static TuneResult tune_matrix(DevMat& M, int K, int reps, bool fp32, hipStream_t s) {
  NMG_CHECK(M.layout != L_CSR, NMG_EINVAL, "autotuning applies to the blocked layouts");
  int idx = 0;
  while ((1 << idx) != K) ++idx;
  const int64_t nk = M.nrow * K;
  DVec x, y;
  x.ensure(nk);
  y.ensure(nk);
  k_fill(nk, 1.0, x.p, s);
  // The wide layouts (kinds 4 and 5: one entry per lane, the K columns in
  // registers) are chosen only when they win by 3%. Timed alone they can tie
  // the row layouts within noise, but beside a second stream's products they
  // run slower (51 540 DoF, K = 4, two-stream B-solve: 12.8 against 11.9 us
  // per launch), and the split B-solve is how k-column products run.
  TuneResult best;
  double best_score = 1e30;
  auto consider = [&](double t, int kind, int group, int rows, int cap, int64_t nt, int block = 1024) {
    const double score = (kind == 4 || kind == 5) ? 1.03 * t : t;
    if (score < best_score) {
      best_score = score;
      best = TuneResult{t, kind, group, rows, cap, nt, block};
    }
  };
  // untiled layouts (the single-column route is K = 1's choice alone: tuning
  // another width leaves it as it is)
  build_tiling(M, K, 0, 1);
  M.wide[idx] = 0;
  if (M.layout == L_MASS3 && K >= 2 && K <= 16) {
    for (int G : {4, 8, 16, 32}) {
      M.wide[idx] = G;
      consider(time_launches(s, reps, [&] { k_block_spmv(K, M, x.p, y.p, s, fp32); }), 4, 0, G, 0, 0);
    }
    M.wide[idx] = 0;
  }
  if (K == 1) {
    M.single_blocked = false;
    M.pipe = 0;
    for (int pw : {1, 2, 4, 8}) {
      build_row_groups(M, pw);
      consider(time_launches(s, reps, [&] { k_spmv(M, x.p, y.p, s, fp32); }), 3, 0, pw, 0, 0);
    }
    build_row_groups(M, 0);
    for (int g : {8, 16, 32, 64}) {
      M.group = g;
      consider(time_launches(s, reps, [&] { k_spmv(M, x.p, y.p, s, fp32); }), 0, g, 1, 0, 0);
    }
    if (M.layout == L_BSR3)
      for (int g : {16, 32, 64})
        for (int pr : {2, 4, 8}) {
          M.group = g;
          M.pipe = pr;
          consider(time_launches(s, reps, [&] { k_spmv(M, x.p, y.p, s, fp32); }), 6, g, pr, 0, 0);
        }
    M.pipe = 0;
    M.group = 0;
  }
  for (int R : {1, 2, 4, 8}) {
    if (K * R > 64) continue;
    M.block_map[idx] = R;
    consider(time_launches(s, reps, [&] { k_block_spmv(K, M, x.p, y.p, s, fp32); }), 1, 0, R, 0, 0);
  }
  // tiled layouts: tile capacities on a geometric ladder from the longest
  // row (a tile must hold one row) up to 64 KB of LDS per workgroup
  int64_t longest = 0;
  for (int64_t r = 0; r < M.nb; ++r) longest = std::max(longest, M.h_browptr[r + 1] - M.h_browptr[r]);
  const int limit = K <= 16 ? std::min(65535, tile_cap_limit(K)) : 0;  // no LDS tiles beyond K = 16
  std::vector<int> caps;
  for (double f = 1.0; f * (double)longest <= (double)limit * 1.0001; f *= 1.28)
    caps.push_back((int)std::ceil(f * (double)longest));
  if (limit > 0 && (caps.empty() || caps.back() < limit)) caps.push_back(limit);
  for (int cap : caps) {
    if (cap < longest || cap > limit) continue;
    for (int R : {1, 2, 4, 8}) {
      if (K * R > 64) continue;
      for (int tb : {1024, 256}) {
        build_tiling(M, K, cap, R, tb);
        consider(time_launches(s, reps, [&] { k_block_spmv(K, M, x.p, y.p, s, fp32); }), 2, 0, R, cap,
                 M.tiling[idx] ? M.tiling[idx]->ntiles : 0, tb);
      }
    }
    if (M.layout == L_MASS3 && K >= 2)
      for (int G : {4, 8, 16}) {
        build_tiling(M, K, cap, 1, 1024, G);
        consider(time_launches(s, reps, [&] { k_block_spmv(K, M, x.p, y.p, s, fp32); }), 5, 0, G, cap,
                 M.tiling[idx] ? M.tiling[idx]->ntiles : 0, 1024);
      }
  }
  // apply the winner
  build_tiling(M, K, 0, 1);
  M.block_map[idx] = 0;
  if (best.kind == 0) {
    M.group = best.group;
  } else if (best.kind == 6) {
    M.group = best.group;
    M.pipe = best.rows;
  } else if (best.kind == 3) {
    build_row_groups(M, best.rows);
  } else if (best.kind == 4) {
    M.wide[idx] = best.rows;
  } else if (best.kind == 5) {
    build_tiling(M, K, best.cap, 1, 1024, best.rows);
  } else if (best.kind == 1) {
    M.block_map[idx] = best.rows;
    if (K == 1) M.single_blocked = true;
  } else {
    build_tiling(M, K, best.cap, best.rows, best.block);
    if (K == 1) M.single_blocked = true;
  }
  NMG_HIP(hipStreamSynchronize(s));
  return best;
}
}  // namespace nmg

extern "C" {

int nmg_csr_autotune(nmg_handle mat, int K, int reps, int fp32, double* seconds, int* kind, int* rows,
                     int* cap, int* threads) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    NMG_CHECK(K == 1 || K == 2 || K == 4 || K == 8 || K == 16 || K == 32 || K == 64, NMG_EINVAL,
              "K must be 1, 2, 4, 8, 16, 32 or 64");
    NMG_CHECK(reps >= 1, NMG_EINVAL, "reps must be >= 1");
    hipStream_t s = default_stream();
    if (fp32 && !M->bval32) make_fp32(*M, s);
    TuneResult r = tune_matrix(*M, K, reps, fp32 != 0, s);
    M->touch();
    if (seconds) *seconds = r.seconds;
    if (kind) *kind = r.kind;
    if (rows) *rows = r.kind == 0 ? r.group : r.rows;
    if (cap) *cap = r.kind == 6 ? r.group : r.cap;
    if (threads) *threads = r.kind == 2 ? r.block : 0;
  });
}

int nmg_bench_block_chebstep(nmg_handle cheb, int K, int reps, int fp32, double* seconds, double* bytes) {
  return guarded([&] {
    Cheb* C = get<Cheb>(cheb, K_CHEB);
    NMG_CHECK(reps >= 1, NMG_EINVAL, "reps must be >= 1");
    NMG_CHECK(K == 1 || K == 2 || K == 4 || K == 8 || K == 16 || K == 32 || K == 64, NMG_EINVAL,
              "K must be 1, 2, 4, 8, 16, 32 or 64");
    NMG_CHECK(fp32 >= 0 && fp32 <= 2, NMG_EINVAL, "fp32 must be 0 (FP64), 1 (FP32 values) or 2 (and FP32 vectors)");
    DevMat* M = C->mat;
    hipStream_t s = default_stream();
    if (fp32 && !(M->val32 || M->bval32)) make_fp32(*M, s);
    const int64_t nk = M->nrow * K;
    const double th = 0.5 * (C->lmax + C->lmin);
    DVec b, xk, xm;
    FVec f;
    std::function<void()> step;
    if (fp32 == 2) {
      f.ensure(3 * nk);
      float* fb = f.p;
      float* fk = f.p + nk;
      float* fm = f.p + 2 * nk;
      b.ensure(nk);
      k_fill(nk, 1.0, b.p, s);
      k_cheb32_start(nk, 1.0 / th, b.p, fb, fk, fm, s);
      step = [=, &s]() mutable {
        k_block_cheb_step32(K, *M, fb, fk, fm, 1.5, 1.0 / th, s, true);
        std::swap(fk, fm);
      };
    } else {
      b.ensure(nk);
      xk.ensure(nk);
      xm.ensure(nk);
      k_fill(nk, 1.0, b.p, s);
      k_fill(nk, 0.5, xk.p, s);
      k_fill(nk, 0.0, xm.p, s);
      double* pk = xk.p;
      double* pm = xm.p;
      const double* pb = b.p;
      step = [=, &s]() mutable {
        k_block_cheb_step(K, *M, pb, pk, pm, 1.5, 1.0 / th, s, fp32 != 0);
        std::swap(pk, pm);
      };
    }
    step();
    hipEvent_t e0, e1;
    NMG_HIP(hipEventCreate(&e0));
    NMG_HIP(hipEventCreate(&e1));
    NMG_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) step();
    NMG_HIP(hipEventRecord(e1, s));
    NMG_HIP(hipEventSynchronize(e1));
    float ms = 0;
    NMG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (seconds) *seconds = ms * 1e-3 / reps;
    if (bytes) *bytes = M->matrix_bytes(fp32 != 0) + 4.0 * (fp32 == 2 ? 4.0 : 8.0) * (double)nk;
  });
}

int nmg_tiling_host(int64_t nb, const int64_t* rowptr, const int32_t* col, int max_nodes, int64_t* ntiles,
                    int* max_used, int64_t* trp, int64_t* tcp, int32_t* tcols, uint16_t* lcol) {
  return guarded([&] {
    NMG_CHECK(nb > 0 && rowptr && col, NMG_EINVAL, "null pattern");
    HostTiling H;
    compute_tiling(nb, rowptr, col, max_nodes, H);
    if (ntiles) *ntiles = (int64_t)H.trp.size() - 1;
    if (max_used) *max_used = H.max_nodes;
    if (trp) std::memcpy(trp, H.trp.data(), sizeof(int64_t) * H.trp.size());
    if (tcp) std::memcpy(tcp, H.tcp.data(), sizeof(int64_t) * H.tcp.size());
    if (tcols) std::memcpy(tcols, H.tcols.data(), sizeof(int32_t) * H.tcols.size());
    if (lcol) std::memcpy(lcol, H.lcol.data(), sizeof(uint16_t) * H.lcol.size());
  });
}

int nmg_set_tiling(nmg_handle mat, int K, int max_nodes, int rows_per_wave, int threads) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    NMG_CHECK(rows_per_wave >= 1 && rows_per_wave <= 8 && K * rows_per_wave <= 64, NMG_EINVAL,
              "bad rows per wavefront");
    NMG_CHECK(max_nodes <= tile_cap_limit(K), NMG_EINVAL, "tile exceeds 64 KB of LDS");
    build_tiling(*M, K, max_nodes, rows_per_wave, threads > 0 ? threads : 1024);
    for (int i = 0; i < 5; ++i)
      if ((1 << i) == K) M->wide[i] = 0;
    if (K == 1) {
      M->single_blocked = max_nodes > 0 || M->block_map[0] > 0;
      build_row_groups(*M, 0);
    }
  });
}

int nmg_set_tiled_wide(nmg_handle mat, int K, int max_nodes, int lanes) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    NMG_CHECK(K == 2 || K == 4 || K == 8 || K == 16, NMG_EINVAL, "K must be 2, 4, 8 or 16");
    NMG_CHECK(max_nodes * 24 * K <= 64 * 1024, NMG_EINVAL, "tile exceeds 64 KB of LDS");
    build_tiling(*M, K, max_nodes, 1, 1024, max_nodes > 0 ? lanes : 0);
    for (int i = 0; i < 5; ++i)
      if ((1 << i) == K) M->wide[i] = 0;
  });
}

int nmg_set_wide(nmg_handle mat, int K, int lanes) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    int idx = -1;
    for (int i = 1; i < 5; ++i)
      if ((1 << i) == K) idx = i;
    NMG_CHECK(idx >= 1, NMG_EINVAL, "K must be 2, 4, 8 or 16");
    NMG_CHECK(lanes == 0 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32, NMG_EINVAL,
              "lanes per row must be 0, 4, 8, 16 or 32");
    NMG_CHECK(lanes == 0 || M->layout == L_MASS3, NMG_EINVAL, "the wide layout applies to scalar-mass matrices");
    M->wide[idx] = lanes;
    M->touch();
  });
}

int nmg_set_row_pipeline(nmg_handle mat, int lanes, int rows) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    NMG_CHECK(rows == 0 || rows == 2 || rows == 4 || rows == 8, NMG_EINVAL, "rows per lane group must be 0, 2, 4 or 8");
    NMG_CHECK(rows == 0 || lanes == 16 || lanes == 32 || lanes == 64, NMG_EINVAL, "lanes per group must be 16, 32 or 64");
    NMG_CHECK(rows == 0 || M->layout == L_BSR3, NMG_EINVAL, "pipelined rows apply to 3x3-block matrices");
    M->pipe = rows;
    if (rows > 0) {  // the single-column products take the lanes-per-row route
      M->group = lanes;
      M->single_blocked = false;
      build_row_groups(*M, 0);
    }
    M->touch();
  });
}

int nmg_set_row_groups(nmg_handle mat, int per_wave) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    NMG_CHECK(per_wave == 0 || per_wave == 1 || per_wave == 2 || per_wave == 4 || per_wave == 8, NMG_EINVAL,
              "groups per wavefront must be 0, 1, 2, 4 or 8");
    NMG_CHECK(per_wave == 0 || M->layout != L_CSR, NMG_EINVAL, "row groups apply to the blocked layouts");
    build_row_groups(*M, per_wave);
  });
}

int nmg_bench_spmv(nmg_handle mat, int reps, int fp32, double* seconds, double* bytes) {
  return guarded([&] {
    DevMat* M = get<DevMat>(mat, K_MAT);
    NMG_CHECK(reps >= 1, NMG_EINVAL, "reps must be >= 1");
    hipStream_t s = default_stream();
    if (fp32 && !(M->val32 || M->bval32)) make_fp32(*M, s);
    DVec dx, dy;
    dx.ensure(M->ncol);
    dy.ensure(M->nrow);
    k_fill(M->ncol, 1.0, dx.p, s);
    k_spmv(*M, dx.p, dy.p, s, fp32 != 0);  // warm-up
    hipEvent_t e0, e1;
    NMG_HIP(hipEventCreate(&e0));
    NMG_HIP(hipEventCreate(&e1));
    NMG_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) k_spmv(*M, dx.p, dy.p, s, fp32 != 0);
    NMG_HIP(hipEventRecord(e1, s));
    NMG_HIP(hipEventSynchronize(e1));
    float ms = 0;
    NMG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (seconds) *seconds = ms * 1e-3 / reps;
    if (bytes) *bytes = M->algorithmic_bytes(fp32 != 0);
  });
}

int nmg_gemm_tn(int64_t n, int m, int kc, const double* A, const double* B, double* C) {
  return guarded([&] {
    NMG_CHECK(n > 0 && m > 0 && kc > 0 && A && B && C, NMG_EINVAL, "bad argument");
    Solver S;
    S.s = default_stream();
    S.borrowed = true;
    DVec dA, dB, dC;
    dA.ensure(n * m);
    dB.ensure(n * kc);
    dC.ensure((int64_t)m * kc);
    NMG_HIP(hipMemcpyAsync(dA.p, A, sizeof(double) * n * m, hipMemcpyHostToDevice, S.s));
    NMG_HIP(hipMemcpyAsync(dB.p, B, sizeof(double) * n * kc, hipMemcpyHostToDevice, S.s));
    gemm_tn_any(S, n, m, kc, dA.p, n, dB.p, n, dC.p, m);
    NMG_HIP(hipMemcpyAsync(C, dC.p, sizeof(double) * m * kc, hipMemcpyDeviceToHost, S.s));
    NMG_HIP(hipStreamSynchronize(S.s));
  });
}

int nmg_bench_gemm_tn(int64_t n, int m, int kc, int reps, double* seconds, double* bytes) {
  return guarded([&] {
    NMG_CHECK(n > 0 && m > 0 && kc > 0 && reps >= 1, NMG_EINVAL, "bad argument");
    Solver S;
    S.s = default_stream();
    S.borrowed = true;
    DVec dA, dB, dC;
    dA.ensure(n * m);
    dB.ensure(n * kc);
    dC.ensure((int64_t)m * kc);
    k_fill(n * m, 0.5, dA.p, S.s);
    k_fill(n * kc, 0.25, dB.p, S.s);
    const double sec = time_launches(S.s, reps, [&] { gemm_tn_any(S, n, m, kc, dA.p, n, dB.p, n, dC.p, m); });
    if (seconds) *seconds = sec;
    if (bytes) *bytes = 8.0 * ((double)n * m + (double)n * kc + (double)m * kc);
  });
}

// Dense products on host arrays (tests) and their timing: op 0
// C (n x K) = alpha A (n x m) B (m x K) + beta C, op 1 C (m x K) = A^T B
// (A n x m, B n x K); kernel 1 = the MFMA kernels, 0 = rocBLAS DGEMM (op 0)
// or the FMA split reduction (op 1).
}  // extern "C"

template <class F>
static void with_dense_solver(int kernel, F&& f) {
  Solver S;
  S.s = default_stream();
  S.borrowed = true;
  S.mfma = kernel;
  struct Blas {
    rocblas_handle h = nullptr;
    ~Blas() {
      if (h) rocblas_destroy_handle(h);
    }
  } bh;
  if (!kernel) {
    NMG_LIB(rocblas_create_handle(&bh.h), "rocblas handle");
    NMG_LIB(rocblas_set_stream(bh.h, S.s), "rocblas stream");
    S.blas = bh.h;
  }
  f(S);
  S.blas = nullptr;
}

extern "C" {

int nmg_dense_product(int op, int kernel, int64_t n, int m, int K, double alpha, const double* A, const double* B,
                      double beta, double* C) {
  return guarded([&] {
    NMG_CHECK(op == 0 || op == 1, NMG_EINVAL, "op must be 0 (A B) or 1 (A^T B)");
    NMG_CHECK(n > 0 && m > 0 && K > 0 && A && B && C, NMG_EINVAL, "bad argument");
    NMG_CHECK(op == 0 || (alpha == 1.0 && beta == 0.0), NMG_EINVAL, "A^T B takes alpha 1, beta 0");
    with_dense_solver(kernel, [&](Solver& S) {
      const int64_t nc = op == 0 ? n : m;
      DVec dA, dB, dC;
      dA.ensure(n * m);
      dB.ensure(op == 0 ? (int64_t)m * K : n * K);
      dC.ensure(nc * K);
      NMG_HIP(hipMemcpyAsync(dA.p, A, sizeof(double) * n * m, hipMemcpyHostToDevice, S.s));
      NMG_HIP(hipMemcpyAsync(dB.p, B, sizeof(double) * (op == 0 ? (int64_t)m * K : n * K), hipMemcpyHostToDevice,
                             S.s));
      if (op == 0) {
        NMG_HIP(hipMemcpyAsync(dC.p, C, sizeof(double) * nc * K, hipMemcpyHostToDevice, S.s));
        dense_nn(S, n, K, m, alpha, dA.p, n, dB.p, m, beta, dC.p, n, "rocblas_dgemm failed");
      } else {
        dense_tn(S, n, m, K, dA.p, n, dB.p, n, dC.p, m);
      }
      NMG_HIP(hipMemcpyAsync(C, dC.p, sizeof(double) * nc * K, hipMemcpyDeviceToHost, S.s));
      NMG_HIP(hipStreamSynchronize(S.s));
    });
  });
}

int nmg_bench_dense(int op, int kernel, int64_t n, int m, int K, int reps, double* seconds, double* bytes,
                    double* flops) {
  return guarded([&] {
    NMG_CHECK(op == 0 || op == 1, NMG_EINVAL, "op must be 0 (A B) or 1 (A^T B)");
    NMG_CHECK(n > 0 && m > 0 && K > 0 && reps >= 1, NMG_EINVAL, "bad argument");
    with_dense_solver(kernel, [&](Solver& S) {
      const int64_t nc = op == 0 ? n : m;
      DVec dA, dB, dC;
      dA.ensure(n * m);
      dB.ensure(op == 0 ? (int64_t)m * K : n * K);
      dC.ensure(nc * K);
      k_fill(n * m, 0.5, dA.p, S.s);
      k_fill(op == 0 ? (int64_t)m * K : n * K, 0.25, dB.p, S.s);
      k_fill(nc * K, 0.0, dC.p, S.s);
      const double sec = time_launches(S.s, reps, [&] {
        if (op == 0) dense_nn(S, n, K, m, -1.0, dA.p, n, dB.p, m, 1.0, dC.p, n, "rocblas_dgemm failed");
        else dense_tn(S, n, m, K, dA.p, n, dB.p, n, dC.p, m);
      });
      if (seconds) *seconds = sec;
      // algorithmic bytes: A once, B once, C read and written (op 0) or written (op 1)
      if (bytes)
        *bytes = op == 0 ? 8.0 * ((double)n * m + (double)m * K + 2.0 * n * K)
                         : 8.0 * ((double)n * m + (double)n * K + (double)m * K);
      if (flops) *flops = 2.0 * (double)n * m * K;
    });
  });
}

int nmg_set_dense_kernels(nmg_handle h, int mfma) {
  return guarded([&] {
    Solver* S = get<Solver>(h, K_SOLVER);
    NMG_CHECK(mfma >= 0 && mfma <= 2, NMG_EINVAL, "mode must be 0, 1 or 2");
    S->mfma = mfma;
  });
}

int nmg_bench_copy(int64_t n, int reps, double* seconds) {
  return guarded([&] {
    NMG_CHECK(n > 0 && reps >= 1, NMG_EINVAL, "n and reps must be positive");
    hipStream_t s = default_stream();
    DVec x, y;
    x.ensure(n);
    y.ensure(n);
    k_fill(n, 1.0, x.p, s);
    const double sec = time_launches(s, reps, [&] { k_copy_kernel(n, x.p, y.p, s); });
    if (seconds) *seconds = sec;
  });
}

int nmg_chebiter_set_mode(nmg_handle cheb, int mode) {
  return guarded([&] {
    Cheb* C = get<Cheb>(cheb, K_CHEB);
    NMG_CHECK(mode == 0 || mode == 1, NMG_EINVAL, "mode must be 0 (one launch per step) or 1 (persistent)");
    C->mode = mode;
  });
}

int nmg_assemble_elastic(int64_t n_nodes, int64_t n_elem, int np, const double* nodes, const int32_t* elems,
                         const int32_t* vslot, const double* lam, const double* mu, const double* rho,
                         const double* g, const double* refM, const double* refD, const int64_t* adj_ptr,
                         const int32_t* adj, const int64_t* a_rowptr, double* a_val, const int64_t* b_rowptr,
                         double* b_val) {
  return guarded([&] {
    NMG_CHECK(nodes && elems && vslot && lam && mu && rho && refM && refD && adj_ptr && adj && a_rowptr && a_val &&
                  b_rowptr && b_val,
              NMG_EINVAL, "null argument");
    const int rc = assemble_elastic_device(n_nodes, n_elem, np, nodes, elems, vslot, lam, mu, rho, g, refM, refD,
                                           adj_ptr, adj, a_rowptr, a_val, b_rowptr, b_val, default_stream());
    NMG_CHECK(rc == 0, NMG_EINVAL, "flat element " + std::to_string(-rc - 1));
  });
}

int nmg_assemble_fs(int64_t n_nodes, int64_t n_elem, int np, int nfp, const double* nodes, const int32_t* elems,
                    const int32_t* vslot, const double* lam, const double* mu, const double* rho, const double* g,
                    const int8_t* efluid, const int32_t* neigh, int purefluid, const double* refM, const double* refD,
                    const double* massF, const int32_t* fmask, const int64_t* adj_ptr, const int32_t* adj,
                    const int8_t* vstat, const int64_t* vstt, const int64_t* pstt, const int64_t* ad_ptr,
                    int32_t* ad_col, double* ad_val, const int64_t* b_ptr, int32_t* b_col, double* b_val,
                    const int64_t* e_ptr, int32_t* e_col, double* e_val, const int64_t* et_ptr, int32_t* et_col,
                    double* et_val, const int64_t* ap_ptr, int32_t* ap_col, double* ap_val, int64_t* stats) {
  return guarded([&] {
    NMG_CHECK(nodes && elems && vslot && lam && mu && rho && efluid && neigh && refM && refD && massF && fmask &&
                  adj_ptr && adj && vstat && vstt && pstt && ad_ptr && ad_col && ad_val && b_ptr && b_col && b_val &&
                  e_ptr && e_col && e_val && et_ptr && et_col && et_val && ap_ptr && ap_col && ap_val,
              NMG_EINVAL, "null argument");
    assemble_fs_device(n_nodes, n_elem, np, nfp, nodes, elems, vslot, lam, mu, rho, g, efluid, neigh, purefluid,
                       refM, refD, massF, fmask, adj_ptr, adj, vstat, vstt, pstt, ad_ptr, ad_col, ad_val, b_ptr,
                       b_col, b_val, e_ptr, e_col, e_val, et_ptr, et_col, et_val, ap_ptr, ap_col, ap_val, stats,
                       default_stream());
  });
}

int nmg_backward_fs(int64_t n_nodes, int64_t n_elem, int np, int nfp, const double* nodes, const int32_t* elems,
                    const int32_t* vslot, const double* lam, const double* rho, const double* g,
                    const int8_t* efluid, const int32_t* neigh, int purefluid, const double* refM, const double* refD,
                    const double* massF, const int32_t* fmask, const int8_t* vstat, const int64_t* vstt,
                    const int64_t* pstt, int k, const double* eig, const double* U, const double* P, double* d_rho,
                    double* d_lam, double* d_mu) {
  return guarded([&] {
    NMG_CHECK(nodes && elems && vslot && lam && rho && efluid && neigh && refM && refD && massF && fmask && vstat &&
                  vstt && pstt && eig && U && P && d_rho && d_lam && d_mu,
              NMG_EINVAL, "null argument");
    backward_fs_device(n_nodes, n_elem, np, nfp, nodes, elems, vslot, lam, rho, g, efluid, neigh, purefluid, refM,
                       refD, massF, fmask, vstat, vstt, pstt, k, eig, U, P, d_rho, d_lam, d_mu, default_stream());
  });
}

int nmg_chebiter_set_refine(nmg_handle cheb, int passes, int inner_deg) {
  return guarded([&] {
    Cheb* C = get<Cheb>(cheb, K_CHEB);
    NMG_CHECK(passes == 0 || (passes >= 1 && passes <= 8 && inner_deg >= 1 && inner_deg <= 10000), NMG_EINVAL,
              "passes must be 0 (off) or 1..8 with an inner degree >= 1");
    if (passes > 0 && !(C->mat->val32 || C->mat->bval32)) {
      make_fp32(*C->mat, default_stream());
      NMG_HIP(hipStreamSynchronize(default_stream()));
    }
    C->refine = passes;
    C->inner_deg = passes > 0 ? inner_deg : 0;
    // the single-column solve's FP32 work, allocated here so that a solve
    // replayed from a captured graph never allocates inside the capture
    if (passes > 0) C->own32.ensure(3 * C->mat->nrow);
  });
}

int nmg_chebiter_info(nmg_handle cheb, int fp32, int* persistent, int* grid, int* lds_bytes) {
  return guarded([&] {
    Cheb* C = get<Cheb>(cheb, K_CHEB);
    if (fp32 && !(C->mat->val32 || C->mat->bval32)) make_fp32(*C->mat, default_stream());
    const int saved = C->mode;
    C->mode = 1;
    const PersistPlan* P = C->persistent(fp32 != 0);
    C->mode = saved;
    if (persistent) *persistent = (P && saved == 1) ? 1 : (P ? 2 : 0);
    if (grid) *grid = P ? P->grid : 0;
    if (lds_bytes) *lds_bytes = P ? (int)P->lds_bytes : 0;
  });
}

int nmg_bench_bsolve(nmg_handle cheb, int reps, int fp32, double* seconds, double* bytes) {
  return guarded([&] {
    Cheb* C = get<Cheb>(cheb, K_CHEB);
    NMG_CHECK(reps >= 1, NMG_EINVAL, "reps must be >= 1");
    DevMat* M = C->mat;
    hipStream_t s = default_stream();
    if (fp32 && !(M->val32 || M->bval32)) make_fp32(*M, s);
    const int64_t n = M->nrow;
    DVec b, x, t;
    b.ensure(n);
    x.ensure(n);
    t.ensure(n);
    k_fill(n, 1.0, b.p, s);
    const double sec = time_launches(s, reps, [&] { cheb_solve(*C, b.p, x.p, t.p, s, fp32 != 0); });
    if (seconds) *seconds = sec;
    if (bytes) *bytes = (M->matrix_bytes(fp32 != 0) + 4.0 * 8.0 * (double)n) * C->deg;
  });
}

int nmg_bench_chebstep(nmg_handle cheb, int reps, int fp32, double* seconds, double* bytes) {
  return guarded([&] {
    Cheb* C = get<Cheb>(cheb, K_CHEB);
    NMG_CHECK(reps >= 1, NMG_EINVAL, "reps must be >= 1");
    DevMat* M = C->mat;
    hipStream_t s = default_stream();
    if (fp32 && !(M->val32 || M->bval32)) make_fp32(*M, s);
    const int64_t n = M->nrow;
    DVec b, xk, xm;
    b.ensure(n);
    xk.ensure(n);
    xm.ensure(n);
    k_fill(n, 1.0, b.p, s);
    k_fill(n, 0.5, xk.p, s);
    k_fill(n, 0.0, xm.p, s);
    const double th = 0.5 * (C->lmax + C->lmin);
    k_cheb_step(*M, b.p, xk.p, xm.p, 1.5, 1.0 / th, s, fp32 != 0);
    hipEvent_t e0, e1;
    NMG_HIP(hipEventCreate(&e0));
    NMG_HIP(hipEventCreate(&e1));
    NMG_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) {
      k_cheb_step(*M, b.p, xk.p, xm.p, 1.5, 1.0 / th, s, fp32 != 0);
      std::swap(xk.p, xm.p);
    }
    NMG_HIP(hipEventRecord(e1, s));
    NMG_HIP(hipEventSynchronize(e1));
    float ms = 0;
    NMG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (seconds) *seconds = ms * 1e-3 / reps;
    // product (matrix + x + y) with y replaced by the update: + b read + xm read
    if (bytes) *bytes = M->algorithmic_bytes(fp32 != 0) + 2.0 * 8.0 * (double)n;
  });
}

}  // extern "C"
