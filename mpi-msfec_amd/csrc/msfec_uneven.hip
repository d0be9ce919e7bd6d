// Uneven batches (msfec_qu_*): coarse cells with their own refinement level.
//
// The reference shares one n_refine_local between all cells of the map
// (source/Q/q_parameters.cc:229, source/Q/q_basis.cc:101).  Here one handle
// keeps a msfec_q batch per level present; per level it picks the local
// solver (the band solvers where the interior band fits the window, timed
// against matrix-free CG, CG beyond), and compute forks the levels onto
// streams of their own so that the short band levels overlap the long CG
// level.  Data in batch order is gathered into the levels' contiguous
// buffers, and their outputs are scattered back, by the two copy kernels
// below.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "msfec.h"

namespace msfec_detail {
extern thread_local std::string g_last_error;
hipError_t coarse_solve_device(int dim, int nc, const double* dM, const double* db, double* cK,
                               double* crhs, double* cvec, int* cit, double abs_tol,
                               double rel_tol, int max_iter, hipStream_t s);
}  // namespace msfec_detail

namespace {

msfec_status qu_fail(msfec_status st, const std::string& msg) {
  msfec_detail::g_last_error = msg;
  return st;
}

#define QU_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return qu_fail(MSFEC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define QU_TRY(call)                      \
  do {                                    \
    msfec_status s_ = (call);             \
    if (s_ != MSFEC_OK) return s_;        \
  } while (0)

// dst[i][k] = src[off[i] + k]: the blocks of a level's cells, contiguous
__global__ void qu_gather_kernel(const double* __restrict__ src, const size_t* __restrict__ off,
                                 int len, int n, double* __restrict__ dst) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * len) return;
  const size_t i = t / len, k = t % len;
  dst[t] = src[off[i] + k];
}

// dst[off[i] * mul + k] = src[i][k]: a level's outputs back to batch order
// (off in doubles with mul = 1, or cell indices with mul = the block length)
__global__ void qu_scatter_kernel(const double* __restrict__ src, const size_t* __restrict__ off,
                                  size_t mul, int len, int n, double* __restrict__ dst) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * len) return;
  const size_t i = t / len, k = t % len;
  dst[off[i] * mul + k] = src[t];
}

// the band solvers' condition: the half-bandwidth and the 4 rows entering at
// each super-step fit the 64-slot window
bool band_fits(int dim, int level) {
  const int ni = (1 << level) - 1;
  const int bw = (dim == 2) ? ni + 1 : ni * ni + ni + 1;
  return bw + 4 <= 64;
}

int ipow_qu(int b, int e) {
  int r = 1;
  while (e-- > 0) r *= b;
  return r;
}

// candidates timed for a level inside the band window, in tie-break order:
// the matrix-core band with a 16-, 32- and 64-slot window (the tile shape:
// 1, 2 or 4 tile rows of 16), the vector-ALU band, CG
struct Cand {
  int solver, window;
};
constexpr int kNCand = 5;
constexpr Cand kCand[kNCand] = {{MSFEC_SOLVER_BAND_FP64_MFMA, 16},
                                {MSFEC_SOLVER_BAND_FP64_MFMA, 32},
                                {MSFEC_SOLVER_BAND_FP64_MFMA, 64},
                                {MSFEC_SOLVER_BAND_FP64, 0},
                                {MSFEC_SOLVER_CG, 0}};
constexpr int kTuneCells = 256;

struct Level {
  int level = 0;
  int nfc = 0, nn = 0;          // fine cells, fine nodes per cell
  std::vector<int> cells;       // batch indices, ascending
  std::vector<double> verts;    // host copy of the cells' vertices
  msfec_q_handle h = nullptr;
  int solver = -1, window = 0;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  double* d_coef = nullptr;     // [cells][nfc] staged
  double* d_f = nullptr;
  double* d_reac = nullptr;
  size_t* d_off_fc = nullptr;   // per cell: offset of its fine-cell block in batch order
  size_t* d_off_psi = nullptr;  // per cell: offset of its Psi block in batch order
  size_t* d_off_cell = nullptr; // per cell: batch index (for M, b rows)
  float ms[kNCand] = {-1.f, -1.f, -1.f, -1.f, -1.f};
};

}  // namespace

struct msfec_qu_batch {
  msfec_q_params p{};
  int n_cells = 0, nv = 0;
  std::vector<int> levels;
  std::vector<size_t> coef_off, psi_off;  // n_cells + 1
  std::vector<Level> lv;                  // ascending level
  hipStream_t stream = nullptr;           // the caller's stream of the last compute
  hipEvent_t fork = nullptr;
  bool coef_set = false, rhs_set = false, reac_set = false, computed = false, tuned = false;
  double* d_stage = nullptr;              // scatter staging (largest level output)
  size_t stage_doubles = 0;
  double* d_M = nullptr;                  // batch-order M, b for the coarse solve
  double* d_b = nullptr;
  double *d_cK = nullptr, *d_crhs = nullptr, *d_cvec = nullptr;
  int* d_cit = nullptr;
  int coarse_nc = -1;
};

namespace {

void qu_free(msfec_qu_batch* h) {
  for (Level& L : h->lv) {
    if (L.h) msfec_q_destroy(L.h);
    if (L.stream) (void)hipStreamDestroy(L.stream);
    if (L.done) (void)hipEventDestroy(L.done);
    for (void* p : {(void*)L.d_coef, (void*)L.d_f, (void*)L.d_reac, (void*)L.d_off_fc,
                    (void*)L.d_off_psi, (void*)L.d_off_cell})
      if (p) (void)hipFree(p);
  }
  for (void* p : {(void*)h->d_stage, (void*)h->d_M, (void*)h->d_b, (void*)h->d_cK,
                  (void*)h->d_crhs, (void*)h->d_cvec, (void*)h->d_cit})
    if (p) (void)hipFree(p);
  if (h->fork) (void)hipEventDestroy(h->fork);
}

msfec_q_params level_params(const msfec_qu_batch* h, const Level& L, int solver, int window) {
  msfec_q_params q = h->p;
  q.n_refine_local = L.level;
  q.solver = solver;
  q.band_window = (solver == MSFEC_SOLVER_BAND_FP64_MFMA) ? window : 0;
  return q;
}

// (Re)create the batch of a level with a solver (and band window) and load
// its staged data.
msfec_status make_level_handle(msfec_qu_batch* h, Level& L, int solver, int window = 0) {
  if (L.h && L.solver == solver && L.window == window) return MSFEC_OK;
  msfec_q_handle nh = nullptr;
  const msfec_q_params q = level_params(h, L, solver, window);
  QU_TRY(msfec_q_create(&q, (int)L.cells.size(), L.verts.data(), MSFEC_HOST, &nh));
  const size_t cnt = L.cells.size() * (size_t)L.nfc;
  msfec_status st = MSFEC_OK;
  if (h->coef_set) st = msfec_q_set_coefficient(nh, L.d_coef, cnt, MSFEC_DEVICE);
  if (st == MSFEC_OK && h->rhs_set) st = msfec_q_set_rhs(nh, L.d_f, cnt, MSFEC_DEVICE);
  if (st == MSFEC_OK && h->reac_set) st = msfec_q_set_reaction(nh, L.d_reac, cnt, MSFEC_DEVICE);
  if (st != MSFEC_OK) {
    msfec_q_destroy(nh);
    return st;
  }
  if (L.h) msfec_q_destroy(L.h);
  L.h = nh;
  L.solver = solver;
  L.window = window;
  return MSFEC_OK;
}

// Host or device data in batch order -> each level's staged buffer (on the
// level's stream after the handle's stream, waited for).
msfec_status stage_in(msfec_qu_batch* h, const double* data, size_t count, int loc,
                      double* Level::*buf) {
  if (!data) return qu_fail(MSFEC_ERR_INVALID, "null data pointer");
  if (count != h->coef_off[h->n_cells])
    return qu_fail(MSFEC_ERR_INVALID, "count " + std::to_string(count) + " != expected " +
                                          std::to_string(h->coef_off[h->n_cells]));
  QU_HIP(hipStreamSynchronize(h->stream));  // a compute in flight reads the old data to the end
  for (Level& L : h->lv) {
    const size_t n = L.cells.size(), len = (size_t)L.nfc;
    if (loc == MSFEC_DEVICE) {
      const unsigned blocks = (unsigned)((n * len + 255) / 256);
      hipLaunchKernelGGL(qu_gather_kernel, dim3(blocks), dim3(256), 0, L.stream, data, L.d_off_fc,
                         (int)len, (int)n, L.*buf);
      QU_HIP(hipGetLastError());
    } else {
      std::vector<double> tmp(n * len);
      for (size_t i = 0; i < n; ++i)
        std::memcpy(tmp.data() + i * len, data + h->coef_off[L.cells[i]], len * sizeof(double));
      QU_HIP(hipMemcpyAsync(L.*buf, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice,
                            L.stream));
    }
    QU_HIP(hipStreamSynchronize(L.stream));
  }
  return MSFEC_OK;
}

// Time one solver on the first cells of a level; -1 if a cell fails.
msfec_status time_candidate(msfec_qu_batch* h, Level& L, const Cand& cd, float* ms) {
  *ms = -1.f;
  const int n = std::min((int)L.cells.size(), kTuneCells);
  msfec_q_params q = level_params(h, L, cd.solver, cd.window);
  q.flags = 0;
  msfec_q_handle t = nullptr;
  msfec_status st = msfec_q_create(&q, n, L.verts.data(), MSFEC_HOST, &t);
  if (st == MSFEC_ERR_UNSUPPORTED || st == MSFEC_ERR_INVALID) return MSFEC_OK;  // not a candidate here
  if (st != MSFEC_OK) return st;
  const size_t cnt = (size_t)n * L.nfc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    msfec_q_destroy(t);
  };
  st = msfec_q_set_coefficient(t, L.d_coef, cnt, MSFEC_DEVICE);
  if (st == MSFEC_OK && h->rhs_set) st = msfec_q_set_rhs(t, L.d_f, cnt, MSFEC_DEVICE);
  if (st == MSFEC_OK && h->reac_set) st = msfec_q_set_reaction(t, L.d_reac, cnt, MSFEC_DEVICE);
  if (st == MSFEC_OK) st = msfec_q_compute(t, L.stream);  // warm-up
  if (st == MSFEC_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    st = qu_fail(MSFEC_ERR_HIP, "hipEventCreate failed");
  constexpr int reps = 3;
  if (st == MSFEC_OK && hipEventRecord(e0, L.stream) != hipSuccess)
    st = qu_fail(MSFEC_ERR_HIP, "hipEventRecord failed");
  for (int r = 0; st == MSFEC_OK && r < reps; ++r) st = msfec_q_compute(t, L.stream);
  if (st == MSFEC_OK && hipEventRecord(e1, L.stream) != hipSuccess)
    st = qu_fail(MSFEC_ERR_HIP, "hipEventRecord failed");
  std::vector<int> status(n);
  if (st == MSFEC_OK) st = msfec_q_get_cell_status(t, status.data());  // synchronizes
  float el = 0.f;
  if (st == MSFEC_OK && hipEventElapsedTime(&el, e0, e1) != hipSuccess)
    st = qu_fail(MSFEC_ERR_HIP, "hipEventElapsedTime failed");
  cleanup();
  if (st != MSFEC_OK) return st;
  bool ok = true;
  for (int v : status) ok = ok && (v == MSFEC_CELL_OK);
  if (ok) *ms = el / reps;
  return MSFEC_OK;
}

msfec_status qu_tune(msfec_qu_batch* h) {
  if (!h->coef_set) return qu_fail(MSFEC_ERR_STATE, "coefficient not set");
  if (h->rhs_set == false && h->p.rhs_mode == MSFEC_RHS_CELL)
    return qu_fail(MSFEC_ERR_STATE, "rhs data not set");
  if (h->reac_set == false && h->p.reaction_mode == MSFEC_REACTION_CELL)
    return qu_fail(MSFEC_ERR_STATE, "reaction data not set");
  for (Level& L : h->lv) {
    if (!band_fits(h->p.dim, L.level)) {
      QU_TRY(make_level_handle(h, L, MSFEC_SOLVER_CG));
      continue;
    }
    int best = -1;
    for (int k = 0; k < kNCand; ++k) {
      QU_TRY(time_candidate(h, L, kCand[k], &L.ms[k]));
      if (L.ms[k] >= 0.f && (best < 0 || L.ms[k] < L.ms[best])) best = k;
    }
    if (best < 0) return qu_fail(MSFEC_ERR_CELL, "no solver succeeded on the cells of level " +
                                                   std::to_string(L.level));
    QU_TRY(make_level_handle(h, L, kCand[best].solver, kCand[best].window));
  }
  h->tuned = true;
  return MSFEC_OK;
}

Level* find_level(msfec_qu_batch* h, int level) {
  for (Level& L : h->lv)
    if (L.level == level) return &L;
  return nullptr;
}

template <typename T>
msfec_status upload(T** dst, const std::vector<T>& src) {
  QU_HIP(hipMalloc((void**)dst, std::max<size_t>(1, src.size()) * sizeof(T)));
  QU_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return MSFEC_OK;
}

}  // namespace

extern "C" {

msfec_status msfec_qu_create(const msfec_q_params* params, int n_cells, const double* vertices,
                             int vertices_loc, const int* levels, msfec_qu_handle* out) {
  if (!params || !vertices || !levels || !out) return qu_fail(MSFEC_ERR_INVALID, "null argument");
  *out = nullptr;
  const msfec_q_params& p = *params;
  if (p.dim != 2 && p.dim != 3) return qu_fail(MSFEC_ERR_INVALID, "dim must be 2 or 3");
  if (n_cells <= 0) return qu_fail(MSFEC_ERR_INVALID, "n_cells must be positive");
  if (p.coef_mode != MSFEC_COEF_SCALAR_CELL || p.rhs_mode == MSFEC_RHS_QP)
    return qu_fail(MSFEC_ERR_UNSUPPORTED,
                   "uneven batches take a scalar coefficient per fine cell and a constant or "
                   "per-cell load");
  if (p.solver < MSFEC_SOLVER_BAND_FP64 || p.solver > MSFEC_SOLVER_AUTO)
    return qu_fail(MSFEC_ERR_INVALID, "unknown solver");
  const int maxl = (p.dim == 3) ? 4 : 6;
  for (int c = 0; c < n_cells; ++c)
    if (levels[c] < 1 || levels[c] > maxl)
      return qu_fail(MSFEC_ERR_UNSUPPORTED, "cell " + std::to_string(c) + ": level " +
                                                std::to_string(levels[c]) + " outside 1.." +
                                                std::to_string(maxl));
  const int nv = 1 << p.dim;
  std::vector<double> hv((size_t)n_cells * nv * p.dim);
  if (vertices_loc == MSFEC_DEVICE) {
    if (hipMemcpy(hv.data(), vertices, hv.size() * sizeof(double), hipMemcpyDeviceToHost) !=
        hipSuccess)
      return qu_fail(MSFEC_ERR_HIP, "copy of device vertices failed");
  } else {
    std::memcpy(hv.data(), vertices, hv.size() * sizeof(double));
  }
  msfec_qu_batch* h = new (std::nothrow) msfec_qu_batch();
  if (!h) return qu_fail(MSFEC_ERR_INVALID, "out of host memory");
  h->p = p;
  h->n_cells = n_cells;
  h->nv = nv;
  h->levels.assign(levels, levels + n_cells);
  h->coef_off.assign(n_cells + 1, 0);
  h->psi_off.assign(n_cells + 1, 0);
  for (int c = 0; c < n_cells; ++c) {
    const int N = 1 << levels[c];
    h->coef_off[c + 1] = h->coef_off[c] + (size_t)ipow_qu(N, p.dim);
    h->psi_off[c + 1] = h->psi_off[c] + (size_t)nv * ipow_qu(N + 1, p.dim);
  }
  std::vector<int> present(h->levels);
  std::sort(present.begin(), present.end());
  present.erase(std::unique(present.begin(), present.end()), present.end());
  h->lv.resize(present.size());
  msfec_status st = MSFEC_OK;
  auto bail = [&](msfec_status s) {
    qu_free(h);
    delete h;
    return s;
  };
  if (hipEventCreateWithFlags(&h->fork, hipEventDisableTiming) != hipSuccess)
    return bail(qu_fail(MSFEC_ERR_HIP, "hipEventCreate failed"));
  for (size_t k = 0; k < present.size(); ++k) {
    Level& L = h->lv[k];
    L.level = present[k];
    const int N = 1 << L.level;
    L.nfc = ipow_qu(N, p.dim);
    L.nn = ipow_qu(N + 1, p.dim);
    std::vector<size_t> off_fc, off_psi, off_cell;
    for (int c = 0; c < n_cells; ++c)
      if (levels[c] == L.level) {
        L.cells.push_back(c);
        L.verts.insert(L.verts.end(), hv.begin() + (size_t)c * nv * p.dim,
                       hv.begin() + (size_t)(c + 1) * nv * p.dim);
        off_fc.push_back(h->coef_off[c]);
        off_psi.push_back(h->psi_off[c]);
        off_cell.push_back((size_t)c);
      }
    const size_t cnt = L.cells.size() * (size_t)L.nfc;
    if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&L.done, hipEventDisableTiming) != hipSuccess ||
        hipMalloc((void**)&L.d_coef, cnt * sizeof(double)) != hipSuccess ||
        (p.rhs_mode == MSFEC_RHS_CELL && hipMalloc((void**)&L.d_f, cnt * sizeof(double)) != hipSuccess) ||
        (p.reaction_mode == MSFEC_REACTION_CELL &&
         hipMalloc((void**)&L.d_reac, cnt * sizeof(double)) != hipSuccess))
      return bail(qu_fail(MSFEC_ERR_HIP, "allocation of the level buffers failed"));
    if ((st = upload(&L.d_off_fc, off_fc)) != MSFEC_OK) return bail(st);
    if ((st = upload(&L.d_off_psi, off_psi)) != MSFEC_OK) return bail(st);
    if ((st = upload(&L.d_off_cell, off_cell)) != MSFEC_OK) return bail(st);
    h->stage_doubles = std::max(h->stage_doubles, L.cells.size() * (size_t)nv * L.nn);
    // the solver before any tuning: the matrix-core band where it fits, CG beyond
    int s0 = p.solver;
    if (s0 == MSFEC_SOLVER_AUTO) s0 = MSFEC_SOLVER_BAND_FP64_MFMA;
    if (!band_fits(p.dim, L.level)) s0 = MSFEC_SOLVER_CG;
    if ((st = make_level_handle(h, L, s0)) != MSFEC_OK) return bail(st);
  }
  if (hipMalloc((void**)&h->d_stage, std::max<size_t>(1, h->stage_doubles) * sizeof(double)) !=
          hipSuccess ||
      hipMalloc((void**)&h->d_M, (size_t)n_cells * nv * nv * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&h->d_b, (size_t)n_cells * nv * sizeof(double)) != hipSuccess)
    return bail(qu_fail(MSFEC_ERR_HIP, "allocation of the output staging failed"));
  // tuning is only needed where several solvers apply
  h->tuned = (p.solver != MSFEC_SOLVER_AUTO);
  *out = h;
  return MSFEC_OK;
}

msfec_status msfec_qu_offsets(msfec_qu_handle h, size_t* coef_offsets, size_t* basis_offsets) {
  if (!h) return qu_fail(MSFEC_ERR_INVALID, "null handle");
  if (coef_offsets) std::copy(h->coef_off.begin(), h->coef_off.end(), coef_offsets);
  if (basis_offsets) std::copy(h->psi_off.begin(), h->psi_off.end(), basis_offsets);
  return MSFEC_OK;
}

msfec_status msfec_qu_set_coefficient(msfec_qu_handle h, const double* data, size_t count, int loc) {
  if (!h) return qu_fail(MSFEC_ERR_INVALID, "null handle");
  QU_TRY(stage_in(h, data, count, loc, &Level::d_coef));
  for (Level& L : h->lv)
    QU_TRY(msfec_q_set_coefficient(L.h, L.d_coef, L.cells.size() * (size_t)L.nfc, MSFEC_DEVICE));
  h->coef_set = true;
  return MSFEC_OK;
}

msfec_status msfec_qu_set_rhs(msfec_qu_handle h, const double* data, size_t count, int loc) {
  if (!h) return qu_fail(MSFEC_ERR_INVALID, "null handle");
  if (h->p.rhs_mode != MSFEC_RHS_CELL)
    return qu_fail(MSFEC_ERR_STATE, "rhs_mode is not per cell; no data expected");
  QU_TRY(stage_in(h, data, count, loc, &Level::d_f));
  for (Level& L : h->lv)
    QU_TRY(msfec_q_set_rhs(L.h, L.d_f, L.cells.size() * (size_t)L.nfc, MSFEC_DEVICE));
  h->rhs_set = true;
  return MSFEC_OK;
}

msfec_status msfec_qu_set_reaction(msfec_qu_handle h, const double* data, size_t count, int loc) {
  if (!h) return qu_fail(MSFEC_ERR_INVALID, "null handle");
  if (h->p.reaction_mode != MSFEC_REACTION_CELL)
    return qu_fail(MSFEC_ERR_STATE, "reaction_mode takes no per-cell data");
  QU_TRY(stage_in(h, data, count, loc, &Level::d_reac));
  for (Level& L : h->lv)
    QU_TRY(msfec_q_set_reaction(L.h, L.d_reac, L.cells.size() * (size_t)L.nfc, MSFEC_DEVICE));
  h->reac_set = true;
  return MSFEC_OK;
}

msfec_status msfec_qu_tune(msfec_qu_handle h) {
  if (!h) return qu_fail(MSFEC_ERR_INVALID, "null handle");
  if (h->p.solver != MSFEC_SOLVER_AUTO) return MSFEC_OK;
  QU_HIP(hipStreamSynchronize(h->stream));
  return qu_tune(h);
}

msfec_status msfec_qu_level_info(msfec_qu_handle h, int level, int* solver, int* window,
                                 int* n_cells, float* ms5) {
  if (!h) return qu_fail(MSFEC_ERR_INVALID, "null handle");
  Level* L = find_level(h, level);
  if (!L) return qu_fail(MSFEC_ERR_INVALID, "no cell of level " + std::to_string(level));
  if (solver) *solver = L->solver;
  if (window) {
    // the window the level's batch runs with (0 where its solver has none)
    msfec_q_info info{};
    QU_TRY(msfec_q_get_info(L->h, &info));
    *window = (L->solver == MSFEC_SOLVER_BAND_FP64_MFMA) ? info.band_window : 0;
  }
  if (n_cells) *n_cells = (int)L->cells.size();
  if (ms5)
    for (int k = 0; k < kNCand; ++k) ms5[k] = L->ms[k];
  return MSFEC_OK;
}

msfec_status msfec_qu_level_handle(msfec_qu_handle h, int level, msfec_q_handle* out) {
  if (!h || !out) return qu_fail(MSFEC_ERR_INVALID, "null argument");
  Level* L = find_level(h, level);
  if (!L) return qu_fail(MSFEC_ERR_INVALID, "no cell of level " + std::to_string(level));
  *out = L->h;
  return MSFEC_OK;
}

msfec_status msfec_qu_compute(msfec_qu_handle h, void* stream) {
  if (!h) return qu_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->coef_set) return qu_fail(MSFEC_ERR_STATE, "coefficient not set");
  hipStream_t s = (hipStream_t)stream;
  if (!h->tuned) {
    QU_HIP(hipStreamSynchronize(s));
    QU_TRY(qu_tune(h));
  }
  h->stream = s;
  QU_HIP(hipEventRecord(h->fork, s));
  // the costliest levels first: the finest level starts before the others
  for (auto it = h->lv.rbegin(); it != h->lv.rend(); ++it) {
    Level& L = *it;
    QU_HIP(hipStreamWaitEvent(L.stream, h->fork, 0));
    QU_TRY(msfec_q_compute(L.h, L.stream));
    QU_HIP(hipEventRecord(L.done, L.stream));
  }
  for (Level& L : h->lv) QU_HIP(hipStreamWaitEvent(s, L.done, 0));
  h->computed = true;
  return MSFEC_OK;
}

msfec_status msfec_qu_synchronize(msfec_qu_handle h) {
  if (!h) return qu_fail(MSFEC_ERR_INVALID, "null handle");
  QU_HIP(hipStreamSynchronize(h->stream));
  for (Level& L : h->lv) QU_HIP(hipStreamSynchronize(L.stream));
  return MSFEC_OK;
}

msfec_status msfec_qu_get_element(msfec_qu_handle h, double* M, double* b, int loc) {
  if (!h) return qu_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->computed) return qu_fail(MSFEC_ERR_STATE, "compute not called");
  QU_TRY(msfec_qu_synchronize(h));
  const int nv = h->nv;
  for (Level& L : h->lv) {
    const size_t n = L.cells.size();
    if (loc == MSFEC_DEVICE) {
      // M rows of cell i at cell index * nv * nv
      for (int part = 0; part < 2; ++part) {
        double* dst = part == 0 ? M : b;
        if (!dst) continue;
        const int len = part == 0 ? nv * nv : nv;
        QU_TRY(msfec_q_get_element(L.h, part == 0 ? h->d_stage : nullptr,
                                   part == 0 ? nullptr : h->d_stage, MSFEC_DEVICE));
        const unsigned blocks = (unsigned)((n * len + 255) / 256);
        hipLaunchKernelGGL(qu_scatter_kernel, dim3(blocks), dim3(256), 0, L.stream, h->d_stage,
                           L.d_off_cell, (size_t)len, len, (int)n, dst);
        QU_HIP(hipGetLastError());
        QU_HIP(hipStreamSynchronize(L.stream));
      }
    } else {
      std::vector<double> tM(n * nv * nv), tb(n * nv);
      QU_TRY(msfec_q_get_element(L.h, tM.data(), tb.data(), MSFEC_HOST));
      for (size_t i = 0; i < n; ++i) {
        const size_t c = (size_t)L.cells[i];
        if (M) std::memcpy(M + c * nv * nv, tM.data() + i * nv * nv, nv * nv * sizeof(double));
        if (b) std::memcpy(b + c * nv, tb.data() + i * nv, nv * sizeof(double));
      }
    }
  }
  return MSFEC_OK;
}

msfec_status msfec_qu_get_basis(msfec_qu_handle h, double* Psi, int loc) {
  if (!h || !Psi) return qu_fail(MSFEC_ERR_INVALID, "null argument");
  if (!h->computed) return qu_fail(MSFEC_ERR_STATE, "compute not called");
  QU_TRY(msfec_qu_synchronize(h));
  const int nv = h->nv;
  for (Level& L : h->lv) {
    const size_t n = L.cells.size(), len = (size_t)nv * L.nn;
    if (loc == MSFEC_DEVICE) {
      QU_TRY(msfec_q_get_basis(L.h, h->d_stage, MSFEC_DEVICE));
      const unsigned blocks = (unsigned)((n * len + 255) / 256);
      hipLaunchKernelGGL(qu_scatter_kernel, dim3(blocks), dim3(256), 0, L.stream, h->d_stage,
                         L.d_off_psi, (size_t)1, (int)len, (int)n, Psi);
      QU_HIP(hipGetLastError());
      QU_HIP(hipStreamSynchronize(L.stream));
    } else {
      std::vector<double> t(n * len);
      QU_TRY(msfec_q_get_basis(L.h, t.data(), MSFEC_HOST));
      for (size_t i = 0; i < n; ++i)
        std::memcpy(Psi + h->psi_off[L.cells[i]], t.data() + i * len, len * sizeof(double));
    }
  }
  return MSFEC_OK;
}

msfec_status msfec_qu_get_cell_status(msfec_qu_handle h, int* status) {
  if (!h || !status) return qu_fail(MSFEC_ERR_INVALID, "null argument");
  if (!h->computed) return qu_fail(MSFEC_ERR_STATE, "compute not called");
  QU_TRY(msfec_qu_synchronize(h));
  for (Level& L : h->lv) {
    std::vector<int> t(L.cells.size());
    QU_TRY(msfec_q_get_cell_status(L.h, t.data()));
    for (size_t i = 0; i < t.size(); ++i) status[L.cells[i]] = t[i];
  }
  return MSFEC_OK;
}

msfec_status msfec_qu_coarse_solve(msfec_qu_handle h, int nc, double abs_tol, double rel_tol,
                                   int max_iter, double* u, int u_loc, double* weights,
                                   int weights_loc, int* iterations, double* residual) {
  if (!h) return qu_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->computed) return qu_fail(MSFEC_ERR_STATE, "compute not called");
  const int dim = h->p.dim, nv = h->nv;
  if (nc <= 0 || ipow_qu(nc, dim) != h->n_cells)
    return qu_fail(MSFEC_ERR_INVALID, "nc^dim must equal the number of cells of the batch");
  QU_TRY(msfec_qu_get_element(h, h->d_M, h->d_b, MSFEC_DEVICE));
  const int nnode = ipow_qu(nc + 1, dim), nst = ipow_qu(3, dim);
  if (h->coarse_nc != nc) {
    for (void* p : {(void*)h->d_cK, (void*)h->d_crhs, (void*)h->d_cvec, (void*)h->d_cit})
      if (p) (void)hipFree(p);
    h->d_cK = h->d_crhs = h->d_cvec = nullptr;
    h->d_cit = nullptr;
    h->coarse_nc = -1;
    QU_HIP(hipMalloc((void**)&h->d_cK, (size_t)nnode * nst * sizeof(double)));
    QU_HIP(hipMalloc((void**)&h->d_crhs, (size_t)nnode * sizeof(double)));
    QU_HIP(hipMalloc((void**)&h->d_cvec,
                     ((size_t)5 * nnode + 1 + (size_t)h->n_cells * nv) * sizeof(double)));
    QU_HIP(hipMalloc((void**)&h->d_cit, sizeof(int)));
    h->coarse_nc = nc;
  }
  hipStream_t s = h->stream;
  QU_HIP(msfec_detail::coarse_solve_device(dim, nc, h->d_M, h->d_b, h->d_cK, h->d_crhs, h->d_cvec,
                                           h->d_cit, abs_tol, rel_tol, max_iter, s));
  QU_HIP(hipStreamSynchronize(s));
  int it = 0;
  double rr = 0.0;
  const double* du = h->d_cvec;
  const double* res = du + 5 * (size_t)nnode;
  const double* dw = res + 1;
  QU_HIP(hipMemcpy(&it, h->d_cit, sizeof(int), hipMemcpyDeviceToHost));
  QU_HIP(hipMemcpy(&rr, res, sizeof(double), hipMemcpyDeviceToHost));
  if (iterations) *iterations = it;
  if (residual) *residual = rr;
  auto out = [&](double* dst, const double* src, size_t count, int loc) -> msfec_status {
    QU_HIP(hipMemcpyAsync(dst, src, count * sizeof(double),
                          loc == MSFEC_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    QU_HIP(hipStreamSynchronize(s));
    return MSFEC_OK;
  };
  if (u) QU_TRY(out(u, du, nnode, u_loc));
  if (weights) QU_TRY(out(weights, dw, (size_t)h->n_cells * nv, weights_loc));
  if (it < 0) return qu_fail(MSFEC_ERR_CELL, "coarse CG did not converge");
  return MSFEC_OK;
}

void msfec_qu_destroy(msfec_qu_handle h) {
  if (!h) return;
  (void)msfec_qu_synchronize(h);
  qu_free(h);
  delete h;
}

}  // extern "C"
