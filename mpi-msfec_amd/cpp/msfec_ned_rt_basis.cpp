// NedRTBasis / NedRTBatch shim over the C-ABI (see msfec_ned_rt_basis.hpp).
#include "msfec_ned_rt_basis.hpp"

#include <string>

#include "msfec_vtu.hpp"

namespace msfec {

namespace {
void check(msfec_status st, const char* what) {
  if (st != MSFEC_OK)
    throw std::runtime_error(std::string(what) + ": " + msfec_last_error());
}
constexpr int kBases = 18, kEdgeBases = 12;
}  // namespace

NedRTBatch::NedRTBatch(const ParametersNedRT& prm) : prm_(prm) {
  const int m = 1 << prm.n_refine_local;
  n_qp_ = m * m * m * 8;
  n_edges_ = 3 * m * (m + 1) * (m + 1);
  n_dofs_ = n_edges_ + 3 * m * m * (m + 1);
}

NedRTBatch::~NedRTBatch() {
  if (handle_) msfec_nrt_destroy(handle_);
}

std::size_t NedRTBatch::add_cell(const std::vector<double>& vertices, const std::vector<double>& ainv,
                                 const std::vector<double>& b, const std::vector<double>& r,
                                 const std::vector<double>& f) {
  const std::size_t nq = (std::size_t)n_qp_;
  if (vertices.size() != 24) throw std::invalid_argument("a cell has 8 vertices x 3 coordinates");
  if (ainv.size() != nq * 6 || b.size() != nq || (!r.empty() && r.size() != nq) ||
      (!f.empty() && f.size() != nq * 3))
    throw std::invalid_argument("quadrature-point data of the wrong size");
  verts_.insert(verts_.end(), vertices.begin(), vertices.end());
  ainv_.insert(ainv_.end(), ainv.begin(), ainv.end());
  b_.insert(b_.end(), b.begin(), b.end());
  // absent r or f are zero; the batch passes NULL when no cell has one
  if (r.empty()) r_.insert(r_.end(), nq, 0.0); else r_.insert(r_.end(), r.begin(), r.end());
  if (f.empty()) f_.insert(f_.end(), nq * 3, 0.0); else f_.insert(f_.end(), f.begin(), f.end());
  any_r_ = any_r_ || !r.empty();
  any_f_ = any_f_ || !f.empty();
  computed_ = false;
  return n_cells_++;
}

void NedRTBatch::compute() {
  if (computed_) return;
  if (n_cells_ == 0) throw std::logic_error("no cells registered");
  if (handle_) {
    msfec_nrt_destroy(handle_);
    handle_ = nullptr;
  }
  msfec_nrt_params p{};
  p.n_refine_local = (int)prm_.n_refine_local;
  p.flags = 0;
  p.solver = prm_.solver;
  p.max_iter = prm_.max_iter;
  p.rtol = prm_.rtol;
  check(msfec_nrt_create(&p, (int)n_cells_, verts_.data(), MSFEC_HOST, &handle_), "msfec_nrt_create");
  check(msfec_nrt_set_coefficients(handle_, ainv_.data(), b_.data(), any_r_ ? r_.data() : nullptr,
                                   any_f_ ? f_.data() : nullptr, MSFEC_HOST),
        "msfec_nrt_set_coefficients");
  check(msfec_nrt_compute(handle_, nullptr), "msfec_nrt_compute");
  M_.resize(n_cells_ * kBases * kBases);
  rhs_.resize(n_cells_ * kBases);
  X_.resize(n_cells_ * kBases * (std::size_t)n_dofs_);
  status_.resize(n_cells_);
  check(msfec_nrt_get_element(handle_, M_.data(), rhs_.data(), MSFEC_HOST), "msfec_nrt_get_element");
  check(msfec_nrt_get_basis(handle_, X_.data(), MSFEC_HOST), "msfec_nrt_get_basis");
  check(msfec_nrt_get_cell_status(handle_, status_.data()), "msfec_nrt_get_cell_status");
  computed_ = true;
}

const double* NedRTBatch::element_matrix(std::size_t slot) const { return &M_[slot * kBases * kBases]; }
const double* NedRTBatch::element_rhs(std::size_t slot) const { return &rhs_[slot * kBases]; }
const double* NedRTBatch::basis(std::size_t slot, int i) const {
  return &X_[(slot * kBases + (std::size_t)i) * n_dofs_];
}
int NedRTBatch::cell_status(std::size_t slot) const { return status_[slot]; }
const double* NedRTBatch::vertices(std::size_t slot) const { return &verts_[slot * 24]; }

namespace {
// The Nedelec field (from the fine edge values: line integrals along +x, +y,
// +z) and the Raviart-Thomas field (from the fine face fluxes) at the fine
// cell centres, 3 components each: the mean of the 4 parallel edges over the
// edge length, and of the 2 opposite faces over the face area.
void cell_centre_fields(const double* v, const double* verts, int l, std::vector<double>& sig,
                        std::vector<double>& u) {
  const int m = 1 << l, m1 = m + 1;
  const int nx = m * m1 * m1, ne = 3 * nx, nf1 = m * m * m1;
  const double hx = (verts[3] - verts[0]) / m, hy = (verts[7] - verts[1]) / m,
               hz = (verts[14] - verts[2]) / m;
  sig.assign((size_t)m * m * m * 3, 0.0);
  u.assign((size_t)m * m * m * 3, 0.0);
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) {
        double* s = &sig[(((size_t)k * m + j) * m + i) * 3];
        double* w = &u[(((size_t)k * m + j) * m + i) * 3];
        for (int a = 0; a < 2; ++a)
          for (int c = 0; c < 2; ++c) {
            s[0] += v[((k + c) * m1 + j + a) * m + i];
            s[1] += v[nx + ((k + c) * m + j) * m1 + i + a];
            s[2] += v[2 * nx + (k * m1 + j + c) * m1 + i + a];
          }
        s[0] *= 0.25 / hx;
        s[1] *= 0.25 / hy;
        s[2] *= 0.25 / hz;
        const double* f = v + ne;
        w[0] = 0.5 * (f[(k * m + j) * m1 + i] + f[(k * m + j) * m1 + i + 1]) / (hy * hz);
        w[1] = 0.5 * (f[nf1 + (k * m1 + j) * m + i] + f[nf1 + (k * m1 + j + 1) * m + i]) / (hx * hz);
        w[2] = 0.5 * (f[2 * nf1 + (k * m + j) * m + i] + f[2 * nf1 + ((k + 1) * m + j) * m + i]) /
               (hx * hy);
      }
}
}  // namespace

NedRTBasis::NedRTBasis(std::shared_ptr<NedRTBatch> batch, const std::vector<double>& corner_points,
                       const std::vector<double>& ainv, const std::vector<double>& b,
                       const std::vector<double>& r, const std::vector<double>& f,
                       const std::string& cell_id)
    : batch_(std::move(batch)) {
  if (!batch_) throw std::invalid_argument("null batch");
  slot_ = batch_->add_cell(corner_points, ainv, b, r, f);
  cell_id_ = cell_id.empty() ? std::to_string(slot_) : cell_id;
}

void NedRTBasis::run() {
  run_called_ = true;
  const ParametersNedRT& prm = batch_->parameters();
  filename_global_ = prm.filename_output + "." + int_to_string(prm.local_subdomain, 5) +
                     ".cell-" + cell_id_ + ".vtu";
}

void NedRTBasis::set_output_flag() {
  if (!batch_->parameters().prevent_output) output_flag_ = (slot_ == 0);
}

void NedRTBasis::output_basis() const {
  ensure();
  const ParametersNedRT& prm = batch_->parameters();
  const int l = (int)prm.n_refine_local;
  std::vector<double> sig, u;
  for (int i = 0; i < kBases; ++i) {
    cell_centre_fields(batch_->basis(slot_, i), batch_->vertices(slot_), l, sig, u);
    const bool edge = i < kEdgeBases;
    const std::string name = std::string("basis_ned-rt.") + (edge ? "curl." : "div.") +
                             int_to_string(prm.local_subdomain, 5) + ".cell-" + cell_id_ +
                             ".index-" + int_to_string((unsigned)(edge ? i : i - kEdgeBases), 2) +
                             ".vtu";
    write_vtu_box(prm.dirname_output + "/" + name, 3, batch_->vertices(slot_), l, {}, {},
                  {{"sigma", 3, sig.data()}, {"u", 3, u.data()}});
  }
}

void NedRTBasis::output_global_solution_in_cell() const {
  if (solution_.empty()) throw std::logic_error("global weights must be set first");
  const ParametersNedRT& prm = batch_->parameters();
  const int l = (int)prm.n_refine_local;
  std::vector<double> sig, u;
  cell_centre_fields(solution_.data(), batch_->vertices(slot_), l, sig, u);
  write_vtu_box(prm.dirname_output + "/" + filename_global_, 3, batch_->vertices(slot_), l, {}, {},
                {{"sigma", 3, sig.data()}, {"u", 3, u.data()}});
}

void NedRTBasis::ensure() const {
  if (!run_called_) throw std::logic_error("run() was not called on this cell");
  batch_->compute();
  if (batch_->cell_status(slot_) != MSFEC_CELL_OK)
    throw std::runtime_error("the local solve of this cell failed (status " +
                             std::to_string(batch_->cell_status(slot_)) + ")");
}

const std::vector<double>& NedRTBasis::get_global_element_matrix() const {
  ensure();
  const double* m = batch_->element_matrix(slot_);
  M_.assign(m, m + kBases * kBases);
  return M_;
}

const std::vector<double>& NedRTBasis::get_global_element_rhs() const {
  ensure();
  const double* b = batch_->element_rhs(slot_);
  b_.assign(b, b + kBases);
  return b_;
}

std::vector<double> NedRTBasis::get_basis(int i) const {
  ensure();
  if (i < 0 || i >= kBases) throw std::out_of_range("basis index");
  const double* p = batch_->basis(slot_, i);
  return std::vector<double>(p, p + batch_->n_dofs());
}

void NedRTBasis::set_global_weights(const std::vector<double>& weights) {
  ensure();
  if (weights.size() != (size_t)kBases) throw std::invalid_argument("18 weights expected");
  const int ne = batch_->n_edges(), n = batch_->n_dofs();
  solution_.assign((std::size_t)n, 0.0);
  // the sigma block from the edge bases, the u block from the face bases
  for (int i = 0; i < kBases; ++i) {
    const double* p = batch_->basis(slot_, i);
    const int lo = i < kEdgeBases ? 0 : ne, hi = i < kEdgeBases ? ne : n;
    for (int e = lo; e < hi; ++e) solution_[e] += weights[i] * p[e];
  }
}

const std::vector<double>& NedRTBasis::get_global_solution() const { return solution_; }

int NedRTBasis::status() const {
  batch_->compute();
  return batch_->cell_status(slot_);
}

}  // namespace msfec
