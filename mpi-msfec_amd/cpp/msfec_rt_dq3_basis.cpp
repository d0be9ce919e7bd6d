// RTDQ3Basis / RTDQ3Batch shim over the C-ABI (see msfec_rt_dq3_basis.hpp).
#include "msfec_rt_dq3_basis.hpp"

#include <string>

#include "msfec_vtu.hpp"

namespace msfec {

namespace {
void check(msfec_status st, const char* what) {
  if (st != MSFEC_OK)
    throw std::runtime_error(std::string(what) + ": " + msfec_last_error());
}
constexpr int kFaceBases = 6, kRows = 7;

// The flux at the fine cell centres (3 components per cell, x fastest): the
// mean of the two opposite face fluxes over the face area.
std::vector<double> cell_centre_flux(const double* s, const double* verts, int l) {
  const int m = 1 << l, m1 = m + 1, nf1 = m * m * m1;
  const double hx = (verts[3] - verts[0]) / m, hy = (verts[7] - verts[1]) / m,
               hz = (verts[14] - verts[2]) / m;
  std::vector<double> out((size_t)m * m * m * 3);
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) {
        double* w = &out[(((size_t)k * m + j) * m + i) * 3];
        w[0] = 0.5 * (s[(k * m + j) * m1 + i] + s[(k * m + j) * m1 + i + 1]) / (hy * hz);
        w[1] = 0.5 * (s[nf1 + (k * m1 + j) * m + i] + s[nf1 + (k * m1 + j + 1) * m + i]) / (hx * hz);
        w[2] = 0.5 * (s[2 * nf1 + (k * m + j) * m + i] + s[2 * nf1 + ((k + 1) * m + j) * m + i]) /
               (hx * hy);
      }
  return out;
}
}  // namespace

RTDQ3Batch::RTDQ3Batch(const ParametersRTDQ3& prm) : prm_(prm) {
  const int m = 1 << prm.n_refine_local;
  n_fine_cells_ = m * m * m;
  n_qp_ = n_fine_cells_ * 8;
  n_faces_ = 3 * m * m * (m + 1);
}

RTDQ3Batch::~RTDQ3Batch() {
  if (handle_) msfec_rt3_destroy(handle_);
}

std::size_t RTDQ3Batch::add_cell(const std::vector<double>& vertices,
                                 const std::vector<double>& ainv, const std::vector<double>& f) {
  const std::size_t nq = (std::size_t)n_qp_;
  if (vertices.size() != 24) throw std::invalid_argument("a cell has 8 vertices x 3 coordinates");
  if (ainv.size() != nq * 6 || (!f.empty() && f.size() != nq))
    throw std::invalid_argument("quadrature-point data of the wrong size");
  verts_.insert(verts_.end(), vertices.begin(), vertices.end());
  ainv_.insert(ainv_.end(), ainv.begin(), ainv.end());
  // an absent f is zero; the batch passes NULL when no cell has one
  if (f.empty()) f_.insert(f_.end(), nq, 0.0); else f_.insert(f_.end(), f.begin(), f.end());
  any_f_ = any_f_ || !f.empty();
  computed_ = false;
  return n_cells_++;
}

void RTDQ3Batch::compute() {
  if (computed_) return;
  if (n_cells_ == 0) throw std::logic_error("no cells registered");
  if (handle_) {
    msfec_rt3_destroy(handle_);
    handle_ = nullptr;
  }
  msfec_rt3_params p{};
  p.n_refine_local = (int)prm_.n_refine_local;
  p.flags = 0;
  p.max_iter = prm_.max_iter;
  p.rtol = prm_.rtol;
  check(msfec_rt3_create(&p, (int)n_cells_, verts_.data(), MSFEC_HOST, &handle_), "msfec_rt3_create");
  check(msfec_rt3_set_coefficients(handle_, ainv_.data(), any_f_ ? f_.data() : nullptr, MSFEC_HOST),
        "msfec_rt3_set_coefficients");
  check(msfec_rt3_compute(handle_, nullptr), "msfec_rt3_compute");
  M_.resize(n_cells_ * kRows * kRows);
  rhs_.resize(n_cells_ * kRows);
  S_.resize(n_cells_ * kFaceBases * (std::size_t)n_faces_);
  status_.resize(n_cells_);
  check(msfec_rt3_get_element(handle_, M_.data(), rhs_.data(), MSFEC_HOST), "msfec_rt3_get_element");
  check(msfec_rt3_get_basis(handle_, S_.data(), MSFEC_HOST), "msfec_rt3_get_basis");
  check(msfec_rt3_get_cell_status(handle_, status_.data()), "msfec_rt3_get_cell_status");
  computed_ = true;
}

const double* RTDQ3Batch::element_matrix(std::size_t slot) const { return &M_[slot * kRows * kRows]; }
const double* RTDQ3Batch::element_rhs(std::size_t slot) const { return &rhs_[slot * kRows]; }
const double* RTDQ3Batch::flux(std::size_t slot, int i) const {
  return &S_[(slot * kFaceBases + (std::size_t)i) * n_faces_];
}
int RTDQ3Batch::cell_status(std::size_t slot) const { return status_[slot]; }
const double* RTDQ3Batch::vertices(std::size_t slot) const { return &verts_[slot * 24]; }

RTDQ3Basis::RTDQ3Basis(std::shared_ptr<RTDQ3Batch> batch, const std::vector<double>& corner_points,
                       const std::vector<double>& ainv, const std::vector<double>& f,
                       const std::string& cell_id)
    : batch_(std::move(batch)) {
  if (!batch_) throw std::invalid_argument("null batch");
  slot_ = batch_->add_cell(corner_points, ainv, f);
  cell_id_ = cell_id.empty() ? std::to_string(slot_) : cell_id;
}

void RTDQ3Basis::run() {
  run_called_ = true;
  const ParametersRTDQ3& prm = batch_->parameters();
  filename_global_ = prm.filename_output + "." + int_to_string(prm.local_subdomain, 5) +
                     ".cell-" + cell_id_ + ".vtu";
}

void RTDQ3Basis::set_output_flag() {
  if (!batch_->parameters().prevent_output) output_flag_ = (slot_ == 0);
}

void RTDQ3Basis::output_basis() const {
  ensure();
  const ParametersRTDQ3& prm = batch_->parameters();
  const int l = (int)prm.n_refine_local;
  const std::vector<double> ones((size_t)batch_->n_fine_cells(), 1.0);
  for (int i = 0; i < kFaceBases; ++i) {
    const std::vector<double> sig = cell_centre_flux(batch_->flux(slot_, i), batch_->vertices(slot_), l);
    const std::string name = "basis_rt-dq.div." + int_to_string(prm.local_subdomain, 5) + ".cell-" +
                             cell_id_ + ".index-" + int_to_string((unsigned)i, 2) + ".vtu";
    write_vtu_box(prm.dirname_output + "/" + name, 3, batch_->vertices(slot_), l, {}, {},
                  {{"sigma", 3, sig.data()}, {"u", 1, ones.data()}});
  }
}

void RTDQ3Basis::output_global_solution_in_cell() const {
  if (solution_.empty()) throw std::logic_error("global weights must be set first");
  const ParametersRTDQ3& prm = batch_->parameters();
  const int l = (int)prm.n_refine_local;
  const std::vector<double> sig = cell_centre_flux(solution_.data(), batch_->vertices(slot_), l);
  write_vtu_box(prm.dirname_output + "/" + filename_global_, 3, batch_->vertices(slot_), l, {}, {},
                {{"sigma", 3, sig.data()}, {"u", 1, solution_.data() + batch_->n_faces()}});
}

void RTDQ3Basis::ensure() const {
  if (!run_called_) throw std::logic_error("run() was not called on this cell");
  batch_->compute();
  if (batch_->cell_status(slot_) != MSFEC_CELL_OK)
    throw std::runtime_error("the local solve of this cell failed (status " +
                             std::to_string(batch_->cell_status(slot_)) + ")");
}

const std::vector<double>& RTDQ3Basis::get_global_element_matrix() const {
  ensure();
  const double* m = batch_->element_matrix(slot_);
  M_.assign(m, m + kRows * kRows);
  return M_;
}

const std::vector<double>& RTDQ3Basis::get_global_element_rhs() const {
  ensure();
  const double* b = batch_->element_rhs(slot_);
  b_.assign(b, b + kRows);
  return b_;
}

std::vector<double> RTDQ3Basis::get_flux(int i) const {
  ensure();
  if (i < 0 || i >= kFaceBases) throw std::out_of_range("face basis index");
  const double* p = batch_->flux(slot_, i);
  return std::vector<double>(p, p + batch_->n_faces());
}

void RTDQ3Basis::set_global_weights(const std::vector<double>& weights) {
  ensure();
  if (weights.size() != (size_t)kRows) throw std::invalid_argument("7 weights expected");
  const int nf = batch_->n_faces();
  // the fluxes from the 6 face bases, u = w_6 (the unit u of set_u_to_std)
  solution_.assign((std::size_t)nf + batch_->n_fine_cells(), weights[6]);
  for (int e = 0; e < nf; ++e) solution_[e] = 0.0;
  for (int i = 0; i < kFaceBases; ++i) {
    const double* p = batch_->flux(slot_, i);
    for (int e = 0; e < nf; ++e) solution_[e] += weights[i] * p[e];
  }
}

const std::vector<double>& RTDQ3Basis::get_global_solution() const { return solution_; }

int RTDQ3Basis::status() const {
  batch_->compute();
  return batch_->cell_status(slot_);
}

}  // namespace msfec
