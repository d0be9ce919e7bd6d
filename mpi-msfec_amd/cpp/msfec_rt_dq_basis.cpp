// RTDQBasis / RTDQBatch shim over the C-ABI (see msfec_rt_dq_basis.hpp).
#include "msfec_rt_dq_basis.hpp"

#include <string>

#include "msfec_vtu.hpp"

namespace msfec {

namespace {
void check(msfec_status st, const char* what) {
  if (st != MSFEC_OK)
    throw std::runtime_error(std::string(what) + ": " + msfec_last_error());
}
}  // namespace

RTDQBatch::RTDQBatch(const ParametersRTDQ& prm) : prm_(prm) {
  const int m = 1 << prm.n_refine_local;
  n_fine_cells_ = m * m;
  n_edges_ = 2 * m * (m + 1);
}

RTDQBatch::~RTDQBatch() {
  if (handle_) msfec_rt_destroy(handle_);
}

std::size_t RTDQBatch::add_cell(const std::vector<double>& vertices, const std::vector<double>& a,
                                const std::vector<double>& f) {
  const std::size_t nf = (std::size_t)n_fine_cells_;
  const std::size_t na = nf * (prm_.coef_mode == MSFEC_COEF_TENSOR_QP ? 12 : 1);
  if (vertices.size() != 8) throw std::invalid_argument("a cell has 4 vertices x 2 coordinates");
  if (a.size() != na || f.size() != nf) throw std::invalid_argument("per-fine-cell data of the wrong size");
  verts_.insert(verts_.end(), vertices.begin(), vertices.end());
  a_.insert(a_.end(), a.begin(), a.end());
  f_.insert(f_.end(), f.begin(), f.end());
  computed_ = false;
  return n_cells_++;
}

void RTDQBatch::compute() {
  if (computed_) return;
  if (n_cells_ == 0) throw std::logic_error("no cells registered");
  if (handle_) {
    msfec_rt_destroy(handle_);
    handle_ = nullptr;
  }
  msfec_rt_params p{};
  p.n_refine_local = (int)prm_.n_refine_local;
  p.flags = 0;
  p.coef_mode = prm_.coef_mode;
  check(msfec_rt_create(&p, (int)n_cells_, verts_.data(), MSFEC_HOST, &handle_), "msfec_rt_create");
  check(msfec_rt_set_coefficients(handle_, a_.data(), f_.data(), MSFEC_HOST),
        "msfec_rt_set_coefficients");
  check(msfec_rt_compute(handle_, nullptr), "msfec_rt_compute");
  M_.resize(n_cells_ * 25);
  b_.resize(n_cells_ * 5);
  sigma_.resize(n_cells_ * 4 * (std::size_t)n_edges_);
  status_.resize(n_cells_);
  check(msfec_rt_get_element(handle_, M_.data(), b_.data(), MSFEC_HOST), "msfec_rt_get_element");
  check(msfec_rt_get_basis(handle_, sigma_.data(), MSFEC_HOST), "msfec_rt_get_basis");
  check(msfec_rt_get_cell_status(handle_, status_.data()), "msfec_rt_get_cell_status");
  computed_ = true;
}

const double* RTDQBatch::element_matrix(std::size_t slot) const { return &M_[slot * 25]; }
const double* RTDQBatch::element_rhs(std::size_t slot) const { return &b_[slot * 5]; }
const double* RTDQBatch::flux(std::size_t slot, int i) const {
  return &sigma_[(slot * 4 + (std::size_t)i) * n_edges_];
}
int RTDQBatch::cell_status(std::size_t slot) const { return status_[slot]; }
const double* RTDQBatch::vertices(std::size_t slot) const { return &verts_[slot * 8]; }

namespace {
// The flux field at the fine cell centres (3 components, z = 0) from the
// fine edge fluxes: the mean of the two opposite fluxes over the edge length.
std::vector<double> cell_centre_flux(const double* flux, const double* verts, int l) {
  const int m = 1 << l, nv = m * (m + 1);
  const double hx = (verts[2] - verts[0]) / m, hy = (verts[5] - verts[1]) / m;
  std::vector<double> out((size_t)m * m * 3, 0.0);
  for (int q = 0; q < m; ++q)
    for (int p = 0; p < m; ++p) {
      double* o = &out[((size_t)q * m + p) * 3];
      o[0] = 0.5 * (flux[q * (m + 1) + p] + flux[q * (m + 1) + p + 1]) / hy;
      o[1] = 0.5 * (flux[nv + q * m + p] + flux[nv + (q + 1) * m + p]) / hx;
    }
  return out;
}
}  // namespace

RTDQBasis::RTDQBasis(std::shared_ptr<RTDQBatch> batch, const std::vector<double>& corner_points,
                     const std::vector<double>& a, const std::vector<double>& f,
                     const std::string& cell_id)
    : batch_(std::move(batch)) {
  if (!batch_) throw std::invalid_argument("null batch");
  slot_ = batch_->add_cell(corner_points, a, f);
  cell_id_ = cell_id.empty() ? std::to_string(slot_) : cell_id;
}

void RTDQBasis::run() {
  run_called_ = true;
  const ParametersRTDQ& prm = batch_->parameters();
  filename_global_ = prm.filename_output + "." + int_to_string(prm.local_subdomain, 5) +
                     ".cell-" + cell_id_ + ".vtu";
}

void RTDQBasis::set_output_flag() {
  if (!batch_->parameters().prevent_output) output_flag_ = (slot_ == 0);
}

void RTDQBasis::output_basis() const {
  ensure();
  const ParametersRTDQ& prm = batch_->parameters();
  const int l = (int)prm.n_refine_local;
  const std::vector<double> ones((size_t)batch_->n_fine_cells(), 1.0);
  for (int i = 0; i < 4; ++i) {
    const std::vector<double> sig = cell_centre_flux(batch_->flux(slot_, i), batch_->vertices(slot_), l);
    const std::string name = "basis_rt-dq.div." + int_to_string(prm.local_subdomain, 5) + ".cell-" +
                             cell_id_ + ".index-" + int_to_string((unsigned)i, 2) + ".vtu";
    write_vtu_box(prm.dirname_output + "/" + name, 2, batch_->vertices(slot_), l, {}, {},
                  {{"sigma", 3, sig.data()}, {"u", 1, ones.data()}});
  }
}

void RTDQBasis::output_global_solution_in_cell() const {
  if (solution_.empty()) throw std::logic_error("global weights must be set first");
  const ParametersRTDQ& prm = batch_->parameters();
  const int l = (int)prm.n_refine_local;
  const std::vector<double> sig = cell_centre_flux(solution_.data(), batch_->vertices(slot_), l);
  write_vtu_box(prm.dirname_output + "/" + filename_global_, 2, batch_->vertices(slot_), l, {}, {},
                {{"sigma", 3, sig.data()}, {"u", 1, solution_.data() + batch_->n_fluxes()}});
}

void RTDQBasis::ensure() const {
  if (!run_called_) throw std::logic_error("run() was not called on this cell");
  batch_->compute();
  if (batch_->cell_status(slot_) != MSFEC_CELL_OK)
    throw std::runtime_error("the local solve of this cell failed (status " +
                             std::to_string(batch_->cell_status(slot_)) + ")");
}

const std::vector<double>& RTDQBasis::get_global_element_matrix() const {
  ensure();
  const double* m = batch_->element_matrix(slot_);
  M_.assign(m, m + 25);
  return M_;
}

const std::vector<double>& RTDQBasis::get_global_element_rhs() const {
  ensure();
  const double* b = batch_->element_rhs(slot_);
  b_.assign(b, b + 5);
  return b_;
}

std::vector<double> RTDQBasis::get_flux(int i) const {
  ensure();
  if (i < 0 || i >= 4) throw std::out_of_range("basis index");
  const double* p = batch_->flux(slot_, i);
  return std::vector<double>(p, p + batch_->n_fluxes());
}

void RTDQBasis::set_global_weights(const std::vector<double>& weights) {
  ensure();
  if (weights.size() != 5) throw std::invalid_argument("5 weights expected");
  const int ne = batch_->n_fluxes(), nf = batch_->n_fine_cells();
  solution_.assign((std::size_t)ne + nf, weights[4]);
  for (int e = 0; e < ne; ++e) solution_[e] = 0.0;
  for (int i = 0; i < 4; ++i) {
    const double* p = batch_->flux(slot_, i);
    for (int e = 0; e < ne; ++e) solution_[e] += weights[i] * p[e];
  }
}

const std::vector<double>& RTDQBasis::get_global_solution() const { return solution_; }

int RTDQBasis::status() const {
  batch_->compute();
  return batch_->cell_status(slot_);
}

}  // namespace msfec
