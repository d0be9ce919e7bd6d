// QNedBasis / QNedBatch shim over the C-ABI (see msfec_q_ned_basis.hpp).
#include "msfec_q_ned_basis.hpp"

#include <string>

#include "msfec_vtu.hpp"

namespace msfec {

namespace {
void check(msfec_status st, const char* what) {
  if (st != MSFEC_OK)
    throw std::runtime_error(std::string(what) + ": " + msfec_last_error());
}
}  // namespace

QNedBatch::QNedBatch(const ParametersNed& prm) : prm_(prm) {
  const int m = 1 << prm.n_refine_local;
  n_fine_cells_ = m * m;
  n_sigma_ = (m + 1) * (m + 1);
  n_dofs_ = n_sigma_ + 2 * m * (m + 1);
}

QNedBatch::~QNedBatch() {
  if (handle_) msfec_ned_destroy(handle_);
}

std::size_t QNedBatch::add_cell(const std::vector<double>& vertices, const std::vector<double>& a,
                                const std::vector<double>& binv, const std::vector<double>& r,
                                const std::vector<double>& f) {
  const std::size_t nf = (std::size_t)n_fine_cells_;
  if (vertices.size() != 8) throw std::invalid_argument("a cell has 4 vertices x 2 coordinates");
  if (a.size() != nf || binv.size() != nf || r.size() != nf || f.size() != 2 * nf)
    throw std::invalid_argument("per-fine-cell data of the wrong size");
  verts_.insert(verts_.end(), vertices.begin(), vertices.end());
  a_.insert(a_.end(), a.begin(), a.end());
  binv_.insert(binv_.end(), binv.begin(), binv.end());
  r_.insert(r_.end(), r.begin(), r.end());
  f_.insert(f_.end(), f.begin(), f.end());
  computed_ = false;
  return n_cells_++;
}

void QNedBatch::compute() {
  if (computed_) return;
  if (n_cells_ == 0) throw std::logic_error("no cells registered");
  if (handle_) {
    msfec_ned_destroy(handle_);
    handle_ = nullptr;
  }
  msfec_ned_params p{};
  p.n_refine_local = (int)prm_.n_refine_local;
  p.full_rhs = prm_.full_rhs ? 1 : 0;
  p.flags = 0;
  p.solver = prm_.solver;
  p.max_refine = prm_.max_refine;
  check(msfec_ned_create(&p, (int)n_cells_, verts_.data(), MSFEC_HOST, &handle_),
        "msfec_ned_create");
  check(msfec_ned_set_coefficients(handle_, a_.data(), binv_.data(), r_.data(), f_.data(),
                                   MSFEC_HOST),
        "msfec_ned_set_coefficients");
  check(msfec_ned_compute(handle_, nullptr), "msfec_ned_compute");
  M_.resize(n_cells_ * 64);
  b_.resize(n_cells_ * 8);
  psi_.resize(n_cells_ * 8 * (std::size_t)n_dofs_);
  status_.resize(n_cells_);
  check(msfec_ned_get_element(handle_, M_.data(), b_.data(), MSFEC_HOST), "msfec_ned_get_element");
  check(msfec_ned_get_basis(handle_, psi_.data(), MSFEC_HOST), "msfec_ned_get_basis");
  check(msfec_ned_get_cell_status(handle_, status_.data()), "msfec_ned_get_cell_status");
  computed_ = true;
}

const double* QNedBatch::element_matrix(std::size_t slot) const { return &M_[slot * 64]; }
const double* QNedBatch::element_rhs(std::size_t slot) const { return &b_[slot * 8]; }
const double* QNedBatch::basis(std::size_t slot, int i) const {
  return &psi_[(slot * 8 + (std::size_t)i) * n_dofs_];
}
int QNedBatch::cell_status(std::size_t slot) const { return status_[slot]; }
const double* QNedBatch::vertices(std::size_t slot) const { return &verts_[slot * 8]; }

namespace {
// The Nedelec field at the fine cell centres (3 components, z = 0) from the
// tangential line integrals: the mean of the two parallel edges over their
// length.  Edge layout: horizontal edge (i, j) at ns + j m + i, vertical edge
// (i, j) at ns + nh + j (m + 1) + i.
std::vector<double> cell_centre_u(const double* dofs, const double* verts, int l) {
  const int m = 1 << l, ns = (m + 1) * (m + 1), nh = m * (m + 1);
  const double hx = (verts[2] - verts[0]) / m, hy = (verts[5] - verts[1]) / m;
  std::vector<double> out((size_t)m * m * 3, 0.0);
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      double* o = &out[((size_t)j * m + i) * 3];
      o[0] = 0.5 * (dofs[ns + j * m + i] + dofs[ns + (j + 1) * m + i]) / hx;
      o[1] = 0.5 * (dofs[ns + nh + j * (m + 1) + i] + dofs[ns + nh + j * (m + 1) + i + 1]) / hy;
    }
  return out;
}
}  // namespace

QNedBasis::QNedBasis(std::shared_ptr<QNedBatch> batch, const std::vector<double>& corner_points,
                     const std::vector<double>& a, const std::vector<double>& binv,
                     const std::vector<double>& r, const std::vector<double>& f,
                     const std::string& cell_id)
    : batch_(std::move(batch)) {
  if (!batch_) throw std::invalid_argument("null batch");
  slot_ = batch_->add_cell(corner_points, a, binv, r, f);
  cell_id_ = cell_id.empty() ? std::to_string(slot_) : cell_id;
}

void QNedBasis::run() {
  run_called_ = true;
  const ParametersNed& prm = batch_->parameters();
  filename_global_ = prm.filename_output + "." + int_to_string(prm.local_subdomain, 5) +
                     ".cell-" + cell_id_ + ".vtu";
}

void QNedBasis::set_output_flag() {
  if (!batch_->parameters().prevent_output) output_flag_ = (slot_ == 0);
}

void QNedBasis::output_basis() const {
  ensure();
  const ParametersNed& prm = batch_->parameters();
  const int l = (int)prm.n_refine_local;
  for (int i = 0; i < 8; ++i) {
    const double* p = batch_->basis(slot_, i);
    const std::vector<double> u = cell_centre_u(p, batch_->vertices(slot_), l);
    const std::string name = std::string("basis_q-ned") + (i < 4 ? ".h1." : ".curl.") +
                             int_to_string(prm.local_subdomain, 5) + ".cell-" + cell_id_ +
                             ".index-" + int_to_string((unsigned)(i % 4), 2) + ".vtu";
    write_vtu_box(prm.dirname_output + "/" + name, 2, batch_->vertices(slot_), l, {"sigma"}, {p},
                  {{"u", 3, u.data()}});
  }
}

void QNedBasis::output_global_solution_in_cell() const {
  if (solution_.empty()) throw std::logic_error("global weights must be set first");
  const ParametersNed& prm = batch_->parameters();
  const int l = (int)prm.n_refine_local;
  const std::vector<double> u = cell_centre_u(solution_.data(), batch_->vertices(slot_), l);
  write_vtu_box(prm.dirname_output + "/" + filename_global_, 2, batch_->vertices(slot_), l,
                {"sigma"}, {solution_.data()}, {{"u", 3, u.data()}});
}

void QNedBasis::ensure() const {
  if (!run_called_) throw std::logic_error("run() was not called on this cell");
  batch_->compute();
  if (batch_->cell_status(slot_) != MSFEC_CELL_OK)
    throw std::runtime_error("the local solve of this cell failed (status " +
                             std::to_string(batch_->cell_status(slot_)) + ")");
}

const std::vector<double>& QNedBasis::get_global_element_matrix() const {
  ensure();
  const double* m = batch_->element_matrix(slot_);
  M_.assign(m, m + 64);
  return M_;
}

const std::vector<double>& QNedBasis::get_global_element_rhs() const {
  ensure();
  const double* b = batch_->element_rhs(slot_);
  b_.assign(b, b + 8);
  return b_;
}

std::vector<double> QNedBasis::get_basis(int i) const {
  ensure();
  if (i < 0 || i >= 8) throw std::out_of_range("basis index");
  const double* p = batch_->basis(slot_, i);
  return std::vector<double>(p, p + batch_->n_dofs());
}

// the sigma block from the vertex bases, the u block from the edge bases
void QNedBasis::set_global_weights(const std::vector<double>& weights) {
  ensure();
  if (weights.size() != 8) throw std::invalid_argument("8 weights expected");
  const int nd = batch_->n_dofs(), ns = batch_->n_sigma();
  solution_.assign(nd, 0.0);
  for (int i = 0; i < 4; ++i) {
    const double* p = batch_->basis(slot_, i);
    for (int d = 0; d < ns; ++d) solution_[d] += weights[i] * p[d];
  }
  for (int i = 4; i < 8; ++i) {
    const double* p = batch_->basis(slot_, i);
    for (int d = ns; d < nd; ++d) solution_[d] += weights[i] * p[d];
  }
}

const std::vector<double>& QNedBasis::get_global_solution() const { return solution_; }

int QNedBasis::status() const {
  batch_->compute();
  return batch_->cell_status(slot_);
}

}  // namespace msfec
