// QBasis / QBatch shim over the C-ABI (see msfec_q_basis.hpp).
#include "msfec_q_basis.hpp"

#include "msfec_vtu.hpp"

#include <string>

namespace msfec {

namespace {
void check(msfec_status st, const char* what) {
  if (st != MSFEC_OK)
    throw std::runtime_error(std::string(what) + ": " + msfec_last_error());
}
}  // namespace

QBatch::QBatch(const ParametersBasis& prm) : prm_(prm) {
  if (prm.dim != 2 && prm.dim != 3) throw std::invalid_argument("dim must be 2 or 3");
}

QBatch::~QBatch() {
  if (handle_) msfec_q_destroy(handle_);
  if (uhandle_) msfec_qu_destroy(uhandle_);
}

int QBatch::n_fine_nodes(std::size_t slot) const {
  const int S = (1 << (slot < levels_.size() ? levels_[slot] : (int)prm_.n_refine_local)) + 1;
  return (prm_.dim == 2) ? S * S : S * S * S;
}

bool QBatch::uneven() const {
  for (int l : levels_)
    if (l != levels_.front()) return true;
  return false;
}

std::size_t QBatch::add_cell(const std::vector<double>& vertices, const std::vector<double>& coef,
                             const std::vector<double>& f, const std::vector<double>& reac,
                             int level) {
  const std::size_t nvd = (std::size_t)(1 << prm_.dim) * prm_.dim;
  if (vertices.size() != nvd) throw std::invalid_argument("wrong number of vertex coordinates");
  if (level < 0) level = (int)prm_.n_refine_local;
  if (level < 1) throw std::invalid_argument("refinement level must be >= 1");
  std::size_t nfc = 1;
  for (int d = 0; d < prm_.dim; ++d) nfc <<= level;
  if (prm_.coef_mode == MSFEC_COEF_SCALAR_CELL && coef.size() != nfc)
    throw std::invalid_argument("coefficient block size does not match the cell's level");
  if (n_cells_ > 0 && prm_.coef_mode != MSFEC_COEF_SCALAR_CELL &&
      coef.size() * n_cells_ != coef_.size())
    throw std::invalid_argument("coefficient block size differs between cells");
  verts_.insert(verts_.end(), vertices.begin(), vertices.end());
  coef_.insert(coef_.end(), coef.begin(), coef.end());
  f_.insert(f_.end(), f.begin(), f.end());
  reac_.insert(reac_.end(), reac.begin(), reac.end());
  levels_.push_back(level);
  computed_ = false;
  return n_cells_++;
}

void QBatch::compute() {
  if (computed_) return;
  if (n_cells_ == 0) throw std::logic_error("no cells registered");
  if (handle_) {
    msfec_q_destroy(handle_);
    handle_ = nullptr;
  }
  if (uhandle_) {
    msfec_qu_destroy(uhandle_);
    uhandle_ = nullptr;
  }
  msfec_q_params p{};
  p.dim = prm_.dim;
  p.n_refine_local = levels_.front();
  p.coef_mode = prm_.coef_mode;
  p.rhs_mode = prm_.rhs_mode;
  p.rhs_const = prm_.rhs_const;
  p.reaction_mode = prm_.reaction_mode;
  p.reaction_const = prm_.reaction_const;
  p.solver = prm_.solver;
  p.max_refine = prm_.max_refine;
  p.cg_tol = prm_.cg_tol;
  p.cg_max_iter = prm_.cg_max_iter;
  p.flags = 0;
  const int nb = n_basis();
  psi_off_.assign(n_cells_ + 1, 0);
  for (std::size_t c = 0; c < n_cells_; ++c) psi_off_[c + 1] = psi_off_[c] + (std::size_t)nb * n_fine_nodes(c);
  M_.resize(n_cells_ * nb * nb);
  b_.resize(n_cells_ * nb);
  psi_.resize(psi_off_.back());
  status_.resize(n_cells_);
  if (uneven()) {
    check(msfec_qu_create(&p, (int)n_cells_, verts_.data(), MSFEC_HOST, levels_.data(), &uhandle_),
          "msfec_qu_create");
    check(msfec_qu_set_coefficient(uhandle_, coef_.data(), coef_.size(), MSFEC_HOST),
          "msfec_qu_set_coefficient");
    if (!f_.empty()) check(msfec_qu_set_rhs(uhandle_, f_.data(), f_.size(), MSFEC_HOST), "msfec_qu_set_rhs");
    if (!reac_.empty())
      check(msfec_qu_set_reaction(uhandle_, reac_.data(), reac_.size(), MSFEC_HOST),
            "msfec_qu_set_reaction");
    check(msfec_qu_compute(uhandle_, nullptr), "msfec_qu_compute");
    check(msfec_qu_get_element(uhandle_, M_.data(), b_.data(), MSFEC_HOST), "msfec_qu_get_element");
    check(msfec_qu_get_basis(uhandle_, psi_.data(), MSFEC_HOST), "msfec_qu_get_basis");
    check(msfec_qu_get_cell_status(uhandle_, status_.data()), "msfec_qu_get_cell_status");
  } else {
    if (p.solver == MSFEC_SOLVER_AUTO) p.solver = MSFEC_SOLVER_BAND_FP64_MFMA;
    check(msfec_q_create(&p, (int)n_cells_, verts_.data(), MSFEC_HOST, &handle_), "msfec_q_create");
    check(msfec_q_set_coefficient(handle_, coef_.data(), coef_.size(), MSFEC_HOST),
          "msfec_q_set_coefficient");
    if (!f_.empty()) check(msfec_q_set_rhs(handle_, f_.data(), f_.size(), MSFEC_HOST), "msfec_q_set_rhs");
    if (!reac_.empty())
      check(msfec_q_set_reaction(handle_, reac_.data(), reac_.size(), MSFEC_HOST),
            "msfec_q_set_reaction");
    check(msfec_q_compute(handle_, nullptr), "msfec_q_compute");
    check(msfec_q_get_element(handle_, M_.data(), b_.data(), MSFEC_HOST), "msfec_q_get_element");
    check(msfec_q_get_basis(handle_, psi_.data(), MSFEC_HOST), "msfec_q_get_basis");
    check(msfec_q_get_cell_status(handle_, status_.data()), "msfec_q_get_cell_status");
  }
  computed_ = true;
}

const double* QBatch::element_matrix(std::size_t slot) const {
  return M_.data() + slot * n_basis() * n_basis();
}
const double* QBatch::element_rhs(std::size_t slot) const { return b_.data() + slot * n_basis(); }
const double* QBatch::basis(std::size_t slot, int i) const {
  return psi_.data() + psi_off_[slot] + (std::size_t)i * n_fine_nodes(slot);
}
int QBatch::cell_status(std::size_t slot) const { return status_[slot]; }
const double* QBatch::vertices(std::size_t slot) const {
  return verts_.data() + slot * (std::size_t)(n_basis() * prm_.dim);
}

std::vector<double> QBatch::coarse_solve(int nc, double abs_tol, double rel_tol, int max_iter,
                                         int* iterations, double* residual) {
  compute();
  std::vector<double> w(n_cells_ * n_basis());
  if (uhandle_)
    check(msfec_qu_coarse_solve(uhandle_, nc, abs_tol, rel_tol, max_iter, nullptr, MSFEC_HOST,
                                w.data(), MSFEC_HOST, iterations, residual),
          "msfec_qu_coarse_solve");
  else
    check(msfec_coarse_solve(handle_, nc, abs_tol, rel_tol, max_iter, nullptr, MSFEC_HOST, w.data(),
                             MSFEC_HOST, iterations, residual),
          "msfec_coarse_solve");
  return w;
}

QBasis::QBasis(std::shared_ptr<QBatch> batch, const std::vector<double>& corner_points,
               const std::vector<double>& coef, const std::vector<double>& f,
               const std::vector<double>& reac, const std::string& cell_id, int level)
    : batch_(std::move(batch)) {
  if (!batch_) throw std::invalid_argument("null batch");
  slot_ = batch_->add_cell(corner_points, coef, f, reac, level);
  cell_id_ = cell_id.empty() ? std::to_string(slot_) : cell_id;
}

void QBasis::run() {
  run_called_ = true;
  const ParametersBasis& prm = batch_->parameters();
  filename_global_ = prm.filename_output + "." + int_to_string(prm.local_subdomain, 5) +
                     ".cell-" + cell_id_ + ".vtu";
}

void QBasis::set_output_flag() {
  if (!batch_->parameters().prevent_output) output_flag_ = (slot_ == 0);
}

void QBasis::output_basis() const {
  ensure();
  const ParametersBasis& prm = batch_->parameters();
  for (int i = 0; i < batch_->n_basis(); ++i) {
    const std::string name = "basis_q." + int_to_string(prm.local_subdomain, 5) + ".cell-" +
                             cell_id_ + ".index-" + int_to_string((unsigned)i, 2) + ".vtu";
    write_vtu_box(prm.dirname_output + "/" + name, prm.dim, batch_->vertices(slot_),
                  batch_->level(slot_), {"u"}, {batch_->basis(slot_, i)});
  }
}

void QBasis::output_global_solution_in_cell() const {
  if (solution_.empty()) throw std::logic_error("global weights must be set first");
  const ParametersBasis& prm = batch_->parameters();
  write_vtu_box(prm.dirname_output + "/" + filename_global_, prm.dim, batch_->vertices(slot_),
                batch_->level(slot_), {"u"}, {solution_.data()});
}

void QBasis::ensure() const {
  if (!run_called_) throw std::logic_error("QBasis::run() must be called first");
  batch_->compute();
  if (batch_->cell_status(slot_) != MSFEC_CELL_OK)
    throw std::runtime_error("local solve failed in cell slot " + std::to_string(slot_));
  if (M_.empty()) {
    const int nb = batch_->n_basis();
    M_.assign(batch_->element_matrix(slot_), batch_->element_matrix(slot_) + nb * nb);
    b_.assign(batch_->element_rhs(slot_), batch_->element_rhs(slot_) + nb);
  }
}

const std::vector<double>& QBasis::get_global_element_matrix() const {
  ensure();
  return M_;
}

const std::vector<double>& QBasis::get_global_element_rhs() const {
  ensure();
  return b_;
}

std::vector<double> QBasis::get_basis(int i) const {
  ensure();
  const double* p = batch_->basis(slot_, i);
  return std::vector<double>(p, p + batch_->n_fine_nodes(slot_));
}

void QBasis::set_global_weights(const std::vector<double>& weights) {
  ensure();
  const int nb = batch_->n_basis();
  if ((int)weights.size() != nb) throw std::invalid_argument("expected 2^dim weights");
  weights_ = weights;
  const int nn = batch_->n_fine_nodes(slot_);
  solution_.assign(nn, 0.0);
  for (int i = 0; i < nb; ++i) {
    const double* p = batch_->basis(slot_, i);
    for (int n = 0; n < nn; ++n) solution_[n] = solution_[n] + weights_[i] * p[n];
  }
}

const std::vector<double>& QBasis::get_global_solution() const { return solution_; }

int QBasis::status() const {
  batch_->compute();
  return batch_->cell_status(slot_);
}

}  // namespace msfec
