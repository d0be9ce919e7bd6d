// C++ drop-in for MsFEC's per-cell Q::QBasis (include/Q/q_basis.h:66-309)
// over the batched C-ABI of libmsfec_hip.so (include/msfec.h).
//
// The reference builds one QBasis per coarse cell, copies it into a
// std::map<CellId, QBasis> (source/Q/q_global.cc:50-71) and calls run() on
// each entry in turn (q_global.cc:83-86).  Here every QBasis registers its
// cell with a shared QBatch; run() only marks the cell, and the first getter
// called on any cell of the batch computes all registered cells in one GPU
// launch sequence.  Copies share the batch slot (the reference's copy
// constructor copies the object state, q_basis.cc:60-90), so map.insert works
// unchanged.  Errors are reported as std::runtime_error on the C++ side; no
// exception crosses the C-ABI.
#pragma once

#include <array>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "msfec.h"

namespace msfec {

// The fields of Q::ParametersBasis the basis reads (include/Q/q_parameters.h:
// 78-101), plus the coefficient/load data the reference reads from the
// parameter file.
struct ParametersBasis {
  int dim = 3;
  unsigned n_refine_local = 3;
  bool verbose = false;
  int coef_mode = MSFEC_COEF_SCALAR_CELL;
  int rhs_mode = MSFEC_RHS_CONST;
  double rhs_const = 1.0;
  int reaction_mode = MSFEC_REACTION_NONE;
  double reaction_const = 0.0;
  int solver = MSFEC_SOLVER_BAND_FP64;  // MSFEC_SOLVER_AUTO: per-level timing (uneven batches)
  int max_refine = 0;  // BAND_MIXED: refinement steps at most (0 = library default)
  double cg_tol = 0.0;  // CG: relative residual (0 = library default)
  int cg_max_iter = 0;  // CG: iterations at most (0 = library default)
  // output (ParametersBasis, include/Q/q_parameters.h:78-101)
  std::string dirname_output = "Q";            // directory of the VTU files
  std::string filename_output = "solution";    // stem of filename_global
  bool prevent_output = false;
  unsigned local_subdomain = 0;                // the rank in the file names
};

class QBatch {
 public:
  explicit QBatch(const ParametersBasis& prm);
  ~QBatch();
  QBatch(const QBatch&) = delete;
  QBatch& operator=(const QBatch&) = delete;

  // Register a cell: 2^dim vertices (deal.II order) and its coefficient
  // block (layout of msfec_q_set_coefficient), optional load / reaction
  // blocks, and its refinement level (-1: prm.n_refine_local).  Cells of
  // different levels make an uneven batch (msfec_qu_*), whose per-level
  // solver is chosen by timing unless prm.solver fixes it.  Returns the
  // batch slot.
  std::size_t add_cell(const std::vector<double>& vertices, const std::vector<double>& coef,
                       const std::vector<double>& f = {}, const std::vector<double>& reac = {},
                       int level = -1);

  // Compute every registered cell (idempotent until a new cell is added).
  void compute();
  bool computed() const { return computed_; }

  std::size_t n_cells() const { return n_cells_; }
  int n_basis() const { return 1 << prm_.dim; }
  // fine nodes of the cell in a slot ((2^level + 1)^dim)
  int n_fine_nodes(std::size_t slot = 0) const;
  int level(std::size_t slot) const { return levels_[slot]; }
  bool uneven() const;
  const ParametersBasis& parameters() const { return prm_; }

  const double* element_matrix(std::size_t slot) const;   // nb x nb
  const double* element_rhs(std::size_t slot) const;      // nb
  const double* basis(std::size_t slot, int i) const;     // n_fine_nodes
  const double* vertices(std::size_t slot) const;         // 2^dim x dim
  int cell_status(std::size_t slot) const;

  // Coarse problem on an nc^dim structured grid whose cells were added in
  // lexicographic order (QMultiscale::assemble_system + solve_iterative,
  // q_global.cc:198-363).  Returns the per-cell vertex weights
  // (send_global_weights_to_cell, q_global.cc:368-392), n_cells x nb.
  std::vector<double> coarse_solve(int nc, double abs_tol, double rel_tol, int max_iter,
                                   int* iterations = nullptr, double* residual = nullptr);

 private:
  ParametersBasis prm_;
  std::size_t n_cells_ = 0;
  std::vector<int> levels_;
  std::vector<double> verts_, coef_, f_, reac_;
  std::vector<double> M_, b_, psi_;
  std::vector<std::size_t> psi_off_;  // per slot: offset of its Psi block
  std::vector<int> status_;
  msfec_q_handle handle_ = nullptr;   // one level
  msfec_qu_handle uhandle_ = nullptr; // mixed levels
  bool computed_ = false;
};

class QBasis {
 public:
  QBasis() = delete;
  // Construct from a coarse cell (corner points + its coefficient block);
  // replaces QBasis(parameters_ms, prm_file, cell, first_cell, subdomain, comm)
  // (q_basis.cc:7-56).
  // cell_id names the cell in output file names (CellId::to_string in the
  // reference); empty means the batch slot.  level: the cell's own
  // n_refine_local (-1: the batch parameter's).
  QBasis(std::shared_ptr<QBatch> batch, const std::vector<double>& corner_points,
         const std::vector<double>& coef, const std::vector<double>& f = {},
         const std::vector<double>& reac = {}, const std::string& cell_id = "", int level = -1);
  QBasis(const QBasis& other) = default;
  QBasis& operator=(const QBasis& other) = default;

  // q_basis.cc:518-603: marks the cell; the batch is computed on first use.
  void run();

  // q_basis.cc:361-373 (row = test basis, nb x nb row-major)
  const std::vector<double>& get_global_element_matrix() const;
  const std::vector<double>& get_global_element_rhs() const;

  // The fine basis coefficients (solution_vector[i], q_basis.h:253) on the
  // lexicographic fine nodes of the cell.
  std::vector<double> get_basis(int i) const;

  // q_basis.cc:394-416: global_solution = sum_i w_i psi_i.
  void set_global_weights(const std::vector<double>& weights);
  const std::vector<double>& get_global_solution() const;

  // q_basis.cc:377-389, 420-426: filename_global = filename_output
  // + "." + subdomain (5 digits) + ".cell-" + cell id + ".vtu", set by run();
  // the output flag is raised on the batch's first cell unless
  // prevent_output (ParametersBasis::set_output_flag, q_parameters.cc:276-281).
  const std::string& get_filename_global() const { return filename_global_; }
  void set_output_flag();
  bool output_flag() const { return output_flag_; }
  // q_basis.cc:431-476: basis_q.<subdomain>.cell-<id>.index-<i>.vtu, one file
  // per basis, in dirname_output (the reference writes them from run() when
  // the output flag is set; here run() precedes the batch computation, so the
  // caller asks for them).
  void output_basis() const;
  // q_basis.cc:481-513: the global solution of set_global_weights as the
  // point field "u" in dirname_output / filename_global.
  void output_global_solution_in_cell() const;

  int status() const;
  std::size_t slot() const { return slot_; }

 private:
  void ensure() const;
  std::shared_ptr<QBatch> batch_;
  std::size_t slot_ = 0;
  bool run_called_ = false;
  mutable std::vector<double> M_, b_;
  std::vector<double> weights_, solution_;
  std::string cell_id_, filename_global_;
  bool output_flag_ = false;
};

}  // namespace msfec
