// C++ drop-in for MsFEC's per-cell QNedBasis (include/Q_Ned/q_ned_basis.h:
// 84-155) in its two-dimensional form, over the msfec_ned_* entry points of
// libmsfec_hip.so (include/msfec.h).  Same pattern as msfec::QBasis: every
// QNedBasis registers its cell with a shared QNedBatch; run() marks the
// cell and the first getter computes the whole batch.  Copies share the
// batch slot, so std::map<CellId, QNedBasis>::insert works unchanged.
#pragma once

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "msfec.h"

namespace msfec {

struct ParametersNed {
  unsigned n_refine_local = 4;  // ParametersMs::n_refine_local
  bool full_rhs = true;         // ParametersMs::full_rhs (q_ned_basis.cc:432-444)
  // local solver: MSFEC_SOLVER_BAND_FP64_MFMA (FP64 LDL^T) or
  // MSFEC_SOLVER_BAND_MIXED (FP32 LDL^T + FP64 refinement, at most
  // max_refine passes, 0 = 3); both replace solve_direct (q_ned_basis.cc:570-625)
  int solver = MSFEC_SOLVER_BAND_FP64_MFMA;
  int max_refine = 0;
  // output (the ParametersBasis fields of include/Q_Ned/q_ned_parameters.h)
  std::string dirname_output = "Q_Ned";
  std::string filename_output = "solution";
  bool prevent_output = false;
  unsigned local_subdomain = 0;
};

class QNedBatch {
 public:
  explicit QNedBatch(const ParametersNed& prm);
  ~QNedBatch();
  QNedBatch(const QNedBatch&) = delete;
  QNedBatch& operator=(const QNedBatch&) = delete;

  // Register a cell: 4 vertices x 2 coordinates (deal.II order) and its
  // per-fine-cell data (4^l values each; f: 2 per fine cell).
  std::size_t add_cell(const std::vector<double>& vertices, const std::vector<double>& a,
                       const std::vector<double>& binv, const std::vector<double>& r,
                       const std::vector<double>& f);
  void compute();
  bool computed() const { return computed_; }
  std::size_t n_cells() const { return n_cells_; }
  int n_dofs() const { return n_dofs_; }    // fine sigma + u coefficients of one basis
  int n_sigma() const { return n_sigma_; }  // the sigma ones, first

  const double* element_matrix(std::size_t slot) const;  // 8 x 8
  const double* element_rhs(std::size_t slot) const;     // 8
  const double* basis(std::size_t slot, int i) const;    // n_dofs
  const double* vertices(std::size_t slot) const;        // 4 x 2
  const ParametersNed& parameters() const { return prm_; }
  int cell_status(std::size_t slot) const;

 private:
  ParametersNed prm_;
  std::size_t n_cells_ = 0;
  int n_fine_cells_ = 0, n_sigma_ = 0, n_dofs_ = 0;
  std::vector<double> verts_, a_, binv_, r_, f_;
  std::vector<double> M_, b_, psi_;
  std::vector<int> status_;
  msfec_ned_handle handle_ = nullptr;
  bool computed_ = false;
};

class QNedBasis {
 public:
  QNedBasis() = delete;
  // Replaces QNedBasis(parameters_ms, prm_file, cell, first_cell, subdomain,
  // comm) (q_ned_basis.cc:27-90): the cell's data is handed over explicitly.
  // cell_id names the cell in output file names (CellId::to_string in the
  // reference); empty means the batch slot.
  QNedBasis(std::shared_ptr<QNedBatch> batch, const std::vector<double>& corner_points,
            const std::vector<double>& a, const std::vector<double>& binv,
            const std::vector<double>& r, const std::vector<double>& f,
            const std::string& cell_id = "");
  QNedBasis(const QNedBasis& other) = default;
  QNedBasis& operator=(const QNedBasis& other) = default;

  void run();  // q_ned_basis.cc:1281-1429
  // 8 x 8 row-major (row = test basis) and 8 (q_ned_basis.cc:1183-1201)
  const std::vector<double>& get_global_element_matrix() const;
  const std::vector<double>& get_global_element_rhs() const;
  // sigma then u fine coefficients of basis i (basis_h1_v / basis_curl_v)
  std::vector<double> get_basis(int i) const;
  // global_solution = sum_i w_i psi_i (q_ned_basis.cc:1204-1230)
  void set_global_weights(const std::vector<double>& weights);
  const std::vector<double>& get_global_solution() const;
  int status() const;

  // q_ned_basis.cc:1152-1158, 1255-1280: filename_global = filename_output +
  // "." + subdomain (5 digits) + ".cell-" + cell id + ".vtu", set by run();
  // the output flag is raised on the batch's first cell unless prevent_output.
  const std::string& get_filename_global() const { return filename_global_; }
  void set_output_flag();
  bool output_flag() const { return output_flag_; }
  // q_ned_basis.cc:942-1010: basis_q-ned.h1.<subdomain>.cell-<id>.index-<i>.vtu
  // for the vertex bases and basis_q-ned.curl.<...>.index-<i>.vtu for the edge
  // bases, in dirname_output: sigma at the fine nodes, u (the Nedelec field
  // at the fine cell centres) as cell data.
  void output_basis() const;
  // q_ned_basis.cc:1082-1149: the global solution of set_global_weights in
  // dirname_output / filename_global.
  void output_global_solution_in_cell() const;

 private:
  void ensure() const;
  std::shared_ptr<QNedBatch> batch_;
  std::size_t slot_ = 0;
  bool run_called_ = false;
  mutable std::vector<double> M_, b_;
  std::vector<double> solution_;
  std::string cell_id_, filename_global_;
  bool output_flag_ = false;
};

}  // namespace msfec
