// C++ drop-in for MsFEC's per-cell RTDQBasis (include/RT_DQ/rt_dq_basis.h:
// 78-160) in its two-dimensional form, over the msfec_rt_* entry points of
// libmsfec_hip.so (include/msfec.h).  Same pattern as msfec::QBasis: every
// RTDQBasis registers its cell with a shared RTDQBatch; run() marks the
// cell and the first getter computes the whole batch.  Copies share the
// batch slot, so std::map<CellId, RTDQBasis>::insert works unchanged.
#pragma once

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "msfec.h"

namespace msfec {

struct ParametersRTDQ {
  unsigned n_refine_local = 4;  // ParametersMs::n_refine_local
  // MSFEC_COEF_SCALAR_CELL (a per fine cell) or MSFEC_COEF_TENSOR_QP
  // ((xx, xy, yy) of A at the 4 Gauss points of every fine cell)
  int coef_mode = MSFEC_COEF_SCALAR_CELL;
  // output (the ParametersBasis fields of include/RT_DQ/rt_dq_parameters.h)
  std::string dirname_output = "RT_DQ";
  std::string filename_output = "solution";
  bool prevent_output = false;
  unsigned local_subdomain = 0;
};

class RTDQBatch {
 public:
  explicit RTDQBatch(const ParametersRTDQ& prm);
  ~RTDQBatch();
  RTDQBatch(const RTDQBatch&) = delete;
  RTDQBatch& operator=(const RTDQBatch&) = delete;

  // Register a cell: 4 vertices x 2 coordinates (deal.II order), its
  // diffusion a (4^l values, or 4^l x 4 x 3 in tensor mode) and load f per
  // fine cell (4^l values).
  std::size_t add_cell(const std::vector<double>& vertices, const std::vector<double>& a,
                       const std::vector<double>& f);
  void compute();
  std::size_t n_cells() const { return n_cells_; }
  int n_fluxes() const { return n_edges_; }           // fine edge fluxes of one basis
  int n_fine_cells() const { return n_fine_cells_; }  // fine cell values (u)

  const double* element_matrix(std::size_t slot) const;  // 5 x 5
  const double* element_rhs(std::size_t slot) const;     // 5
  const double* flux(std::size_t slot, int i) const;     // n_fluxes, basis i < 4
  const double* vertices(std::size_t slot) const;        // 4 x 2
  const ParametersRTDQ& parameters() const { return prm_; }
  int cell_status(std::size_t slot) const;

 private:
  ParametersRTDQ prm_;
  std::size_t n_cells_ = 0;
  int n_fine_cells_ = 0, n_edges_ = 0;
  std::vector<double> verts_, a_, f_;
  std::vector<double> M_, b_, sigma_;
  std::vector<int> status_;
  msfec_rt_handle handle_ = nullptr;
  bool computed_ = false;
};

class RTDQBasis {
 public:
  RTDQBasis() = delete;
  // Replaces RTDQBasis(parameters_ms, prm_file, cell, first_cell, subdomain,
  // comm) (rt_dq_basis.cc:7-68): the cell's data is handed over explicitly.
  // cell_id names the cell in output file names (CellId::to_string in the
  // reference); empty means the batch slot.
  RTDQBasis(std::shared_ptr<RTDQBatch> batch, const std::vector<double>& corner_points,
            const std::vector<double>& a, const std::vector<double>& f,
            const std::string& cell_id = "");
  RTDQBasis(const RTDQBasis& other) = default;
  RTDQBasis& operator=(const RTDQBasis& other) = default;

  void run();  // rt_dq_basis.cc:1114-1170
  // 5 x 5 row-major (row = test basis) and 5 (rt_dq_basis.cc:1096-1106)
  const std::vector<double>& get_global_element_matrix() const;
  const std::vector<double>& get_global_element_rhs() const;
  // the fine fluxes of face basis i < 4 (the sigma block of basis_div_v[i];
  // its u block is the constant 1)
  std::vector<double> get_flux(int i) const;
  // global_solution = (sum_{i<4} w_i sigma_i, w_4) (rt_dq_basis.cc:1011-1036):
  // the fluxes, then one value per fine cell
  void set_global_weights(const std::vector<double>& weights);
  const std::vector<double>& get_global_solution() const;
  int status() const;

  // rt_dq_basis.cc:1088-1110: filename_global = filename_output + "." +
  // subdomain (5 digits) + ".cell-" + cell id + ".vtu", set by run(); the
  // output flag is raised on the batch's first cell unless prevent_output.
  const std::string& get_filename_global() const { return filename_global_; }
  void set_output_flag();
  bool output_flag() const { return output_flag_; }
  // rt_dq_basis.cc:916-970: basis_rt-dq.div.<subdomain>.cell-<id>.index-<i>.vtu
  // per face basis in dirname_output, with the cell fields "sigma" (the flux
  // field at the fine cell centres) and "u" (1, set_u_to_std).
  void output_basis() const;
  // rt_dq_basis.cc:973-1002: the global solution of set_global_weights
  // ("sigma", "u") in dirname_output / filename_global.
  void output_global_solution_in_cell() const;

 private:
  void ensure() const;
  std::shared_ptr<RTDQBatch> batch_;
  std::size_t slot_ = 0;
  bool run_called_ = false;
  mutable std::vector<double> M_, b_;
  std::vector<double> solution_;
  std::string cell_id_, filename_global_;
  bool output_flag_ = false;
};

}  // namespace msfec
