// C++ drop-in for MsFEC's per-cell RTDQBasis in three dimensions
// (include/RT_DQ/rt_dq_basis.h:84-215, the Raviart-Thomas x DQ0 pairing as the
// reference builds it), over the msfec_rt3_* entry points of libmsfec_hip.so
// (include/msfec.h).  Same pattern as msfec::NedRTBasis: every basis
// registers its cell with a shared RTDQ3Batch; run() marks the cell and the
// first getter computes the whole batch.  Copies share the batch slot, so
// std::map<CellId, RTDQ3Basis>::insert works unchanged.
#pragma once

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "msfec.h"

namespace msfec {

struct ParametersRTDQ3 {
  unsigned n_refine_local = 2;  // ParametersMs::n_refine_local (1 to 4)
  // the CG solve of every face basis: relative tolerance (0: 1e-13) and
  // iteration limit (0: 20000)
  int max_iter = 0;
  double rtol = 0.0;
  // output (the ParametersBasis fields of include/RT_DQ/rt_dq_parameters.h)
  std::string dirname_output = "RT_DQ";
  std::string filename_output = "solution";
  bool prevent_output = false;
  unsigned local_subdomain = 0;
};

class RTDQ3Batch {
 public:
  explicit RTDQ3Batch(const ParametersRTDQ3& prm);
  ~RTDQ3Batch();
  RTDQ3Batch(const RTDQ3Batch&) = delete;
  RTDQ3Batch& operator=(const RTDQ3Batch&) = delete;

  // Register a cell: 8 vertices x 3 coordinates (deal.II order, an
  // axis-aligned box), A^-1 at the 8 QGauss(2) points of every fine cell
  // (8^l x 8 x 6: xx, xy, xz, yy, yz, zz) and the right-hand side at the same
  // points (8^l x 8, empty for zero).
  std::size_t add_cell(const std::vector<double>& vertices, const std::vector<double>& ainv,
                       const std::vector<double>& f);
  void compute();

  std::size_t n_cells() const { return n_cells_; }
  int n_faces() const { return n_faces_; }            // fine faces (the sigma block)
  int n_fine_cells() const { return n_fine_cells_; }  // fine cells (the u block)
  const double* element_matrix(std::size_t slot) const;  // 7 x 7
  const double* element_rhs(std::size_t slot) const;     // 7
  const double* flux(std::size_t slot, int i) const;     // n_faces, face basis i < 6
  const double* vertices(std::size_t slot) const;        // 8 x 3
  const ParametersRTDQ3& parameters() const { return prm_; }
  int cell_status(std::size_t slot) const;

 private:
  ParametersRTDQ3 prm_;
  std::size_t n_cells_ = 0;
  int n_qp_ = 0, n_faces_ = 0, n_fine_cells_ = 0;
  bool any_f_ = false;
  std::vector<double> verts_, ainv_, f_;
  std::vector<double> M_, rhs_, S_;
  std::vector<int> status_;
  msfec_rt3_handle handle_ = nullptr;
  bool computed_ = false;
};

class RTDQ3Basis {
 public:
  RTDQ3Basis() = delete;
  // Replaces RTDQBasis(parameters_ms, parameter_filename, cell, first_cell,
  // local_subdomain, mpi_communicator) (rt_dq_basis.cc:7-62): the cell's data
  // is handed over explicitly.  cell_id names the cell in output file names
  // (CellId::to_string in the reference); empty means the batch slot.
  RTDQ3Basis(std::shared_ptr<RTDQ3Batch> batch, const std::vector<double>& corner_points,
             const std::vector<double>& ainv, const std::vector<double>& f,
             const std::string& cell_id = "");
  RTDQ3Basis(const RTDQ3Basis& other) = default;
  RTDQ3Basis& operator=(const RTDQ3Basis& other) = default;

  void run();  // rt_dq_basis.cc:1114
  // 7 x 7 row-major (the 6 face bases in face order, then u) and 7
  const std::vector<double>& get_global_element_matrix() const;
  const std::vector<double>& get_global_element_rhs() const;
  // the fine fluxes of face basis i < 6 (x-, y-, then z-normal fine faces,
  // +axis orientation); u is 1 in every fine cell (set_u_to_std)
  std::vector<double> get_flux(int i) const;
  // global_solution = sum_{i<6} w_i sigma_i on the faces and w_6 (times the
  // unit u) on the fine cells (rt_dq_basis.cc:1011-1036)
  void set_global_weights(const std::vector<double>& weights);
  const std::vector<double>& get_global_solution() const;
  int status() const;

  // filename_global = filename_output + "." + subdomain (5 digits) +
  // ".cell-" + cell id + ".vtu", set by run(); the output flag is raised on
  // the batch's first cell unless prevent_output.
  const std::string& get_filename_global() const { return filename_global_; }
  void set_output_flag();
  bool output_flag() const { return output_flag_; }
  // rt_dq_basis.cc:916-970: basis_rt-dq.div.<subdomain>.cell-<id>.index-<i>.vtu
  // for the 6 face bases in dirname_output, with the cell fields "sigma"
  // (the flux at the fine cell centres) and "u".
  void output_basis() const;
  // rt_dq_basis.cc:973-1003: the global solution of set_global_weights in
  // dirname_output / filename_global.
  void output_global_solution_in_cell() const;

 private:
  void ensure() const;
  std::shared_ptr<RTDQ3Batch> batch_;
  std::size_t slot_ = 0;
  bool run_called_ = false;
  mutable std::vector<double> M_, b_;
  std::vector<double> solution_;
  std::string cell_id_, filename_global_;
  bool output_flag_ = false;
};

}  // namespace msfec
