// C++ drop-in for MsFEC's per-cell NedRTBasis (include/Ned_RT/ned_rt_basis.h:
// 81-153), the three-dimensional Nedelec x Raviart-Thomas pairing, over the
// msfec_nrt_* entry points of libmsfec_hip.so (include/msfec.h).  Same
// pattern as msfec::QBasis: every NedRTBasis registers its cell with a shared
// NedRTBatch; run() marks the cell and the first getter computes the whole
// batch.  Copies share the batch slot, so std::map<CellId, NedRTBasis>::insert
// works unchanged.
#pragma once

#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "msfec.h"

namespace msfec {

struct ParametersNedRT {
  unsigned n_refine_local = 2;  // ParametersMs::n_refine_local (1 to 4)
  // the basis solver (use_direct_solver of the reference's ParametersBasis):
  // 0 solves l <= 2 directly and runs MINRES above, MSFEC_SOLVER_CG runs
  // MINRES at every l, to the relative tolerance rtol (0: 1e-14) within
  // max_iter iterations (0: 20000)
  int solver = 0;
  int max_iter = 0;
  double rtol = 0.0;
  // output (the ParametersBasis fields of include/Ned_RT/ned_rt_parameters.h)
  std::string dirname_output = "Ned_RT";
  std::string filename_output = "solution";
  bool prevent_output = false;
  unsigned local_subdomain = 0;
};

class NedRTBatch {
 public:
  explicit NedRTBatch(const ParametersNedRT& prm);
  ~NedRTBatch();
  NedRTBatch(const NedRTBatch&) = delete;
  NedRTBatch& operator=(const NedRTBatch&) = delete;

  // Register a cell: 8 vertices x 3 coordinates (deal.II order, an
  // axis-aligned box) and its coefficients at the 8 QGauss(2) points of every
  // fine cell (8^l cells): ainv 8^l x 8 x 6 (xx, xy, xz, yy, yz, zz of A^-1),
  // b 8^l x 8, r 8^l x 8 (empty for zero), f 8^l x 8 x 3 (empty for zero).
  std::size_t add_cell(const std::vector<double>& vertices, const std::vector<double>& ainv,
                       const std::vector<double>& b, const std::vector<double>& r,
                       const std::vector<double>& f);
  void compute();
  std::size_t n_cells() const { return n_cells_; }
  int n_edges() const { return n_edges_; }  // fine edges (the sigma block)
  int n_dofs() const { return n_dofs_; }    // fine edges + fine faces

  const double* element_matrix(std::size_t slot) const;  // 18 x 18
  const double* element_rhs(std::size_t slot) const;     // 18
  const double* basis(std::size_t slot, int i) const;    // n_dofs, basis i < 18
  const double* vertices(std::size_t slot) const;        // 8 x 3
  const ParametersNedRT& parameters() const { return prm_; }
  int cell_status(std::size_t slot) const;

 private:
  ParametersNedRT prm_;
  std::size_t n_cells_ = 0;
  int n_qp_ = 0, n_edges_ = 0, n_dofs_ = 0;
  bool any_r_ = false, any_f_ = false;
  std::vector<double> verts_, ainv_, b_, r_, f_;
  std::vector<double> M_, rhs_, X_;
  std::vector<int> status_;
  msfec_nrt_handle handle_ = nullptr;
  bool computed_ = false;
};

class NedRTBasis {
 public:
  NedRTBasis() = delete;
  // Replaces NedRTBasis(parameters_ms, parameter_filename, cell, first_cell,
  // local_subdomain, mpi_communicator) (ned_rt_basis.cc:7-75): the cell's data
  // is handed over explicitly.  cell_id names the cell in output file names
  // (CellId::to_string in the reference); empty means the batch slot.
  NedRTBasis(std::shared_ptr<NedRTBatch> batch, const std::vector<double>& corner_points,
             const std::vector<double>& ainv, const std::vector<double>& b,
             const std::vector<double>& r, const std::vector<double>& f,
             const std::string& cell_id = "");
  NedRTBasis(const NedRTBasis& other) = default;
  NedRTBasis& operator=(const NedRTBasis& other) = default;

  void run();  // ned_rt_basis.cc:1280-1352
  // 18 x 18 row-major (row = test basis, the 12 edge bases first) and 18
  // (ned_rt_basis.cc:1261-1271)
  const std::vector<double>& get_global_element_matrix() const;
  const std::vector<double>& get_global_element_rhs() const;
  // the fine vector of basis i < 18: i < 12 the edge bases (basis_curl_v), i
  // >= 12 the face bases (basis_div_v); fine edges first, then fine faces
  std::vector<double> get_basis(int i) const;
  // global_solution = (sum_{i<12} w_i sigma_i, sum_{i>=12} w_i u_i)
  // (ned_rt_basis.cc:1156-1181): the fine edge values, then the fine faces
  void set_global_weights(const std::vector<double>& weights);
  const std::vector<double>& get_global_solution() const;
  int status() const;

  // ned_rt_basis.cc:1253-1259: filename_global = filename_output + "." +
  // subdomain (5 digits) + ".cell-" + cell id + ".vtu", set by run(); the
  // output flag is raised on the batch's first cell unless prevent_output.
  const std::string& get_filename_global() const { return filename_global_; }
  void set_output_flag();
  bool output_flag() const { return output_flag_; }
  // ned_rt_basis.cc:950-1030: basis_ned-rt.curl.<subdomain>.cell-<id>.index-<i>
  // .vtu for the 12 edge bases and basis_ned-rt.div...index-<i - 12>.vtu for
  // the 6 face bases, in dirname_output, with the cell fields "sigma" (the
  // Nedelec field) and "u" (the Raviart-Thomas field) at the fine cell centres.
  void output_basis() const;
  // ned_rt_basis.cc:1091-1147: the global solution of set_global_weights
  // ("sigma", "u") in dirname_output / filename_global.
  void output_global_solution_in_cell() const;

 private:
  void ensure() const;
  std::shared_ptr<NedRTBatch> batch_;
  std::size_t slot_ = 0;
  bool run_called_ = false;
  mutable std::vector<double> M_, b_;
  std::vector<double> solution_;
  std::string cell_id_, filename_global_;
  bool output_flag_ = false;
};

}  // namespace msfec
