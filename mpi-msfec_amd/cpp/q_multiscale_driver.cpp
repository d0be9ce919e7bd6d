// A QMultiscale-shaped driver over the QBasis shim, following the call order
// of Q::QMultiscale::run (source/Q/q_global.cc:520-583):
//   make_grid -> initialize_and_compute_basis (std::map<CellId, QBasis> filled
//   by copy-insertion, then run() on every entry) -> assemble_system +
//   solve_iterative -> send_global_weights_to_cell (set_global_weights).
//
// Usage: q_multiscale_driver <in.bin> <out.bin> [output directory]
//   in.bin : int32 dim, int32 n_refine_local, int32 nc, then the coefficient
//            blocks (nc^dim cells x (2^l)^dim fine cells, FP64, per-cell
//            lexicographic as in msfec_q_set_coefficient).  n_refine_local
//            = 0 marks an uneven batch: int32 levels of the nc^dim cells
//            follow the header, and cell c's block has (2^levels[c])^dim
//            values (the solver of each level is then chosen by timing)
//   out.bin: M (cells x nb x nb), b (cells x nb), weights (cells x nb),
//            fine solution (cells x their fine nodes), all FP64, then int32
//            CG iterations
//   output directory (optional): the VTU files of output_basis (first cell)
//            and output_global_solution_in_cell (every cell), as
//            QMultiscale::output_results has the cells write them
//            (q_global.cc:397-428)
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <vector>

#include "msfec_q_basis.hpp"

int main(int argc, char** argv) {
  if (argc != 3 && argc != 4) {
    std::fprintf(stderr, "usage: %s in.bin out.bin [output directory]\n", argv[0]);
    return 2;
  }
  std::FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  int32_t hdr[3];
  if (std::fread(hdr, sizeof(int32_t), 3, in) != 3) return 2;
  const int dim = hdr[0], l = hdr[1], nc = hdr[2];
  const int nv = 1 << dim;
  int ncell = 1;
  for (int d = 0; d < dim; ++d) ncell *= nc;
  std::vector<int32_t> levels(ncell, l);
  if (l <= 0 && std::fread(levels.data(), sizeof(int32_t), ncell, in) != (size_t)ncell) return 2;
  std::vector<size_t> coff(ncell + 1, 0);
  for (int c = 0; c < ncell; ++c) {
    size_t nfc = 1;
    for (int d = 0; d < dim; ++d) nfc <<= levels[c];
    coff[c + 1] = coff[c] + nfc;
  }
  std::vector<double> coef(coff[ncell]);
  if (std::fread(coef.data(), sizeof(double), coef.size(), in) != coef.size()) return 2;
  std::fclose(in);

  try {
    msfec::ParametersBasis prm;
    prm.dim = dim;
    prm.n_refine_local = (unsigned)(l > 0 ? l : levels[0]);
    if (l <= 0) prm.solver = MSFEC_SOLVER_AUTO;
    if (argc == 4) prm.dirname_output = argv[3];
    auto batch = std::make_shared<msfec::QBatch>(prm);

    // make_grid + initialize_and_compute_basis: one QBasis per coarse cell,
    // copied into the map, then run() in map order
    std::map<int, msfec::QBasis> cell_basis_map;
    const double h = 1.0 / nc;
    for (int c = 0; c < ncell; ++c) {
      int cc[3] = {0, 0, 0}, r = c;
      for (int d = 0; d < dim; ++d) {
        cc[d] = r % nc;
        r /= nc;
      }
      std::vector<double> corners((size_t)nv * dim);
      for (int v = 0; v < nv; ++v)
        for (int d = 0; d < dim; ++d) corners[v * dim + d] = (cc[d] + ((v >> d) & 1)) * h;
      std::vector<double> cblock(coef.begin() + coff[c], coef.begin() + coff[c + 1]);
      msfec::QBasis current_cell_problem(batch, corners, cblock, {}, {}, "", levels[c]);
      auto res = cell_basis_map.insert(std::make_pair(c, current_cell_problem));
      if (!res.second) throw std::runtime_error("insertion into the cell map failed");
    }
    for (auto& kv : cell_basis_map) kv.second.run();

    // assemble_system + solve_iterative + send_global_weights_to_cell
    int its = 0;
    double resid = 0.0;
    std::vector<double> w = batch->coarse_solve(nc, 1e-12, 1e-12, 100000, &its, &resid);
    for (auto& kv : cell_basis_map) {
      std::vector<double> wc(w.begin() + (size_t)kv.first * nv, w.begin() + (size_t)(kv.first + 1) * nv);
      kv.second.set_global_weights(wc);
    }

    std::FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    for (auto& kv : cell_basis_map) {
      const auto& M = kv.second.get_global_element_matrix();
      std::fwrite(M.data(), sizeof(double), M.size(), out);
    }
    for (auto& kv : cell_basis_map) {
      const auto& b = kv.second.get_global_element_rhs();
      std::fwrite(b.data(), sizeof(double), b.size(), out);
    }
    std::fwrite(w.data(), sizeof(double), w.size(), out);
    for (auto& kv : cell_basis_map) {
      const auto& u = kv.second.get_global_solution();
      std::fwrite(u.data(), sizeof(double), u.size(), out);
    }
    int32_t it32 = its;
    std::fwrite(&it32, sizeof(int32_t), 1, out);
    std::fclose(out);
    if (argc == 4)
      for (auto& kv : cell_basis_map) {
        kv.second.set_output_flag();
        if (kv.second.output_flag()) kv.second.output_basis();
        kv.second.output_global_solution_in_cell();
      }
    std::printf("cells %d, coarse CG iterations %d, residual %.3e\n", ncell, its, resid);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
