// The Q_Ned cell loop over the QNedBasis shim: a std::map of bases filled by
// copy-insertion, run() on every entry, then the getters and
// set_global_weights (the per-cell calls the reference's Q_Ned driver makes).
//
// Usage: q_ned_driver <in.bin> <out.bin> [output directory]
//   in.bin : int32 n_refine_local, int32 nc, then a (nc^2 cells x 4^l fine
//            cells, FP64, per-cell lexicographic); B = 1, R = 0, f = (1, 1)
//            on the unit square
//   out.bin: M (cells x 8 x 8), b (cells x 8), Psi (cells x 8 x dofs), and the
//            cell solution for the weights w_i = i + 1 (cells x dofs), FP64
//   output directory (optional): the VTU files of output_basis (first cell)
//            and output_global_solution_in_cell (every cell)
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <vector>

#include "msfec_q_ned_basis.hpp"

int main(int argc, char** argv) {
  if (argc != 3 && argc != 4) {
    std::fprintf(stderr, "usage: %s in.bin out.bin [output directory]\n", argv[0]);
    return 2;
  }
  std::FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  int32_t hdr[2];
  if (std::fread(hdr, sizeof(int32_t), 2, in) != 2) return 2;
  const int l = hdr[0], nc = hdr[1], m = 1 << l, nf = m * m, ncell = nc * nc;
  std::vector<double> a((size_t)ncell * nf);
  if (std::fread(a.data(), sizeof(double), a.size(), in) != a.size()) return 2;
  std::fclose(in);

  msfec::ParametersNed prm;
  prm.n_refine_local = (unsigned)l;
  if (argc == 4) prm.dirname_output = argv[3];
  auto batch = std::make_shared<msfec::QNedBatch>(prm);
  std::map<int, msfec::QNedBasis> bases;
  const double h = 1.0 / nc;
  const std::vector<double> binv(nf, 1.0), r(nf, 0.0), f(2 * nf, 1.0);
  try {
    for (int cy = 0; cy < nc; ++cy)
      for (int cx = 0; cx < nc; ++cx) {
        const int c = cy * nc + cx;
        std::vector<double> corners = {cx * h, cy * h, (cx + 1) * h, cy * h,
                                       cx * h, (cy + 1) * h, (cx + 1) * h, (cy + 1) * h};
        std::vector<double> ac(a.begin() + (size_t)c * nf, a.begin() + (size_t)(c + 1) * nf);
        bases.insert(std::make_pair(c, msfec::QNedBasis(batch, corners, ac, binv, r, f)));
      }
    for (auto& e : bases) e.second.run();
    std::FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    for (auto& e : bases) std::fwrite(e.second.get_global_element_matrix().data(), 8, 64, out);
    for (auto& e : bases) std::fwrite(e.second.get_global_element_rhs().data(), 8, 8, out);
    for (auto& e : bases)
      for (int i = 0; i < 8; ++i) {
        const std::vector<double> p = e.second.get_basis(i);
        std::fwrite(p.data(), 8, p.size(), out);
      }
    const std::vector<double> w = {1, 2, 3, 4, 5, 6, 7, 8};
    for (auto& e : bases) {
      e.second.set_global_weights(w);
      const std::vector<double>& s = e.second.get_global_solution();
      std::fwrite(s.data(), 8, s.size(), out);
    }
    std::fclose(out);
    if (argc == 4)
      for (auto& e : bases) {
        e.second.set_output_flag();
        if (e.second.output_flag()) e.second.output_basis();
        e.second.output_global_solution_in_cell();
      }
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
  return 0;
}
