// The 3-D RT_DQ cell loop over the RTDQ3Basis shim: a std::map of bases
// filled by copy-insertion, run() on every entry, then the getters and
// set_global_weights (the per-cell calls the reference's RT_DQ driver makes).
//
// Usage: rt_dq3_driver <in.bin> <out.bin> [output directory]
//   in.bin : int32 n_refine_local, int32 nc, then per cell of the nc^3 boxes
//            of the unit cube (x fastest) ainv (8^l x 8 x 6) and f (8^l x 8),
//            FP64
//   out.bin: M (cells x 7 x 7), b (cells x 7), the fluxes (cells x 6 x
//            n_faces) and the cell solution for the weights w_i = i + 1
//            (cells x (n_faces + 8^l)), FP64
//   output directory (optional): the VTU files of output_basis (first cell)
//            and output_global_solution_in_cell (every cell)
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <vector>

#include "msfec_rt_dq3_basis.hpp"

int main(int argc, char** argv) {
  if (argc != 3 && argc != 4) {
    std::fprintf(stderr, "usage: %s in.bin out.bin [output directory]\n", argv[0]);
    return 2;
  }
  std::FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  int32_t hdr[2];
  if (std::fread(hdr, sizeof(int32_t), 2, in) != 2) return 2;
  const int l = hdr[0], nc = hdr[1], m = 1 << l, nq = m * m * m * 8, ncell = nc * nc * nc;
  std::vector<double> ainv((size_t)ncell * nq * 6), f((size_t)ncell * nq);
  if (std::fread(ainv.data(), sizeof(double), ainv.size(), in) != ainv.size() ||
      std::fread(f.data(), sizeof(double), f.size(), in) != f.size())
    return 2;
  std::fclose(in);

  msfec::ParametersRTDQ3 prm;
  prm.n_refine_local = (unsigned)l;
  if (argc == 4) prm.dirname_output = argv[3];
  auto batch = std::make_shared<msfec::RTDQ3Batch>(prm);
  std::map<int, msfec::RTDQ3Basis> bases;
  const double h = 1.0 / nc;
  try {
    for (int cz = 0; cz < nc; ++cz)
      for (int cy = 0; cy < nc; ++cy)
        for (int cx = 0; cx < nc; ++cx) {
          const int c = (cz * nc + cy) * nc + cx;
          std::vector<double> corners;
          for (int v = 0; v < 8; ++v) {
            corners.push_back((cx + (v & 1)) * h);
            corners.push_back((cy + ((v >> 1) & 1)) * h);
            corners.push_back((cz + (v >> 2)) * h);
          }
          auto slice = [c](const std::vector<double>& x, size_t per) {
            return std::vector<double>(x.begin() + c * per, x.begin() + (c + 1) * per);
          };
          bases.insert(std::make_pair(
              c, msfec::RTDQ3Basis(batch, corners, slice(ainv, (size_t)nq * 6), slice(f, nq))));
        }
    for (auto& e : bases) e.second.run();
    std::FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    for (auto& e : bases) std::fwrite(e.second.get_global_element_matrix().data(), 8, 49, out);
    for (auto& e : bases) std::fwrite(e.second.get_global_element_rhs().data(), 8, 7, out);
    for (auto& e : bases)
      for (int i = 0; i < 6; ++i) {
        const std::vector<double> p = e.second.get_flux(i);
        std::fwrite(p.data(), 8, p.size(), out);
      }
    std::vector<double> w(7);
    for (int i = 0; i < 7; ++i) w[i] = i + 1;
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
