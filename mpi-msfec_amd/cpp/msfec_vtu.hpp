// VTU (VTK XML unstructured grid, ASCII) output of fields on the fine grid
// of one axis-aligned coarse cell: the counterpart of the DataOut::write_vtu
// calls of QBasis::output_basis and output_global_solution_in_cell
// (source/Q/q_basis.cc:431-513).  deal.II writes one patch per fine cell;
// this writer shares the fine nodes between cells instead, which a viewer
// shows identically.
#pragma once

#include <string>
#include <vector>

namespace msfec {

// Writes the (2^l)^dim fine cells of the box with the given deal.II-order
// corners (2^dim x dim doubles) and one point field per name, each with
// (2^l + 1)^dim values on the lexicographic fine nodes (x fastest).
// Throws std::runtime_error when the file cannot be written.
void write_vtu_box(const std::string& path, int dim, const double* corners, int n_refine,
                   const std::vector<std::string>& names, const std::vector<const double*>& fields);

// A field given per fine cell (lexicographic, x fastest) with ncomp
// components per cell (1: scalar, 3: vector).
struct CellField {
  std::string name;
  int ncomp;
  const double* values;
};

// The same with cell fields beside (or instead of) the point fields.
void write_vtu_box(const std::string& path, int dim, const double* corners, int n_refine,
                   const std::vector<std::string>& names, const std::vector<const double*>& fields,
                   const std::vector<CellField>& cell_fields);

// deal.II's Utilities::int_to_string(value, digits): zero-padded decimal.
std::string int_to_string(unsigned value, unsigned digits);

}  // namespace msfec
