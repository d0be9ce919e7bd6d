"""The VTU writer of the C++ shims (mpi-msfec_amd/cpp/msfec_vtu.cpp), host code
only: compiled here with g++ into a small program that writes a 2-D and a 3-D
box, then parsed back (points, VTK cell order, point and cell fields)."""
import os
import shutil
import subprocess
import tempfile
import xml.etree.ElementTree as ET

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(REPO, "mpi-msfec_amd", "cpp")

MAIN = r"""
#include <cstdio>
#include <vector>
#include "msfec_vtu.hpp"
int main(int argc, char** argv) {
  const std::string dir = argv[1];
  // 2-D box [1, 3] x [-1, 0], l = 2: point field x + 10 y, cell field (c, 2c, 0)
  const double c2[8] = {1, -1, 3, -1, 1, 0, 3, 0};
  std::vector<double> u2(25), v2(16 * 3);
  for (int j = 0; j < 5; ++j)
    for (int i = 0; i < 5; ++i) u2[j * 5 + i] = (1 + 0.5 * i) + 10 * (-1 + 0.25 * j);
  for (int c = 0; c < 16; ++c) { v2[3 * c] = c; v2[3 * c + 1] = 2 * c; v2[3 * c + 2] = 0; }
  msfec::write_vtu_box(dir + "/b2.vtu", 2, c2, 2, {"u"}, {u2.data()}, {{"v", 3, v2.data()}});
  // 3-D unit cube, l = 1: point field = node index
  const double c3[24] = {0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0, 0, 0, 1, 1, 0, 1, 0, 1, 1, 1, 1, 1};
  std::vector<double> u3(27);
  for (int k = 0; k < 27; ++k) u3[k] = k;
  msfec::write_vtu_box(dir + "/b3.vtu", 3, c3, 1, {"u"}, {u3.data()});
  std::printf("%s\n", msfec::int_to_string(7, 5).c_str());
  return 0;
}
"""


def parse(path):
    piece = ET.parse(path).getroot().find("UnstructuredGrid").find("Piece")
    arr = lambda e, t: np.array(e.text.split(), dtype=t)
    pts = arr(piece.find("Points").find("DataArray"), np.float64).reshape(-1, 3)
    cells = {e.get("Name"): e for e in piece.find("Cells").findall("DataArray")}
    types = arr(cells["types"], np.int64)
    conn = arr(cells["connectivity"], np.int64).reshape(len(types), -1)
    offs = arr(cells["offsets"], np.int64)
    pd = {e.get("Name"): arr(e, np.float64) for e in piece.find("PointData").findall("DataArray")}
    cd = {}
    if piece.find("CellData") is not None:
        for e in piece.find("CellData").findall("DataArray"):
            cd[e.get("Name")] = arr(e, np.float64).reshape(len(types), int(e.get("NumberOfComponents")))
    return pts, conn, types, offs, pd, cd


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_vtu_writer_boxes():
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "main.cpp")
        open(src, "w").write(MAIN)
        exe = os.path.join(td, "vtu_main")
        subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", f"-I{CPP}", src,
                        os.path.join(CPP, "msfec_vtu.cpp"), "-o", exe], check=True)
        r = subprocess.run([exe, td], capture_output=True, text=True, check=True)
        assert r.stdout.strip() == "00007"
        pts, conn, types, offs, pd, cd = parse(os.path.join(td, "b2.vtu"))
        assert pts.shape == (25, 3) and np.all(types == 9) and np.all(offs == 4 * np.arange(1, 17))
        np.testing.assert_allclose(pd["u"], pts[:, 0] + 10 * pts[:, 1], rtol=0, atol=1e-14)
        np.testing.assert_allclose(cd["v"][:, 0], np.arange(16))
        # counter-clockwise quads of side 0.5 x 0.25
        q = pts[conn][:, :, :2]
        area = 0.5 * (q[:, :, 0] * np.roll(q[:, :, 1], -1, 1) - np.roll(q[:, :, 0], -1, 1) * q[:, :, 1]).sum(1)
        np.testing.assert_allclose(area, 0.125)
        pts, conn, types, offs, pd, cd = parse(os.path.join(td, "b3.vtu"))
        assert pts.shape == (27, 3) and np.all(types == 12) and conn.shape == (8, 8)
        np.testing.assert_array_equal(pd["u"], np.arange(27))
        # VTK hexahedron: bottom face counter-clockwise, top face above it
        for c in conn:
            p = pts[c]
            np.testing.assert_allclose(p[4:, :2], p[:4, :2])
            np.testing.assert_allclose(p[4:, 2] - p[:4, 2], 0.5)
            e = np.diff(np.vstack([p[:4], p[:1]]), axis=0)
            turn = e[:-1, 0] * e[1:, 1] - e[:-1, 1] * e[1:, 0]
            assert np.all(turn > 0)
