"""The C++ host layer's host-only helpers (m2c_amd/host/m2c_host.hpp), compiled and run without a
GPU: ComputeMesh1D (MeshGenerator's axes) and SpaceOperator::FindEdgeSurfaceIntersections give what
the C ABI gives through the Python front end, bit for bit."""
import os
import subprocess

import numpy as np

import m2c_amd as M
from m2c_amd import build

PROG = r'''
#include <cstdio>
#include "m2c_host.hpp"
using namespace m2c_host;
int main() {
  std::vector<double> x, dx;
  ComputeMesh1D(0.0, 1.2, 0, {{0.2, 0.02}, {0.5, 0.01}, {0.9, 0.04}}, x, dx);
  std::printf("%zu %.17g %.17g %.17g\n", x.size(), x[0], x[x.size() / 2], dx.back());
  const double udot[] = {1, 0, 0, 0, 2, 0, 0, 0, 4};
  const int elems[] = {0, 1, 2};
  const double nrm[] = {0, 0, 1};
  const int tid[] = {0, 0};
  const double xi[] = {0.2, 0.3, 0.5, 0.6, 0.3, 0.1};
  const double dist[] = {0.25, 0.75};
  m2c_surface s{3, 1, 2, udot, elems, nrm, tid, xi, dist};
  std::vector<int> cut;
  std::vector<double> wall;
  SpaceOperator::FindEdgeSurfaceIntersections({s}, {0, -1, 1}, {1, -1, 0}, cut, wall);
  std::printf("%d %d %d", cut[0], cut[1], cut[2]);
  for (int q = 0; q < 12; q++) std::printf(" %.17g", wall[q]);
  std::printf("\n");
  return 0;
}
'''


def test_host_helpers_match_the_abi(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(PROG)
    exe = tmp_path / "t"
    here = os.path.dirname(build.OUT)
    r = subprocess.run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", str(src), "-I", os.path.join(here, "host"),
                        "-L" + here, "-lm2c_hip", "-Wl,-rpath," + here, "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    l1, l2 = out.stdout.strip().split("\n")
    x, dx = M.mesh_1d(0.0, 1.2, points=[(0.2, 0.02), (0.5, 0.01), (0.9, 0.04)])
    n, a, b, c = l1.split()
    assert int(n) == len(x) and float(a) == x[0] and float(b) == x[len(x) // 2] and float(c) == dx[-1]
    surf = dict(udot=[[1, 0, 0], [0, 2, 0], [0, 0, 4]], elems=[[0, 1, 2]], elem_norm=[[0, 0, 1]], tid=[0, 0],
                xi=[[0.2, 0.3, 0.5], [0.6, 0.3, 0.1]], dist=[0.25, 0.75])
    cut, wall = M.edge_surface_walls([surf], [[0], [-1], [1]], [[1], [-1], [0]])
    vals = l2.split()
    assert [int(v) for v in vals[:3]] == [int(c) for c in cut]
    assert np.array_equal(np.array([float(v) for v in vals[3:]]), wall[0])
