// A constrained Delaunay triangulation built by gcdt_build_cdt (host) or
// gcdt_build_cdt_device: triangles, neighbours and the final segments.
#pragma once

#include <vector>

struct gcdt_cdt {
  int np = 0;
  std::vector<int> T, N, S;
  double stats[4] = {0, 0, 0, 0};
};
