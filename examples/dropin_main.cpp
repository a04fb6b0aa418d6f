// Drop-in example: a caller written against the reference's interface
// (GPU_Refine/main.cpp:1820-1869 style) that links libgcdt.so instead.
// Reads a CDT from a binary file (int np, nt, ns; double pts[2np];
// int tris[3nt]; int nbrs[3nt]; int segs[2ns]), calls GPU_Refine_Quality
// exactly as the reference's main() does, writes the result in the same
// format and prints the counts.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gpu_cdt_refine.h"

static bool read_all(FILE *f, void *p, size_t n) { return fread(p, 1, n, f) == n; }

int main(int argc, char **argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s in.bin out.bin theta mode [circumcenter]\n", argv[0]);
    return 2;
  }
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 3;
  int np, nt, ns;
  if (!read_all(f, &np, 4) || !read_all(f, &nt, 4) || !read_all(f, &ns, 4)) return 4;
  std::vector<double> pts(2 * (size_t)np);
  std::vector<int> tris(3 * (size_t)nt), nbrs(3 * (size_t)nt), segs(2 * (size_t)ns);
  if (!read_all(f, pts.data(), 16 * (size_t)np) || !read_all(f, tris.data(), 12 * (size_t)nt) ||
      !read_all(f, nbrs.data(), 12 * (size_t)nt) || !read_all(f, segs.data(), 8 * (size_t)ns))
    return 5;
  std::fclose(f);

  triangulateio input, result;
  std::memset(&input, 0, sizeof(input));
  std::memset(&result, 0, sizeof(result));
  input.pointlist = pts.data();
  input.numberofpoints = np;
  input.trianglelist = tris.data();
  input.neighborlist = nbrs.data();
  input.numberoftriangles = nt;
  input.numberofcorners = 3;
  input.segmentlist = segs.data();
  input.numberofsegments = ns;

  double theta = std::atof(argv[3]);
  int mode = std::atoi(argv[4]);
  InsertPolicy ip = (argc > 5 && std::strcmp(argv[5], "circumcenter") == 0) ? Circumcenter : Offcenter;
  PStatus *ps = nullptr;
  TStatus *ts = nullptr;
  GPU_Refine_Quality(&input, &result, theta, ip, Connected, mode, -1, &ps, &ts);

  FILE *o = std::fopen(argv[2], "wb");
  if (!o) return 6;
  std::fwrite(&result.numberofpoints, 4, 1, o);
  std::fwrite(&result.numberoftriangles, 4, 1, o);
  std::fwrite(&result.numberofsegments, 4, 1, o);
  std::fwrite(result.pointlist, 16, result.numberofpoints, o);
  std::fwrite(result.trianglelist, 12, result.numberoftriangles, o);
  std::fwrite(result.neighborlist, 12, result.numberoftriangles, o);
  std::fwrite(result.segmentlist, 8, result.numberofsegments, o);
  std::fclose(o);
  int steiner = 0;
  for (int i = 0; i < result.numberofpoints; i++) steiner += (ps[i]._status >> 1) & 1;
  std::printf("points %d triangles %d segments %d steiner %d\n", result.numberofpoints,
              result.numberoftriangles, result.numberofsegments, steiner);
  delete[] result.pointlist;
  delete[] result.trianglelist;
  delete[] result.neighborlist;
  delete[] result.segmentlist;
  delete[] ps;
  delete[] ts;
  return 0;
}
