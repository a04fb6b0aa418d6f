// CDT files on the host side of the path.
//
//   text    the reference's format (saveCDT / readCDT, GPU_Refine/mesh.cpp:
//           402-532): three count lines (points, triangles, segments), then
//           one line per point "x y", per triangle "a b c", per triangle's
//           neighbours "n0 n1 n2" (-1 on the boundary) and per segment "a b".
//           Coordinates are written with 17 significant digits (the
//           reference prints "%lf", six decimals, which loses the points);
//           every line stays under the reference reader's 100-character
//           buffer, so either reader takes either writer's files.
//   binary  "GCDTBIN1", three little-endian int64 counts, then the point
//           doubles and the triangle, neighbour and segment int32 arrays.
//   Triangle's .node / .ele / .neigh / .poly files (zero-based numbering, as
//           with Triangle's -z switch) for viewers and Triangle itself.
#include <cerrno>
#include <cinttypes>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "gpu_cdt_refine.h"

struct gcdt_file_mesh {
  std::vector<double> P;
  std::vector<int> T, N, S;
};

namespace {

thread_local std::string io_err;

constexpr char BIN_MAGIC[8] = {'G', 'C', 'D', 'T', 'B', 'I', 'N', '1'};
constexpr long long MAX_COUNT = (1LL << 31) - 1;

struct File {
  FILE *f;
  explicit File(const char *path, const char *mode) : f(std::fopen(path, mode)) {
    if (!f) throw std::runtime_error(std::string("cannot open ") + path + ": " + std::strerror(errno));
  }
  ~File() {
    if (f) std::fclose(f);
  }
  void close() {
    if (f && std::fclose(f) != 0) {
      f = nullptr;
      throw std::runtime_error("write failed");
    }
    f = nullptr;
  }
};

void check_arrays(const double *points, int np, const int *tris, const int *nbrs, int nt,
                  const int *segs, int ns) {
  if (np < 0 || nt < 0 || ns < 0) throw std::runtime_error("negative count");
  if ((np && !points) || (nt && (!tris || !nbrs)) || (ns && !segs))
    throw std::runtime_error("missing array");
}

// The next line (without its newline); false at the end of the file.
bool read_line(FILE *f, std::string &line) {
  line.clear();
  int c;
  while ((c = std::fgetc(f)) != EOF) {
    if (c == '\n') return true;
    line.push_back((char)c);
  }
  return !line.empty();
}

long long parse_count(FILE *f, const char *what) {
  std::string line;
  if (!read_line(f, line)) throw std::runtime_error(std::string("missing ") + what + " count");
  char *end = nullptr;
  errno = 0;
  long long v = std::strtoll(line.c_str(), &end, 10);
  if (end == line.c_str() || errno || v < 0 || v > MAX_COUNT)
    throw std::runtime_error(std::string("bad ") + what + " count");
  return v;
}

template <int K, class T>
void parse_rows(FILE *f, long long n, T *out, const char *what) {
  std::string line;
  for (long long i = 0; i < n; i++) {
    if (!read_line(f, line)) throw std::runtime_error(std::string("file ends inside the ") + what);
    const char *p = line.c_str();
    for (int k = 0; k < K; k++) {
      char *end = nullptr;
      errno = 0;
      if constexpr (std::is_same<T, double>::value) {
        out[K * i + k] = std::strtod(p, &end);
      } else {
        long long v = std::strtoll(p, &end, 10);
        if (v < -1 || v > MAX_COUNT) throw std::runtime_error(std::string("index out of range in the ") + what);
        out[K * i + k] = (T)v;
      }
      if (end == p || errno == ERANGE) throw std::runtime_error(std::string("bad number in the ") + what);
      p = end;
    }
  }
}

void load_text(FILE *f, gcdt_file_mesh &m) {
  long long np = parse_count(f, "point"), nt = parse_count(f, "triangle"),
            ns = parse_count(f, "segment");
  m.P.resize(2 * (size_t)np);
  m.T.resize(3 * (size_t)nt);
  m.N.resize(3 * (size_t)nt);
  m.S.resize(2 * (size_t)ns);
  parse_rows<2>(f, np, m.P.data(), "points");
  parse_rows<3>(f, nt, m.T.data(), "triangles");
  parse_rows<3>(f, nt, m.N.data(), "neighbours");
  parse_rows<2>(f, ns, m.S.data(), "segments");
}

template <class T>
void read_exact(FILE *f, T *p, size_t n) {
  if (n && std::fread(p, sizeof(T), n, f) != n) throw std::runtime_error("binary file too short");
}

void load_binary(FILE *f, gcdt_file_mesh &m) {
  int64_t c[3];
  read_exact(f, c, 3);
  for (int64_t v : c)
    if (v < 0 || v > MAX_COUNT) throw std::runtime_error("bad count in binary file");
  m.P.resize(2 * (size_t)c[0]);
  m.T.resize(3 * (size_t)c[1]);
  m.N.resize(3 * (size_t)c[1]);
  m.S.resize(2 * (size_t)c[2]);
  read_exact(f, m.P.data(), m.P.size());
  read_exact(f, m.T.data(), m.T.size());
  read_exact(f, m.N.data(), m.N.size());
  read_exact(f, m.S.data(), m.S.size());
}

void put(FILE *f, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void put(FILE *f, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  int r = std::vfprintf(f, fmt, ap);
  va_end(ap);
  if (r < 0) throw std::runtime_error("write failed");
}

}  // namespace

extern "C" {

const char *gcdt_io_last_error(void) { return io_err.c_str(); }

int gcdt_save_cdt_text(const char *path, const double *points, int np, const int *tris,
                       const int *nbrs, int nt, const int *segs, int ns) {
  io_err.clear();
  try {
    check_arrays(points, np, tris, nbrs, nt, segs, ns);
    File f(path, "w");
    put(f.f, "%d\n%d\n%d\n", np, nt, ns);
    for (int i = 0; i < np; i++) put(f.f, "%.17g %.17g\n", points[2 * i], points[2 * i + 1]);
    for (int i = 0; i < nt; i++) put(f.f, "%d %d %d\n", tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]);
    for (int i = 0; i < nt; i++) put(f.f, "%d %d %d\n", nbrs[3 * i], nbrs[3 * i + 1], nbrs[3 * i + 2]);
    for (int i = 0; i < ns; i++) put(f.f, "%d %d\n", segs[2 * i], segs[2 * i + 1]);
    f.close();
    return 0;
  } catch (const std::exception &e) {
    io_err = e.what();
    return -1;
  }
}

int gcdt_save_cdt_binary(const char *path, const double *points, int np, const int *tris,
                         const int *nbrs, int nt, const int *segs, int ns) {
  io_err.clear();
  try {
    check_arrays(points, np, tris, nbrs, nt, segs, ns);
    File f(path, "wb");
    const int64_t c[3] = {np, nt, ns};
    auto w = [&](const void *p, size_t sz, size_t n) {
      if (n && std::fwrite(p, sz, n, f.f) != n) throw std::runtime_error("write failed");
    };
    w(BIN_MAGIC, 1, sizeof(BIN_MAGIC));
    w(c, sizeof(int64_t), 3);
    w(points, sizeof(double), 2 * (size_t)np);
    w(tris, sizeof(int), 3 * (size_t)nt);
    w(nbrs, sizeof(int), 3 * (size_t)nt);
    w(segs, sizeof(int), 2 * (size_t)ns);
    f.close();
    return 0;
  } catch (const std::exception &e) {
    io_err = e.what();
    return -1;
  }
}

int gcdt_load_cdt(const char *path, gcdt_file_mesh **out) {
  io_err.clear();
  try {
    if (!out) throw std::runtime_error("out is NULL");
    std::unique_ptr<gcdt_file_mesh> m(new gcdt_file_mesh);
    File f(path, "rb");
    char magic[8] = {0};
    size_t got = std::fread(magic, 1, sizeof(magic), f.f);
    if (got == sizeof(magic) && std::memcmp(magic, BIN_MAGIC, sizeof(magic)) == 0) {
      load_binary(f.f, *m);
    } else {
      std::rewind(f.f);
      load_text(f.f, *m);
    }
    *out = m.release();
    return 0;
  } catch (const std::exception &e) {
    io_err = e.what();
    return -1;
  }
}

void gcdt_file_mesh_counts(const gcdt_file_mesh *m, int *np, int *nt, int *ns) {
  *np = (int)(m->P.size() / 2);
  *nt = (int)(m->T.size() / 3);
  *ns = (int)(m->S.size() / 2);
}

void gcdt_file_mesh_copy(const gcdt_file_mesh *m, double *points, int *tris, int *nbrs,
                         int *segs) {
  if (points) std::memcpy(points, m->P.data(), sizeof(double) * m->P.size());
  if (tris) std::memcpy(tris, m->T.data(), sizeof(int) * m->T.size());
  if (nbrs) std::memcpy(nbrs, m->N.data(), sizeof(int) * m->N.size());
  if (segs) std::memcpy(segs, m->S.data(), sizeof(int) * m->S.size());
}

void gcdt_file_mesh_free(gcdt_file_mesh *m) { delete m; }

int gcdt_save_triangle_files(const char *basename, const double *points, int np, const int *tris,
                             const int *nbrs, int nt, const int *segs, int ns) {
  io_err.clear();
  try {
    check_arrays(points, np, tris, nbrs, nt, segs, ns);
    const std::string b(basename ? basename : "");
    if (b.empty()) throw std::runtime_error("empty base name");
    {
      File f((b + ".node").c_str(), "w");
      put(f.f, "%d 2 0 0\n", np);
      for (int i = 0; i < np; i++) put(f.f, "%d %.17g %.17g\n", i, points[2 * i], points[2 * i + 1]);
      f.close();
    }
    {
      File f((b + ".ele").c_str(), "w");
      put(f.f, "%d 3 0\n", nt);
      for (int i = 0; i < nt; i++)
        put(f.f, "%d %d %d %d\n", i, tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]);
      f.close();
    }
    {
      File f((b + ".neigh").c_str(), "w");
      put(f.f, "%d 3\n", nt);
      for (int i = 0; i < nt; i++)
        put(f.f, "%d %d %d %d\n", i, nbrs[3 * i], nbrs[3 * i + 1], nbrs[3 * i + 2]);
      f.close();
    }
    {
      File f((b + ".poly").c_str(), "w");
      put(f.f, "0 2 0 0\n%d 0\n", ns);  // vertices are in the .node file
      for (int i = 0; i < ns; i++) put(f.f, "%d %d %d\n", i, segs[2 * i], segs[2 * i + 1]);
      put(f.f, "0\n");  // no holes
      f.close();
    }
    return 0;
  } catch (const std::exception &e) {
    io_err = e.what();
    return -1;
  }
}

}  // extern "C"
