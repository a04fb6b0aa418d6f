// Host-side input preparation: seeded PSLG generator and constrained
// Delaunay triangulation builder.
//
// The reference builds its input with Triangle ("pzQnc", mesh.cpp:397) after
// generating random points and non-crossing segments (mesh.cpp:49-392,
// GenerateRandomInput).  This file provides the same contract from its own
// code:
//   - gcdt_generate_pslg: points uniform in [0,L]^2, Gaussian clusters, or a
//     lattice + cocircular + uniform precision-stress mix;
//     random segments accepted only if they cross no accepted segment (exact
//     orient2d), are not longer than maxlen, and make at least
//     min_angle_deg with every accepted segment sharing an endpoint
//     (mesh.cpp:262-392).  RNG: xoshiro256** seeded by splitmix64.
//   - gcdt_build_cdt: incremental Bowyer-Watson Delaunay triangulation with a
//     symbolic vertex at infinity (exact convex hull), then each segment is
//     recovered by Sloan's edge-flip method and the CDT restored by Lawson
//     flips around it; segments through collinear points are split there; the
//     hull edges become segments (Triangle's -c).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "gcdt_cdt_result.h"
#include "gpu_cdt_refine.h"
#include "predicates.h"

namespace {

using gcdt::incircle_sign;
using gcdt::orient2d_sign;

// ---------------------------------------------------------------- RNG ----
struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    for (auto &v : s) v = splitmix(seed);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t n) { return next() % n; }
  double normal() {  // Box-Muller
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

inline int orient(const double *P, int a, int b, int c) {
  return orient2d_sign(P[2 * a], P[2 * a + 1], P[2 * b], P[2 * b + 1], P[2 * c],
                       P[2 * c + 1], false, nullptr);
}

thread_local double g_ws[GCDT_INCIRCLE_WS];

inline int incircle(const double *P, int a, int b, int c, int d) {
  return incircle_sign(P[2 * a], P[2 * a + 1], P[2 * b], P[2 * b + 1],
                       P[2 * c], P[2 * c + 1], P[2 * d], P[2 * d + 1], g_ws,
                       false, nullptr);
}

// ------------------------------------------------------------ builder ----
constexpr int INF = -1;

struct Builder {
  const double *P;
  int n;
  std::vector<int> V, N;     // 3 per triangle; N as otri (tri<<2 | edge)
  std::vector<uint8_t> C;    // 3 per triangle: constrained edge
  std::vector<uint8_t> dead;
  std::vector<int> vtri;     // vertex -> a triangle containing it
  std::vector<int> mark;
  int stamp = 0;
  int last = 0;

  int newtri() {
    int t = (int)(V.size() / 3);
    V.insert(V.end(), {0, 0, 0});
    N.insert(N.end(), {-1, -1, -1});
    C.insert(C.end(), {0, 0, 0});
    dead.push_back(0);
    mark.push_back(0);
    return t;
  }
  bool ghost(int t) const { return V[3 * t] == INF || V[3 * t + 1] == INF || V[3 * t + 2] == INF; }

  bool between(int a, int b, int p) const {  // p collinear with a,b: strictly inside (a,b)?
    double ax = P[2 * a], ay = P[2 * a + 1], bx = P[2 * b], by = P[2 * b + 1];
    double px = P[2 * p], py = P[2 * p + 1];
    if (ax != bx) return (px > std::min(ax, bx)) && (px < std::max(ax, bx));
    return (py > std::min(ay, by)) && (py < std::max(ay, by));
  }

  bool conflict(int t, int p) const {
    const int *v = &V[3 * t];
    for (int k = 0; k < 3; k++) {
      if (v[k] == INF) {
        int a = v[(k + 1) % 3], b = v[(k + 2) % 3];
        int o = orient(P, a, b, p);
        if (o > 0) return true;
        if (o == 0) return between(a, b, p);
        return false;
      }
    }
    return incircle(P, v[0], v[1], v[2], p) > 0;
  }

  // Returns a triangle in conflict with p, or -2 if p duplicates a vertex.
  int locate(int p) {
    int t = last;
    if (dead[t]) {
      for (t = 0; dead[t]; t++) {}
    }
    if (ghost(t)) {  // step into the finite neighbour
      for (int j = 0; j < 3; j++) {
        if (V[3 * t + j] == INF) { t = N[3 * t + j] >> 2; break; }
      }
    }
    for (long guard = 0; guard < 4L * (long)(V.size() / 3) + 64; guard++) {
      if (ghost(t)) return t;
      const int *v = &V[3 * t];
      int jn = -1, zeros = 0;
      for (int j = 0; j < 3; j++) {
        int o = orient(P, v[(j + 1) % 3], v[(j + 2) % 3], p);
        if (o < 0) { jn = j; break; }
        if (o == 0) zeros++;
      }
      if (jn >= 0) { t = N[3 * t + jn] >> 2; continue; }
      for (int k = 0; k < 3; k++)
        if (P[2 * v[k]] == P[2 * p] && P[2 * v[k] + 1] == P[2 * p + 1]) return -2;
      return t;
    }
    throw std::runtime_error("point location did not terminate");
  }

  void insert(int p) {
    int t0 = locate(p);
    if (t0 == -2) throw std::runtime_error("duplicate input point " + std::to_string(p));
    stamp++;
    std::vector<int> cav{t0};
    mark[t0] = stamp;
    for (size_t i = 0; i < cav.size(); i++) {
      int t = cav[i];
      for (int j = 0; j < 3; j++) {
        int u = N[3 * t + j] >> 2;
        if (mark[u] == stamp) continue;
        if (conflict(u, p)) { mark[u] = stamp; cav.push_back(u); }
      }
    }
    struct E { int u, v, outer; };
    std::vector<E> bnd;
    for (int t : cav)
      for (int j = 0; j < 3; j++) {
        int nb = N[3 * t + j];
        if (mark[nb >> 2] != stamp)
          bnd.push_back({V[3 * t + (j + 1) % 3], V[3 * t + (j + 2) % 3], nb});
      }
    std::vector<int> slot(bnd.size());
    for (size_t i = 0; i < bnd.size(); i++) slot[i] = i < cav.size() ? cav[i] : newtri();
    // map boundary origin vertex -> star index (origins are distinct)
    std::unordered_map<int, int> from;
    from.reserve(bnd.size() * 2);
    for (size_t i = 0; i < bnd.size(); i++) from[bnd[i].u] = (int)i;
    for (size_t i = 0; i < bnd.size(); i++) {
      int s = slot[i];
      V[3 * s] = p; V[3 * s + 1] = bnd[i].u; V[3 * s + 2] = bnd[i].v;
      C[3 * s] = 0; C[3 * s + 1] = 0; C[3 * s + 2] = 0;
      dead[s] = 0;
      N[3 * s] = bnd[i].outer;
    }
    // star links: edge 1 (v_i, p) meets edge 2 (p, u) of the star triangle
    // whose boundary edge starts at v_i; outer back-pointers to edge 0.
    for (size_t i = 0; i < bnd.size(); i++) {
      int s = slot[i];
      int nx = slot[from.at(bnd[i].v)];
      N[3 * s + 1] = (nx << 2) | 2;
      N[3 * nx + 2] = (s << 2) | 1;
      int o = bnd[i].outer;
      N[3 * (o >> 2) + (o & 3)] = (s << 2) | 0;
      if (bnd[i].u != INF) vtri[bnd[i].u] = s;
      if (bnd[i].v != INF) vtri[bnd[i].v] = s;
    }
    vtri[p] = slot[0];
    for (size_t i = 0; i < bnd.size(); i++)
      if (!ghost(slot[i])) { last = slot[i]; vtri[p] = slot[i]; break; }
  }

  // Triangle/edge holding directed edge a->b, via rotation around a.
  bool find_edge(int a, int b, int &tt, int &jj) {
    int t = vtri[a];
    int start = t;
    for (int guard = 0; guard < 1 << 20; guard++) {
      const int *v = &V[3 * t];
      int i = v[0] == a ? 0 : (v[1] == a ? 1 : 2);
      int x = v[(i + 1) % 3];
      if (x == b) { tt = t; jj = (i + 2) % 3; return true; }
      // rotate: cross edge (a, x) = edge (i+2)%3
      t = N[3 * t + (i + 2) % 3] >> 2;
      if (t == start) return false;
    }
    return false;
  }

  void flip(int t, int j) {
    int nb = N[3 * t + j];
    int t2 = nb >> 2, k = nb & 3;
    int a = V[3 * t + j], b = V[3 * t + (j + 1) % 3], c = V[3 * t + (j + 2) % 3];
    int d = V[3 * t2 + k];
    int n_ca = N[3 * t + (j + 1) % 3], n_ab = N[3 * t + (j + 2) % 3];
    int n_bd = N[3 * t2 + (k + 1) % 3], n_dc = N[3 * t2 + (k + 2) % 3];
    uint8_t c_ca = C[3 * t + (j + 1) % 3], c_ab = C[3 * t + (j + 2) % 3];
    uint8_t c_bd = C[3 * t2 + (k + 1) % 3], c_dc = C[3 * t2 + (k + 2) % 3];
    V[3 * t] = a; V[3 * t + 1] = b; V[3 * t + 2] = d;
    N[3 * t] = n_bd; N[3 * t + 1] = (t2 << 2) | 2; N[3 * t + 2] = n_ab;
    C[3 * t] = c_bd; C[3 * t + 1] = 0; C[3 * t + 2] = c_ab;
    V[3 * t2] = a; V[3 * t2 + 1] = d; V[3 * t2 + 2] = c;
    N[3 * t2] = n_dc; N[3 * t2 + 1] = n_ca; N[3 * t2 + 2] = (t << 2) | 1;
    C[3 * t2] = c_dc; C[3 * t2 + 1] = c_ca; C[3 * t2 + 2] = 0;
    N[3 * (n_bd >> 2) + (n_bd & 3)] = (t << 2) | 0;
    N[3 * (n_ab >> 2) + (n_ab & 3)] = (t << 2) | 2;
    N[3 * (n_dc >> 2) + (n_dc & 3)] = (t2 << 2) | 0;
    N[3 * (n_ca >> 2) + (n_ca & 3)] = (t2 << 2) | 1;
    vtri[a] = t; vtri[b] = t; vtri[d] = t; vtri[c] = t2;
  }

  void set_constrained(int a, int b) {
    int t, j;
    if (!find_edge(a, b, t, j)) throw std::runtime_error("segment not recovered");
    C[3 * t + j] = 1;
    int nb = N[3 * t + j];
    C[3 * (nb >> 2) + (nb & 3)] = 1;
  }

  // Insert segment a-b (split at collinear vertices).  Sloan's method.
  void insert_segment(int a, int b) {
    std::vector<std::pair<int, int>> todo{{a, b}};
    while (!todo.empty()) {
      auto [sa, sb] = todo.back();
      todo.pop_back();
      if (sa == sb) continue;
      int t, j;
      if (find_edge(sa, sb, t, j) || find_edge(sb, sa, t, j)) { set_constrained(sa, sb); continue; }
      // find the triangle around sa whose wedge contains sb
      int ts = vtri[sa], start = ts, found = -1, split = -1;
      for (int guard = 0; guard < 1 << 20; guard++) {
        const int *v = &V[3 * ts];
        int i = v[0] == sa ? 0 : (v[1] == sa ? 1 : 2);
        int x = v[(i + 1) % 3], y = v[(i + 2) % 3];
        if (x != INF && y != INF) {
          int ox = orient(P, sa, x, sb), oy = orient(P, sa, sb, y);
          if (ox == 0 && between(sa, sb, x)) { split = x; break; }
          if (oy == 0 && between(sa, sb, y)) { split = y; break; }
          if (ox > 0 && oy > 0) { found = ts; break; }
        }
        ts = N[3 * ts + (i + 2) % 3] >> 2;
        if (ts == start) break;
      }
      if (split >= 0) { todo.push_back({split, sb}); todo.push_back({sa, split}); continue; }
      if (found < 0) throw std::runtime_error("segment start wedge not found");
      // walk the channel collecting crossing edges
      std::deque<std::pair<int, int>> cross;
      int tc = found;
      {
        const int *v = &V[3 * tc];
        int i = v[0] == sa ? 0 : (v[1] == sa ? 1 : 2);
        int x = v[(i + 1) % 3], y = v[(i + 2) % 3];
        if (C[3 * tc + i]) throw std::runtime_error("segments cross");
        cross.push_back({x, y});
        int nb = N[3 * tc + i];
        tc = nb >> 2;
        int k = nb & 3;
        int u = x, w = y;  // edge just crossed (u,w); triangle tc = (z, w, u) with z at k
        for (int guard = 0; guard < 1 << 26; guard++) {
          int z = V[3 * tc + k];
          if (z == sb) break;
          int oz = orient(P, sa, sb, z);
          if (oz == 0) {
            if (!between(sa, sb, z)) throw std::runtime_error("segment walk left its line");
            split = z;
            break;
          }
          // z on the left of sa->sb: next crossed edge is (u, z) else (z, w)
          int nu, nw, ej;
          // tc = (z, w, u) at positions (k, k+1, k+2): edge (u, z) is
          // opposite w (k+1), edge (z, w) is opposite u (k+2)
          if (oz > 0) { nu = u; nw = z; ej = (k + 1) % 3; }
          else { nu = z; nw = w; ej = (k + 2) % 3; }
          if (C[3 * tc + ej]) throw std::runtime_error("segments cross");
          cross.push_back({nu, nw});
          int nb2 = N[3 * tc + ej];
          tc = nb2 >> 2;
          k = nb2 & 3;
          u = nu;
          w = nw;
        }
      }
      if (split >= 0) { todo.push_back({split, sb}); todo.push_back({sa, split}); continue; }
      // flip crossing edges until the segment appears
      std::vector<std::pair<int, int>> fresh;
      long guard = 0;
      while (!cross.empty()) {
        if (++guard > 100000000L) throw std::runtime_error("segment recovery did not terminate");
        auto [u, w] = cross.front();
        cross.pop_front();
        int te, je;
        if (!find_edge(u, w, te, je)) throw std::runtime_error("crossing edge vanished");
        int nb = N[3 * te + je];
        int t2 = nb >> 2, k2 = nb & 3;
        int x = V[3 * te + je], y = V[3 * t2 + k2];
        if (orient(P, x, y, u) * orient(P, x, y, w) < 0) {
          flip(te, je);
          int o1 = orient(P, sa, sb, x), o2 = orient(P, sa, sb, y);
          if ((x == sa && y == sb) || (x == sb && y == sa)) continue;
          if (o1 * o2 < 0) cross.push_back({x, y});
          else fresh.push_back({x, y});
        } else {
          cross.push_back({u, w});
        }
      }
      set_constrained(sa, sb);
      // Lawson restoration over the new edges
      bool any = true;
      for (int pass = 0; any && pass < 1000000; pass++) {
        any = false;
        for (auto &e : fresh) {
          int te, je;
          if (!find_edge(e.first, e.second, te, je)) continue;
          if (C[3 * te + je]) continue;
          int nb = N[3 * te + je];
          int t2 = nb >> 2, k2 = nb & 3;
          if (ghost(te) || ghost(t2)) continue;
          int d = V[3 * t2 + k2];
          if (incircle(P, V[3 * te], V[3 * te + 1], V[3 * te + 2], d) > 0) {
            int x = V[3 * te + je];
            flip(te, je);
            e = {x, d};
            any = true;
          }
        }
      }
    }
  }
};

struct Segs {
  int a, b;
};

// Points on the boundary of the convex hull, counterclockwise, collinear
// boundary points included (each once).  Only points outside the open
// octagon of the extreme points in eight directions can lie on the hull's
// boundary; Andrew's monotone chain with exact orientation, popping only on
// clockwise turns, orders them.
std::vector<int> hull_boundary(const double *P, int np) {
  int ext[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  auto key = [&](int i, int k) {
    const double x = P[2 * i], y = P[2 * i + 1];
    switch (k) {
      case 0: return x;
      case 1: return -x;
      case 2: return y;
      case 3: return -y;
      case 4: return x + y;
      case 5: return -(x + y);
      case 6: return x - y;
      default: return y - x;
    }
  };
  for (int i = 1; i < np; i++)
    for (int k = 0; k < 8; k++)
      if (key(i, k) > key(ext[k], k)) ext[k] = i;
  auto lex = [&](int a, int b) {
    return P[2 * a] < P[2 * b] || (P[2 * a] == P[2 * b] && P[2 * a + 1] < P[2 * b + 1]);
  };
  auto chain = [&](std::vector<int> pts, bool keep_collinear) {
    std::sort(pts.begin(), pts.end(), lex);
    pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
    if (pts.size() < 3) return pts;
    const int lim = keep_collinear ? -1 : 0;
    std::vector<int> h(2 * pts.size());
    size_t k = 0;
    for (size_t i = 0; i < pts.size(); i++) {
      while (k >= 2 && orient(P, h[k - 2], h[k - 1], pts[i]) <= lim) k--;
      h[k++] = pts[i];
    }
    for (size_t i = pts.size() - 1, t = k + 1; i-- > 0;) {
      while (k >= t && orient(P, h[k - 2], h[k - 1], pts[i]) <= lim) k--;
      h[k++] = pts[i];
    }
    h.resize(k - 1);
    return h;
  };
  std::vector<int> oct = chain(std::vector<int>(ext, ext + 8), false);
  std::vector<int> cand;
  for (int i = 0; i < np; i++) {
    bool inside = oct.size() >= 3;
    for (size_t e = 0; e < oct.size() && inside; e++)
      inside = orient(P, oct[e], oct[(e + 1) % oct.size()], i) > 0;
    if (!inside) cand.push_back(i);
  }
  return chain(cand, true);
}

}  // namespace


extern "C" int gcdt_build_cdt(const double *points, int np, const int *segs,
                              int ns, gcdt_cdt **out) {
  try {
    if (np < 3) throw std::runtime_error("need at least 3 points");
    Builder B;
    B.P = points;
    B.n = np;
    B.vtri.assign(np, -1);
    // spatial insertion order: serpentine grid (bucketed) order
    int g = std::max(1, (int)std::sqrt((double)np / 4.0));
    double xmin = points[0], xmax = points[0], ymin = points[1], ymax = points[1];
    for (int i = 1; i < np; i++) {
      xmin = std::min(xmin, points[2 * i]); xmax = std::max(xmax, points[2 * i]);
      ymin = std::min(ymin, points[2 * i + 1]); ymax = std::max(ymax, points[2 * i + 1]);
    }
    double sx = (xmax - xmin) > 0 ? g / (xmax - xmin) * (1 - 1e-12) : 0;
    double sy = (ymax - ymin) > 0 ? g / (ymax - ymin) * (1 - 1e-12) : 0;
    std::vector<std::pair<long long, int>> ord(np);
    for (int i = 0; i < np; i++) {
      long long cx = std::min<long long>(g - 1, (long long)((points[2 * i] - xmin) * sx));
      long long cy = std::min<long long>(g - 1, (long long)((points[2 * i + 1] - ymin) * sy));
      if (cy & 1) cx = g - 1 - cx;
      ord[i] = {cy * g + cx, i};
    }
    std::sort(ord.begin(), ord.end());
    // initial triangle: first point, then a distinct one, then a non-collinear one
    int a = ord[0].second, b = -1, c = -1;
    size_t ib = 1;
    for (; ib < ord.size(); ib++) {
      int q = ord[ib].second;
      if (points[2 * q] != points[2 * a] || points[2 * q + 1] != points[2 * a + 1]) { b = q; break; }
    }
    if (b < 0) throw std::runtime_error("all points coincide");
    size_t ic = ib + 1;
    for (; ic < ord.size(); ic++) {
      int q = ord[ic].second;
      if (orient(points, a, b, q) != 0) { c = q; break; }
    }
    if (c < 0) throw std::runtime_error("all points are collinear");
    if (orient(points, a, b, c) < 0) std::swap(b, c);
    int t0 = B.newtri(), g0 = B.newtri(), g1 = B.newtri(), g2 = B.newtri();
    auto setv = [&](int t, int x, int y, int z) { B.V[3 * t] = x; B.V[3 * t + 1] = y; B.V[3 * t + 2] = z; };
    setv(t0, a, b, c);
    // ghosts across edges (b,c), (c,a), (a,b): (c,b,INF), (a,c,INF), (b,a,INF)
    setv(g0, c, b, INF);
    setv(g1, a, c, INF);
    setv(g2, b, a, INF);
    B.N[3 * t0 + 0] = (g0 << 2) | 2;
    B.N[3 * t0 + 1] = (g1 << 2) | 2;
    B.N[3 * t0 + 2] = (g2 << 2) | 2;
    B.N[3 * g0 + 2] = (t0 << 2) | 0;
    B.N[3 * g1 + 2] = (t0 << 2) | 1;
    B.N[3 * g2 + 2] = (t0 << 2) | 2;
    // ghost-ghost links: g0=(c,b,INF): edge0=(b,INF) edge1=(INF,c)
    //   (b,INF) is shared with g2=(b,a,INF) edge1=(INF,b)
    B.N[3 * g0 + 0] = (g2 << 2) | 1; B.N[3 * g2 + 1] = (g0 << 2) | 0;
    //   (INF,c) shared with g1=(a,c,INF) edge0=(c,INF)
    B.N[3 * g0 + 1] = (g1 << 2) | 0; B.N[3 * g1 + 0] = (g0 << 2) | 1;
    //   g1 edge1=(INF,a) shared with g2 edge0=(a,INF)
    B.N[3 * g1 + 1] = (g2 << 2) | 0; B.N[3 * g2 + 0] = (g1 << 2) | 1;
    B.vtri[a] = B.vtri[b] = B.vtri[c] = t0;
    B.last = t0;
    for (size_t i = 0; i < ord.size(); i++) {
      int q = ord[i].second;
      if (q == a || q == b || q == c) continue;
      B.insert(q);
    }
    // constraints
    for (int s = 0; s < ns; s++) {
      int u = segs[2 * s], v = segs[2 * s + 1];
      if (u < 0 || u >= np || v < 0 || v >= np) throw std::runtime_error("segment index out of range");
      B.insert_segment(u, v);
    }
    // output
    gcdt_cdt *r = new gcdt_cdt;
    r->np = np;
    int ntri = (int)(B.V.size() / 3);
    std::vector<int> id(ntri, -1);
    int k = 0;
    for (int t = 0; t < ntri; t++)
      if (!B.dead[t] && !B.ghost(t)) id[t] = k++;
    r->T.resize(3 * (size_t)k);
    r->N.resize(3 * (size_t)k);
    std::unordered_set<long long> segset;
    auto key = [](int x, int y) { return ((long long)std::min(x, y) << 32) | (unsigned)std::max(x, y); };
    for (int t = 0; t < ntri; t++) {
      if (id[t] < 0) continue;
      for (int j = 0; j < 3; j++) {
        r->T[3 * id[t] + j] = B.V[3 * t + j];
        int nb = B.N[3 * t + j] >> 2;
        r->N[3 * id[t] + j] = id[nb];
        int x = B.V[3 * t + (j + 1) % 3], y = B.V[3 * t + (j + 2) % 3];
        bool hull = id[nb] < 0;
        if ((B.C[3 * t + j] || hull) && segset.insert(key(x, y)).second) {
          r->S.push_back(hull ? x : x);
          r->S.push_back(y);
        }
      }
    }
    *out = r;
    return 0;
  } catch (const std::exception &e) {
    (void)e;
    return -1;
  }
}

extern "C" void gcdt_cdt_counts(const gcdt_cdt *c, int *np, int *nt, int *ns) {
  *np = c->np;
  *nt = (int)(c->T.size() / 3);
  *ns = (int)(c->S.size() / 2);
}

extern "C" void gcdt_cdt_copy(const gcdt_cdt *c, int *tris, int *nbrs, int *segs) {
  if (tris) std::memcpy(tris, c->T.data(), sizeof(int) * c->T.size());
  if (nbrs) std::memcpy(nbrs, c->N.data(), sizeof(int) * c->N.size());
  if (segs) std::memcpy(segs, c->S.data(), sizeof(int) * c->S.size());
}

extern "C" void gcdt_cdt_free(gcdt_cdt *c) { delete c; }

// ---------------------------------------------------------- generator ----
extern "C" int gcdt_generate_pslg(int dist, int n, int nseg, double L,
                                  double maxlen, double min_angle_deg,
                                  unsigned long long seed, double *points,
                                  int *segs) {
  if (n < 3 || L <= 0 || dist < 0 || dist > 6) return -1;
  Rng rng(seed);
  std::unordered_set<uint64_t> seen;
  seen.reserve((size_t)n * 2);
  std::vector<double> centers;
  int ncl = 0;
  if (dist == 1) {
    ncl = std::max(1, std::min(64, n / 5000 + 8));
    for (int k = 0; k < ncl; k++) {
      centers.push_back(L * (0.1 + 0.8 * rng.uniform()));
      centers.push_back(L * (0.1 + 0.8 * rng.uniform()));
    }
  }
  auto accept = [&](int i, double x, double y) {
    uint64_t hx, hy;
    std::memcpy(&hx, &x, 8);
    std::memcpy(&hy, &y, 8);
    if (!seen.insert(hx * 0x9E3779B97F4A7C15ull ^ hy).second) return false;
    points[2 * i] = x;
    points[2 * i + 1] = y;
    return true;
  };
  int i = 0;
  // lattice bookkeeping for the precision-stress mix (dist == 2)
  int side = 0;
  int n_lattice = 0;     // points [0, n_lattice) are lattice points
  std::vector<int> lat;  // lattice cell -> point index or -1
  if (dist == 2) {
    // Precision stress: thirds of (a) a random third of the cells of a
    // side x side integer lattice (side = 1000 for 10^6 points) scaled to
    // [0,L]^2, with 1e-12 relative jitter, (b) points on random circles, 16
    // per circle, cocircular in exact rational arithmetic before rounding to
    // FP64 (x = cx + r(1-u^2)/(1+u^2), y = cy +- 2ru/(1+u^2), u = k/1000),
    // and (c) uniform background points.
    int nl = n / 3, nc = n / 3;
    side = std::max(2, std::min(1000, (int)std::ceil(std::sqrt(3.0 * nl))));
    nl = std::min(nl, side * side);
    double h = L / side;
    std::vector<int> cells((size_t)side * side);
    for (size_t c = 0; c < cells.size(); c++) cells[c] = (int)c;
    lat.assign(cells.size(), -1);
    for (int t = 0; t < nl; t++) {  // partial Fisher-Yates
      size_t r = t + rng.below(cells.size() - t);
      std::swap(cells[t], cells[r]);
      int c = cells[t];
      double x = (c % side) * h, y = (c / side) * h;
      x += x * 1e-12 * (2 * rng.uniform() - 1);
      y += y * 1e-12 * (2 * rng.uniform() - 1);
      if (accept(i, x, y)) lat[c] = i++;
    }
    n_lattice = i;
    int per = 16;
    int ncirc = std::max(1, nc / per);
    for (int c = 0; c < ncirc && i < nl + nc; c++) {
      double r = L * (0.002 + 0.02 * rng.uniform());
      double cx = r + (L - 2 * r) * rng.uniform(), cy = r + (L - 2 * r) * rng.uniform();
      for (int k = 0; k < per && i < nl + nc; k++) {
        double num = (double)rng.below(2001) - 1000.0, den = 1000.0;
        double u2n = num * num, u2d = den * den;
        double x = cx + r * ((u2d - u2n) / (u2d + u2n));
        double y = cy + r * ((2.0 * num * den) / (u2d + u2n));
        if (rng.below(2)) y = 2 * cy - y;
        if (x < 0 || x > L || y < 0 || y > L) continue;
        if (accept(i, x, y)) i++;
      }
    }
  }
  while (i < n) {
    double x, y;
    if (dist == 1) {
      int k = (int)rng.below((uint64_t)ncl);
      x = centers[2 * k] + 0.05 * L * rng.normal();
      y = centers[2 * k + 1] + 0.05 * L * rng.normal();
      if (x < 0 || x > L || y < 0 || y > L) continue;
    } else if (dist == 3) {
      // disk: uniform in the disk of radius 0.45 L about the domain's centre
      // (mesh.cpp:131-146)
      double u = rng.uniform() - 0.5, v = rng.uniform() - 0.5;
      if (u * u + v * v > 0.45 * 0.45) continue;
      x = L * (u + 0.5);
      y = L * (v + 0.5);
    } else if (dist == 4) {
      // circle: a ring of radii 0.45 L to 0.49 L, the radius drawn uniformly
      // (so denser towards the inside) and the angle uniformly (mesh.cpp:148-161)
      double d = 0.04 * rng.uniform(), a = 6.283185307179586 * rng.uniform();
      x = L * (0.5 + (0.45 + d) * std::cos(a));
      y = L * (0.5 + (0.45 + d) * std::sin(a));
    } else if (dist == 5) {
      // grid: a uniform draw rounded half up to the nearest node of a
      // lattice with 10 n cells a side (mesh.cpp:163-176)
      const double M = 10.0 * n, h = L / M;
      x = h * std::floor(rng.uniform() * M + 0.5);
      y = h * std::floor(rng.uniform() * M + 0.5);
    } else if (dist == 6) {
      // ellipse: on the curve with half-axes 0.25 L (x) and 0.5 L (y)
      // about the domain's centre, the angle uniform (mesh.cpp:178-193)
      double a = 6.283185307179586 * rng.uniform();
      x = L * (0.5 + 0.25 * std::cos(a));
      y = L * (0.5 + 0.5 * std::sin(a));
    } else {
      x = L * rng.uniform();
      y = L * rng.uniform();
    }
    if (x < 0 || x > L || y < 0 || y > L) continue;
    if (accept(i, x, y)) i++;
  }
  if (nseg <= 0) return 0;
  // uniform grid over points and accepted segments
  double cell = maxlen > 0 ? maxlen : L / std::max(1.0, std::sqrt((double)n) / 4);
  int G = std::max(1, std::min(4096, (int)std::ceil(L / cell)));
  cell = L / G;
  auto cidx = [&](double v) { return std::max(0, std::min(G - 1, (int)(v / cell))); };
  std::vector<std::vector<int>> pcell((size_t)G * G);
  // accepted segments by cells of a finer grid (clustered inputs pack
  // thousands of segments into one endpoint cell); only the crossing test's
  // speed depends on it
  const int GS = std::max(1, std::min(8192, 4 * G));
  const double scell_w = L / GS;
  auto sidx = [&](double v) { return std::max(0, std::min(GS - 1, (int)(v / scell_w))); };
  std::vector<std::vector<int>> scell((size_t)GS * GS);
  for (int i = 0; i < n; i++) pcell[(size_t)cidx(points[2 * i + 1]) * G + cidx(points[2 * i])].push_back(i);
  std::vector<std::vector<int>> inc(n);  // accepted segments per vertex
  // The hull's boundary edges join the segment list before the random
  // segments are drawn (mesh.cpp:225-240), so a new segment keeps the
  // minimum angle with them too and never duplicates one.  They are not
  // written out: the CDT builder adds every hull edge itself.
  const std::vector<int> ring = hull_boundary(points, n);
  std::vector<int> ringpos(n, -1);
  for (size_t r = 0; r < ring.size(); r++) ringpos[ring[r]] = (int)r;
  double cosmin = std::cos(min_angle_deg * 3.14159265358979323846 / 180.0);
  double good2 = cosmin * cosmin;
  int k = 0;
  long attempts = 0, max_attempts = 200L * nseg + 1000;
  std::vector<long> tested((size_t)nseg, 0);  // crossing-test stamp per segment
  std::vector<double> sbox(4 * (size_t)nseg);   // accepted segments' x0 x1 y0 y1
  long stamp_now = 0;
  int nlat_seg = dist == 2 ? nseg / 2 : 0;  // segments along lattice lines
  while (k < nseg && attempts++ < max_attempts) {
    int i = (int)rng.below((uint64_t)n), j = -1;
    if (k < nlat_seg) {
      // next lattice point to the right or above a random lattice point
      int c = (int)rng.below((uint64_t)side * side);
      if (lat[c] < 0) continue;
      i = lat[c];
      int step = rng.below(2) ? 1 : side;
      int lim = step == 1 ? side - 1 - c % side : side - 1 - c / side;
      for (int t = 1; t <= lim; t++)
        if (lat[c + t * step] >= 0) { j = lat[c + t * step]; break; }
      if (j < 0) continue;
    } else if (maxlen > 0) {
      // a uniformly drawn point of the 3 x 3 cells around point i (the
      // cells' lists taken in row order, without copying them: dense
      // clusters put 10^5 points in a block)
      int cx = cidx(points[2 * i]), cy = cidx(points[2 * i + 1]);
      size_t total = 0;
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int x = cx + dx, y = cy + dy;
          if (x < 0 || y < 0 || x >= G || y >= G) continue;
          total += pcell[(size_t)y * G + x].size();
        }
      size_t r = rng.below(total);
      for (int dy = -1; dy <= 1 && j < 0; dy++)
        for (int dx = -1; dx <= 1 && j < 0; dx++) {
          int x = cx + dx, y = cy + dy;
          if (x < 0 || y < 0 || x >= G || y >= G) continue;
          const std::vector<int> &pc = pcell[(size_t)y * G + x];
          if (r < pc.size()) j = pc[r];
          else r -= pc.size();
        }
    } else {
      j = (int)rng.below((uint64_t)n);
    }
    if (i == j) continue;
    // lattice points are joined only along lattice lines to their next
    // lattice neighbour: a longer lattice-to-lattice segment would pass
    // within the jitter (~1e-10) of the lattice points between its ends,
    // a feature no refinement resolves in double precision
    if (k >= nlat_seg && i < n_lattice && j < n_lattice) continue;
    double dx = points[2 * j] - points[2 * i], dy = points[2 * j + 1] - points[2 * i + 1];
    double len2 = dx * dx + dy * dy;
    if (maxlen > 0 && len2 > maxlen * maxlen) continue;
    bool ok = true;
    // incident segments and hull edges: duplicate / angle check
    // (mesh.cpp:287-352)
    auto too_close = [&](int e, int o, int f) {
      if (f == o) return true;
      double ux = points[2 * o] - points[2 * e], uy = points[2 * o + 1] - points[2 * e + 1];
      double vx = points[2 * f] - points[2 * e], vy = points[2 * f + 1] - points[2 * e + 1];
      double dot = ux * vx + uy * vy;
      return dot > 0 && dot * dot > good2 * (ux * ux + uy * uy) * (vx * vx + vy * vy);
    };
    for (int e : {i, j}) {
      int o = e == i ? j : i;
      for (int s : inc[e]) {
        int a2 = segs[2 * s], b2 = segs[2 * s + 1];
        if (too_close(e, o, a2 == e ? b2 : a2)) { ok = false; break; }
      }
      const int r = ringpos[e], R = (int)ring.size();
      if (ok && r >= 0 && R >= 3)
        ok = !too_close(e, o, ring[(r + 1) % R]) && !too_close(e, o, ring[(r + R - 1) % R]);
      if (!ok) break;
    }
    if (!ok) continue;
    // crossing test against segments in the cells the new segment overlaps
    int x0 = sidx(std::min(points[2 * i], points[2 * j])), x1 = sidx(std::max(points[2 * i], points[2 * j]));
    int y0 = sidx(std::min(points[2 * i + 1], points[2 * j + 1])), y1 = sidx(std::max(points[2 * i + 1], points[2 * j + 1]));
    ++stamp_now;
    const double bx0 = std::min(points[2 * i], points[2 * j]), bx1 = std::max(points[2 * i], points[2 * j]);
    const double by0 = std::min(points[2 * i + 1], points[2 * j + 1]),
                 by1 = std::max(points[2 * i + 1], points[2 * j + 1]);
    for (int y = y0; y <= y1 && ok; y++)
      for (int x = x0; x <= x1 && ok; x++)
        for (int s : scell[(size_t)y * GS + x]) {
          if (tested[s] == stamp_now) continue;  // met in an earlier cell
          tested[s] = stamp_now;
          int a2 = segs[2 * s], b2 = segs[2 * s + 1];
          if (a2 == i || a2 == j || b2 == i || b2 == j) continue;
          // closed bounding boxes apart: no common point (exact)
          const double *bb = &sbox[4 * (size_t)s];
          if (bb[0] > bx1 || bb[1] < bx0 || bb[2] > by1 || bb[3] < by0) continue;
          int o1 = orient(points, i, j, a2) * orient(points, i, j, b2);
          int o2 = orient(points, a2, b2, i) * orient(points, a2, b2, j);
          if (o1 <= 0 && o2 <= 0) { ok = false; break; }
        }
    if (!ok) continue;
    segs[2 * k] = i;
    segs[2 * k + 1] = j;
    sbox[4 * (size_t)k] = bx0;
    sbox[4 * (size_t)k + 1] = bx1;
    sbox[4 * (size_t)k + 2] = by0;
    sbox[4 * (size_t)k + 3] = by1;
    inc[i].push_back(k);
    inc[j].push_back(k);
    for (int y = y0; y <= y1; y++)
      for (int x = x0; x <= x1; x++) scell[(size_t)y * GS + x].push_back(k);
    k++;
  }
  return k;
}

// Host evaluation of the same predicates the kernels use (for tests).
extern "C" int gcdt_host_predicates(const double *q, int n, int pred_mode,
                                    int *orient_out, int *incircle_out) {
  gcdt::PredStats st = {0, 0, 0, 0};
  for (int i = 0; i < n; i++) {
    const double *p = q + 8 * (size_t)i;
    orient_out[i] = orient2d_sign(p[0], p[1], p[2], p[3], p[4], p[5], pred_mode == 1, &st);
    incircle_out[i] = incircle_sign(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], g_ws,
                                    pred_mode == 1, &st);
  }
  return 0;
}
