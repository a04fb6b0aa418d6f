// Test-only wrapper that compiles the reference's own Triangle 1.6 source
// (GPU_Refine/triangle.cpp, used there as the CPU "-q" baseline and the CDT
// builder) into oracle/_ref/libtriref.so, and exposes a few of its internal
// routines with C linkage so the tests can use them as a differential check:
//   - triangulate()            triangle.cpp:15710 (whole mesher, CPU baseline)
//   - counterclockwise()       triangle.cpp:5244  (adaptive exact orient2d)
//   - incircle()               triangle.cpp:5892  (adaptive exact incircle)
//   - findcircumcenter()       triangle.cpp:6523  (circumcenter / off-center)
// The reference source is included from where it lies; nothing is copied.
#include "triangle.cpp"

static struct mesh g_m;
static struct behavior g_b;
static int g_init = 0;

static void ref_init(void) {
  if (!g_init) {
    memset(&g_m, 0, sizeof(g_m));
    memset(&g_b, 0, sizeof(g_b));
    exactinit();
    g_init = 1;
  }
}

extern "C" {

void triref_triangulate(char *sw, struct triangulateio *in,
                        struct triangulateio *out) {
  triangulate(sw, in, out, (struct triangulateio *)NULL);
}

void triref_free(void *p) { free(p); }

double triref_orient2d(const double *pa, const double *pb, const double *pc) {
  ref_init();
  return counterclockwise(&g_m, &g_b, (vertex)pa, (vertex)pb, (vertex)pc);
}

double triref_incircle(const double *pa, const double *pb, const double *pc,
                       const double *pd) {
  ref_init();
  return incircle(&g_m, &g_b, (vertex)pa, (vertex)pb, (vertex)pc,
                  (vertex)pd);
}

// offconstant as Triangle computes it from the minimum angle (triangle.cpp:3527-3533)
void triref_findcircumcenter(const double *org, const double *dest,
                             const double *apex, double minangle_deg,
                             int offcenter, double *out) {
  ref_init();
  double ga = cos(minangle_deg * PI / 180.0);
  g_b.offconstant = (ga == 1.0) ? 0.0 : 0.475 * sqrt((1.0 + ga) / (1.0 - ga));
  double xi, eta;
  findcircumcenter(&g_m, &g_b, (vertex)org, (vertex)dest, (vertex)apex, out,
                   &xi, &eta, offcenter);
}

}  // extern "C"
