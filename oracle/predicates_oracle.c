/*
 * Oracle exact predicates (test infrastructure only; never linked into the
 * product).  Deliberately a different method from the product's
 * floating-point expansions: every double is decomposed into an integer
 * mantissa and exponent, all coordinates of one predicate are scaled to a
 * common exponent, and the determinant is evaluated in exact big-integer
 * arithmetic.  A loose FP64 filter (|det| > 1e-14 * permanent, about ten
 * times wider than the tightest published bounds) skips the big-integer path
 * for clear cases.
 *
 * Semantics follow the reference's predicate contract:
 *   orient2d  > 0  iff  a,b,c counter-clockwise        (cudaCCW.h:487-524, cuda_ccw)
 *   incircle  > 0  iff  d strictly inside circle(a,b,c) (cudaCCW.h:711-757, cuda_inCircle_exact)
 * The reference's versions are not actually exact because their error-bound
 * constants are never uploaded (cudaCCW.h:41); these are.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#define BI_LIMBS 300 /* 9600 bits: enough for degree-4 products of any doubles */

typedef struct {
  int sign; /* -1, 0, +1 */
  int n;    /* used limbs */
  uint32_t d[BI_LIMBS];
} bigint;

static void bi_zero(bigint *a) { a->sign = 0; a->n = 0; }

static void bi_norm(bigint *a) {
  while (a->n > 0 && a->d[a->n - 1] == 0) a->n--;
  if (a->n == 0) a->sign = 0;
}

/* a = m * 2^shift, m a signed 64-bit integer */
static void bi_from_scaled(bigint *a, int64_t m, int shift) {
  bi_zero(a);
  if (m == 0) return;
  a->sign = m < 0 ? -1 : 1;
  uint64_t u = m < 0 ? (uint64_t)(-m) : (uint64_t)m;
  int limb = shift / 32, bit = shift % 32;
  memset(a->d, 0, sizeof(uint32_t) * (size_t)(limb + 4));
  /* place u << bit at limb offset */
  unsigned __int128 v = ((unsigned __int128)u) << bit;
  for (int i = 0; i < 4; i++) {
    a->d[limb + i] = (uint32_t)(v & 0xffffffffu);
    v >>= 32;
  }
  a->n = limb + 4;
  bi_norm(a);
}

static int bi_cmp_mag(const bigint *a, const bigint *b) {
  if (a->n != b->n) return a->n < b->n ? -1 : 1;
  for (int i = a->n - 1; i >= 0; i--)
    if (a->d[i] != b->d[i]) return a->d[i] < b->d[i] ? -1 : 1;
  return 0;
}

/* |r| = |a| + |b| */
static void mag_add(bigint *r, const bigint *a, const bigint *b) {
  int n = a->n > b->n ? a->n : b->n;
  uint64_t carry = 0;
  for (int i = 0; i < n; i++) {
    uint64_t s = carry;
    if (i < a->n) s += a->d[i];
    if (i < b->n) s += b->d[i];
    r->d[i] = (uint32_t)s;
    carry = s >> 32;
  }
  r->n = n;
  if (carry) r->d[r->n++] = (uint32_t)carry;
}

/* |r| = |a| - |b|, requires |a| >= |b| */
static void mag_sub(bigint *r, const bigint *a, const bigint *b) {
  int64_t borrow = 0;
  for (int i = 0; i < a->n; i++) {
    int64_t s = (int64_t)a->d[i] - borrow - (i < b->n ? (int64_t)b->d[i] : 0);
    if (s < 0) {
      s += ((int64_t)1 << 32);
      borrow = 1;
    } else {
      borrow = 0;
    }
    r->d[i] = (uint32_t)s;
  }
  r->n = a->n;
}

static void bi_add(bigint *r, const bigint *a, const bigint *b) {
  bigint t;
  if (a->sign == 0) { *r = *b; return; }
  if (b->sign == 0) { *r = *a; return; }
  if (a->sign == b->sign) {
    mag_add(&t, a, b);
    t.sign = a->sign;
  } else {
    int c = bi_cmp_mag(a, b);
    if (c == 0) { bi_zero(r); return; }
    if (c > 0) { mag_sub(&t, a, b); t.sign = a->sign; }
    else { mag_sub(&t, b, a); t.sign = b->sign; }
  }
  bi_norm(&t);
  *r = t;
}

static void bi_neg(bigint *a) { a->sign = -a->sign; }

static void bi_sub(bigint *r, const bigint *a, const bigint *b) {
  bigint nb = *b;
  bi_neg(&nb);
  bi_add(r, a, &nb);
}

static void bi_mul(bigint *r, const bigint *a, const bigint *b) {
  bigint t;
  if (a->sign == 0 || b->sign == 0) { bi_zero(r); return; }
  int n = a->n + b->n;
  memset(t.d, 0, sizeof(uint32_t) * (size_t)n);
  for (int i = 0; i < a->n; i++) {
    uint64_t carry = 0;
    for (int j = 0; j < b->n; j++) {
      uint64_t cur = (uint64_t)t.d[i + j] + (uint64_t)a->d[i] * b->d[j] + carry;
      t.d[i + j] = (uint32_t)cur;
      carry = cur >> 32;
    }
    int k = i + b->n;
    while (carry) {
      uint64_t cur = (uint64_t)t.d[k] + carry;
      t.d[k] = (uint32_t)cur;
      carry = cur >> 32;
      k++;
    }
  }
  t.n = n;
  t.sign = a->sign * b->sign;
  bi_norm(&t);
  *r = t;
}

/* Decompose finite x = m * 2^e with |m| < 2^53 integer. */
static void dec(double x, int64_t *m, int *e) {
  if (x == 0.0) { *m = 0; *e = 0; return; }
  int k;
  double f = frexp(x, &k); /* x = f * 2^k, 0.5 <= |f| < 1 */
  *m = (int64_t)ldexp(f, 53);
  *e = k - 53;
}

static void to_common(const double *v, int nv, bigint *out) {
  int64_t m[8];
  int e[8], emin = 1 << 30;
  for (int i = 0; i < nv; i++) {
    dec(v[i], &m[i], &e[i]);
    if (m[i] != 0 && e[i] < emin) emin = e[i];
  }
  if (emin == (1 << 30)) emin = 0;
  for (int i = 0; i < nv; i++) bi_from_scaled(&out[i], m[i], m[i] ? e[i] - emin : 0);
}

int orc_orient2d_exact_sign(const double *a, const double *b, const double *c) {
  double v[6] = {a[0], a[1], b[0], b[1], c[0], c[1]};
  bigint x[6], l, r, t1, t2, det;
  to_common(v, 6, x);
  bi_sub(&t1, &x[0], &x[4]); /* ax - cx */
  bi_sub(&t2, &x[3], &x[5]); /* by - cy */
  bi_mul(&l, &t1, &t2);
  bi_sub(&t1, &x[1], &x[5]); /* ay - cy */
  bi_sub(&t2, &x[2], &x[4]); /* bx - cx */
  bi_mul(&r, &t1, &t2);
  bi_sub(&det, &l, &r);
  return det.sign;
}

int orc_incircle_exact_sign(const double *a, const double *b, const double *c,
                            const double *d) {
  double v[8] = {a[0], a[1], b[0], b[1], c[0], c[1], d[0], d[1]};
  bigint x[8], adx, ady, bdx, bdy, cdx, cdy, t1, t2, m, lift, term, det;
  to_common(v, 8, x);
  bi_sub(&adx, &x[0], &x[6]);
  bi_sub(&ady, &x[1], &x[7]);
  bi_sub(&bdx, &x[2], &x[6]);
  bi_sub(&bdy, &x[3], &x[7]);
  bi_sub(&cdx, &x[4], &x[6]);
  bi_sub(&cdy, &x[5], &x[7]);
  bi_zero(&det);
  const bigint *P[3][6] = {{&adx, &ady, &bdx, &cdy, &cdx, &bdy},
                           {&bdx, &bdy, &cdx, &ady, &adx, &cdy},
                           {&cdx, &cdy, &adx, &bdy, &bdx, &ady}};
  for (int k = 0; k < 3; k++) {
    bi_mul(&t1, P[k][0], P[k][0]);
    bi_mul(&t2, P[k][1], P[k][1]);
    bi_add(&lift, &t1, &t2);
    bi_mul(&t1, P[k][2], P[k][3]);
    bi_mul(&t2, P[k][4], P[k][5]);
    bi_sub(&m, &t1, &t2);
    bi_mul(&term, &lift, &m);
    bi_add(&det, &det, &term);
  }
  return det.sign;
}

int orc_orient2d(const double *a, const double *b, const double *c) {
  double l = (a[0] - c[0]) * (b[1] - c[1]);
  double r = (a[1] - c[1]) * (b[0] - c[0]);
  double det = l - r;
  double perm = fabs(l) + fabs(r);
  if (fabs(det) > 1e-14 * perm) return det > 0 ? 1 : -1;
  return orc_orient2d_exact_sign(a, b, c);
}

int orc_incircle(const double *a, const double *b, const double *c,
                 const double *d) {
  double adx = a[0] - d[0], ady = a[1] - d[1];
  double bdx = b[0] - d[0], bdy = b[1] - d[1];
  double cdx = c[0] - d[0], cdy = c[1] - d[1];
  double la = adx * adx + ady * ady;
  double lb = bdx * bdx + bdy * bdy;
  double lc = cdx * cdx + cdy * cdy;
  double m1 = bdx * cdy - cdx * bdy, m2 = cdx * ady - adx * cdy,
         m3 = adx * bdy - bdx * ady;
  double det = la * m1 + lb * m2 + lc * m3;
  double perm = la * (fabs(bdx * cdy) + fabs(cdx * bdy)) +
                lb * (fabs(cdx * ady) + fabs(adx * cdy)) +
                lc * (fabs(adx * bdy) + fabs(bdx * ady));
  if (fabs(det) > 1e-14 * perm) return det > 0 ? 1 : -1;
  return orc_incircle_exact_sign(a, b, c, d);
}
