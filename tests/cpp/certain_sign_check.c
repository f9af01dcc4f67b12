/* CPU check of the PIP test's certain-sign rule (cuspatial_amd/csrc/qpip_pip.hip,
 * the fast loop; the line form is line_of in qpip_edges.hpp), against the
 * reference's own arithmetic for one (point, edge): lhs = (px - ax) * rise,
 * rhs = run * (py - ay), rounded as written, and the 4-ULP collinearity test
 * (detail/algorithm/is_point_in_polygon.cuh:66-86,
 * detail/utility/floating_point.cuh:97-129).
 *
 * The rule: the edge scaled exactly by 2^-k (the power of two just above
 * n = |run| + |rise|), c' = fma(ay, run', -(ax * rise')),
 * t = fma(px, rise', fma(-py, run', c')); with s = max(|px|, |py|) + max(|ax|, |ay|)
 * and tol = fma(s, 2^-48, 2^-1000) (infinite for s >= 2^400), |t| > tol must give
 * (lhs < rhs) == (t < 0) and no collinearity. Edges with n outside [2^-70, 2^400],
 * a coordinate that is not finite, or a lossy scaling take no certain decision.
 *
 * Usage: certain_sign_check [iterations] [seed]; prints the counts, exits 1 on a
 * counterexample. Built with -ffp-contract=off (no implicit fma). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* The rule's constants; the test suite also builds this file with weaker ones
 * (a smaller tolerance, a wider range of n) to show that they fail. */
#ifndef TOL_SCALE
#define TOL_SCALE 0x1p-48
#endif
#ifndef N_MIN
#define N_MIN 0x1p-70
#endif

static uint64_t st = 88172645463325252ull;
static uint64_t rnd(void)
{
  st ^= st << 13;
  st ^= st >> 7;
  st ^= st << 17;
  return st;
}
static double ur(void) { return (double)(rnd() >> 11) * 0x1p-53; }

static int float_equal(double a, double b)
{
  if (isnan(a) || isnan(b)) return 0;
  uint64_t const sg = 1ull << 63;
  uint64_t ua, ub;
  memcpy(&ua, &a, 8);
  memcpy(&ub, &b, 8);
  ua = (ua & sg) ? (~ua + 1) : (ua | sg);
  ub = (ub & sg) ? (~ub + 1) : (ub | sg);
  return ua >= ub ? (ua - ub) <= 4 : (ub - ua) <= 4;
}

/* the line form; returns 0 where the edge takes no certain decision */
static int line_of(double ax, double ay, double run, double rise, double* rs, double* ss,
                   double* c)
{
  double const n = fabs(run) + fabs(rise);
  if (!(isfinite(ax) && isfinite(ay) && n >= N_MIN && n <= 0x1p400)) return 0;
  int const k = ilogb(n) + 1;
  *rs         = ldexp(run, -k);
  *ss         = ldexp(rise, -k);
  if (ldexp(*rs, k) != run || ldexp(*ss, k) != rise) return 0;
  *c = fma(ay, *rs, -(ax * *ss));
  return 1;
}

static double pick(double scale)
{
  int const k  = (int)(rnd() % 7);
  double const v = (ur() * 2 - 1) * scale;
  switch (k) {
    case 0: return v;
    case 1: return nextafter(v, 1e300);
    case 2: return ldexp(ur(), -(int)(rnd() % 1080));  /* tiny, down to subnormal */
    case 3: return floor(v);
    case 4: return -v;
    case 5: return v * 1e-9;
    default: return ldexp(1.0, (int)(rnd() % 60) - 30) * scale * 0x1p-20;
  }
}

int main(int argc, char** argv)
{
  long long const iters = argc > 1 ? atoll(argv[1]) : 20000000LL;
  if (argc > 2) st ^= (uint64_t)atoll(argv[2]) * 0x9E3779B97F4A7C15ull;
  long long checked = 0, certain = 0, bad = 0, nolines = 0;
  for (long long it = 0; it < iters; ++it) {
    double const scale = ldexp(1.0, (int)(rnd() % 240) - 120);
    double const ax = pick(scale), ay = pick(scale);
    double bx = pick(scale), by = pick(scale), px, py;
    int const mode = (int)(rnd() % 7);
    if (mode == 5) { bx = nextafter(ax, (rnd() & 1) ? 1e300 : -1e300); }  /* nearly vertical */
    if (mode == 6) { by = nextafter(ay, (rnd() & 1) ? 1e300 : -1e300); }  /* nearly horizontal */
    if (mode == 0 || mode >= 5) {  /* on or next to the segment */
      double const f = ur();
      px = ax + (bx - ax) * f;
      py = ay + (by - ay) * f;
      if (rnd() & 1) px = nextafter(px, (rnd() & 1) ? 1e300 : -1e300);
    } else if (mode == 1) {  /* on the line's extension, one ULP off */
      double const f = ur() * 4 - 2;
      px = nextafter(ax + (bx - ax) * f, (rnd() & 1) ? 1e300 : -1e300);
      py = ay + (by - ay) * f;
    } else if (mode == 2) {
      px = ax;
      py = pick(scale);
    } else if (mode == 3) {
      py = ay;
      px = pick(scale);
    } else {
      px = pick(scale);
      py = pick(scale);
    }
    double const run = bx - ax, rise = by - ay;
    if (float_equal(run, 0.0) && float_equal(rise, 0.0)) continue; /* not an effective edge */
    double const lhs = (px - ax) * rise, rhs = run * (py - ay);
    ++checked;
    double rs, ss, c;
    if (!line_of(ax, ay, run, rise, &rs, &ss, &c)) {
      ++nolines;
      continue;
    }
    double const t   = fma(px, ss, fma(-py, rs, c));
    double const s   = fmax(fabs(px), fabs(py)) + fmax(fabs(ax), fabs(ay));
    double const tol = s < 0x1p400 ? fma(s, TOL_SCALE, 0x1p-1000) : INFINITY;
    if (fabs(t) > tol) {
      ++certain;
      if ((lhs < rhs) != (t < 0.0) || float_equal(rhs, lhs)) {
        ++bad;
        if (bad < 10)
          printf("BAD a=(%a,%a) b=(%a,%a) p=(%a,%a) t=%a tol=%a lhs=%a rhs=%a\n", ax, ay, bx, by,
                 px, py, t, tol, lhs, rhs);
      }
    }
  }
  printf("checked %lld certain %lld no_line %lld bad %lld\n", checked, certain, nolines, bad);
  return bad != 0;
}
