/* CPU restatement of CRMod/CRTomo's 2.5D complex forward + sensitivity path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP product is checked
 * against (through tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg) and nothing else: the product path never links, loads or calls it.
 *
 * Each function restates one reference routine, in the reference's operation
 * order, from its Fortran source (file:line cited per function).  Complex
 * arithmetic is spelled out in real arithmetic, the way the Fortran complex
 * operators define it; the library is built with -ffp-contract=off.
 *
 * Pinning (tests/test_oracle_*.py): every function here is compared with the
 * reference's own routines compiled from its source (oracle/_ref), with the
 * FP32 volt.dat fixtures the reference ships, and with closed forms (numpy
 * Gauss-Legendre/Laguerre rules, element-matrix row sums and areas, dense
 * complex solves, L*L^T = A, the Euler sum rule of the sensitivities, dense
 * numpy algebra for the CG products).
 */
#include "crt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ complex */
static inline orc_c cc(double re, double im) { orc_c z = {re, im}; return z; }
static inline orc_c cadd(orc_c a, orc_c b) { return cc(a.re + b.re, a.im + b.im); }
static inline orc_c csub(orc_c a, orc_c b) { return cc(a.re - b.re, a.im - b.im); }
static inline orc_c cmul(orc_c a, orc_c b) {
  return cc(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
/* real scalar promoted to complex, times complex (dcmplx(r) * z) */
static inline orc_c rmul(double r, orc_c z) { return cc(r * z.re, r * z.im); }
/* a/b with the divisor scaled by a power of two (C99 Annex G, G.5.1): the
 * algorithm the Fortran compiler's runtime uses for COMPLEX division, so the
 * oracle rounds like the reference.  Inputs here are always finite. */
static inline orc_c cdiv(orc_c a, orc_c b) {
  double c = b.re, d = b.im;
  double lg = logb(fmax(fabs(c), fabs(d)));
  int il = 0;
  if (isfinite(lg)) {
    il = (int)lg;
    c = scalbn(c, -il);
    d = scalbn(d, -il);
  }
  double den = c * c + d * d;
  return cc(scalbn((a.re * c + a.im * d) / den, -il), scalbn((a.im * c - a.re * d) / den, -il));
}
static inline double cabs_(orc_c a) { return hypot(a.re, a.im); }
/* principal square root */
static inline orc_c csqrt_(orc_c z) {
  double x = z.re, y = z.im;
  if (x == 0.0 && y == 0.0) return cc(0.0, y);
  double t = sqrt((fabs(x) + hypot(x, y)) * 0.5);
  if (x >= 0.0) return cc(t, y / (2.0 * t));
  return cc(fabs(y) / (2.0 * t), copysign(t, y));
}
static inline orc_c clog_(orc_c z) { return cc(log(cabs_(z)), atan2(z.im, z.re)); }
static inline orc_c cconj(orc_c z) { return cc(z.re, -z.im); }

/* ---------------------------------------------------------------- mesh access */
static int m_elanz(const orc_mesh *m) {
  int n = 0;
  for (int i = 0; i < m->typanz; i++) if (m->typ[i] <= 10) n += m->nelanz[i];
  return n;
}
static int m_relanz(const orc_mesh *m) {
  int n = 0;
  for (int i = 0; i < m->typanz; i++) if (m->typ[i] > 10) n += m->nelanz[i];
  return n;
}
static int m_maxnodes(const orc_mesh *m) {
  int n = 0;
  for (int i = 0; i < m->typanz; i++) if (m->selanz[i] > n) n = m->selanz[i];
  return n;
}
/* 1-based node of element iel (0-based), local node k (0-based) */
static inline int NREL(const orc_mesh *m, int nel, int iel, int k) { return m->nrel[iel + (size_t)k * nel]; }
static inline double SX(const orc_mesh *m, int node1) { return m->sx[m->snr[node1 - 1] - 1]; }
static inline double SY(const orc_mesh *m, int node1) { return m->sy[m->snr[node1 - 1] - 1]; }

/* ------------------------------------------------------- quadrature (rwaven) */
/* gammln: src/gammln.f90:6-25 (Lanczos series) */
static double gammln(double xx) {
  static const double cof[6] = {76.18009172947146, -86.50532032941677, 24.01409824083091,
                                -1.231739572450155, .1208650973866179e-2, -.5395239384953e-5};
  const double stp = 2.5066282746310005;
  double x = xx, y = x, tmp = x + 5.5;
  tmp = (x + 0.5) * log(tmp) - tmp;
  double ser = 1.000000000190015;
  for (int j = 0; j < 6; j++) { y = y + 1.0; ser = ser + cof[j] / y; }
  return tmp + log(stp * ser / x);
}

/* gauleg on [x1,x2]: src/gauleg.f90:9-40 (Newton on P_n, EPS 3e-14) */
static void gauleg(double x1, double x2, double *x, double *w, int n) {
  const double EPS = 3.0e-14;
  int m = (n + 1) / 2;
  double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  for (int i = 1; i <= m; i++) {
    double z = cos(3.141592653589793 * ((double)i - .25) / ((double)n + .5));
    double z1, p1, p2, p3, pp;
    do {
      p1 = 1.0; p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        p3 = p2; p2 = p1;
        p1 = ((2.0 * (double)j - 1.0) * z * p2 - ((double)j - 1.0) * p3) / (double)j;
      }
      pp = (double)n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (fabs(z - z1) > EPS);
    x[i - 1] = xm - xl * z;
    x[n - i] = xm + xl * z;
    w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
}

/* gaulag with exponent alf: src/gaulag.f90:6-44 */
static void gaulag(double *x, double *w, int n, double alf) {
  const double EPS = 3.0e-14;
  const int MAXIT = 10;
  double z = 0.0, z1, p1 = 0, p2 = 0, p3, pp = 0;
  for (int i = 1; i <= n; i++) {
    if (i == 1) {
      z = (1.0 + alf) * (3.0 + .92 * alf) / (1.0 + 2.4 * n + 1.8 * alf);
    } else if (i == 2) {
      z = z + (15.0 + 6.25 * alf) / (1.0 + .9 * alf + 2.5 * n);
    } else {
      double ai = (double)(i - 2);
      z = z + ((1.0 + 2.55 * ai) / (1.9 * ai) + 1.26 * ai * alf / (1.0 + 3.5 * ai)) *
                  (z - x[i - 3]) / (1.0 + .3 * alf);
    }
    for (int its = 1; its <= MAXIT; its++) {
      p1 = 1.0; p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        p3 = p2; p2 = p1;
        p1 = (((double)(2 * j - 1) + alf - z) * p2 - ((double)(j - 1) + alf) * p3) / (double)j;
      }
      pp = ((double)n * p1 - ((double)n + alf) * p2) / z;
      z1 = z;
      z = z1 - p1 / pp;
      if (fabs(z - z1) <= EPS) break;
    }
    x[i - 1] = z;
    w[i - 1] = -exp(gammln(alf + (double)n) - gammln((double)n)) / (pp * (double)n * p2);
  }
}

/* rwaven: src/rwaven.f90:48-151.  kwn/kwnwi need room for 30 entries. */
int orc_rwaven(const orc_mesh *m, int nanz, const int *strnr, const int *vnr,
               int *kwnanz, double *kwn, double *kwnwi, double *amin_o, double *amax_o) {
  double amin = 1e9, amax = 0.0;
  for (int i = 0; i < nanz; i++) {
    int elec[4];
    double xke[4] = {0, 0, 0, 0}, yke[4] = {0, 0, 0, 0};
    elec[0] = strnr[i] % 10000;
    elec[1] = (strnr[i] - elec[0]) / 10000;
    elec[2] = vnr[i] % 10000;
    elec[3] = (vnr[i] - elec[2]) / 10000;
    for (int j = 0; j < 4; j++)
      if (elec[j] > 0) {
        xke[j] = SX(m, m->enr[elec[j] - 1]);
        yke[j] = SY(m, m->enr[elec[j] - 1]);
      }
    for (int q = 2; q < 4; q++) {
      if (elec[q] <= 0) continue;
      for (int j = 0; j < 2; j++) {
        if (elec[j] > 0) {
          double x21 = xke[q] - xke[j], y21 = yke[q] - yke[j];
          double dum = sqrt(x21 * x21 + y21 * y21);
          amin = fmin(dum, amin);
          amax = fmax(dum, amax);
        }
      }
    }
  }
  if (amin < 1e-12 || amax < 1e-12) return 73;
  int lanz = 4;
  /* ganz = int(real(6*log10(amax/amin))): note the single-precision cast */
  int ganz = (int)(float)(6.0 * log10(amax / amin));
  if (ganz < 2) ganz = 2;
  if (ganz > 30 - lanz) ganz = 30 - lanz;
  int K = ganz + lanz;
  double kwn0 = 1.0 / (2.0 * amin);
  gauleg(0.0, 1.0, kwn, kwnwi, ganz);
  for (int i = 0; i < ganz; i++) {
    kwnwi[i] = 2.0 * kwn0 * kwnwi[i] * kwn[i];
    kwn[i] = kwn0 * kwn[i] * kwn[i];
  }
  gaulag(kwn + ganz, kwnwi + ganz, lanz, 0.0);
  for (int i = ganz; i < K; i++) {
    kwnwi[i] = kwn0 * kwnwi[i] * exp(kwn[i]);
    kwn[i] = kwn0 * (kwn[i] + 1.0);
  }
  *kwnanz = K;
  if (amin_o) *amin_o = amin;
  if (amax_o) *amax_o = amax;
  return 0;
}

/* ------------------------------------------------------ Bessel (NR/A&S)     */
/* src/bessi0.f90, bessi1.f90, bessk0.f90, bessk1.f90 (A&S 9.8.x polynomials) */
static double bessi0(double x) {
  if (fabs(x) < 3.75) {
    double y = (x / 3.75) * (x / 3.75);
    return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 +
           y * (0.360768e-1 + y * 0.45813e-2)))));
  }
  double ax = fabs(x), y = 3.75 / ax;
  return (exp(ax) / sqrt(ax)) * (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 +
         y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 + y * (0.2635537e-1 +
         y * (-0.1647633e-1 + y * 0.392377e-2))))))));
}
static double bessi1(double x) {
  if (fabs(x) < 3.75) {
    double y = (x / 3.75) * (x / 3.75);
    return x * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 +
           y * (0.2658733e-1 + y * (0.301532e-2 + y * 0.32411e-3))))));
  }
  double ax = fabs(x), y = 3.75 / ax;
  return (exp(ax) / sqrt(ax)) * (0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 +
         y * (0.163801e-2 + y * (-0.1031555e-1 + y * (0.2282967e-1 + y * (-0.2895312e-1 +
         y * (0.1787654e-1 + y * -0.420059e-2))))))));
}
static double bessk0(double x) {
  if (x <= 2.0) {
    double y = x * x / 4.0;
    return (-log(x / 2.0) * bessi0(x)) + (-0.57721566 + y * (0.42278420 + y * (0.23069756 +
           y * (0.3488590e-1 + y * (0.262698e-2 + y * (0.10750e-3 + y * 0.74e-5))))));
  } else if (x > 5e2) {
    return 0.0;
  }
  double y = 2.0 / x;
  return (exp(-x) / sqrt(x)) * (1.25331414 + y * (-0.7832358e-1 + y * (0.2189568e-1 +
         y * (-0.1062446e-1 + y * (0.587872e-2 + y * (-0.251540e-2 + y * 0.53208e-3))))));
}
static double bessk1(double x) {
  if (x <= 2.0) {
    double y = x * x / 4.0;
    return (log(x / 2.0) * bessi1(x)) + (1.0 / x) * (1.0 + y * (0.15443144 + y * (-0.67278579 +
           y * (-0.18156897 + y * (-0.1919402e-1 + y * (-0.110404e-2 + y * -0.4686e-4))))));
  } else if (x > 5e2) {
    return 0.0;
  }
  double y = 2.0 / x;
  return (exp(-x) / sqrt(x)) * (1.25331414 + y * (0.23498619 + y * (-0.3655620e-1 +
         y * (0.1504268e-1 + y * (-0.780353e-2 + y * (0.325614e-2 + y * -0.68245e-3))))));
}

/* beta: src/beta.f90:38-135.  xk/yk = the 2 nodes of the boundary element. */
static int beta_fn(const orc_mesh *m, const double *xk, const double *yk, int nelec1,
                   double kw, double sytop, double *out) {
  double xs = SX(m, m->enr[nelec1 - 1]);
  double ys = SY(m, m->enr[nelec1 - 1]);
  double xr = (xk[0] + xk[1]) / 2.0, yr = (yk[0] + yk[1]) / 2.0;
  double x3 = xr - xs;
  double y3m = (yr - sytop) - (ys - sytop);
  double y3p = (yr - sytop) + (ys - sytop);
  double r1m = sqrt(x3 * x3 + y3m * y3m);
  double r1p = sqrt(x3 * x3 + y3p * y3p);
  double x4 = yk[0] - yk[1], y4 = xk[1] - xk[0];
  double r2 = sqrt(x4 * x4 + y4 * y4);
  if (r1m < 1e-12 || r1p < 1e-12 || r2 < 1e-12) return 35;
  double cthetm = (x3 * x4 + y3m * y4) / r1m / r2;
  double cthetp = (x3 * x4 + y3p * y4) / r1p / r2;
  double bk0m = bessk0(r1m * kw), bk0p = bessk0(r1p * kw);
  double bk1m = bessk1(r1m * kw), bk1p = bessk1(r1p * kw);
  double b;
  if (bk0m + bk0p == 0.0) b = (cthetm + cthetp) / 2.0;
  else b = (cthetm * bk1m + cthetp * bk1p) / (bk0m + bk0p);
  *out = kw * b;
  return 0;
}

/* ---------------------------------------------------------- element matrices */
static const int S1_3[9] = {1, -1, 0, -1, 1, 0, 0, 0, 0};   /* column major (i,j) */
static const int S2_3[9] = {2, -1, -1, -1, 0, 1, -1, 1, 0};
static const int S3_3[9] = {1, 0, -1, 0, 0, 0, -1, 0, 1};
static const int S4_3[9] = {2, 1, 1, 1, 2, 1, 1, 1, 2};
#define IX3(i, j) ((i) + 3 * (j))

/* elem3: src/elem3.f90:27-99; elmas/elmam are 3x3 (ld 3) */
static int elem3(const double *xk, const double *yk, double *elmas, double *elmam) {
  double x21 = xk[1] - xk[0], x31 = xk[2] - xk[0];
  double y21 = yk[1] - yk[0], y31 = yk[2] - yk[0];
  double det = x21 * y31 - x31 * y21;
  if (det <= 0.0) return 26;
  double a = (x31 * x31 + y31 * y31) / det;
  double b = -(x31 * x21 + y31 * y21) / det;
  double c = (x21 * x21 + y21 * y21) / det;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      elmas[IX3(i, j)] = (a * (double)S1_3[IX3(i, j)] + b * (double)S2_3[IX3(i, j)] +
                          c * (double)S3_3[IX3(i, j)]) / 2.0;
      elmam[IX3(i, j)] = det * (double)S4_3[IX3(i, j)] / 2.4e1;
    }
  return 0;
}

/* elem5: src/elem5.f90:1-72; 4x4 (ld 4) */
static int elem5(const double *xk, const double *yk, double *elmas, double *elmam) {
  static const int s1[16] = {2, -2, -1, 1, -2, 2, 1, -1, -1, 1, 2, -2, 1, -1, -2, 2};
  static const int s2[16] = {1, 0, -1, 0, 0, -1, 0, 1, -1, 0, 1, 0, 0, 1, 0, -1};
  static const int s3[16] = {2, 1, -1, -2, 1, 2, -2, -1, -1, -2, 2, 1, -2, -1, 1, 2};
  static const int s4[16] = {4, 2, 1, 2, 2, 4, 2, 1, 1, 2, 4, 2, 2, 1, 2, 4};
  double x21 = xk[1] - xk[0], x31 = xk[3] - xk[0];
  double y21 = yk[1] - yk[0], y31 = yk[3] - yk[0];
  double det = x21 * y31 - x31 * y21;
  if (det <= 0.0) return 26;
  double a = (x31 * x31 + y31 * y31) / det;
  double b = -(x31 * x21 + y31 * y21) / det;
  double c = (x21 * x21 + y21 * y21) / det;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      elmas[i + 4 * j] = (a * (double)s1[i + 4 * j] + 3.0 * b * (double)s2[i + 4 * j] +
                          c * (double)s3[i + 4 * j]) / 6.0;
      elmam[i + 4 * j] = det * (double)s4[i + 4 * j] / 3.6e1;
    }
  return 0;
}

/* elem8 (4 sub-triangles + condensation of the centre node):
 * src/elem8.f90:18-205; kelmas is 4x4 (ld 4) */
static int elem8(const double *xk, const double *yk, double kwert, double *kelmas) {
  double xkd[5], ykd[5], elmas[25], elmam[25];
  memset(elmas, 0, sizeof(elmas));
  memset(elmam, 0, sizeof(elmam));
  double a1 = xk[2] - xk[0], a2 = xk[3] - xk[1];
  if (fabs(a1) <= 1e-12 || fabs(a2) <= 1e-12) return 26;
  a1 = (yk[2] - yk[0]) / a1;
  a2 = (yk[3] - yk[1]) / a2;
  if (fabs(a1 - a2) <= 1e-12) return 26;
  xkd[4] = (a1 * xk[0] - yk[0] - a2 * xk[1] + yk[1]) / (a1 - a2);
  ykd[4] = yk[0] + a1 * (xkd[4] - xk[0]);
  for (int i = 0; i < 4; i++) { xkd[i] = xk[i]; ykd[i] = yk[i]; }
  static const int IK[4][3] = {{0, 1, 4}, {1, 2, 4}, {4, 2, 3}, {0, 4, 3}};
  for (int k = 0; k < 4; k++) {
    const int *ik = IK[k];
    double x21 = xkd[ik[1]] - xkd[ik[0]], x31 = xkd[ik[2]] - xkd[ik[0]];
    double y21 = ykd[ik[1]] - ykd[ik[0]], y31 = ykd[ik[2]] - ykd[ik[0]];
    double det = x21 * y31 - x31 * y21;
    if (det <= 1e-12) return 26;
    double a = (x31 * x31 + y31 * y31) / det;
    double b = -(x31 * x21 + y31 * y21) / det;
    double c = (x21 * x21 + y21 * y21) / det;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        int p = ik[i] + 5 * ik[j];
        elmas[p] = elmas[p] + (a * (double)S1_3[IX3(i, j)] + b * (double)S2_3[IX3(i, j)] +
                               c * (double)S3_3[IX3(i, j)]) / 2.0;
        elmam[p] = elmam[p] + det * (double)S4_3[IX3(i, j)] / 2.4e1;
      }
  }
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) elmas[i + 5 * j] = elmas[i + 5 * j] + kwert * kwert * elmam[i + 5 * j];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      kelmas[i + 4 * j] = elmas[i + 5 * j] - elmas[i + 5 * 4] * elmas[4 + 5 * j] / elmas[4 + 5 * 4];
  return 0;
}

/* bsytop: src/bsytop.f90:7-71 (running mean of no-flow element centres) */
static double bsytop(const orc_mesh *m) {
  int nel = m_elanz(m) + m_relanz(m);
  double sytop = 0.0;
  int iel = 0;
  for (int i = 0; i < m->typanz; i++) {
    if (m->typ[i] == 12) {
      int nkel = m->selanz[i];
      sytop = 0.0;
      for (int j = 1; j <= m->nelanz[i]; j++) {
        double sp = 0.0;
        for (int ik = 0; ik < nkel; ik++) sp = sp + SY(m, NREL(m, nel, iel, ik));
        iel++;
        sp = sp / (double)nkel;
        if (j == 1) sytop = sp;
        else sytop = (sytop * (double)(j - 1) + sp) / (double)j;
      }
    } else {
      iel += m->nelanz[i];
    }
  }
  return sytop;
}

/* precal: src/precal.f90:13-222.
 * elbg: elanz x nimn x K; relbg: relanz x nimn; kg: relanz x eanz x K (col-major),
 * nimn = maxnodes*(maxnodes+1)/2. */
int orc_precal(const orc_mesh *m, int K, const double *kwn, double *elbg, double *relbg,
               double *kg, int *lbeta_o, int *lrandb_o, double *sytop_o) {
  int elanz = m_elanz(m), relanz = m_relanz(m), nel = elanz + relanz;
  int maxn = m_maxnodes(m), nimn = maxn * (maxn + 1) / 2;
  int lbeta = 0, lrandb = 0;
  double sytop = m->lsytop ? bsytop(m) : 0.0;
  double xk[8], yk[8];
  int iel = 0;
  for (int i = 0; i < m->typanz; i++) {
    int ntyp = m->typ[i], nkel = m->selanz[i];
    for (int j = 0; j < m->nelanz[i]; j++, iel++) {
      for (int q = 0; q < nkel; q++) {
        xk[q] = SX(m, NREL(m, nel, iel, q));
        yk[q] = SY(m, NREL(m, nel, iel, q));
      }
      if (ntyp == 11) {
        if (lrandb) return 101;
        lbeta = 1;
        double x21 = xk[1] - xk[0], y21 = yk[1] - yk[0];
        double l = sqrt(x21 * x21 + y21 * y21);
        static const int s4[4] = {2, 1, 1, 2};
        int rel = iel - elanz;
        if (rel < 0) return 36;
        int imn = 0;
        for (int mm = 0; mm < nkel; mm++)
          for (int n = 0; n <= mm; n++, imn++)
            relbg[rel + (size_t)relanz * imn] = l * (double)s4[mm + 2 * n] / 6.0;
        for (int l1 = 0; l1 < m->eanz; l1++)
          for (int k = 0; k < K; k++) {
            double b;
            int e = beta_fn(m, xk, yk, l1 + 1, kwn[k], sytop, &b);
            if (e) return e;
            kg[rel + (size_t)relanz * (l1 + (size_t)m->eanz * k)] = b;
          }
      } else if (ntyp == 12) {
        continue;
      } else if (ntyp == 13) {
        if (lbeta) return 101;
        lrandb = 1;
      } else if (ntyp == 8) {
        double kel[16];
        for (int k = 0; k < K; k++) {
          int e = elem8(xk, yk, kwn[k], kel);
          if (e) return e;
          int imn = 0;
          for (int mm = 0; mm < nkel; mm++)
            for (int n = 0; n <= mm; n++, imn++)
              elbg[iel + (size_t)elanz * (imn + (size_t)nimn * k)] = kel[mm + 4 * n];
        }
      } else if (ntyp == 3 || ntyp == 5) {
        double es[16], em[16];
        int ld = (ntyp == 3) ? 3 : 4;
        int e = (ntyp == 3) ? elem3(xk, yk, es, em) : elem5(xk, yk, es, em);
        if (e) return e;
        for (int k = 0; k < K; k++) {
          int imn = 0;
          for (int mm = 0; mm < nkel; mm++)
            for (int n = 0; n <= mm; n++, imn++)
              elbg[iel + (size_t)elanz * (imn + (size_t)nimn * k)] =
                  es[mm + ld * n] + em[mm + ld * n] * kwn[k] * kwn[k];
        }
      } else {
        return 18;
      }
    }
  }
  if (lbeta_o) *lbeta_o = lbeta;
  if (lrandb_o) *lrandb_o = lrandb;
  if (sytop_o) *sytop_o = sytop;
  return 0;
}

/* ----------------------------------------------------------- forward solve */
/* kompab: src/kompab.f90:75-183 (lsr off, no electrode capacitances).
 * a: (mb+1)*sanz band (1-based index max*mb+min), b: sanz. l,k 0-based. */
static int kompab(const orc_mesh *m, int K, const double *elbg, const double *relbg,
                  const double *kg, const orc_c *sigma, int l, int k, orc_c *a, orc_c *b) {
  int elanz = m_elanz(m), relanz = m_relanz(m), nel = elanz + relanz;
  int maxn = m_maxnodes(m), nimn = maxn * (maxn + 1) / 2, mb = m->mb;
  (void)K;
  memset(a, 0, sizeof(orc_c) * (size_t)(mb + 1) * m->sanz);
  memset(b, 0, sizeof(orc_c) * (size_t)m->sanz);
  int iel = 0;
  for (int i = 0; i < m->typanz; i++) {
    int ntyp = m->typ[i], nkel = m->selanz[i];
    for (int j = 0; j < m->nelanz[i]; j++, iel++) {
      if (ntyp > 11) continue;
      int ikl = 0;
      for (int kk = 0; kk < nkel; kk++) {
        int nzp = NREL(m, nel, iel, kk);
        for (int ll = 0; ll <= kk; ll++) {
          int nnp = NREL(m, nel, iel, ll);
          int idif = abs(nzp - nnp);
          if (idif > mb) return 19;
          int imax = nzp > nnp ? nzp : nnp, imin = nzp < nnp ? nzp : nnp;
          size_t im = (size_t)imax * mb + imin; /* 1-based */
          double dum;
          orc_c dum2;
          if (ntyp == 11) {
            int rel = iel - elanz;
            dum = relbg[rel + (size_t)relanz * ikl] * kg[rel + (size_t)relanz * (l + (size_t)m->eanz * k)];
            dum2 = sigma[m->rnr[rel] - 1];
          } else {
            dum = elbg[iel + (size_t)elanz * (ikl + (size_t)nimn * k)];
            dum2 = sigma[iel];
          }
          ikl++;
          a[im - 1] = cadd(a[im - 1], rmul(dum, dum2));
        }
      }
    }
  }
  b[m->enr[l] - 1] = cc(-1.0, 0.0);
  if (m->lsink) b[m->nsink - 1] = cc(1.0, 0.0);
  return 0;
}

/* randb (homogeneous Dirichlet on type-13 nodes): src/randb.f90:43-82 */
static void randb(const orc_mesh *m, orc_c *a, orc_c *b) {
  int elanz = m_elanz(m), relanz = m_relanz(m), nel = elanz + relanz;
  int mb = m->mb, m1 = mb + 1, sanz = m->sanz;
  int iel = 0;
  for (int i2 = 0; i2 < m->typanz; i2++) {
    int ntyp = m->typ[i2], nkel = m->selanz[i2];
    for (int j2 = 0; j2 < m->nelanz[i2]; j2++, iel++) {
      if (ntyp != 13) continue;
      for (int ir = 0; ir < nkel; ir++) {
        int k = NREL(m, nel, iel, ir);
        b[k - 1] = cc(0, 0);
        int idk = k * m1;
        a[idk - 1] = cc(1, 0);
        if (k != 1) {
          int ia = (1 > mb + 2 - k) ? 1 : mb + 2 - k;
          for (int i = ia; i <= mb; i++) a[idk + i - m1 - 1] = cc(0, 0);
        }
        if (k == sanz) continue;
        int ia = (1 > k - sanz + m1) ? 1 : k - sanz + m1;
        for (int i = ia; i <= mb; i++) {
          int j = k - i + m1;
          int ji = (j - 1) * m1 + i;
          a[ji - 1] = cc(0, 0);
        }
      }
    }
  }
}

/* randb2 (prescribed boundary values): src/randb2.f90:28-73.  Row and column
 * k are zeroed with unit diagonal, b(k) = -value, and the column's entries
 * move to the right-hand side (b(j) += value * a(k,j)), node by node in the
 * order of the boundary file. */
static void randb2(const orc_mesh *m, orc_c *a, orc_c *b) {
  int mb = m->mb, m1 = mb + 1, sanz = m->sanz;
  for (int ir = 0; ir < m->rwdanz; ir++) {
    int k = m->rwdnr[ir];
    orc_c rw = m->rwd[ir];
    b[k - 1] = cc(-rw.re, -rw.im);
    int idk = k * m1;
    a[idk - 1] = cc(1, 0);
    if (k != 1) {
      int ia = (1 > mb + 2 - k) ? 1 : mb + 2 - k;
      for (int i = ia; i <= mb; i++) {
        int j = k + i - m1, ki = idk + i - m1;
        b[j - 1] = cadd(b[j - 1], cmul(rw, a[ki - 1]));
        a[ki - 1] = cc(0, 0);
      }
    }
    if (k == sanz) continue;
    int ia = (1 > k - sanz + m1) ? 1 : k - sanz + m1;
    for (int i = ia; i <= mb; i++) {
      int j = k - i + m1, ji = (j - 1) * m1 + i;
      b[j - 1] = cadd(b[j - 1], cmul(rw, a[ji - 1]));
      a[ji - 1] = cc(0, 0);
    }
  }
}

/* scalab: src/scalab.f90:52-76 */
static int scalab(int sanz, int mb, orc_c *a, orc_c *b, double *fak) {
  for (int i = 1; i <= sanz; i++) {
    int idi = i * (mb + 1);
    double dum = cabs_(a[idi - 1]);
    if (dum <= 0.0) return 27;
    a[idi - 1] = cdiv(a[idi - 1], cc(dum, 0.0));
    fak[i - 1] = 1.0 / sqrt(dum);
    b[i - 1] = rmul(fak[i - 1], b[i - 1]);
    if (i == 1) continue;
    int i0 = i * mb;
    int ja = (1 > i - mb) ? 1 : i - mb;
    for (int j = ja; j <= i - 1; j++) a[i0 + j - 1] = rmul(fak[i - 1] * fak[j - 1], a[i0 + j - 1]);
  }
  return 0;
}

/* chol: src/chol.f90:44-78 (row-oriented banded L L^T, no conjugation) */
int orc_chol(int sanz, int mb, orc_c *a) {
  int m1 = mb + 1;
  for (int i = 1; i <= sanz; i++) {
    int idi = i * m1;
    int fi = (1 > i - mb) ? 1 : i - mb;
    int i0 = idi - i;
    for (int j = fi; j <= i; j++) {
      int ij = i0 + j, j0 = j * mb;
      orc_c s = a[ij - 1];
      for (int k = fi; k <= j - 1; k++) s = csub(s, cmul(a[i0 + k - 1], a[j0 + k - 1]));
      if (j < i) {
        a[ij - 1] = cdiv(s, a[j * m1 - 1]);
      } else {
        if (cabs_(s) <= 0.0) return 28;
        a[idi - 1] = csqrt_(s);
      }
    }
  }
  return 0;
}

/* vre: src/vre.f90:45-69 (A x + b = 0 with the negation in the back sweep) */
void orc_vre(int sanz, int mb, const orc_c *a, const orc_c *b, orc_c *pot) {
  int m1 = mb + 1, idi = 0;
  for (int i = 1; i <= sanz; i++) {
    idi = i * m1;
    orc_c s = b[i - 1];
    int i0 = idi - i;
    int jlow = (1 > i - mb) ? 1 : i - mb;
    for (int j = jlow; j <= i - 1; j++) s = csub(s, cmul(a[i0 + j - 1], pot[j - 1]));
    pot[i - 1] = cdiv(s, a[idi - 1]);
  }
  for (int i = sanz; i >= 1; i--) {
    orc_c x = cdiv(pot[i - 1], a[idi - 1]);
    pot[i - 1] = cc(-x.re, -x.im);
    int jlow = (1 > i - mb) ? 1 : i - mb;
    int i0 = idi - i;
    for (int j = jlow; j <= i - 1; j++) pot[j - 1] = cadd(pot[j - 1], cmul(a[i0 + j - 1], pot[i - 1]));
    idi = idi - m1;
  }
}

/* Verification aid, not a restatement: the same band system (scaled matrix a,
 * right-hand side b, both as orc_assemble returns them) factored and solved in
 * x87 long double (64-bit significand), pot = x (A x + b = 0).  Comparing the
 * double-precision oracle and the HIP path with this solution measures how far
 * each is from the exact solution of the discrete system, i.e. the
 * conditioning floor that any FP64 evaluation order shares.  Plain complex
 * long double arithmetic; divisions by the pivot use the textbook formula. */
typedef struct { long double re, im; } ld_c;
static inline ld_c ld_mul(ld_c a, ld_c b) {
  ld_c r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
static inline ld_c ld_div(ld_c a, ld_c b) {
  long double den = b.re * b.re + b.im * b.im;
  ld_c r = {(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
  return r;
}
static inline ld_c ld_sqrt(ld_c z) {
  long double t = sqrtl((fabsl(z.re) + hypotl(z.re, z.im)) * 0.5L);
  ld_c r;
  if (t == 0.0L) { r.re = 0.0L; r.im = z.im; return r; }
  if (z.re >= 0.0L) { r.re = t; r.im = z.im / (2.0L * t); }
  else { r.re = fabsl(z.im) / (2.0L * t); r.im = copysignl(t, z.im); }
  return r;
}
int orc_solve_ld(int sanz, int mb, const orc_c *a_in, const orc_c *b, orc_c *pot) {
  const int m1 = mb + 1;
  ld_c *a = (ld_c *)malloc(sizeof(ld_c) * (size_t)m1 * sanz);
  ld_c *x = (ld_c *)malloc(sizeof(ld_c) * (size_t)sanz);
  if (!a || !x) { free(a); free(x); return -1; }
  for (size_t i = 0; i < (size_t)m1 * sanz; i++) { a[i].re = a_in[i].re; a[i].im = a_in[i].im; }
  /* row-oriented band L L^T, row i holds columns i-mb..i (diagonal last) */
  for (int i = 0; i < sanz; i++) {
    const int fi = i - mb > 0 ? i - mb : 0;
    ld_c *ri = a + (size_t)i * m1 + (mb - i);  /* ri[j] = entry (i, j) */
    for (int j = fi; j <= i; j++) {
      ld_c *rj = a + (size_t)j * m1 + (mb - j);
      ld_c s = ri[j];
      for (int k = fi; k < j; k++) {
        ld_c p = ld_mul(ri[k], rj[k]);
        s.re -= p.re; s.im -= p.im;
      }
      if (j < i) ri[j] = ld_div(s, rj[j]);
      else { if (s.re == 0.0L && s.im == 0.0L) { free(a); free(x); return 28; } ri[i] = ld_sqrt(s); }
    }
  }
  for (int i = 0; i < sanz; i++) {
    const int fi = i - mb > 0 ? i - mb : 0;
    const ld_c *ri = a + (size_t)i * m1 + (mb - i);
    ld_c s = {b[i].re, b[i].im};
    for (int j = fi; j < i; j++) { ld_c p = ld_mul(ri[j], x[j]); s.re -= p.re; s.im -= p.im; }
    x[i] = ld_div(s, ri[i]);
  }
  for (int i = sanz - 1; i >= 0; i--) {
    const int fi = i - mb > 0 ? i - mb : 0;
    const ld_c *ri = a + (size_t)i * m1 + (mb - i);
    ld_c q = ld_div(x[i], ri[i]);
    x[i].re = -q.re; x[i].im = -q.im;
    for (int j = fi; j < i; j++) { ld_c p = ld_mul(ri[j], x[i]); x[j].re += p.re; x[j].im += p.im; }
  }
  for (int i = 0; i < sanz; i++) { pot[i].re = (double)x[i].re; pot[i].im = (double)x[i].im; }
  free(a);
  free(x);
  return 0;
}

int orc_assemble(const orc_mesh *m, int K, const double *elbg, const double *relbg,
                 const double *kg, int lrandb, const orc_c *sigma, int l, int k,
                 orc_c *a, orc_c *b, double *fak) {
  int e = kompab(m, K, elbg, relbg, kg, sigma, l, k, a, b);
  if (e) return e;
  if (lrandb) randb(m, a, b);
  if (m->rwdanz > 0) randb2(m, a, b);
  return scalab(m->sanz, m->mb, a, b, fak);
}

/* Forward loop over (k,l): src/fem.f90:328-397, then rtrafo (src/rtrafo.f90:73-88).
 * kpot: sanz x eanz x K, hpot: sanz x eanz.  Systems s = k*eanz + l in
 * [sys_begin, sys_end) are solved (the whole set when sys_end <= 0); the
 * back-transform runs only for the whole set.  Parallel over systems. */
int orc_forward(const orc_mesh *m, int K, const double *kwnwi, int swrtr, const double *elbg,
                const double *relbg, const double *kg, int lbeta, int lrandb,
                const orc_c *sigma, orc_c *kpot, orc_c *hpot, int nthreads, int sys_begin,
                int sys_end) {
  int sanz = m->sanz, mb = m->mb, E = m->eanz;
  int nsys = K * E;
  int whole = sys_end <= 0;
  if (whole) { sys_begin = 0; sys_end = nsys; }
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  if (lbeta) {
#pragma omp parallel
    {
      orc_c *a = malloc(sizeof(orc_c) * (size_t)(mb + 1) * sanz);
      orc_c *b = malloc(sizeof(orc_c) * (size_t)sanz);
      orc_c *pot = malloc(sizeof(orc_c) * (size_t)sanz);
      double *fak = malloc(sizeof(double) * (size_t)sanz);
#pragma omp for schedule(dynamic, 1)
      for (int s = sys_begin; s < sys_end; s++) {
        int k = s / E, l = s % E;
        int e = orc_assemble(m, K, elbg, relbg, kg, lrandb, sigma, l, k, a, b, fak);
        if (!e) e = orc_chol(sanz, mb, a);
        if (e) {
#pragma omp atomic write
          err = e;
          continue;
        }
        orc_vre(sanz, mb, a, b, pot);
        orc_c *dst = kpot + (size_t)sanz * (l + (size_t)E * k);
        for (int j = 0; j < sanz; j++) dst[j] = rmul(fak[j], pot[j]);
      }
      free(a); free(b); free(pot); free(fak);
    }
  } else {
    /* one factor per wavenumber, RHS rebuilt by kompb (src/kompb.f90:39-59) */
    int k_lo = sys_begin / E, k_hi = (sys_end + E - 1) / E;
#pragma omp parallel
    {
      orc_c *a = malloc(sizeof(orc_c) * (size_t)(mb + 1) * sanz);
      orc_c *b = malloc(sizeof(orc_c) * (size_t)sanz);
      orc_c *pot = malloc(sizeof(orc_c) * (size_t)sanz);
      double *fak = malloc(sizeof(double) * (size_t)sanz);
#pragma omp for schedule(dynamic, 1)
      for (int k = k_lo; k < k_hi; k++) {
        int e = orc_assemble(m, K, elbg, relbg, kg, lrandb, sigma, 0, k, a, b, fak);
        if (!e) e = orc_chol(sanz, mb, a);
        if (e) {
#pragma omp atomic write
          err = e;
          continue;
        }
        for (int l = 0; l < E; l++) {
          int s = k * E + l;
          if (s < sys_begin || s >= sys_end) continue;
          if (l > 0) {
            memset(b, 0, sizeof(orc_c) * (size_t)sanz);
            b[m->enr[l] - 1] = cc(-fak[m->enr[l] - 1], 0.0);
            if (m->lsink) b[m->nsink - 1] = cc(fak[m->nsink - 1], 0.0);
          }
          orc_vre(sanz, mb, a, b, pot);
          orc_c *dst = kpot + (size_t)sanz * (l + (size_t)E * k);
          for (int j = 0; j < sanz; j++) dst[j] = rmul(fak[j], pot[j]);
        }
      }
      free(a); free(b); free(pot); free(fak);
    }
  }
  if (err || !whole) return err;
  if (swrtr == 0) {
    memcpy(hpot, kpot, sizeof(orc_c) * (size_t)sanz * E);
  } else {
    double pi = acos(-1.0);
#pragma omp parallel for collapse(2)
    for (int l = 0; l < E; l++)
      for (int j = 0; j < sanz; j++) {
        orc_c s = cc(0, 0);
        for (int k = 0; k < K; k++) s = cadd(s, rmul(kwnwi[k], kpot[j + (size_t)sanz * (l + (size_t)E * k)]));
        hpot[j + (size_t)sanz * l] = cdiv(s, cc(pi, 0.0));
      }
  }
  return 0;
}

/* bvolt (mode 0, sigmaa = kfak*V): src/bvolt.f90:40-80;
 * bvolti (mode 1, sigmaa = -ln V, errnr 82 on V = 0): src/bvolti.f90:48-90 */
int orc_volt(const orc_mesh *m, int nanz, const int *strnr, const int *vnr, const orc_c *hpot,
             int mode, const double *kfak, orc_c *volt, orc_c *sigmaa) {
  int sanz = m->sanz;
  for (int i = 0; i < nanz; i++) {
    int e1 = strnr[i] % 10000, e2 = (strnr[i] - e1) / 10000;
    int e3 = vnr[i] % 10000, e4 = (vnr[i] - e3) / 10000;
#define HP(ep, es) hpot[(m->enr[((ep) > 1 ? (ep) : 1) - 1] - 1) + (size_t)sanz * (((es) > 1 ? (es) : 1) - 1)]
#define MSK(a, b) ((double)(((a) < 1 ? (a) : 1) * ((b) < 1 ? (b) : 1)))
    orc_c d1 = rmul(MSK(e4, e2), HP(e4, e2));
    orc_c d2 = rmul(MSK(e4, e1), HP(e4, e1));
    orc_c d3 = rmul(MSK(e3, e2), HP(e3, e2));
    orc_c d4 = rmul(MSK(e3, e1), HP(e3, e1));
#undef HP
#undef MSK
    volt[i] = csub(csub(d1, d2), csub(d3, d4));
    if (mode == 0) {
      sigmaa[i] = rmul(kfak ? kfak[i] : 1.0, volt[i]);
    } else {
      if (cabs_(volt[i]) == 0.0) return 82;
      orc_c lg = clog_(volt[i]);
      sigmaa[i] = cc(-lg.re, -lg.im);
    }
  }
  return 0;
}

/* bsens (mode 0, sens -= dum): src/bsens.f90:94-193;
 * bsensi (mode 1, sens += dum*sigma/V): src/bsensi.f90:90-188.
 * sens: nanz x elanz col-major (mnr = identity, as CRMod sets it). */
int orc_sens(const orc_mesh *m, int nanz, const int *strnr, const int *vnr, int K,
             const double *kwnwi, int swrtr, const double *elbg, const orc_c *kpot, int mode,
             const orc_c *sigma, const orc_c *volt, orc_c *sens, int nthreads) {
  int elanz = m_elanz(m), relanz = m_relanz(m), nel = elanz + relanz;
  int maxn = m_maxnodes(m), nimn = maxn * (maxn + 1) / 2;
  int sanz = m->sanz, E = m->eanz;
  double pi = acos(-1.0);
  memset(sens, 0, sizeof(orc_c) * (size_t)nanz * elanz);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel
  {
    orc_c *hsens = malloc(sizeof(orc_c) * (size_t)K);
#pragma omp for schedule(guided, 16)
    for (int i = 0; i < nanz; i++) {
      int e1 = strnr[i] % 10000, e2 = (strnr[i] - e1) / 10000;
      int e3 = vnr[i] % 10000, e4 = (vnr[i] - e3) / 10000;
      orc_c sup[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
      int iel = 0;
      for (int ityp = 0; ityp < m->typanz; ityp++) {
        int ntyp = m->typ[ityp], nkel = m->selanz[ityp];
        if (ntyp > 10) continue;
        for (int jn = 0; jn < m->nelanz[ityp]; jn++, iel++) {
          for (int k = 0; k < K; k++) {
            orc_c h = cc(0, 0);
            for (int mi = 0; mi < nkel; mi++) {
              int nzp = NREL(m, nel, iel, mi);
              for (int mj = 0; mj < nkel; mj++) {
                int nnp = NREL(m, nel, iel, mj);
                int imax = (mi > mj ? mi : mj) + 1, imin = (mi < mj ? mi : mj) + 1;
                int imn = imax * (imax - 1) / 2 + imin;
                const orc_c *kp = kpot + (size_t)sanz * (size_t)E * k;
                if (e1 > 0) sup[0] = kp[(nnp - 1) + (size_t)sanz * (e1 - 1)];
                if (e2 > 0) sup[1] = kp[(nnp - 1) + (size_t)sanz * (e2 - 1)];
                if (e3 > 0) sup[2] = kp[(nzp - 1) + (size_t)sanz * (e3 - 1)];
                if (e4 > 0) sup[3] = kp[(nzp - 1) + (size_t)sanz * (e4 - 1)];
                orc_c dum = cmul(csub(sup[1], sup[0]), csub(sup[3], sup[2]));
                h = cadd(h, rmul(elbg[iel + (size_t)elanz * ((imn - 1) + (size_t)nimn * k)], dum));
              }
            }
            hsens[k] = h;
          }
          orc_c dum;
          if (swrtr == 0) {
            dum = hsens[0]; /* strom = 1 */
          } else {
            dum = cc(0, 0);
            for (int k = 0; k < K; k++) dum = cadd(dum, rmul(kwnwi[k], hsens[k]));
            dum = mode == 0 ? rmul(1.0 / pi, dum) : cdiv(dum, cc(pi, 0.0));
          }
          orc_c *dst = &sens[i + (size_t)nanz * iel];
          if (mode == 0) *dst = csub(*dst, dum);
          else *dst = cadd(*dst, cdiv(cmul(dum, sigma[iel]), volt[i]));
        }
      }
    }
    free(hsens);
  }
  return 0;
}

/* ---------------------------------------------------------------- DC twins
 * The real (ldc) variants of the forward path: kompadc
 * (src/kompadc.f90:66-160), randdc (src/randdc.f90, the randb loop on real
 * arrays), scaldc (src/scaldc.f90:49-73), choldc (src/choldc.f90:44-77),
 * vredc (src/vredc.f90:45-77), kompbdc (src/kompbdc.f90:39-59), kpotdc and
 * hpotdc (src/inv.f90:393-399), rtrafo's DC branch (src/rtrafo.f90:48-66) and
 * bsendc (src/bsendc.f90:74-184).  Real arithmetic throughout; sigma enters
 * through its real part (DBLE(sigma)). */
static int kompadc(const orc_mesh *m, const double *elbg, const double *relbg, const double *kg,
                   const orc_c *sigma, int l, int k, double *a, double *b) {
  int elanz = m_elanz(m), relanz = m_relanz(m), nel = elanz + relanz;
  int maxn = m_maxnodes(m), nimn = maxn * (maxn + 1) / 2, mb = m->mb;
  memset(a, 0, sizeof(double) * (size_t)(mb + 1) * m->sanz);
  memset(b, 0, sizeof(double) * (size_t)m->sanz);
  int iel = 0;
  for (int i = 0; i < m->typanz; i++) {
    int ntyp = m->typ[i], nkel = m->selanz[i];
    for (int j = 0; j < m->nelanz[i]; j++, iel++) {
      if (ntyp > 11) continue;
      int ikl = 0;
      for (int kk = 0; kk < nkel; kk++) {
        int nzp = NREL(m, nel, iel, kk);
        for (int ll = 0; ll <= kk; ll++) {
          int nnp = NREL(m, nel, iel, ll);
          if (abs(nzp - nnp) > mb) return 19;
          int imax = nzp > nnp ? nzp : nnp, imin = nzp < nnp ? nzp : nnp;
          size_t im = (size_t)imax * mb + imin; /* 1-based */
          double dum, dum2;
          if (ntyp == 11) {
            int rel = iel - elanz;
            dum = relbg[rel + (size_t)relanz * ikl] * kg[rel + (size_t)relanz * (l + (size_t)m->eanz * k)];
            dum2 = sigma[m->rnr[rel] - 1].re;
          } else {
            dum = elbg[iel + (size_t)elanz * (ikl + (size_t)nimn * k)];
            dum2 = sigma[iel].re;
          }
          ikl++;
          a[im - 1] += dum * dum2;
        }
      }
    }
  }
  b[m->enr[l] - 1] = -1.0;
  if (m->lsink) b[m->nsink - 1] = 1.0;
  return 0;
}

static void randdc(const orc_mesh *m, double *a, double *b) {
  int elanz = m_elanz(m), relanz = m_relanz(m), nel = elanz + relanz;
  int mb = m->mb, m1 = mb + 1, sanz = m->sanz;
  int iel = 0;
  for (int i2 = 0; i2 < m->typanz; i2++) {
    int ntyp = m->typ[i2], nkel = m->selanz[i2];
    for (int j2 = 0; j2 < m->nelanz[i2]; j2++, iel++) {
      if (ntyp != 13) continue;
      for (int ir = 0; ir < nkel; ir++) {
        int k = NREL(m, nel, iel, ir);
        b[k - 1] = 0.0;
        int idk = k * m1;
        a[idk - 1] = 1.0;
        if (k != 1) {
          int ia = (1 > mb + 2 - k) ? 1 : mb + 2 - k;
          for (int i = ia; i <= mb; i++) a[idk + i - m1 - 1] = 0.0;
        }
        if (k == sanz) continue;
        int ia = (1 > k - sanz + m1) ? 1 : k - sanz + m1;
        for (int i = ia; i <= mb; i++) {
          int j = k - i + m1;
          a[(j - 1) * m1 + i - 1] = 0.0;
        }
      }
    }
  }
}

/* randbdc2: src/randbdc2.f90, randb2 on the real arrays with the real values */
static void randbdc2(const orc_mesh *m, double *a, double *b) {
  int mb = m->mb, m1 = mb + 1, sanz = m->sanz;
  for (int ir = 0; ir < m->rwdanz; ir++) {
    int k = m->rwdnr[ir];
    double rw = m->rwd[ir].re;
    b[k - 1] = -rw;
    int idk = k * m1;
    a[idk - 1] = 1.0;
    if (k != 1) {
      int ia = (1 > mb + 2 - k) ? 1 : mb + 2 - k;
      for (int i = ia; i <= mb; i++) {
        int j = k + i - m1, ki = idk + i - m1;
        b[j - 1] = b[j - 1] + rw * a[ki - 1];
        a[ki - 1] = 0.0;
      }
    }
    if (k == sanz) continue;
    int ia = (1 > k - sanz + m1) ? 1 : k - sanz + m1;
    for (int i = ia; i <= mb; i++) {
      int j = k - i + m1, ji = (j - 1) * m1 + i;
      b[j - 1] = b[j - 1] + rw * a[ji - 1];
      a[ji - 1] = 0.0;
    }
  }
}

static int scaldc(int sanz, int mb, double *a, double *b, double *fak) {
  for (int i = 1; i <= sanz; i++) {
    int idi = i * (mb + 1);
    double dum = a[idi - 1];
    if (dum <= 0.0) return 27;
    a[idi - 1] = 1.0;
    fak[i - 1] = 1.0 / sqrt(dum);
    b[i - 1] = b[i - 1] * fak[i - 1];
    if (i == 1) continue;
    int i0 = i * mb;
    int ja = (1 > i - mb) ? 1 : i - mb;
    for (int j = ja; j <= i - 1; j++) a[i0 + j - 1] = a[i0 + j - 1] * fak[i - 1] * fak[j - 1];
  }
  return 0;
}

int orc_choldc(int sanz, int mb, double *a) {
  int m1 = mb + 1;
  for (int i = 1; i <= sanz; i++) {
    int idi = i * m1;
    int fi = (1 > i - mb) ? 1 : i - mb;
    int i0 = idi - i;
    for (int j = fi; j <= i; j++) {
      int ij = i0 + j, j0 = j * mb;
      double s = a[ij - 1];
      for (int k = fi; k <= j - 1; k++) s = s - a[i0 + k - 1] * a[j0 + k - 1];
      if (j < i) {
        a[ij - 1] = s / a[j * m1 - 1];
      } else {
        if (s <= 0.0) return 28;
        a[idi - 1] = sqrt(s);
      }
    }
  }
  return 0;
}

void orc_vredc(int sanz, int mb, const double *a, const double *b, double *pot) {
  int m1 = mb + 1, idi = 0;
  for (int i = 1; i <= sanz; i++) {
    idi = i * m1;
    double s = b[i - 1];
    int i0 = idi - i;
    int jlow = (1 > i - mb) ? 1 : i - mb;
    for (int j = jlow; j <= i - 1; j++) s = s - a[i0 + j - 1] * pot[j - 1];
    pot[i - 1] = s / a[idi - 1];
  }
  for (int i = sanz; i >= 1; i--) {
    pot[i - 1] = -pot[i - 1] / a[idi - 1];
    int jlow = (1 > i - mb) ? 1 : i - mb;
    int i0 = idi - i;
    for (int j = jlow; j <= i - 1; j++) pot[j - 1] = pot[j - 1] + a[i0 + j - 1] * pot[i - 1];
    idi = idi - m1;
  }
}

int orc_assemble_dc(const orc_mesh *m, const double *elbg, const double *relbg, const double *kg,
                    int lrandb, const orc_c *sigma, int l, int k, double *a, double *b, double *fak) {
  int e = kompadc(m, elbg, relbg, kg, sigma, l, k, a, b);
  if (e) return e;
  if (lrandb) randdc(m, a, b);
  if (m->rwdanz > 0) randbdc2(m, a, b);
  return scaldc(m->sanz, m->mb, a, b, fak);
}

/* The DC loop of src/inv.f90:343-400 (kpotdc = pot*fak) and rtrafo's DC
 * branch (hpotdc = sum_k kpotdc*kwnwi / pi, src/rtrafo.f90:54-62; a copy of
 * k = 1 for swrtr = 0).  kpotdc: sanz x eanz x K, hpotdc: sanz x eanz. */
int orc_forward_dc(const orc_mesh *m, int K, const double *kwnwi, int swrtr, const double *elbg,
                   const double *relbg, const double *kg, int lbeta, int lrandb, const orc_c *sigma,
                   double *kpotdc, double *hpotdc, int nthreads) {
  int sanz = m->sanz, mb = m->mb, E = m->eanz;
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel
  {
    double *a = malloc(sizeof(double) * (size_t)(mb + 1) * sanz);
    double *b = malloc(sizeof(double) * (size_t)sanz);
    double *pot = malloc(sizeof(double) * (size_t)sanz);
    double *fak = malloc(sizeof(double) * (size_t)sanz);
#pragma omp for schedule(dynamic, 1)
    for (int k = 0; k < K; k++) {
      int failed = 0;
      for (int l = 0; l < E && !failed; l++) {
        if (lbeta || l == 0) {
          int e = orc_assemble_dc(m, elbg, relbg, kg, lrandb, sigma, l, k, a, b, fak);
          if (!e) e = orc_choldc(sanz, mb, a);
          if (e) {
#pragma omp atomic write
            err = e;
            failed = 1;
            continue;
          }
        } else {
          memset(b, 0, sizeof(double) * (size_t)sanz); /* kompbdc */
          b[m->enr[l] - 1] = -fak[m->enr[l] - 1];
          if (m->lsink) b[m->nsink - 1] = fak[m->nsink - 1];
        }
        orc_vredc(sanz, mb, a, b, pot);
        double *dst = kpotdc + (size_t)sanz * (l + (size_t)E * k);
        for (int j = 0; j < sanz; j++) dst[j] = pot[j] * fak[j];
      }
    }
    free(a); free(b); free(pot); free(fak);
  }
  if (err) return err;
  if (swrtr == 0) {
    memcpy(hpotdc, kpotdc, sizeof(double) * (size_t)sanz * E);
  } else {
    double pi = acos(-1.0);
#pragma omp parallel for collapse(2)
    for (int l = 0; l < E; l++)
      for (int j = 0; j < sanz; j++) {
        double s = 0.0;
        for (int k = 0; k < K; k++) s = s + kpotdc[j + (size_t)sanz * (l + (size_t)E * k)] * kwnwi[k];
        hpotdc[j + (size_t)sanz * l] = s / pi;
      }
  }
  return 0;
}

/* bsendc: sensdc(i, e) += dum * DBLE(sigma(e) / volt(i)) with the real h_k
 * sums of kpotdc; sensdc: nanz x elanz column major. */
int orc_sens_dc(const orc_mesh *m, int nanz, const int *strnr, const int *vnr, int K,
                const double *kwnwi, int swrtr, const double *elbg, const double *kpotdc,
                const orc_c *sigma, const orc_c *volt, double *sensdc, int nthreads) {
  int elanz = m_elanz(m), relanz = m_relanz(m), nel = elanz + relanz;
  int maxn = m_maxnodes(m), nimn = maxn * (maxn + 1) / 2;
  int sanz = m->sanz, E = m->eanz;
  double pi = acos(-1.0);
  memset(sensdc, 0, sizeof(double) * (size_t)nanz * elanz);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel
  {
    double *hsens = malloc(sizeof(double) * (size_t)K);
#pragma omp for schedule(guided, 16)
    for (int i = 0; i < nanz; i++) {
      int e1 = strnr[i] % 10000, e2 = (strnr[i] - e1) / 10000;
      int e3 = vnr[i] % 10000, e4 = (vnr[i] - e3) / 10000;
      double sup[4] = {0, 0, 0, 0};
      int iel = 0;
      for (int ityp = 0; ityp < m->typanz; ityp++) {
        int ntyp = m->typ[ityp], nkel = m->selanz[ityp];
        if (ntyp > 10) continue;
        for (int jn = 0; jn < m->nelanz[ityp]; jn++, iel++) {
          for (int k = 0; k < K; k++) {
            double h = 0.0;
            const double *kp = kpotdc + (size_t)sanz * (size_t)E * k;
            for (int mi = 0; mi < nkel; mi++) {
              int nzp = NREL(m, nel, iel, mi);
              for (int mj = 0; mj < nkel; mj++) {
                int nnp = NREL(m, nel, iel, mj);
                int imax = (mi > mj ? mi : mj) + 1, imin = (mi < mj ? mi : mj) + 1;
                int imn = imax * (imax - 1) / 2 + imin;
                if (e1 > 0) sup[0] = kp[(nnp - 1) + (size_t)sanz * (e1 - 1)];
                if (e2 > 0) sup[1] = kp[(nnp - 1) + (size_t)sanz * (e2 - 1)];
                if (e3 > 0) sup[2] = kp[(nzp - 1) + (size_t)sanz * (e3 - 1)];
                if (e4 > 0) sup[3] = kp[(nzp - 1) + (size_t)sanz * (e4 - 1)];
                double dum = (sup[1] - sup[0]) * (sup[3] - sup[2]);
                h = h + elbg[iel + (size_t)elanz * ((imn - 1) + (size_t)nimn * k)] * dum;
              }
            }
            hsens[k] = h;
          }
          double dum;
          if (swrtr == 0) {
            dum = hsens[0];
          } else {
            dum = 0.0;
            for (int k = 0; k < K; k++) dum = dum + hsens[k] * kwnwi[k];
            dum = dum / pi;
          }
          sensdc[i + (size_t)nanz * iel] += dum * cdiv(sigma[iel], volt[i]).re;
        }
      }
    }
    free(hsens);
  }
  return 0;
}

/* Preconditioner and RHS of update: src/update.f90:179-293 (ldc, lfpi, ldiff
 * and the reference-model term off).  rdiag[j] = the diagonal entry of R that
 * update adds for the active ltri (smatm(j,1) for ltri 0/3/4,
 * smatm(j,maxnodes+1) for the triangular operators); rm = R m (the bvec that
 * update builds first, src/update.f90:63-172), NULL for m = 0. */
void orc_update_rhs(int nanz, int manz, const orc_c *sens, const double *wmatd, const int *wdfak,
                    const orc_c *dat, const orc_c *sigmaa, double lam, int ltri,
                    const double *rdiag, const orc_c *rm, double *cgfac, orc_c *bvec) {
  for (int j = 0; j < manz; j++) {
    double dum = 0.0;
    for (int i = 0; i < nanz; i++) {
      orc_c s = sens[i + (size_t)nanz * j];
      orc_c p = cmul(cconj(s), s);
      dum = dum + p.re * wmatd[i] * (double)wdfak[i];
    }
    dum = dum + lam * rdiag[j];
    cgfac[j] = 1.0 / sqrt(dum);
  }
  for (int j = 0; j < manz; j++) {
    orc_c cdum = cc(0, 0);
    for (int i = 0; i < nanz; i++) {
      orc_c s = sens[i + (size_t)nanz * j];
      cdum = cadd(cdum, cmul(cmul(cconj(s), cc(wmatd[i] * (double)wdfak[i], 0.0)), csub(dat[i], sigmaa[i])));
    }
    orc_c bv = rm ? rm[j] : cc(0, 0);
    if (ltri == 3 || ltri == 4) bvec[j] = cdum;
    else bvec[j] = csub(cdum, rmul(lam, bv));
    bvec[j] = rmul(cgfac[j], bvec[j]);
  }
}

/* Data part of the gradient, sum_i conj(J_ij) w_i r_i (src/update.f90:263-266). */
void orc_adjoint_grad(int nanz, int manz, const orc_c *sens, const double *wmatd,
                      const int *wdfak, const orc_c *resid, orc_c *grad) {
  for (int j = 0; j < manz; j++) {
    orc_c cdum = cc(0, 0);
    for (int i = 0; i < nanz; i++) {
      orc_c s = sens[i + (size_t)nanz * j];
      cdum = cadd(cdum, cmul(cmul(cconj(s), cc(wmatd[i] * (double)wdfak[i], 0.0)), resid[i]));
    }
    grad[j] = cdum;
  }
}

/* Rectangular smoothness operator: src/bsmatm_mod.f90:190-362 with no
 * discontinuities (ndis_x = ndis_z = 0).  smatm: manz x 3 col-major; element
 * k(i,j) = (i-1)*nx + j, i = row from the surface, nodes ordered lower-left,
 * lower-right, upper-right, upper-left. */
void orc_bsmatmreg(const orc_mesh *m, int nx, int nz, double alfx, double alfz, double *smatm) {
  int elanz = m_elanz(m), relanz = m_relanz(m), nel = elanz + relanz, manz = elanz;
  memset(smatm, 0, sizeof(double) * (size_t)manz * 3);
#define KIJ(i, j) (((i) - 1) * nx + (j) - 1)
#define NX_(e, l) SX(m, NREL(m, nel, (e), (l)))
#define NY_(e, l) SY(m, NREL(m, nel, (e), (l)))
  for (int i = 1; i <= nz; i++) {
    for (int j = 1; j <= nx; j++) {
      int e = KIJ(i, j);
      double dzleft = fabs(NY_(e, 3) - NY_(e, 0));
      double dzright = fabs(NY_(e, 2) - NY_(e, 1));
      double xmean = 0.0;
      for (int l = 0; l < 4; l++) xmean = xmean + NX_(e, l);
      xmean = xmean / 4.0;
      if (j > 1) {
        double xleft = 0.0;
        for (int l = 0; l < 4; l++) xleft = xleft + NX_(KIJ(i, j - 1), l);
        xleft = xleft / 4.0;
        smatm[e] = 1.0 * alfx * dzleft / fabs(xmean - xleft);
      }
      if (j < nx) {
        double xright = 0.0;
        for (int l = 0; l < 4; l++) xright = xright + NX_(KIJ(i, j + 1), l);
        xright = xright / 4.0;
        double dum = 1.0 * alfx * dzright / fabs(xright - xmean);
        smatm[e] = smatm[e] + dum;
        smatm[e + (size_t)manz] = -dum;
      }
      double dxup = fabs(NX_(e, 2) - NX_(e, 3));
      double dxdown = fabs(NX_(e, 1) - NX_(e, 0));
      double zmean = 0.0;
      for (int l = 0; l < 4; l++) zmean = zmean + NY_(e, l);
      zmean = zmean / 4.0;
      if (i > 1) {
        double zup = 0.0;
        for (int l = 0; l < 4; l++) zup = zup + NY_(KIJ(i - 1, j), l);
        zup = zup / 4.0;
        double dum = 1.0 * alfz * dxup / fabs(zup - zmean);
        smatm[e] = smatm[e] + dum;
      }
      if (i < nz) {
        double zdown = 0.0;
        for (int l = 0; l < 4; l++) zdown = zdown + NY_(KIJ(i + 1, j), l);
        zdown = zdown / 4.0;
        double dum = 1.0 * alfz * dxdown / fabs(zmean - zdown);
        smatm[e] = smatm[e] + dum;
        smatm[e + 2 * (size_t)manz] = -dum;
      }
    }
  }
#undef KIJ
#undef NX_
#undef NY_
}

/* bnachbar: src/bnachbar.f90:25-66.  nachbar(i,k) = the element that shares
 * edge k (nodes k, k+1 mod maxnodes) of element i (the last one found), and
 * nachbar(i, maxnodes+1) = the number of edges with a neighbour; manz x
 * (maxnodes+1) column major, 1-based element numbers, 0 = none. */
void orc_bnachbar(const orc_mesh *m, int *nachbar) {
  int elanz = m_elanz(m), nel = elanz + m_relanz(m), mx = m_maxnodes(m);
  memset(nachbar, 0, sizeof(int) * (size_t)elanz * (mx + 1));
  for (int i = 0; i < elanz; i++)
    for (int ik = 0; ik < mx; ik++) {
      int ik1 = NREL(m, nel, i, ik), ik2 = NREL(m, nel, i, (ik + 1) % mx);
      for (int j = 0; j < elanz; j++) {
        if (j == i) continue;
        for (int jk = 0; jk < mx; jk++) {
          int jk1 = NREL(m, nel, j, jk), jk2 = NREL(m, nel, j, (jk + 1) % mx);
          if ((ik1 == jk1 && ik2 == jk2) || (ik1 == jk2 && ik2 == jk1)) {
            nachbar[i + (size_t)elanz * ik] = j + 1;
            nachbar[i + (size_t)elanz * mx] += 1;
          }
        }
      }
    }
}

/* bsmatmtri: src/bsmatm_mod.f90:364-461 without decoupled edges.  For each
 * edge k of element i with a neighbour: w = edge length / centroid distance *
 * sqrt((alfx cos a)^2 + (alfz sin a)^2), a the angle of the centroid
 * difference; smatm(i,k) = -w, smatm(i,maxnodes+1) += w.  Centroids as relem
 * sets them (src/relem.f90:134-143).  smatm: manz x (maxnodes+1). */
void orc_bsmatmtri(const orc_mesh *m, double alfx, double alfz, const int *nachbar, int ndec,
                   const int *edecoup, const double *edecstr, double *smatm) {
  int elanz = m_elanz(m), nel = elanz + m_relanz(m), mx = m_maxnodes(m);
  double *ex = calloc((size_t)elanz, sizeof(double)), *ey = calloc((size_t)elanz, sizeof(double));
  int iel = 0;
  for (int ti = 0; ti < m->typanz; ti++)
    for (int j = 0; j < m->nelanz[ti]; j++, iel++) {
      if (m->typ[ti] >= 10) continue;
      int nk = m->selanz[ti];
      for (int k = 0; k < nk; k++) {
        ex[iel] = ex[iel] + SX(m, NREL(m, nel, iel, k));
        ey[iel] = ey[iel] + SY(m, NREL(m, nel, iel, k));
      }
      ex[iel] = ex[iel] / nk;
      ey[iel] = ey[iel] / nk;
    }
  memset(smatm, 0, sizeof(double) * (size_t)elanz * (mx + 1));
  for (int i = 0; i < elanz; i++)
    for (int k = 0; k < mx; k++) {
      int a = NREL(m, nel, i, k), b = NREL(m, nel, i, (k + 1) % mx);
      double dx = SX(m, a) - SX(m, b), dy = SY(m, a) - SY(m, b);
      double edglen = sqrt(dx * dx + dy * dy);
      int nb = nachbar[i + (size_t)elanz * k];
      if (nb > 0) {
        double px = ex[i] - ex[nb - 1], py = ey[i] - ey[nb - 1];
        double dist = sqrt(px * px + py * py);
        double ang = atan2(py, px);
        double cx = alfx * cos(ang), cz = alfz * sin(ang);
        double alfgeo = sqrt(cx * cx + cz * cz);
        double w = edglen / dist * alfgeo;
        /* decoupled edges, src/bsmatm_mod.f90:434-445: every listed pair
         * (either orientation) scales the weight, in list order */
        for (int d = 0; d < ndec; d++)
          if ((i + 1 == edecoup[d] && nb == edecoup[d + ndec]) ||
              (nb == edecoup[d] && i + 1 == edecoup[d + ndec]))
            w = w * edecstr[d];
        smatm[i + (size_t)elanz * k] = -w;
        smatm[i + (size_t)elanz * mx] = smatm[i + (size_t)elanz * mx] + w;
      }
    }
  free(ex);
  free(ey);
}

/* R m for the rectangular operator (the ltri = 0 branch of update,
 * src/update.f90:63-84, no prior model). */
void orc_rm_reg(int manz, int nx, const double *smatm, const orc_c *par, orc_c *out) {
  for (int i = 0; i < manz; i++) {
    orc_c cdum = cc(0, 0);
    if (i > 0) cdum = rmul(smatm[(i - 1) + (size_t)manz], par[i - 1]);
    if (i < manz - 1) cdum = cadd(cdum, rmul(smatm[i + (size_t)manz], par[i + 1]));
    if (i + 1 > nx) cdum = cadd(cdum, rmul(smatm[(i - nx) + 2 * (size_t)manz], par[i - nx]));
    if (i + 1 < manz - nx + 1) cdum = cadd(cdum, rmul(smatm[i + 2 * (size_t)manz], par[i + nx]));
    out[i] = cadd(cdum, rmul(smatm[i], par[i]));
  }
}

/* CG with Fletcher-Reeves, real alpha/beta: src/cg_mod.f90:426-518, bap :520-545,
 * bp :547-577, bptri :579-613, bb :699-726. */
int orc_cjggra(int nanz, int manz, const orc_c *sens, const double *wmatd, const int *wdfak,
               const double *cgfac, const orc_c *bvec_in, double lam, int ltri, int nx,
               const double *smatm, int ld, const int *nachbar, int maxnodes, int ncgmax,
               double eps, orc_c *dpar, double *cgres) {
  orc_c *rvec = malloc(sizeof(orc_c) * manz), *pvec = malloc(sizeof(orc_c) * manz);
  orc_c *bvec = malloc(sizeof(orc_c) * manz), *ap = malloc(sizeof(orc_c) * nanz);
  double alpha, dr, dr0 = 0, dr1 = 0, beta;
  int ncg = 0, k;
  for (int j = 0; j < manz; j++) { dpar[j] = cc(0, 0); rvec[j] = bvec_in[j]; pvec[j] = cc(0, 0); }
  for (k = 1; k <= ncgmax; k++) {
    ncg = k - 1;
    dr = 0.0;
    for (int j = 0; j < manz; j++) dr = dr + cmul(cconj(rvec[j]), rvec[j]).re;
    if (k == 1) {
      dr0 = dr * eps;
      beta = 0.0;
    } else {
      if (dr <= dr0) goto done;
      beta = dr / dr1;
    }
    for (int j = 0; j < manz; j++) pvec[j] = cadd(rvec[j], rmul(beta, pvec[j]));
    /* bap */
    for (int i = 0; i < nanz; i++) {
      orc_c s = cc(0, 0);
      for (int j = 0; j < manz; j++)
        s = cadd(s, cmul(cmul(pvec[j], sens[i + (size_t)nanz * j]), cc(cgfac[j], 0.0)));
      ap[i] = s;
    }
    /* R p */
    if (ltri == 0) {
      for (int i = 0; i < manz; i++) {
        orc_c cdum = cc(0, 0);
        if (i > 0) cdum = rmul(smatm[(i - 1) + (size_t)ld * 1] * cgfac[i - 1], pvec[i - 1]);
        if (i < manz - 1) cdum = cadd(cdum, rmul(smatm[i + (size_t)ld * 1] * cgfac[i + 1], pvec[i + 1]));
        if (i + 1 > nx) cdum = cadd(cdum, rmul(smatm[(i - nx) + (size_t)ld * 2] * cgfac[i - nx], pvec[i - nx]));
        if (i + 1 < manz - nx + 1) cdum = cadd(cdum, rmul(smatm[i + (size_t)ld * 2] * cgfac[i + nx], pvec[i + nx]));
        bvec[i] = cadd(cdum, rmul(smatm[i] * cgfac[i], pvec[i]));
      }
    } else if (ltri == 3 || ltri == 4) {
      for (int i = 0; i < manz; i++) bvec[i] = rmul(smatm[i], rmul(cgfac[i], pvec[i]));
    } else {
      for (int i = 0; i < manz; i++) {
        orc_c cdum = cc(0, 0);
        for (int j = 0; j < maxnodes; j++) {
          int idum = nachbar[i + (size_t)manz * j];
          if (idum != 0)
            cdum = cadd(cdum, rmul(cgfac[idum - 1], rmul(smatm[i + (size_t)ld * j], pvec[idum - 1])));
        }
        bvec[i] = cadd(cdum, rmul(cgfac[i], rmul(smatm[i + (size_t)ld * maxnodes], pvec[i])));
      }
    }
    /* bb */
    for (int j = 0; j < manz; j++) {
      orc_c cdum = cc(0, 0);
      for (int i = 0; i < nanz; i++)
        cdum = cadd(cdum, cmul(cmul(cconj(sens[i + (size_t)nanz * j]), cc(wmatd[i] * (double)wdfak[i], 0.0)), ap[i]));
      bvec[j] = cadd(cdum, rmul(lam, bvec[j]));
      bvec[j] = rmul(cgfac[j], bvec[j]);
    }
    dr1 = 0.0;
    for (int j = 0; j < manz; j++) dr1 = dr1 + cmul(cconj(pvec[j]), bvec[j]).re;
    alpha = dr / dr1;
    for (int j = 0; j < manz; j++) {
      dpar[j] = cadd(dpar[j], rmul(alpha, pvec[j]));
      rvec[j] = csub(rvec[j], rmul(alpha, bvec[j]));
    }
    dr1 = dr;
    if (cgres) cgres[k] = (double)(float)(eps * dr / dr0);
  }
  ncg = ncgmax;
done:
  if (cgres) cgres[0] = (double)ncg;
  free(rvec); free(pvec); free(bvec); free(ap);
  return ncg;
}

/* ------------------------------------------------------------------------
 * Inversion-loop pieces around the path: data misfit with polarity folding,
 * model parameters from and to conductivities, the model update step.
 * ------------------------------------------------------------------------ */

/* chkpo2 (src/dmisft.f90:213-300): phase-range classes of datum and model
 * response (imag parts are phases in rad), folded by pi as there; wdlok = 0
 * marks a polarity mismatch. */
void orc_chkpo2(orc_c dati, orc_c sigi, int ldum, orc_c *cdat, orc_c *csig, int *wdlok) {
  const double pi = acos(-1.0);
  double redat = dati.re, imdat = dati.im, resig = sigi.re, imsig = sigi.im;
  int idat, isig;
  if (imdat > pi / 2.0) idat = -1;          /* :254-260 */
  else if (imdat <= -pi / 2.0) idat = 1;
  else idat = 0;
  if (imsig > pi / 2.0) isig = -1;          /* :262-268 */
  else if (imsig <= -pi / 2.0) isig = 1;
  else isig = 0;
  if (idat == 0 && isig != 0) {             /* :270-277 */
    imsig = imsig + (double)isig * pi;
    if (!ldum) imdat = imdat - copysign(pi, imdat);
    *wdlok = 0;
  } else if (idat != 0 && isig == 0) {      /* :279-285 */
    if (ldum) imdat = imdat + (double)idat * pi;
    *wdlok = 0;
  } else if (idat != 0 && isig != 0) {      /* :287-291 */
    imsig = imsig + (double)isig * pi;
    imdat = imdat + (double)idat * pi;
  }
  *cdat = cc(redat, imdat);
  *csig = cc(resig, imsig);
}

/* dmisft (src/dmisft.f90:17-208) without its text output.  flags: 1 lpol,
 * 2 lfpi, 4 ldiff (d0, fm0 used), 8 lrobust.  stats: nrmsd, betrms, pharms,
 * rmssum, l1rat (0 unless robust), npol.  wmatd is updated in place by the
 * robust reweighting when l1rat > l1min (:140-191).  wdlok (optional) gets
 * the per-datum polarity flags.  Returns 99 when every datum is excluded. */
int orc_dmisft(int nanz, const orc_c *dat, const orc_c *sigmaa, double *wmatd,
               const double *wmatdr, const double *wmatdp, const orc_c *d0, const orc_c *fm0,
               int flags, double l1min, double *stats, int *wdlok_out) {
  const int lpol = flags & 1, lfpi = flags & 2, ldiff = flags & 4, lrobust = flags & 8;
  double *psi = (double *)malloc(sizeof(double) * (size_t)(nanz > 0 ? nanz : 1));
  int *wdlok = (int *)malloc(sizeof(int) * (size_t)(nanz > 0 ? nanz : 1));
  double nrmsd = 0.0, betrms = 0.0, pharms = 0.0, l1rat = 0.0;
  int idum = 0;
  for (int i = 0; i < nanz; ++i) {          /* :85-123 */
    orc_c cdat, csig, cdum;
    wdlok[i] = 1;
    orc_chkpo2(dat[i], sigmaa[i], lpol, &cdat, &csig, &wdlok[i]);
    if (!ldiff) cdum = csub(cdat, csig);
    else cdum = csub(csub(cdat, csig), csub(d0[i], fm0[i]));
    psi[i] = lfpi ? sqrt(wmatd[i]) * fabs(cdum.im) : sqrt(wmatd[i]) * cabs_(cdum);
    idum += wdlok[i];
    nrmsd += psi[i] * psi[i] * (double)wdlok[i];
    betrms += wmatdr[i] * (double)wdlok[i] * cdum.re * cdum.re;
    pharms += wmatdp[i] * (double)wdlok[i] * cdum.im * cdum.im;
  }
  if (idum == 0) {
    free(psi);
    free(wdlok);
    return 99;
  }
  stats[3] = nrmsd;                         /* rmssum, :133-137 */
  stats[5] = (double)(nanz - idum);         /* npol */
  nrmsd = sqrt(nrmsd / (double)idum);
  betrms = sqrt(betrms / (double)idum);
  pharms = sqrt(pharms / (double)idum);
  if (lrobust) {                            /* :140-191 */
    double *eps2 = (double *)malloc(sizeof(double) * (size_t)nanz);
    double norm = 0.0, norm2 = 0.0;
    for (int i = 0; i < nanz; ++i) {
      double dum = 1.0 / sqrt(wmatd[i]);
      eps2[i] = dum * sqrt(psi[i]);
      norm = norm + psi[i] * (double)wdlok[i];
      norm2 = norm2 + psi[i] * (double)wdlok[i] * dum / eps2[i];
    }
    for (int i = 0; i < nanz; ++i) eps2[i] = eps2[i] * norm2 / norm;
    norm2 = 0.0;
    for (int i = 0; i < nanz; ++i) {
      double dum = 1.0 / sqrt(wmatd[i]);
      eps2[i] = fmax(dum, eps2[i]);
      norm2 = norm2 + psi[i] * (double)wdlok[i] * dum / eps2[i];
    }
    l1rat = norm / norm2;
    if (l1rat > l1min)
      for (int i = 0; i < nanz; ++i) wmatd[i] = 1.0 / (eps2[i] * eps2[i]);
    free(eps2);
  }
  stats[0] = nrmsd;
  stats[1] = betrms;
  stats[2] = pharms;
  stats[4] = l1rat;
  if (wdlok_out) memcpy(wdlok_out, wdlok, sizeof(int) * (size_t)nanz);
  free(psi);
  free(wdlok);
  return 0;
}

/* bpar (src/bpar.f90): par(mnr(i)) = log(sigma(i)) for every element, in
 * element order (mnr 1-based; NULL = identity). */
void orc_bpar(int elanz, const int *mnr, const orc_c *sigma, orc_c *par) {
  for (int i = 0; i < elanz; ++i) par[(mnr ? mnr[i] : i + 1) - 1] = clog_(sigma[i]);
}

/* bsigma (src/bsigma.f90): sigma(i) = exp(par(mnr(i))). */
void orc_bsigma(int elanz, const int *mnr, const orc_c *par, orc_c *sigma) {
  for (int i = 0; i < elanz; ++i) {
    const orc_c p = par[(mnr ? mnr[i] : i + 1) - 1];
    const double e = exp(p.re);
    sigma[i] = cc(e * cos(p.im), e * sin(p.im));
  }
}

/* The model update of src/update.f90:332-351: par += step dpar, with the
 * lphi0 rule (a negative conductivity phase is set to zero, counted), and
 * bdpar = step * sum |dpar|^2.  out: bdpar, count. */
void orc_update_step(int manz, orc_c *par, const orc_c *dpar, double step, int lphi0, double *out) {
  double bdpar = 0.0;
  int in = 0;
  for (int j = 0; j < manz; ++j) {
    par[j] = cadd(par[j], cmul(cc(step, 0.0), dpar[j]));
    if (lphi0 && par[j].im < 0.0) {
      par[j] = cc(par[j].re, 0.0);
      in = in + 1;
    }
    bdpar = bdpar + cmul(dpar[j], cconj(dpar[j])).re;
  }
  out[0] = bdpar * step;
  out[1] = (double)in;
}

/* blam0 (src/blam0.f90:69-170): the starting regularisation weight from the
 * column sums of |J^H W J| row j, in the reference's loop order (j, then i,
 * then k: O(N M^2)).  mode 0: complex J (cdabs of the sum); mode 1: real
 * parts of J (the FPI branch, and the DC branch on sensdc).  sens is N x M
 * column major; w = wmatd * wdfak.  Presets of a fixed lambda (:47-62) are
 * the caller's. */
double orc_blam0(int nanz, int manz, const orc_c *sens, const double *w, int mode, double alfx,
                 double alfz) {
  double *jtj = (double *)malloc(sizeof(double) * (size_t)manz);
#pragma omp parallel for schedule(guided)
  for (int j = 0; j < manz; ++j) {
    if (mode == 0) {
      orc_c cdum = cc(0.0, 0.0);
      for (int i = 0; i < nanz; ++i) {
        const orc_c sij = cconj(sens[(size_t)j * nanz + i]);
        const orc_c wi = cc(w[i], 0.0);
        for (int k = 0; k < manz; ++k)
          cdum = cadd(cdum, cmul(cmul(sij, sens[(size_t)k * nanz + i]), wi));
      }
      jtj[j] = cabs_(cdum);
    } else {
      double dum = 0.0;
      for (int i = 0; i < nanz; ++i)
        for (int k = 0; k < manz; ++k)
          dum = dum + sens[(size_t)j * nanz + i].re * sens[(size_t)k * nanz + i].re * w[i];
      jtj[j] = fabs(dum);
    }
  }
  double s = 0.0;
  for (int j = 0; j < manz; ++j) s += jtj[j];
  free(jtj);
  double lammax = s / (double)manz;
  lammax = lammax * 2.0 / (alfx + alfz);
  return lammax * 5.0;
}
