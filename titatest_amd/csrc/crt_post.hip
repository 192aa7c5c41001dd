// Element pre-computation, back-transform, voltage pick-off, sensitivities,
// adjoint gradient and CG normal-equation products (gfx950, complex FP64).
#include "crt_device.h"
#include "crt_post.h"

// ---------------------------------------------------------------- precal --
// Per-element matrices (src/precal.f90:99-205): elem3 (src/elem3.f90:27-99),
// elem5 (src/elem5.f90:1-72), elem8 with centre-node condensation
// (src/elem8.f90:18-205).  One thread per (area element, wavenumber); output
// elbg[k][e][imn], packed lower triangle imn = m(m+1)/2 + n (0-based, n <= m).
__constant__ int c_s1[9] = {1, -1, 0, -1, 1, 0, 0, 0, 0};
__constant__ int c_s2[9] = {2, -1, -1, -1, 0, 1, -1, 1, 0};
__constant__ int c_s3[9] = {1, 0, -1, 0, 0, 0, -1, 0, 1};
__constant__ int c_s4[9] = {2, 1, 1, 1, 2, 1, 1, 1, 2};

__device__ int elem8_dev(const double* xk, const double* yk, double kw, double* kel /*4x4 col*/) {
  double xkd[5], ykd[5], elmas[25], elmam[25];
  for (int i = 0; i < 25; ++i) { elmas[i] = 0.0; elmam[i] = 0.0; }
  double a1 = xk[2] - xk[0], a2 = xk[3] - xk[1];
  if (fabs(a1) <= 1e-12 || fabs(a2) <= 1e-12) return 26;
  a1 = (yk[2] - yk[0]) / a1;
  a2 = (yk[3] - yk[1]) / a2;
  if (fabs(a1 - a2) <= 1e-12) return 26;
  xkd[4] = (a1 * xk[0] - yk[0] - a2 * xk[1] + yk[1]) / (a1 - a2);
  ykd[4] = yk[0] + a1 * (xkd[4] - xk[0]);
  for (int i = 0; i < 4; ++i) { xkd[i] = xk[i]; ykd[i] = yk[i]; }
  const int IK[4][3] = {{0, 1, 4}, {1, 2, 4}, {4, 2, 3}, {0, 4, 3}};
  for (int k = 0; k < 4; ++k) {
    const int* ik = IK[k];
    double x21 = xkd[ik[1]] - xkd[ik[0]], x31 = xkd[ik[2]] - xkd[ik[0]];
    double y21 = ykd[ik[1]] - ykd[ik[0]], y31 = ykd[ik[2]] - ykd[ik[0]];
    double det = x21 * y31 - x31 * y21;
    if (det <= 1e-12) return 26;
    double a = (x31 * x31 + y31 * y31) / det;
    double b = -(x31 * x21 + y31 * y21) / det;
    double c = (x21 * x21 + y21 * y21) / det;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        int p = ik[i] + 5 * ik[j];
        elmas[p] = elmas[p] + (a * (double)c_s1[i + 3 * j] + b * (double)c_s2[i + 3 * j] +
                               c * (double)c_s3[i + 3 * j]) / 2.0;
        elmam[p] = elmam[p] + det * (double)c_s4[i + 3 * j] / 2.4e1;
      }
  }
  for (int i = 0; i < 25; ++i) elmas[i] = elmas[i] + kw * kw * elmam[i];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) kel[i + 4 * j] = elmas[i + 5 * j] - elmas[i + 20] * elmas[4 + 5 * j] / elmas[24];
  return 0;
}

__global__ void crt_precal_area_kernel(PrecalArgs P) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P.elanz * P.K) return;
  const int e = idx % P.elanz, k = idx / P.elanz;
  const int ty = P.el_type[e];
  const int nk = (ty == 3) ? 3 : 4;
  double xk[4], yk[4];
  for (int q = 0; q < nk; ++q) {
    int node = P.el_nodes[e * 4 + q];
    xk[q] = P.nx[node];
    yk[q] = P.ny[node];
  }
  const double kw = P.kwn[k];
  double* out = P.elbg + ((size_t)k * P.elanz + e) * P.nimn;
  if (ty == 8) {
    double kel[16];
    int err = elem8_dev(xk, yk, kw, kel);
    if (err) { atomicCAS(P.err, 0, err); return; }
    int imn = 0;
    for (int m = 0; m < 4; ++m)
      for (int n = 0; n <= m; ++n) out[imn++] = kel[m + 4 * n];
  } else if (ty == 3) {
    double x21 = xk[1] - xk[0], x31 = xk[2] - xk[0];
    double y21 = yk[1] - yk[0], y31 = yk[2] - yk[0];
    double det = x21 * y31 - x31 * y21;
    if (det <= 0.0) { atomicCAS(P.err, 0, 26); return; }
    double a = (x31 * x31 + y31 * y31) / det;
    double b = -(x31 * x21 + y31 * y21) / det;
    double c = (x21 * x21 + y21 * y21) / det;
    int imn = 0;
    for (int m = 0; m < 3; ++m)
      for (int n = 0; n <= m; ++n) {
        double es = (a * (double)c_s1[m + 3 * n] + b * (double)c_s2[m + 3 * n] + c * (double)c_s3[m + 3 * n]) / 2.0;
        double em = det * (double)c_s4[m + 3 * n] / 2.4e1;
        out[imn++] = es + em * kw * kw;
      }
  } else if (ty == 5) {
    const int s1[16] = {2, -2, -1, 1, -2, 2, 1, -1, -1, 1, 2, -2, 1, -1, -2, 2};
    const int s2[16] = {1, 0, -1, 0, 0, -1, 0, 1, -1, 0, 1, 0, 0, 1, 0, -1};
    const int s3[16] = {2, 1, -1, -2, 1, 2, -2, -1, -1, -2, 2, 1, -2, -1, 1, 2};
    const int s4[16] = {4, 2, 1, 2, 2, 4, 2, 1, 1, 2, 4, 2, 2, 1, 2, 4};
    double x21 = xk[1] - xk[0], x31 = xk[3] - xk[0];
    double y21 = yk[1] - yk[0], y31 = yk[3] - yk[0];
    double det = x21 * y31 - x31 * y21;
    if (det <= 0.0) { atomicCAS(P.err, 0, 26); return; }
    double a = (x31 * x31 + y31 * y31) / det;
    double b = -(x31 * x21 + y31 * y21) / det;
    double c = (x21 * x21 + y21 * y21) / det;
    int imn = 0;
    for (int m = 0; m < 4; ++m)
      for (int n = 0; n <= m; ++n) {
        double es = (a * (double)s1[m + 4 * n] + 3.0 * b * (double)s2[m + 4 * n] + c * (double)s3[m + 4 * n]) / 6.0;
        double em = det * (double)s4[m + 4 * n] / 3.6e1;
        out[imn++] = es + em * kw * kw;
      }
  } else {
    atomicCAS(P.err, 0, 18);
  }
}

// Bessel K0/K1 via the Abramowitz & Stegun 9.8.1-9.8.8 polynomials with the
// reference's coefficients (src/bessk0.f90:11-24, bessk1.f90:10-23,
// bessi0.f90, bessi1.f90); K = 0 above x = 500.
__device__ double bessi0_dev(double x) {
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
__device__ double bessi1_dev(double x) {
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
__device__ double bessk0_dev(double x) {
  if (x <= 2.0) {
    double y = x * x / 4.0;
    return (-log(x / 2.0) * bessi0_dev(x)) + (-0.57721566 + y * (0.42278420 + y * (0.23069756 +
           y * (0.3488590e-1 + y * (0.262698e-2 + y * (0.10750e-3 + y * 0.74e-5))))));
  } else if (x > 5e2) {
    return 0.0;
  }
  double y = 2.0 / x;
  return (exp(-x) / sqrt(x)) * (1.25331414 + y * (-0.7832358e-1 + y * (0.2189568e-1 +
         y * (-0.1062446e-1 + y * (0.587872e-2 + y * (-0.251540e-2 + y * 0.53208e-3))))));
}
__device__ double bessk1_dev(double x) {
  if (x <= 2.0) {
    double y = x * x / 4.0;
    return (log(x / 2.0) * bessi1_dev(x)) + (1.0 / x) * (1.0 + y * (0.15443144 + y * (-0.67278579 +
           y * (-0.18156897 + y * (-0.1919402e-1 + y * (-0.110404e-2 + y * -0.4686e-4))))));
  } else if (x > 5e2) {
    return 0.0;
  }
  double y = 2.0 / x;
  return (exp(-x) / sqrt(x)) * (1.25331414 + y * (0.23498619 + y * (-0.3655620e-1 +
         y * (0.1504268e-1 + y * (-0.780353e-2 + y * (0.325614e-2 + y * -0.68245e-3))))));
}

// Mixed-boundary factor kg[k][l][rel] = beta(l, k) (src/beta.f90:38-135) and the
// boundary mass matrix relbg = L/6 [[2,1],[1,2]] (src/elem1.f90:53-64) for
// type-11 boundary elements.  One thread per (boundary element, electrode, k).
__global__ void crt_precal_bnd_kernel(PrecalArgs P) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)P.relanz * P.eanz * P.K;
  if (idx >= total) return;
  const int rel = (int)(idx % P.relanz);
  const int l = (int)((idx / P.relanz) % P.eanz);
  const int k = (int)(idx / ((long)P.relanz * P.eanz));
  double* kgo = P.kg + ((size_t)k * P.eanz + l) * P.relanz + rel;
  if (P.bnd_type[rel] != 11) { *kgo = 0.0; return; }
  const int n1 = P.bnd_nodes[2 * rel], n2 = P.bnd_nodes[2 * rel + 1];
  const double xk0 = P.nx[n1], xk1 = P.nx[n2], yk0 = P.ny[n1], yk1 = P.ny[n2];
  if (l == 0 && k == 0) {
    double x21 = xk1 - xk0, y21 = yk1 - yk0;
    double len = sqrt(x21 * x21 + y21 * y21);
    double* rb = P.relbg + (size_t)rel * P.nimn;
    rb[0] = len * 2.0 / 6.0;
    rb[1] = len * 1.0 / 6.0;
    rb[2] = len * 2.0 / 6.0;
  }
  const int en = P.enr[l];
  const double xs = P.nx[en], ys = P.ny[en];
  const double xr = (xk0 + xk1) / 2.0, yr = (yk0 + yk1) / 2.0;
  const double sytop = P.sytop;
  const double x3 = xr - xs;
  const double y3m = (yr - sytop) - (ys - sytop);
  const double y3p = (yr - sytop) + (ys - sytop);
  const double r1m = sqrt(x3 * x3 + y3m * y3m);
  const double r1p = sqrt(x3 * x3 + y3p * y3p);
  const double x4 = yk0 - yk1, y4 = xk1 - xk0;
  const double r2 = sqrt(x4 * x4 + y4 * y4);
  if (r1m < 1e-12 || r1p < 1e-12 || r2 < 1e-12) { atomicCAS(P.err, 0, 35); *kgo = 0.0; return; }
  const double cthetm = (x3 * x4 + y3m * y4) / r1m / r2;
  const double cthetp = (x3 * x4 + y3p * y4) / r1p / r2;
  const double kw = P.kwn[k];
  const double bk0m = bessk0_dev(r1m * kw), bk0p = bessk0_dev(r1p * kw);
  const double bk1m = bessk1_dev(r1m * kw), bk1p = bessk1_dev(r1p * kw);
  double b;
  if (bk0m + bk0p == 0.0) b = (cthetm + cthetp) / 2.0;
  else b = (cthetm * bk1m + cthetp * bk1p) / (bk0m + bk0p);
  *kgo = kw * b;
}

// ---------------------------------------------------------- back-transform --
// hpot(j,l) = sum_k kpot(j,l,k) * kwnwi(k) / pi  (src/rtrafo.f90:73-88), or
// hpot = kpot (2D, swrtr = 0).  kpot[k][l][n], hpot[l][n].
__global__ void crt_rtrafo_kernel(const double2* __restrict__ kpot, const double* __restrict__ kwnwi,
                                  int n, int eanz, int K, int swrtr, double2* __restrict__ hpot) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * eanz) return;
  if (swrtr == 0) { hpot[idx] = kpot[idx]; return; }
  cplx s = mkc(0.0, 0.0);
  const size_t stride = (size_t)n * eanz;
  for (int k = 0; k < K; ++k) s = cadd(s, rmul(kwnwi[k], ldc(kpot + idx + k * stride)));
  stc(hpot + idx, cdiv(s, mkc(M_PI, 0.0)));
}

// The DC twins assemble with the real part of sigma (kompadc :120-124,
// DBLE(sigma)): out = (Re sigma, 0).
__global__ void crt_real_sigma_kernel(const double2* __restrict__ in, int n, double2* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_double2(in[i].x, 0.0);
}

// Real parts of a complex array (the DC outputs kpotdc, hpotdc, sensdc, whose
// imaginary parts are zero or, for sensdc, dropped as DBLE does in bsendc :184).
__global__ void crt_real_part_kernel(const double2* __restrict__ in, size_t n, double* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i].x;
}

// Electrode slice of the back-transform: he[l][m] = hpot(enr(m), l), the
// sum over wavenumbers in ascending k.  kpot holds wavenumbers k0 .. k0 + kc
// - 1; a forward that holds them one pass at a time carries the partial sums
// in acc between passes (first: start from zero; last: finish into he), with
// the same operations in the same order as one pass over all K.
__global__ void crt_rtrafo_elec_kernel(const double2* __restrict__ kpot, const double* __restrict__ kwnwi,
                                       const int* __restrict__ enr, int n, int eanz, int k0, int kc, int swrtr,
                                       int first, int last, double2* __restrict__ acc, double2* __restrict__ he) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= eanz * eanz) return;
  const int m = idx % eanz, l = idx / eanz;
  const size_t base = (size_t)l * n + enr[m];
  if (swrtr == 0) { he[idx] = kpot[base]; return; }
  cplx s = first ? mkc(0.0, 0.0) : ldc(acc + idx);
  const size_t stride = (size_t)n * eanz;
  for (int k = 0; k < kc; ++k) s = cadd(s, rmul(kwnwi[k0 + k], ldc(kpot + base + k * stride)));
  if (last) stc(he + idx, cdiv(s, mkc(M_PI, 0.0)));
  else stc(acc + idx, s);
}

// Voltage superposition (src/bvolt.f90:54-63) from the electrode slice;
// mode 0: sigmaa = kfak*V (bvolt); mode 1: sigmaa = -ln V (bvolti :88).
__global__ void crt_volt_kernel(const double2* __restrict__ he, int eanz, int nanz,
                                const int* __restrict__ strnr, const int* __restrict__ vnr, int mode,
                                const double* __restrict__ kfak, double2* __restrict__ volt,
                                double2* __restrict__ sigmaa, int* err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nanz) return;
  const int e1 = strnr[i] % 10000, e2 = (strnr[i] - e1) / 10000;
  const int e3 = vnr[i] % 10000, e4 = (vnr[i] - e3) / 10000;
  auto H = [&](int ep, int es) {
    const double msk = (double)(min(ep, 1) * min(es, 1));
    return rmul(msk, ldc(he + (size_t)(max(es, 1) - 1) * eanz + (max(ep, 1) - 1)));
  };
  const cplx d1 = H(e4, e2), d2 = H(e4, e1), d3 = H(e3, e2), d4 = H(e3, e1);
  const cplx v = csub(csub(d1, d2), csub(d3, d4));
  stc(volt + i, v);
  if (mode == 0) {
    stc(sigmaa + i, rmul(kfak ? kfak[i] : 1.0, v));
  } else {
    if (cabs_(v) == 0.0) { atomicCAS(err, 0, 82); stc(sigmaa + i, mkc(0.0, 0.0)); return; }
    cplx lg = clog_(v);
    stc(sigmaa + i, mkc(-lg.re, -lg.im));
  }
}

// ---------------------------------------------------------- sensitivities --
// bsens (mode 0: sens(i,e) = -dV_i/dsigma_e summed the reference's way,
// src/bsens.f90:124-187) and bsensi (mode 1: +dum*sigma_e/V_i,
// src/bsensi.f90:119-184).  One workgroup per area element: the element's
// node potentials for all electrodes at one wavenumber are staged in LDS and
// every thread evaluates h_k for its measurements in the reference's (mi, mj)
// order, accumulating sum_k w_k h_k in registers.
constexpr int kSensMaxQ = 8;  // measurements per thread per pass

// kpot[K][E][n] -> kpt[K][n][E]: the sensitivity kernel reads one element's
// nodes for every electrode, which the transposed copy makes contiguous.
// Tiles of 64 nodes x 16 electrodes through LDS, both sides coalesced.
__global__ __launch_bounds__(256) void crt_kpot_transpose_kernel(const double2* __restrict__ kpot, int n, int E,
                                                                  double2* __restrict__ kpt) {
  __shared__ double2 tile[16][65];
  const int k = blockIdx.z;
  const int n0 = blockIdx.x * 64, a0 = blockIdx.y * 16;
  const int t = threadIdx.x;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int a = a0 + r * 4 + t / 64, node = n0 + (t & 63);
    if (a < E && node < n) tile[r * 4 + t / 64][t & 63] = kpot[((size_t)k * E + a) * n + node];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int node = n0 + r * 16 + t / 16, a = a0 + (t & 15);
    if (a < E && node < n) kpt[((size_t)k * n + node) * E + a] = tile[t & 15][r * 16 + t / 16];
  }
}

// Explicit Jacobian, one workgroup per area element e, threads over the
// measurements i: for every wavenumber k
//   h_k = sum_{a,b} S_ab(e,k) (phi_e2 - phi_e1)(n_b) (phi_e4 - phi_e3)(n_a)
// (bsens :137-167, electrode number 0 contributing zero), evaluated as
// u_a = sum_b S_ab d12_b, h_k = sum_a d34_a u_a, then sum_k w_k h_k / pi
// (:170-184) and the mode's final form: -that (bsens :187, dV/dsigma) or
// that * sigma_e / V_i (bsensi :184, d(-ln V)/d(ln sigma)).  The element's
// potentials of wavenumber k (its <= 4 nodes x all electrodes) and S(e,k)
// are staged in LDS; rows beyond the element's node count are zero.
__global__ __launch_bounds__(256) void crt_sens_kernel(SensArgs S) {
  extern __shared__ double2 sm[];
  const int e = blockIdx.x;
  const int nk = S.el_nk[e];
  const int E = S.eanz;
  cplx* P = reinterpret_cast<cplx*>(sm);  // [4][E]
  __shared__ double sel[4][4];
  __shared__ int nodes[4];
  if (threadIdx.x < 4) nodes[threadIdx.x] = S.el_nodes[e * 4 + threadIdx.x];
  const cplx sig = ldc(S.sigma + e);
  for (int i0 = 0; i0 < S.nanz; i0 += kSensMaxQ * blockDim.x) {
    cplx acc[kSensMaxQ];
    int e1[kSensMaxQ], e2[kSensMaxQ], e3[kSensMaxQ], e4[kSensMaxQ];
#pragma unroll
    for (int q = 0; q < kSensMaxQ; ++q) {
      acc[q] = mkc(0.0, 0.0);
      const int i = min(i0 + (int)threadIdx.x + q * (int)blockDim.x, S.nanz - 1);
      const int st = S.strnr[i], vn = S.vnr[i];
      e1[q] = st % 10000;
      e2[q] = st / 10000;
      e3[q] = vn % 10000;
      e4[q] = vn / 10000;
    }
    // the element's potentials of wavenumber k + 1 are loaded into registers
    // while k is evaluated (kSensPre per thread covers 4 E for E <= 128 at
    // 256 threads); larger E loads the rest when staging
    constexpr int kSensPre = 2;
    cplx pre[kSensPre];
    auto fetch = [&](int k) {
#pragma unroll
      for (int r = 0; r < kSensPre; ++r) {
        const int t = threadIdx.x + r * blockDim.x;
        const int q = t / E, a = t - q * E;
        pre[r] = (t < 4 * E && q < nk) ? ldc(S.kpt + ((size_t)k * S.n + nodes[q]) * E + a) : mkc(0.0, 0.0);
      }
    };
    __syncthreads();  // nodes[]
    fetch(0);
    for (int k = 0; k < S.K; ++k) {
      __syncthreads();
#pragma unroll
      for (int r = 0; r < kSensPre; ++r) {
        const int t = threadIdx.x + r * blockDim.x;
        if (t < 4 * E) P[t] = pre[r];
      }
      for (int t = threadIdx.x + kSensPre * blockDim.x; t < 4 * E; t += blockDim.x) {
        const int q = t / E, a = t - q * E;
        P[t] = q < nk ? ldc(S.kpt + ((size_t)k * S.n + nodes[q]) * E + a) : mkc(0.0, 0.0);
      }
      if (k + 1 < S.K) fetch(k + 1);
      if (threadIdx.x < 16) {
        const int a = threadIdx.x >> 2, b = threadIdx.x & 3;
        const int imax = max(a, b), imin = min(a, b);
        sel[a][b] = (a < nk && b < nk) ? S.elbg[((size_t)k * S.elanz + e) * S.nimn + (imax * (imax + 1)) / 2 + imin]
                                       : 0.0;
      }
      __syncthreads();
      const double wk = S.kwnwi[k];
#pragma unroll
      for (int q = 0; q < kSensMaxQ; ++q) {
        cplx d12[4], d34[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const cplx s1 = csel(e1[q] > 0, P[m * E + max(e1[q] - 1, 0)], mkc(0.0, 0.0));
          const cplx s2 = csel(e2[q] > 0, P[m * E + max(e2[q] - 1, 0)], mkc(0.0, 0.0));
          const cplx s3 = csel(e3[q] > 0, P[m * E + max(e3[q] - 1, 0)], mkc(0.0, 0.0));
          const cplx s4 = csel(e4[q] > 0, P[m * E + max(e4[q] - 1, 0)], mkc(0.0, 0.0));
          d12[m] = csub(s2, s1);
          d34[m] = csub(s4, s3);
        }
        cplx h = mkc(0.0, 0.0);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          cplx u = mkc(0.0, 0.0);
#pragma unroll
          for (int b = 0; b < 4; ++b) u = cadd(u, rmul(sel[a][b], d12[b]));
          h = cadd(h, cmul(d34[a], u));
        }
        acc[q] = S.swrtr == 0 ? h : cadd(acc[q], rmul(wk, h));
      }
    }
#pragma unroll
    for (int q = 0; q < kSensMaxQ; ++q) {
      const int i = i0 + threadIdx.x + q * blockDim.x;
      if (i >= S.nanz) continue;
      cplx out;
      if (S.mode == 0) {
        cplx dum = (S.swrtr == 0) ? acc[q] : rmul(1.0 / M_PI, acc[q]);
        out = mkc(0.0 - dum.re, 0.0 - dum.im);
      } else {
        cplx dum = (S.swrtr == 0) ? acc[q] : cdiv(acc[q], mkc(M_PI, 0.0));
        out = cdiv(cmul(dum, sig), ldc(S.volt + i));
      }
      stc(S.sens + (size_t)e * S.nanz + i, out);
    }
  }
}

// The explicit Jacobian with the per-pair work hoisted: d12 = P(:,e2) -
// P(:,e1) and u_a = sum_b S_ab d12_b depend only on the measurement's current
// pair, d34 = P(:,e4) - P(:,e3) only on its potential pair.  Per element and
// wavenumber, u of every distinct current pair and d34 of every distinct
// potential pair are formed once in LDS (the same operations, in the same
// order, as the per-measurement form of crt_sens_kernel); each measurement then
// takes h_k = sum_a d34_a u_a from two LDS rows.  The measurements are taken in
// (current pair, potential pair) order, interleaved over the threads or in
// contiguous runs per thread (RUNS), and written to their own rows.  LDS layout, complex values: P [4][E + 1] (column
// E stays zero: the potential of an absent electrode), U [ncp][kSensRow] and D
// [npp][kSensRow].  A pair's four values sit at constant offsets from its row
// (immediate offsets of the LDS loads), and the row of kSensRow = 5 entries
// puts 16 consecutive rows on distinct banks.
template <bool RUNS>
__global__ __launch_bounds__(512) void crt_sens_pairs_kernel(SensArgs S, SensPairs Q) {
  extern __shared__ double2 sm[];
  const int e = blockIdx.x;
  const int nk = S.el_nk[e];
  const int E = S.eanz, E1 = E + 1, ncp = Q.ncp, npp = Q.npp;
  cplx* P = reinterpret_cast<cplx*>(sm);  // [4][E + 1]
  cplx* U = P + 4 * E1;                   // [ncp][kSensRow]
  cplx* D = U + kSensRow * ncp;           // [npp][kSensRow]
  __shared__ double sel[4][4];
  __shared__ int nodes[4];
  if (threadIdx.x < 4) nodes[threadIdx.x] = S.el_nodes[e * 4 + threadIdx.x];
  const cplx sig = ldc(S.sigma + e);
  // column of electrode el (1-based, 0 = none) in P
  auto col = [&](int el) { return el > 0 ? el - 1 : E; };
  for (int i0 = 0; i0 < S.nanz; i0 += kSensPairQ * blockDim.x) {
    // this pass's measurements in (current pair, potential pair) order,
    // interleaved over the threads: the lanes of a wavefront take consecutive
    // measurements, so their potential pairs' rows of D are mostly
    // consecutive, their current pairs' rows of U mostly one broadcast, and
    // their stores one run of the sensitivity row.
    // RUNS: a contiguous run of measurements per thread instead, which mostly
    // share their current pair, so that u is read once per change
    const int npass = min(S.nanz - i0, kSensPairQ * (int)blockDim.x);
    const int rp = (npass + (int)blockDim.x - 1) / (int)blockDim.x;
    const int jb = RUNS ? i0 + (int)threadIdx.x * rp : i0 + (int)threadIdx.x;
    const int js = RUNS ? 1 : (int)blockDim.x;  // step between a thread's measurements
    const int nq = RUNS ? max(0, min(rp, i0 + npass - jb)) : (i0 + npass - jb + js - 1) / js;
    cplx acc[kSensPairQ];
    int uo[kSensPairQ], dq[kSensPairQ];
#pragma unroll
    for (int q = 0; q < kSensPairQ; ++q) {
      acc[q] = mkc(0.0, 0.0);
      const int i = Q.ord[min(jb + q * js, S.nanz - 1)];
      const int2 m = Q.mp[i];
      uo[q] = kSensRow * m.x;
      dq[q] = kSensRow * m.y;
    }
    // the element's potentials of wavenumber k + 1 are loaded into registers
    // while k is evaluated (kSensPre per thread covers 4 E for E <= 128 at
    // 256 threads); larger E loads the rest when staging
    constexpr int kSensPre = 2;
    cplx pre[kSensPre];
    auto fetch = [&](int k) {
#pragma unroll
      for (int r = 0; r < kSensPre; ++r) {
        const int t = threadIdx.x + r * blockDim.x;
        const int q = t / E, a = t - q * E;
        pre[r] = (t < 4 * E && q < nk) ? ldc(S.kpt + ((size_t)k * S.n + nodes[q]) * E + a) : mkc(0.0, 0.0);
      }
    };
    __syncthreads();  // nodes[]
    fetch(0);
    if (threadIdx.x < 4) P[threadIdx.x * E1 + E] = mkc(0.0, 0.0);
    for (int k = 0; k < S.K; ++k) {
      __syncthreads();
#pragma unroll
      for (int r = 0; r < kSensPre; ++r) {
        const int t = threadIdx.x + r * blockDim.x;
        const int q = t / E, a = t - q * E;
        if (t < 4 * E) P[q * E1 + a] = pre[r];
      }
      for (int t = threadIdx.x + kSensPre * blockDim.x; t < 4 * E; t += blockDim.x) {
        const int q = t / E, a = t - q * E;
        P[q * E1 + a] = q < nk ? ldc(S.kpt + ((size_t)k * S.n + nodes[q]) * E + a) : mkc(0.0, 0.0);
      }
      if (k + 1 < S.K) fetch(k + 1);
      if (threadIdx.x < 16) {
        const int a = threadIdx.x >> 2, b = threadIdx.x & 3;
        const int imax = max(a, b), imin = min(a, b);
        sel[a][b] = (a < nk && b < nk) ? S.elbg[((size_t)k * S.elanz + e) * S.nimn + (imax * (imax + 1)) / 2 + imin]
                                       : 0.0;
      }
      __syncthreads();
      for (int c = threadIdx.x; c < ncp; c += blockDim.x) {
        const int2 pr = Q.cp[c];
        const int c1 = col(pr.x), c2 = col(pr.y);
        cplx d12[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) d12[m] = csub(P[m * E1 + c2], P[m * E1 + c1]);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          cplx u = mkc(0.0, 0.0);
#pragma unroll
          for (int b = 0; b < 4; ++b) u = cadd(u, rmul(sel[a][b], d12[b]));
          U[c * kSensRow + a] = u;
        }
      }
      for (int p = threadIdx.x; p < npp; p += blockDim.x) {
        const int2 pr = Q.pp[p];
        const int c1 = col(pr.x), c2 = col(pr.y);
#pragma unroll
        for (int m = 0; m < 4; ++m) D[p * kSensRow + m] = csub(P[m * E1 + c2], P[m * E1 + c1]);
      }
      __syncthreads();
      // h of every measurement; without the wavenumber sum (swrtr = 0) the
      // sum is h itself (one branch for the whole loop, uniform)
      const double wk = S.kwnwi[k];
      auto eval = [&](auto accumulate) __attribute__((always_inline)) {
        int cur = -1;
        cplx u[4];
#pragma unroll
        for (int q = 0; q < kSensPairQ; ++q) {
          if (q < nq) {
            if (!RUNS || uo[q] != cur) {
              cur = uo[q];
#pragma unroll
              for (int a = 0; a < 4; ++a) u[a] = U[cur + a];
            }
            const cplx* dr = D + dq[q];
            cplx h = mkc(0.0, 0.0);
#pragma unroll
            for (int a = 0; a < 4; ++a) h = fmacc(h, dr[a], u[a]);
            acc[q] = accumulate(acc[q], h);
          }
        }
      };
      if (S.swrtr == 0)
        eval([](cplx, cplx h) { return h; });
      else
        eval([wk](cplx a, cplx h) { return mkc(fma(wk, h.re, a.re), fma(wk, h.im, a.im)); });
    }
#pragma unroll
    for (int q = 0; q < kSensPairQ; ++q) {
      const int j = jb + q * js;
      if (q >= nq) continue;
      const int i = Q.ord[j];
      cplx out;
      if (S.mode == 0) {
        cplx dum = (S.swrtr == 0) ? acc[q] : rmul(1.0 / M_PI, acc[q]);
        out = mkc(0.0 - dum.re, 0.0 - dum.im);
      } else {
        cplx dum = (S.swrtr == 0) ? acc[q] : cdiv(acc[q], mkc(M_PI, 0.0));
        out = cdiv(cmul(dum, sig), ldc(S.volt + i));
      }
      stc(S.sens + (size_t)e * S.nanz + i, out);
    }
  }
}

template __global__ void crt_sens_pairs_kernel<false>(SensArgs, SensPairs);
template __global__ void crt_sens_pairs_kernel<true>(SensArgs, SensPairs);

// ------------------------------------------------------- adjoint gradient --
// g_e = sum_i conj(J_ie) w_i r_i without forming J.  With J_ie = alpha_i beta_e
// (1/pi) sum_k w_k h_ik(e) (bsens: alpha=-1, beta=1; bsensi: alpha=1/V_i,
// beta=sigma_e) and h_ik(e) = sum_{mi,mj} S_k[mi,mj] v_i(n_mi) u_i(n_mj):
//   g_e = conj(beta_e (1/pi) sum_k w_k sum_{mi,mj} S_k[mi,mj]
//                 sum_b Y_k(n_mi, b) P_k(n_mj, b)),   Y_k = P_k C,
//   C[a][b] = sum_i gamma_i (d_{a,e4} - d_{a,e3}) (d_{b,e2} - d_{b,e1}),
//   gamma_i = alpha_i conj(w_i r_i).
__global__ void crt_adj_gamma_kernel(int nanz, int mode, const double* __restrict__ w,
                                     const double2* __restrict__ resid, const double2* __restrict__ volt,
                                     double2* __restrict__ gamma) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nanz) return;
  cplx r = ldc(resid + i);
  cplx wr = cconj(rmul(w[i], r));
  cplx g = (mode == 0) ? mkc(-wr.re, -wr.im) : cdiv(wr, ldc(volt + i));
  stc(gamma + i, g);
}

// C[a][b] (a = receiver electrode, b = source electrode), one thread per entry,
// summing over measurements in order (deterministic).
__global__ void crt_adj_cmat_kernel(int nanz, int eanz, const int* __restrict__ strnr,
                                    const int* __restrict__ vnr, const int* __restrict__ qptr,
                                    const int* __restrict__ qlist, const double2* __restrict__ gamma,
                                    double2* __restrict__ C) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= eanz * eanz) return;
  const int a = idx / eanz + 1, b = idx % eanz + 1;
  cplx s = mkc(0.0, 0.0);
  // only the measurements with a among their potential electrodes, in
  // ascending order (qlist is sorted per electrode)
  for (int p = qptr[a - 1]; p < qptr[a]; ++p) {
    const int i = qlist[p];
    const int e1 = strnr[i] % 10000, e2 = strnr[i] / 10000;
    const int e3 = vnr[i] % 10000, e4 = vnr[i] / 10000;
    const int ca = (a == e4) - (a == e3);
    const int cb = (b == e2) - (b == e1);
    const int c = ca * cb;
    if (c) s = cadd(s, rmul((double)c, ldc(gamma + i)));
  }
  (void)nanz;
  stc(C + (size_t)(a - 1) * eanz + (b - 1), s);
}

// Y_k[j][b] = sum_a P_k[j][a] C[a][b] for all k (P_k = kpot[k][.][.], n x E with
// leading dimension n).  Output layout Y[k][j][b] (row per node).  Block: 256
// rows j x a 32-column slab of C held in LDS.
constexpr int kAdjCols = 32;
// Y[k] = P[k]^T C on the matrix cores: v_mfma_f64_16x16x4f64, whose operand
// A[i][kk] sits in lane i + 16 kk, B[kk][j] in lane j + 16 kk, and whose
// result D[i][j] sits in lane j + 16 (i mod 4), register i / 4 (lane l,
// register v: row l / 16 + 4 v, column l % 16).  One wavefront owns 16 nodes and
// all electrode columns (up to kAdjTiles tiles of 16), so every potential is
// read once; the complex product takes four real products per 4-wide step
// (re += Pr Cr - Pi Ci, im += Pr Ci + Pi Cr).  C is read through the cache.
constexpr int kAdjTiles = 8;
typedef double v4d __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void crt_adj_gemm_mfma_kernel(const double2* __restrict__ kpot,
                                                              const double2* __restrict__ C, int n, int eanz,
                                                              double2* __restrict__ Y) {
  const int k = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int j0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
  if (j0 >= n) return;
  const int ntile = (eanz + 15) / 16;
  const int col = lane & 15, kk = lane >> 4;
  v4d dr[kAdjTiles], di[kAdjTiles];
#pragma unroll
  for (int t = 0; t < kAdjTiles; ++t) {
    dr[t] = v4d{0.0, 0.0, 0.0, 0.0};
    di[t] = v4d{0.0, 0.0, 0.0, 0.0};
  }
  const double2* Pk = kpot + (size_t)k * eanz * n;
  const int jA = min(j0 + col, n - 1);  // A operand: node j0 + (lane % 16)
  for (int a0 = 0; a0 < eanz; a0 += 4) {
    const int a = a0 + kk;
    const bool aok = a < eanz;
    const double2 p = Pk[(size_t)min(a, eanz - 1) * n + jA];
    const double pr = aok && j0 + col < n ? p.x : 0.0, pi = aok && j0 + col < n ? p.y : 0.0;
#pragma unroll
    for (int t = 0; t < kAdjTiles; ++t) {
      if (t < ntile) {
        const int b = 16 * t + col;
        const double2 c = C[(size_t)min(a, eanz - 1) * eanz + min(b, eanz - 1)];
        const bool ok = aok && b < eanz;
        const double cr = ok ? c.x : 0.0, ci = ok ? c.y : 0.0;
        dr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(pr, cr, dr[t], 0, 0, 0);
        dr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(pi, -ci, dr[t], 0, 0, 0);
        di[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(pr, ci, di[t], 0, 0, 0);
        di[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(pi, cr, di[t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < kAdjTiles; ++t) {
    if (t < ntile) {
      const int b = 16 * t + col;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = j0 + kk + 4 * v;
        if (j < n && b < eanz) Y[((size_t)k * n + j) * eanz + b] = make_double2(dr[t][v], di[t][v]);
      }
    }
  }
}

// The same product with C stationary: wavefront w of the workgroup keeps the
// 16 columns [16 w, 16 w + 16) of C in registers for the ceil(eanz / 4)
// 4-electrode steps (operand B), the workgroup stages 16-node tiles of P[k]
// (every electrode) in LDS, double buffered, and each wavefront forms its
// 16 x 16 block of Y per node tile (operand A from LDS).  A complex step
// takes three real products (Gauss): Pr Cr, Pi Ci and (Pr + Pi)(Cr + Ci) in
// three accumulators, Y = (Pr Cr - Pi Ci) + i ((Pr + Pi)(Cr + Ci) - Pr Cr -
// Pi Ci) at the end; three independent chains share the matrix core.  C is
// read once per workgroup; the next tile's global loads are in flight during
// the products.
template <int NS>
__global__ __launch_bounds__(512) void crt_adj_gemm_cs_kernel(const double2* __restrict__ kpot,
                                                            const double2* __restrict__ C, int n, int eanz,
                                                            double2* __restrict__ Y) {
  extern __shared__ double2 ptile[];  // 2 x [4 NS electrodes][16 nodes]
  constexpr int NA = 4 * NS, TILE = NA * 16, MAXSTG = 4;
  const int k = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int col = lane & 15, kk = lane >> 4;
  const int b = 16 * w + col;
  const int nsr = (eanz + 3) / 4, tile = 64 * nsr;  // steps and staged values in use
  double cr[NS], ci[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int a = 4 * s + kk;
    const bool ok = a < eanz && b < eanz;
    const double2 c = C[(size_t)min(a, eanz - 1) * eanz + min(b, eanz - 1)];
    cr[s] = ok ? c.x : 0.0;
    ci[s] = ok ? c.y : 0.0;
  }
  const double2* Pk = kpot + (size_t)k * eanz * n;
  const int ntiles = (n + 15) / 16;
  double2 st[MAXSTG];
  auto fetch = [&](int t) __attribute__((always_inline)) {
    const int j0 = 16 * t;
#pragma unroll
    for (int r = 0; r < MAXSTG; ++r) {
      const int x = threadIdx.x + r * blockDim.x, a = x >> 4, j = j0 + (x & 15);
      st[r] = (x < tile && a < eanz && j < n) ? Pk[(size_t)a * n + j] : make_double2(0.0, 0.0);
    }
  };
  auto put = [&](double2* buf) __attribute__((always_inline)) {
#pragma unroll
    for (int r = 0; r < MAXSTG; ++r) {
      const int x = threadIdx.x + r * blockDim.x;
      if (x < tile) buf[x] = st[r];
    }
  };
  int t = blockIdx.x;
  if (t < ntiles) {
    fetch(t);
    put(ptile);
  }
  __syncthreads();
  for (int it = 0; t < ntiles; t += gridDim.x, ++it) {
    const double2* cur = ptile + (it & 1) * TILE;
    double2* nxt = ptile + ((it + 1) & 1) * TILE;
    const int tn = t + gridDim.x;
    if (tn < ntiles) fetch(tn);
    v4d rr = v4d{0.0, 0.0, 0.0, 0.0}, ii = rr, ss = rr;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      if (s < nsr) {
        const double2 p = cur[(4 * s + kk) * 16 + col];
        rr = __builtin_amdgcn_mfma_f64_16x16x4f64(p.x, cr[s], rr, 0, 0, 0);
        ii = __builtin_amdgcn_mfma_f64_16x16x4f64(p.y, ci[s], ii, 0, 0, 0);
        ss = __builtin_amdgcn_mfma_f64_16x16x4f64(p.x + p.y, cr[s] + ci[s], ss, 0, 0, 0);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int j = 16 * t + kk + 4 * v;
      if (j < n && b < eanz) Y[((size_t)k * n + j) * eanz + b] = make_double2(rr[v] - ii[v], ss[v] - rr[v] - ii[v]);
    }
    if (tn < ntiles) put(nxt);
    __syncthreads();
  }
}
template __global__ void crt_adj_gemm_cs_kernel<16>(const double2*, const double2*, int, int, double2*);
template __global__ void crt_adj_gemm_cs_kernel<24>(const double2*, const double2*, int, int, double2*);
template __global__ void crt_adj_gemm_cs_kernel<32>(const double2*, const double2*, int, int, double2*);

// Gram form of the Jacobian (bsens/bsensi, src/bsens.f90:137-187): for the
// element e and wavenumber k, with P_k the potentials of every source
// electrode at the element's nodes (4 x E) and S_k the element matrix
// (elbg), every measurement's term is a double difference of one E x E
// matrix, h = G[p2][c2] - G[p2][c1] - G[p1][c2] + G[p1][c1] with G = P_k^T
// S_k P_k (the reference's sum over node pairs of elbg (sup2 - sup1)
// (sup4 - sup3), expanded; an absent electrode's terms are 0).  G is
// symmetric, so its lower triangle of 16 x 16 tiles holds it; the tiles
// accumulate sum_k w_k G_k as rank-4 products on the FP64 matrix cores
// (A = w_k S_k P_k, B = P_k: one v_mfma_f64_16x16x4f64 per real product,
// three or four per complex tile and wavenumber), one wavefront per tile set.  Each
// wavenumber's P_k and w_k S_k P_k are staged in LDS by one thread per
// electrode (double-buffered, one barrier per wavenumber); at the end, per
// pass of kGramQC measurements per thread, the tiles go to LDS, real then
// imaginary part, and every thread reads its measurements' four terms at
// offsets the host computed (a zero slot for absent electrodes).  MT: tiles
// per wavefront, at least a quarter of the formed tiles (at most TB (TB + 1)
// / 2 for TB = ceil(E / 16) tile rows, E <= 16 TB <= 128).
template <int MT, bool REAL>
__global__ __launch_bounds__(kGramNT, MT > 6 ? 2 : 3) void crt_sens_gram_kernel(SensArgs S,
                                                                           const uint2* __restrict__ off,
                                                                           const int* __restrict__ tiles,
                                                                           int nt) {
  constexpr int NT = kGramNT, NW = NT / 64;
  extern __shared__ double2 sm[];
  // XCD-major element order: the dispatcher deals blocks to the 8 XCDs in
  // turn, so block b takes element (b mod 8) grid / 8 + b / 8 and each XCD
  // works through one contiguous element range, whose elements share nodes
  // and read them through the same L2 (grid: elanz rounded up to 8)
  const int e = (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
  if (e >= S.elanz) return;
  const int E = S.eanz, TB = (E + 15) / 16, EP = 16 * TB, ntl = TB * (TB + 1) / 2;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, col = lane & 15, kk = lane >> 4;
  const int nk = S.el_nk[e];
  // this wavefront's tiles tiles[w + NW q] of the lower triangle (row ti,
  // column tj, x = ti (ti + 1) / 2 + tj; only the tiles that some
  // measurement reads): operand offsets in a staging buffer of P rows
  // [4][EP] followed by the A rows [4][EP]
  int oa[MT], ob[MT], xq[MT];
  const int nq = nt > w ? (nt - w + NW - 1) / NW : 0;
#pragma unroll
  for (int q = 0; q < MT; ++q) {
    const int x = tiles[min(w + NW * q, nt - 1)];
    xq[q] = x;
    int ti = 0;
    while ((ti + 1) * (ti + 2) / 2 <= x) ++ti;
    const int tj = x - ti * (ti + 1) / 2;
    oa[q] = 4 * EP + kk * EP + 16 * ti + col;
    ob[q] = kk * EP + 16 * tj + col;
  }
  // complex tile products: up to five tiles per wavefront Gauss's three real
  // products (Pr Qr, Pi Qi, (Pr + Pi)(Qr + Qi)) in three accumulators,
  // which the registers hold at three workgroups per CU; more tiles the four
  // products of the plain form in two
  constexpr bool kGauss = MT <= 5 && !REAL;
  v4d gr[MT], gi[MT], gs[kGauss ? MT : 1];
#pragma unroll
  for (int q = 0; q < MT; ++q) gr[q] = gi[q] = gs[kGauss ? q : 0] = v4d{0.0, 0.0, 0.0, 0.0};
  // without the wavenumber sum (swrtr = 0) the reference keeps one
  // wavenumber's h (the last one, as the pair form)
  const int k0 = S.swrtr == 0 ? S.K - 1 : 0;
  int node[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) node[b] = S.el_nodes[e * 4 + b];
  cplx pc[4];
  auto fetch = [&](int k) __attribute__((always_inline)) {
#pragma unroll
    for (int b = 0; b < 4; ++b)
      pc[b] = (tid < E && b < nk) ? ldc(S.kpt + ((size_t)k * S.n + node[b]) * E + tid) : mkc(0.0, 0.0);
  };
  if (tid < EP) fetch(k0);
  for (int k = k0; k < S.K; ++k) {
    double2* buf = sm + (k & 1) * 8 * EP;
    if (tid < EP) {
      const double wk = S.swrtr == 0 ? 1.0 : S.kwnwi[k];
      const double* el = S.elbg + ((size_t)k * S.elanz + e) * S.nimn;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        cplx s = mkc(0.0, 0.0);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int imax = a > b ? a : b, imin = a > b ? b : a;
          const double sab = (a < nk && b < nk) ? el[(imax * (imax + 1)) / 2 + imin] : 0.0;
          s = mkc(fma(sab, pc[b].re, s.re), fma(sab, pc[b].im, s.im));
        }
        buf[a * EP + tid] = make_double2(pc[a].re, pc[a].im);
        buf[4 * EP + a * EP + tid] = make_double2(wk * s.re, wk * s.im);
      }
      if (k + 1 < S.K) fetch(k + 1);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MT; ++q) {
      if (q < nq) {
        const double2 a = buf[oa[q]], b = buf[ob[q]];
        if constexpr (REAL) {
          // the DC twin's potentials are real: G is real
          gr[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, gr[q], 0, 0, 0);
        } else if constexpr (kGauss) {
          gr[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, gr[q], 0, 0, 0);
          gi[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.y, gi[q], 0, 0, 0);
          gs[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x + a.y, b.x + b.y, gs[q], 0, 0, 0);
        } else {
          gr[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, gr[q], 0, 0, 0);
          gr[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.y, gr[q], 0, 0, 0);
          gi[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, gi[q], 0, 0, 0);
          gi[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, gi[q], 0, 0, 0);
        }
      }
    }
  }
  // every wavefront's last staging reads are done before the tiles overwrite them
  __syncthreads();
  if constexpr (kGauss) {
    // re = Pr Qr - Pi Qi, im = (Pr + Pi)(Qr + Qi) - Pr Qr - Pi Qi
#pragma unroll
    for (int q = 0; q < MT; ++q) {
      gs[q] = gs[q] - gr[q] - gi[q];
      gr[q] = gr[q] - gi[q];
      gi[q] = gs[q];
    }
  }
  double* G = reinterpret_cast<double*>(sm);  // [ntl][16][16], then the zero slot
  if (tid == 0) G[ntl * 256] = 0.0;
  auto put = [&](const v4d (&t)[MT]) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < MT; ++q)
      if (q < nq) {
        double* g = G + xq[q] * 256 + kk * 16 + col;
#pragma unroll
        for (int v = 0; v < 4; ++v) g[64 * v] = t[q][v];
      }
  };
  // h from its four terms' offsets (crt_gram_offset, two 16-bit offsets per word)
  auto hsum = [&](uint2 o) __attribute__((always_inline)) {
    return (G[o.x & 0xffff] - G[o.x >> 16]) - (G[o.y & 0xffff] - G[o.y >> 16]);
  };
  const cplx sig = ldc(S.sigma + e);
  double2* out = S.sens + (size_t)e * S.nanz;
  // kGramQC measurements per thread per pass: the real parts of the pass,
  // then the imaginary parts (offsets read again: fewer registers held)
  for (int i0 = 0; i0 < S.nanz; i0 += kGramQC * NT) {
    double hr[kGramQC];
    put(gr);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kGramQC; ++j) hr[j] = hsum(off[min(i0 + j * NT + tid, S.nanz - 1)]);
    if constexpr (!REAL) {
      __syncthreads();
      put(gi);
      __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < kGramQC; ++j) {
      const int i = i0 + j * NT + tid;
      if (i >= S.nanz) continue;
      const cplx acc = mkc(hr[j], REAL ? 0.0 : hsum(off[i]));
      cplx r;
      if (S.mode == 0) {
        cplx dum = (S.swrtr == 0) ? acc : rmul(1.0 / M_PI, acc);
        r = mkc(0.0 - dum.re, 0.0 - dum.im);
      } else {
        cplx dum = (S.swrtr == 0) ? acc : cdiv(acc, mkc(M_PI, 0.0));
        r = cdiv(cmul(dum, sig), ldc(S.volt + i));
      }
      stc(out + i, r);
    }
    __syncthreads();
  }
}
#define GRAM(MT, R) template __global__ void crt_sens_gram_kernel<MT, R>(SensArgs, const uint2*, const int*, int);
GRAM(3, false)
GRAM(5, false)
GRAM(6, false)
GRAM(9, false)
GRAM(3, true)
GRAM(6, true)
GRAM(9, true)
#undef GRAM

__global__ __launch_bounds__(256) void crt_adj_gemm_kernel(const double2* __restrict__ kpot,
                                                         const double2* __restrict__ C, int n, int eanz,
                                                         double2* __restrict__ Y) {
  extern __shared__ double2 csm[];
  const int k = blockIdx.z;
  const int b0 = blockIdx.y * kAdjCols;
  const int nb = min(kAdjCols, eanz - b0);
  for (int t = threadIdx.x; t < eanz * kAdjCols; t += blockDim.x) {
    const int a = t / kAdjCols, bb = t - a * kAdjCols;
    csm[t] = (bb < nb) ? C[(size_t)a * eanz + b0 + bb] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double ar[kAdjCols], ai[kAdjCols];
#pragma unroll
  for (int bb = 0; bb < kAdjCols; ++bb) { ar[bb] = 0.0; ai[bb] = 0.0; }
  const double2* Pk = kpot + (size_t)k * eanz * n;
  for (int a = 0; a < eanz; ++a) {
    const double2 p = Pk[(size_t)a * n + j];
#pragma unroll
    for (int bb = 0; bb < kAdjCols; ++bb) {
      const double2 c = csm[a * kAdjCols + bb];
      ar[bb] += p.x * c.x - p.y * c.y;
      ai[bb] += p.x * c.y + p.y * c.x;
    }
  }
  double2* out = Y + ((size_t)k * n + j) * eanz + b0;
  for (int bb = 0; bb < nb; ++bb) out[bb] = make_double2(ar[bb], ai[bb]);
}

// Element contraction: half a wavefront (32 lanes) per element, lanes over
// source electrodes b; each lane accumulates sum_k w_k sum_{mi,mj} S Y P for
// its b, then a reduction over the 32 lanes.  Blocks take elements
// XCD-major: the blocks the dispatcher places on one XCD (block index mod 8)
// take one contiguous element range, so neighbouring elements, which share
// nodes, read them through the same L2.  The grid is a multiple of 8 blocks.
// The wavenumbers of a pass (A.k0, A.kc) add to the sums carried in A.acc;
// the last pass forms the gradient.
constexpr int kAdjElemPerBlock = 8;
__global__ __launch_bounds__(256) void crt_adj_elem_kernel(AdjArgs A) {
  constexpr int kXcd = 8;
  const int per = gridDim.x / kXcd;
  const int blk = (blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
  const int hl = threadIdx.x & 31;
  const int e0 = blk * kAdjElemPerBlock + (threadIdx.x >> 5);
  const int e = min(e0, A.elanz - 1);
  const int nk = A.el_nk[e];
  int nodes[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) nodes[q] = A.el_nodes[e * 4 + min(q, nk - 1)];
  const int E = A.eanz;
  double accr = 0.0, acci = 0.0;
  for (int kr = 0; kr < A.kc; ++kr) {
    const int k = A.k0 + kr;  // wavenumber; kr its index in kpt and Y
    const double* sel = A.elbg + ((size_t)k * A.elanz + e) * A.nimn;
    double sm[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int imax = max(a, b), imin = min(a, b);
        sm[a][b] = (a < nk && b < nk) ? sel[(imax * (imax + 1)) / 2 + imin] : 0.0;
      }
    const double wk = (A.swrtr == 0) ? 1.0 : A.kwnwi[k];
    // Y and the transposed potentials are both [K][n][E]: every load below
    // reads a contiguous run of electrodes
    const double2* Yk = A.Y + (size_t)kr * A.n * E;
    const double2* Pk = A.kpt + (size_t)kr * A.n * E;
    for (int b = hl; b < E; b += 32) {
      cplx yv[4], pv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        yv[q] = ldc(Yk + (size_t)nodes[q] * E + b);
        pv[q] = ldc(Pk + (size_t)nodes[q] * E + b);
      }
      // sum_{a,b} S_ab Y_a P_b as sum_a Y_a (S P)_a
      cplx t = mkc(0.0, 0.0);
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        cplx u = mkc(0.0, 0.0);
#pragma unroll
        for (int c = 0; c < 4; ++c) u = cadd(u, rmul(sm[a][c], pv[c]));
        t = cadd(t, cmul(yv[a], u));
      }
      accr += wk * t.re;
      acci += wk * t.im;
    }
  }
  for (int off = 16; off > 0; off >>= 1) {
    accr += __shfl_xor(accr, off, 32);
    acci += __shfl_xor(acci, off, 32);
  }
  if (hl == 0 && e0 < A.elanz) {
    cplx s = mkc(accr, acci);
    if (!A.first) s = cadd(ldc(A.acc + e), s);
    if (!A.last) {
      stc(A.acc + e, s);
      return;
    }
    if (A.swrtr != 0) s = rmul(1.0 / M_PI, s);
    const cplx beta = (A.mode == 0) ? mkc(1.0, 0.0) : ldc(A.sigma + e);
    stc(A.grad + e, cconj(cmul(beta, s)));
  }
}

// Data term of the model update from the explicit Jacobian:
// g_j = sum_i conj(J_ij) w_i r_i (src/update.f90:263-266 without cgfac).
// One wavefront per column j.
__global__ __launch_bounds__(256) void crt_jhwr_kernel(const double2* __restrict__ J, int nanz, int manz,
                                                      const double* __restrict__ w,
                                                      const double2* __restrict__ r, double2* __restrict__ g) {
  const int j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (j >= manz) return;
  const double2* col = J + (size_t)j * nanz;
  double sr = 0.0, si = 0.0;
  for (int i = lane; i < nanz; i += 64) {
    const cplx s = cconj(ldc(col + i));
    const cplx a = rmul(w[i], ldc(r + i));
    sr += s.re * a.re - s.im * a.im;
    si += s.re * a.im + s.im * a.re;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off, 64);
    si += __shfl_xor(si, off, 64);
  }
  if (lane == 0) stc(g + j, mkc(sr, si));
}

// ------------------------------------------------------- CG products ------
// Preconditioner cgfac_j = 1/sqrt(sum_i |J_ij|^2 w_i + lam*rdiag_j) and the
// scaled RHS (src/update.f90:179-293).  One wavefront per column j (J is
// column major N x M).  Reductions use a fixed lane order (deterministic).
__global__ __launch_bounds__(256) void crt_cg_rhs_kernel(CgArgs G, const double2* __restrict__ dat,
                                                        const double2* __restrict__ sigmaa,
                                                        const double2* __restrict__ rm) {
  const int j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (j >= G.manz) return;
  const double2* col = G.J + (size_t)j * G.nanz;
  double d = 0.0, br = 0.0, bi = 0.0;
  for (int i = lane; i < G.nanz; i += 64) {
    const cplx s = ldc(col + i);
    const double w = G.w[i];
    d += (s.re * s.re + s.im * s.im) * w;
    const cplx r = csub(ldc(dat + i), ldc(sigmaa + i));
    const cplx t = cmul(rmul(w, cconj(s)), r);
    br += t.re;
    bi += t.im;
  }
  for (int off = 32; off > 0; off >>= 1) {
    d += __shfl_xor(d, off, 64);
    br += __shfl_xor(br, off, 64);
    bi += __shfl_xor(bi, off, 64);
  }
  if (lane == 0) {
    const double cg = 1.0 / sqrt(d + G.lam * G.rdiag[j]);
    G.cgfac[j] = cg;
    cplx b = mkc(br, bi);
    if (rm && !(G.ltri == 3 || G.ltri == 4)) b = csub(b, rmul(G.lam, ldc(rm + j)));
    stc(G.bvec + j, rmul(cg, b));
  }
}

// ap_i = sum_j p_j J_ij cgfac_j (bap, src/cg_mod.f90:520-545).  Thread per
// row i, coalesced over i for each column.
// in two stages: slice q of the columns (chunk consecutive j, ascending) for
// rows i in stage 1, the slices added in order per row in stage 2, so that a
// small nanz still fills the device
__global__ void crt_cg_bap_kernel(CgArgs G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (i >= G.nanz) return;
  const int j0 = q * G.chunk, j1 = min(G.manz, j0 + G.chunk);
  double sr = 0.0, si = 0.0;
  for (int j = j0; j < j1; ++j) {
    const cplx p = rmul(G.cgfac[j], ldc(G.pvec + j));
    const cplx s = ldc(G.J + (size_t)j * G.nanz + i);
    sr += p.re * s.re - p.im * s.im;
    si += p.re * s.im + p.im * s.re;
  }
  stc(G.part + (size_t)q * G.nanz + i, mkc(sr, si));
}

__global__ void crt_cg_bap_sum_kernel(CgArgs G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.nanz) return;
  double sr = 0.0, si = 0.0;
  for (int q = 0; q < G.nslice; ++q) {
    const cplx v = ldc(G.part + (size_t)q * G.nanz + i);
    sr += v.re;
    si += v.im;
  }
  stc(G.ap + i, mkc(sr, si));
}

// bvec = R (cgfac o p) (bp / bptri / bplma, src/cg_mod.f90:547-640), then
// bvec_j = cgfac_j (sum_i conj(J_ij) w_i ap_i + lam bvec_j) (bb, :699-726).
__global__ __launch_bounds__(256) void crt_cg_bb_kernel(CgArgs G) {
  const int j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (j >= G.manz) return;
  const double2* col = G.J + (size_t)j * G.nanz;
  double sr = 0.0, si = 0.0;
  for (int i = lane; i < G.nanz; i += 64) {
    const cplx s = cconj(ldc(col + i));
    const cplx a = rmul(G.w[i], ldc(G.ap + i));
    sr += s.re * a.re - s.im * a.im;
    si += s.re * a.im + s.im * a.re;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off, 64);
    si += __shfl_xor(si, off, 64);
  }
  if (lane != 0) return;
  // R * (cgfac o p)
  cplx rp = mkc(0.0, 0.0);
  if (G.ltri == 0) {
    const int M = G.manz, nx = G.nx;
    const double* sm = G.smatm;
    if (j > 0) rp = rmul(sm[(j - 1) + M] * G.cgfac[j - 1], ldc(G.pvec + j - 1));
    if (j < M - 1) rp = cadd(rp, rmul(sm[j + M] * G.cgfac[j + 1], ldc(G.pvec + j + 1)));
    if (j + 1 > nx) rp = cadd(rp, rmul(sm[(j - nx) + 2 * M] * G.cgfac[j - nx], ldc(G.pvec + j - nx)));
    if (j + 1 < M - nx + 1) rp = cadd(rp, rmul(sm[j + 2 * M] * G.cgfac[j + nx], ldc(G.pvec + j + nx)));
    rp = cadd(rp, rmul(sm[j] * G.cgfac[j], ldc(G.pvec + j)));
  } else if (G.ltri == 3 || G.ltri == 4) {
    rp = rmul(G.smatm[j], rmul(G.cgfac[j], ldc(G.pvec + j)));
  } else {
    const int M = G.manz;
    for (int q = 0; q < G.maxnodes; ++q) {
      const int nb = G.nachbar[j + (size_t)M * q];
      if (nb != 0) rp = cadd(rp, rmul(G.cgfac[nb - 1], rmul(G.smatm[j + (size_t)M * q], ldc(G.pvec + nb - 1))));
    }
    rp = cadd(rp, rmul(G.cgfac[j], rmul(G.smatm[j + (size_t)M * G.maxnodes], ldc(G.pvec + j))));
  }
  cplx b = cadd(mkc(sr, si), rmul(G.lam, rp));
  stc(G.bvec + j, rmul(G.cgfac[j], b));
}

// Dot products of the CG recurrence and the vector updates, on one workgroup
// so the scalars never leave the device (the loop can be graph-captured).
// stage 0: dr = sum |r|^2; beta; p = r + beta p
// stage 1: dr1 = Re sum conj(p) b; alpha = dr/dr1; dpar += alpha p; r -= alpha b
__global__ __launch_bounds__(1024) void crt_cg_scalar_kernel(CgArgs G, int stage, int iter) {
  __shared__ double red[1024];
  const int tid = threadIdx.x;
  double s = 0.0;
  if (stage == 0) {
    for (int j = tid; j < G.manz; j += blockDim.x) {
      const cplx r = ldc(G.rvec + j);
      s += r.re * r.re + r.im * r.im;
    }
  } else {
    for (int j = tid; j < G.manz; j += blockDim.x) {
      const cplx p = ldc(G.pvec + j), b = ldc(G.bvec + j);
      s += p.re * b.re + p.im * b.im;
    }
  }
  red[tid] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  const double tot = red[0];
  double* sc = G.scal;  // [0]=dr [1]=dr0 [2]=dr1 [3]=alpha [4]=beta [5]=done [6]=ncg
  if (stage == 0) {
    if (sc[5] != 0.0) return;  // uniform: every thread reads the same flag
    const double dr0 = (iter == 0) ? tot * G.eps : sc[1];
    if (iter > 0 && tot <= dr0) {
      __syncthreads();
      if (tid == 0) { sc[5] = 1.0; sc[6] = (double)iter; }
      return;
    }
    const double beta = (iter == 0) ? 0.0 : tot / sc[2];
    __syncthreads();
    if (tid == 0) { sc[0] = tot; sc[1] = dr0; sc[4] = beta; }
    for (int j = tid; j < G.manz; j += blockDim.x) {
      const cplx p = ldc(G.pvec + j), r = ldc(G.rvec + j);
      stc(G.pvec + j, cadd(r, rmul(beta, p)));
    }
  } else {
    if (sc[5] != 0.0) return;
    const double dr = sc[0];
    const double alpha = dr / tot;
    __syncthreads();
    for (int j = tid; j < G.manz; j += blockDim.x) {
      const cplx p = ldc(G.pvec + j), b = ldc(G.bvec + j);
      stc(G.dpar + j, cadd(ldc(G.dpar + j), rmul(alpha, p)));
      stc(G.rvec + j, csub(ldc(G.rvec + j), rmul(alpha, b)));
    }
    if (tid == 0) {
      sc[3] = alpha;
      sc[2] = dr;
      G.cgres[iter + 1] = (double)(float)(G.eps * dr / sc[1]);
      sc[6] = (double)(iter + 1);
    }
  }
}
