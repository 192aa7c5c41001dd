// Inversion-loop pieces around the forward/sensitivity path, on the device so
// that a Gauss-Newton iteration needs no host round trip (and can be captured
// in a graph): the data misfit with polarity folding (src/dmisft.f90:17-300),
// bpar (src/bpar.f90), bsigma (src/bsigma.f90), the model update step
// (src/update.f90:332-351) and the row sums for blam0's starting lambda
// (src/blam0.f90).  Reductions run on one workgroup: each thread sums
// its strided share in index order, then a fixed tree over the workgroup, so
// results are deterministic (their summation order is not the reference's
// sequential one; the difference is rounding).
#include "crt_device.h"
#include "crt_inv.h"

namespace {

// chkpo2 (src/dmisft.f90:213-300): the imaginary parts are phases in rad;
// each of datum and response is classed by its phase range, and the pair
// folded by pi as there.  wdlok = 0 marks a polarity mismatch.
__device__ __forceinline__ void chkpo2(cplx dati, cplx sigi, bool ldum, cplx& cdat, cplx& csig, int& wdlok) {
  const double pi = 3.141592653589793;  // dacos(-1d0)
  double imdat = dati.im, imsig = sigi.im;
  const int idat = imdat > pi / 2.0 ? -1 : (imdat <= -pi / 2.0 ? 1 : 0);
  const int isig = imsig > pi / 2.0 ? -1 : (imsig <= -pi / 2.0 ? 1 : 0);
  if (idat == 0 && isig != 0) {
    imsig = imsig + (double)isig * pi;
    if (!ldum) imdat = imdat - copysign(pi, imdat);
    wdlok = 0;
  } else if (idat != 0 && isig == 0) {
    if (ldum) imdat = imdat + (double)idat * pi;
    wdlok = 0;
  } else if (idat != 0 && isig != 0) {
    imsig = imsig + (double)isig * pi;
    imdat = imdat + (double)idat * pi;
  }
  cdat = mkc(dati.re, imdat);
  csig = mkc(sigi.re, imsig);
}

// Sum of one value per thread over the workgroup, fixed tree order; every
// thread receives the result.
__device__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
#pragma unroll 1
  for (int s = kInvThreads / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

}  // namespace

// The misfit of dmisft without its text output: per datum chkpo2, the
// residual cdum = cdat - csig (minus d0 - f(m0) for difference inversion)
// and psi = sqrt(w) |cdum| (FPI: sqrt(w) |Im cdum|); nrmsd, betrms, pharms
// over the data of matching polarity (:85-137).  With lrobust, the estimated
// standard deviations of the robust (L1-like) reweighting and, when their
// norm ratio exceeds l1min, the new weights (:140-191).  Every datum excluded
// leaves npol = nanz (the reference's errnr 99) and the other statistics NaN.
__global__ __launch_bounds__(kInvThreads) void crt_misfit_kernel(MisfitArgs A) {
#pragma clang fp contract(off)
  __shared__ double sh[kInvThreads];
  const int t = threadIdx.x, N = A.nanz;
  const bool lpol = A.flags & kMisfitLpol, lfpi = A.flags & kMisfitLfpi, ldiff = A.flags & kMisfitLdiff;
  double s_id = 0.0, s_rms = 0.0, s_bet = 0.0, s_pha = 0.0;
  for (int i = t; i < N; i += kInvThreads) {
    cplx cdat, csig;
    int wd = 1;
    chkpo2(ldc(A.dat + i), ldc(A.sigmaa + i), lpol, cdat, csig, wd);
    cplx cdum = mkc(cdat.re - csig.re, cdat.im - csig.im);
    if (ldiff) {
      const cplx d0 = ldc(A.d0 + i), f0 = ldc(A.fm0 + i);
      cdum = mkc(cdum.re - (d0.re - f0.re), cdum.im - (d0.im - f0.im));
    }
    const double psi = lfpi ? sqrt(A.wmatd[i]) * fabs(cdum.im) : sqrt(A.wmatd[i]) * hypot(cdum.re, cdum.im);
    A.psi[i] = psi;
    A.wdlok[i] = wd;
    const double w = (double)wd;
    s_id += w;
    s_rms += psi * psi * w;
    s_bet += A.wmatdr[i] * w * cdum.re * cdum.re;
    s_pha += A.wmatdp[i] * w * cdum.im * cdum.im;
  }
  const double idum = block_sum(s_id, sh);
  const double rmssum = block_sum(s_rms, sh);
  const double bet = block_sum(s_bet, sh);
  const double pha = block_sum(s_pha, sh);
  double l1rat = 0.0;
  if (idum > 0.0 && (A.flags & kMisfitLrobust)) {
    // estimated weights eps2 = sqrt(psi)/sqrt(w) and their 1-norms (:150-158)
    double s_n = 0.0, s_n2 = 0.0;
    for (int i = t; i < N; i += kInvThreads) {
      const double psi = A.psi[i], w = (double)A.wdlok[i];
      const double dum = 1.0 / sqrt(A.wmatd[i]);
      const double e = dum * sqrt(psi);
      s_n += psi * w;
      s_n2 += psi * w * dum / e;
    }
    const double norm = block_sum(s_n, sh);
    const double norm2 = block_sum(s_n2, sh);
    // normalised, bounded below by the data's own deviation (:161-172)
    double s_m = 0.0;
    for (int i = t; i < N; i += kInvThreads) {
      const double psi = A.psi[i], w = (double)A.wdlok[i];
      const double dum = 1.0 / sqrt(A.wmatd[i]);
      const double e = fmax(dum, dum * sqrt(psi) * norm2 / norm);
      s_m += psi * w * dum / e;
    }
    l1rat = norm / block_sum(s_m, sh);
    if (l1rat > A.l1min)  // new weights (:176-188)
      for (int i = t; i < N; i += kInvThreads) {
        const double psi = A.psi[i];
        const double dum = 1.0 / sqrt(A.wmatd[i]);
        const double e = fmax(dum, dum * sqrt(psi) * norm2 / norm);
        A.wmatd[i] = 1.0 / (e * e);
      }
  }
  if (t == 0) {
    const double nan = __builtin_nan("");
    A.stats[0] = idum > 0.0 ? sqrt(rmssum / idum) : nan;
    A.stats[1] = idum > 0.0 ? sqrt(bet / idum) : nan;
    A.stats[2] = idum > 0.0 ? sqrt(pha / idum) : nan;
    A.stats[3] = idum > 0.0 ? rmssum : nan;
    A.stats[4] = l1rat;
    A.stats[5] = (double)N - idum;
  }
}

// log and exp of complex conductivities: log|z| + i arg z (the principal
// branch, as CDLOG) and e^re (cos im + i sin im) (CDEXP)
__global__ void crt_bpar_kernel(int manz, const int* last, const double2* sigma, double2* par) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= manz) return;
  const int e = last ? last[j] : j;
  if (e < 0) return;  // a parameter without elements keeps its value
  const cplx s = ldc(sigma + e);
  stc(par + j, mkc(log(hypot(s.re, s.im)), atan2(s.im, s.re)));
}

__global__ void crt_bsigma_kernel(int elanz, const int* mnr, const double2* par, double2* sigma) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= elanz) return;
  const cplx p = ldc(par + (mnr ? mnr[i] - 1 : i));
  const double e = exp(p.re);
  stc(sigma + i, mkc(e * cos(p.im), e * sin(p.im)));
}

// par_j = par_j + (step, 0) dpar_j, rounded as the reference's complex
// product then sum (no fused multiply-add); with lphi0 a negative imaginary
// part (conductivity phase) is set to zero and counted; bdpar = step *
// sum_j Re(dpar_j conj(dpar_j)).
__global__ __launch_bounds__(kInvThreads) void crt_update_step_kernel(int manz, double2* par, const double2* dpar,
                                                                      double step, int lphi0, double* out) {
#pragma clang fp contract(off)
  __shared__ double sh[kInvThreads];
  const int t = threadIdx.x;
  double s_b = 0.0, s_n = 0.0;
  for (int j = t; j < manz; j += kInvThreads) {
    const cplx d = ldc(dpar + j);
    cplx p = ldc(par + j);
    p = mkc(p.re + (step * d.re - 0.0 * d.im), p.im + (step * d.im + 0.0 * d.re));
    if (lphi0 && p.im < 0.0) {
      p.im = 0.0;
      s_n += 1.0;
    }
    stc(par + j, p);
    s_b += d.re * d.re - d.im * (-d.im);
  }
  const double bd = block_sum(s_b, sh);
  const double nf = block_sum(s_n, sh);
  if (t == 0) {
    out[0] = bd * step;
    out[1] = nf;
  }
}

// Row sums of J for blam0's J^H W (J 1): s_i = sum_j J_ij (mode 1: real parts
// only).  J is N x M column major.  Stage 1: slice q of kRowSlices sums the
// columns j = q, q + kRowSlices, ... for rows i (coalesced over i); stage 2
// adds the slices in order (deterministic).
__global__ void crt_rowsum_partial_kernel(const double2* __restrict__ J, int nanz, int manz, int mode,
                                          double2* __restrict__ part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  if (i >= nanz) return;
  double sr = 0.0, si = 0.0;
  for (int j = q; j < manz; j += kRowSlices) {
    const cplx v = ldc(J + (size_t)j * nanz + i);
    sr += v.re;
    si += v.im;
  }
  stc(part + (size_t)q * nanz + i, mkc(sr, mode ? 0.0 : si));
}

__global__ void crt_rowsum_final_kernel(const double2* __restrict__ part, int nanz, double2* __restrict__ s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nanz) return;
  double sr = 0.0, si = 0.0;
  for (int q = 0; q < kRowSlices; ++q) {
    const cplx v = ldc(part + (size_t)q * nanz + i);
    sr += v.re;
    si += v.im;
  }
  stc(s + i, mkc(sr, si));
}
