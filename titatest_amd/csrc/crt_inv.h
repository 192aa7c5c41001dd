// Inversion-loop pieces on the device (gfx950): data misfit with polarity
// folding, model parameters from and to conductivities, model update step.
#pragma once
#include <hip/hip_runtime.h>

constexpr int kInvThreads = 1024;  // one workgroup: fixed-order reductions
constexpr int kRowSlices = 64;     // column slices of the row-sum first stage

// dmisft switches (the reference's lpol, lfpi, ldiff, lrobust)
constexpr int kMisfitLpol = 1, kMisfitLfpi = 2, kMisfitLdiff = 4, kMisfitLrobust = 8;

struct MisfitArgs {
  int nanz, flags;
  double l1min;
  const double2* dat;      // measured -ln V (log magnitude, phase)
  const double2* sigmaa;   // modelled -ln V
  double* wmatd;           // in/out: the robust reweighting updates it
  const double* wmatdr;
  const double* wmatdp;
  const double2* d0;       // ldiff only
  const double2* fm0;      // ldiff only
  double* psi;             // scratch [nanz]
  int* wdlok;              // out [nanz]: 0 for a polarity mismatch
  double* stats;           // out [6]: nrmsd, betrms, pharms, rmssum, l1rat, npol
};

__global__ void crt_misfit_kernel(MisfitArgs A);
// par[j] = log(sigma[last[j]]) (last = the last element of parameter j in
// element order, -1 for none), or par[j] = log(sigma[j]) without a map
__global__ void crt_bpar_kernel(int manz, const int* last, const double2* sigma, double2* par);
// sigma[i] = exp(par[mnr[i] - 1]), or exp(par[i]) without a map
__global__ void crt_bsigma_kernel(int elanz, const int* mnr, const double2* par, double2* sigma);
// par += step dpar, lphi0 rule, out = {bdpar, forced count}
__global__ void crt_update_step_kernel(int manz, double2* par, const double2* dpar, double step, int lphi0,
                                       double* out);
// blam0's row sums s_i = sum_j J_ij (mode 1: real parts), two stages
__global__ void crt_rowsum_partial_kernel(const double2* J, int nanz, int manz, int mode, double2* part);
__global__ void crt_rowsum_final_kernel(const double2* part, int nanz, double2* s);
