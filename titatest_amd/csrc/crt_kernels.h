// Kernel argument blocks and launch declarations shared by the HIP sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "crt_device.h"

// Assembly term marking "this row is a Dirichlet row: unit diagonal".
constexpr int kDirichletUnit = -0x7fffffff;

// Arguments of the factor + forward-sweep + back-sweep kernel.  Device
// pointers; all node / element indices here are 0-based.
struct FactorArgs {
  int n, mb, eanz, elanz, relanz, nimn;
  int nsys;                 // systems to solve
  const int* sys_list;      // system ids (k*eanz + l), or null for 0..nsys-1
  // row assembly plan (CSR over rows, terms sorted by band offset then element)
  const int* rowptr;        // n+1
  const int16_t* term_t;    // band offset t = c - r + mb
  const int8_t* term_imn;   // packed element-matrix index
  const int* term_e;        // >= 0 area element, < 0: -(boundary element + 1)
  const double* elbg;       // [K][elanz][nimn]
  const double* relbg;      // [relanz][nimn]
  const double* kg;         // [K][eanz][relanz]
  const double2* sigma;     // elanz
  const int* rnr;           // relanz, adjacent area element
  const int* enr;           // eanz, electrode nodes
  const uint8_t* isdir;     // n Dirichlet flags, or null
  int lsink, nsink;
  // per-slot scratch (slot = workgroup)
  double2* L;               // [slots][n][mb+1]
  double2* y;               // [slots][n]
  double* fak;              // [slots][n]
  // output potentials kpot[k][l][n] (the reference's kpot(sanz,eanz,kwnanz))
  double2* kpot;
  int* err;                 // first error code (0 = ok)
};

template <int NB>
__global__ void crt_factor_kernel(FactorArgs A);

__device__ void crt_backsolve_block(const FactorArgs& A, int sys, int slot, cplx* smem);

// LDS bytes the factor kernel needs for bandwidth mb.
inline size_t crt_factor_lds_bytes(int mb, int nb) {
  size_t W = (size_t)mb + nb;
  size_t fac = (W * (W + 1) / 2 + W) * 16;
  size_t P = 1;
  while (P < (size_t)mb + 16 + 1) P <<= 1;
  size_t back = (P + 16 * ((size_t)mb + 1)) * 16;
  return fac > back ? fac : back;
}
