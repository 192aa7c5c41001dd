// Kernel argument blocks and launch declarations shared by the HIP sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "crt_device.h"

// Assembly term marking "this row is a Dirichlet row: unit diagonal".
constexpr int kDirichletUnit = -0x7fffffff;
// Stored band entries per matrix row (pattern slots) of the scaled-row records.
constexpr int kMaxR = 8;
// Columns per stage of the factor kernel (panel width).
constexpr int kNB = 8;

// Arguments of the assembly, scaling and factor + forward-sweep kernels.
// Device pointers; node / element indices are 0-based.
struct FactorArgs {
  int n, mb, eanz, elanz, relanz, nimn;
  int nsys;                 // systems in this batch
  int sys_first;            // batch entry s is system (sys_first + s) * sys_stride
  int sys_stride;           // 1, or eanz when one factor serves every source (k*eanz + l)
  // area-term plan (CSR over rows, terms sorted by band offset then element)
  const int* rowptr;        // n+1
  const int16_t* term_t;    // band offset t = c - r + mb
  const int8_t* term_imn;   // packed element-matrix index
  const int* term_e;        // >= 0 area element, kDirichletUnit for Dirichlet rows
  // row pattern: R stored entries per row (slots in ascending t, -1 = unused)
  int R;
  const int16_t* pat_t;     // n x R
  const int8_t* pat_diag;   // n: slot of the diagonal
  // mixed-boundary terms per row (CSR, element order)
  const int* bptr;          // n+1
  const int8_t* bslot;      // pattern slot
  const int* bcode;         // -(boundary element + 1)
  const int8_t* bimn;       // packed boundary-matrix index
  const double* elbg;       // [K][elanz][nimn]
  const double* relbg;      // [relanz][nimn]
  const double* kg;         // [K][eanz][relanz]
  const double2* sigma;     // elanz
  const int* rnr;           // relanz, adjacent area element
  const int* enr;           // eanz, electrode nodes
  const uint8_t* isdir;     // n Dirichlet flags, or null
  int lsink, nsink;
  const double2* av;        // [K][n][R] area sums (crt_area_values_kernel)
  // per-system arrays of the batch (index = system - sys_first)
  double* fak;              // [nsys][n] scaling factors (crt_scale_kernel)
  double2* S;               // [nsys][n][R] scaled row entries (crt_rows_kernel)
  double2* L;               // [nsys][n][mb+1]
  double2* y;               // [nsys][n]
  int dc;                   // DC twins (scaldc, choldc): real sigma, a_ii <= 0 -> 27, pivot <= 0 -> 28
  // prescribed boundary values (randb2), all null when there are none
  const int* mv_ptr;        // n+1: moved terms of each free row (CSR)
  const int* mv_ir;         // position of the prescribed node in the boundary file
  const int* mv_e;          // area element, or -(boundary element + 1)
  const int8_t* mv_imn;     // packed element-matrix index
  const int* rb_last;       // n: last file position of a prescribed node, -1 otherwise
  const double2* rwd;       // file values
  double2* rhs;             // [nsys][n] scaled right-hand sides (crt_rhs_kernel), or null
  double2* ring;            // [grid][ring_stride] trailing-window rings in global memory (wide bands), or null
  size_t ring_stride;       // entries per workgroup's ring: at least half x (mb + 1)
  int* err;                 // first error code (0 = ok) and the failing system
  unsigned long long* prof; // optional: 8 phase-time accumulators (wall clock ticks)
};

// Substitution sweeps over right-hand sides: rhs r of the batch uses factor
// f = r / rhs_per_fac of the batch and is system (fac_first + f) * fstride +
// r % rhs_per_fac (one rhs per factor, or eanz sources sharing one factor).
struct BacksolveArgs {
  int n, mb, eanz, nrhs, rhs_per_fac, fac_first, fstride;
  const double2* L;         // [factor][n][mb+1]
  const double2* y;         // [rhs][n]
  const double* fak;        // [factor][n]
  double2* kpot;            // [K'][eanz][n]: wavenumbers kpot_k0 .. kpot_k0 + K' - 1
  int kpot_k0;
};

// Forward sweep L y = b for every source of a shared factor (vre :45-56).
struct FwdArgs {
  int n, mb, eanz, nrhs, fac_first;
  const double2* Lc;        // [factor][n][mb+1] column-band copy: Lc[j][s] = L_{j+s, j}
  const double* fak;        // [factor][n]
  const int* enr;
  const uint8_t* isdir;
  int lsink, nsink;
  const double2* rhs;       // [factor][n] scaled right-hand side of source 0 (randb2), or null
  double2* y;               // [rhs][n], rhs r = factor r / eanz, source r % eanz
};

__global__ void crt_area_values_kernel(FactorArgs A, int K, double2* av);
__global__ void crt_scale_kernel(FactorArgs A);
template <class T = cplx>
__global__ void crt_rows_kernel(FactorArgs A);
__global__ void crt_rhs_kernel(FactorArgs A);
// The factor kernel on complex (T = cplx) or the DC twin's real (T = double)
// entries (choldc, src/choldc.f90, with vredc's forward sweep): the scaled
// rows, L and y hold entries of T's size.
template <int ROWS, int MAXT, int NT, bool GR, class T>
__global__ void crt_factor_kernel(FactorArgs A);
template <int ROWS, int MAXT, int NT, bool GR>
__global__ void crt_factor_many_kernel(const FactorArgs* args, const int2* items, int nitems, int* counter);
// Substitutions and the column-band copy on complex (T = cplx) or real (T =
// double, the DC twin) factors.  kpot stays complex (imaginary part 0 for DC).
template <int DEPTH, int G, int NS, class T = cplx>
__global__ void crt_backsolve_kernel(BacksolveArgs B);
template <int DEPTH, int G, int NS>
__global__ void crt_backsolve_many_kernel(const BacksolveArgs* args, const int2* items, int nitems);
template <class T = cplx>
__global__ void crt_lcol_kernel(const void* L, void* Lc, int n, int mb, int nfac);
template <int DEPTH, int G, int NS, class T = cplx>
__global__ void crt_fwdsweep_kernel(FwdArgs F);

// Entries of one workgroup's trailing-window ring: (mb + 1) / 2 rows of mb + 1
// slots, then a slot that stays zero (read for entries outside the window)
// and a discard slot (their stores).
__host__ __device__ inline int crt_ring_entries(int mb) { return ((mb + 1) / 2) * (mb + 1) + 2; }
// Entries of the ring in LDS for bandwidth mb and entry size es bytes (16
// complex, 8 real), rounded up to an even count for 8-byte entries so that
// what follows stays 16-byte aligned.
__host__ __device__ inline int crt_ring_lds_entries(int mb, size_t es) {
  const int n = crt_ring_entries(mb);
  return es == 16 ? n : (n + 1) & ~1;
}
// Entries of one panel buffer of the factor kernel: mb + kNB rows and 4 pad
// rows of kNB columns, with one pad entry after every 4-row group.
__host__ __device__ inline int crt_panel_entries(int mb) {
  const int WP = mb + kNB + 4;
  return WP * kNB + (WP + 3) / 4;
}
// LDS bytes of the factor kernel for bandwidth mb, entry size es (16 complex,
// 8 real) and the ring in LDS (gr false) or global memory: the ring, two
// panel buffers, and two staged row batches with their band offsets.
inline size_t crt_factor_lds_bytes(int mb, size_t es, bool gr) {
  const size_t ring = gr ? 0 : (size_t)crt_ring_lds_entries(mb, es);
  return (ring + 2 * (size_t)crt_panel_entries(mb) + 2 * kNB * kMaxR) * es + 2 * kNB * kMaxR * 2;
}
// Widest band the factorisation and substitutions take: the wide-band factor
// form holds the mb panel rows below the diagonal block in four rows per
// lane, and the
// substitutions' 256-row slot window needs 8-row groups below 257 - mb.
constexpr int kMaxBand = 248;
