// Argument blocks of the post-solve kernels (precal, sensitivities, adjoint, CG).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct PrecalArgs {
  int elanz, relanz, eanz, K, nimn;
  const int* el_type;     // elanz (3, 5 or 8)
  const int* el_nodes;    // elanz x 4, 0-based
  const int* bnd_type;    // relanz (11, 12, 13)
  const int* bnd_nodes;   // relanz x 2, 0-based
  const double* nx;       // node coordinates (through snr)
  const double* ny;
  const int* enr;         // eanz, 0-based
  const double* kwn;      // K
  double sytop;
  double* elbg;           // [K][elanz][nimn]
  double* relbg;          // [relanz][nimn]
  double* kg;             // [K][eanz][relanz]
  int* err;
};

struct SensArgs {
  int n, elanz, eanz, K, nimn, nanz, swrtr, mode;
  const int* el_nk;
  const int* el_nodes;
  const double* elbg;
  const double* kwnwi;
  const double2* kpot;    // [K][eanz][n]
  const double2* kpt;     // [K][n][eanz] (crt_kpot_transpose_kernel)
  const double2* sigma;
  const double2* volt;
  const int* strnr;
  const int* vnr;
  double2* sens;          // nanz x elanz, column major
};

// Distinct electrode pairs of the measurements for crt_sens_pairs_kernel:
// current pairs cp (one per distinct strnr), potential pairs pp (one per
// distinct vnr), and per measurement its (cp, pp) indices.
struct SensPairs {
  int ncp, npp;
  const int2* cp;  // (e1, e2), 1-based electrode numbers, 0 = none
  const int2* pp;  // (e3, e4)
  const int2* mp;  // [nanz] (cp index, pp index)
  const int* ord;  // [nanz] measurements in (cp, pp) order: the lanes of a
                   // wavefront then mostly share u and read neighbouring d34
};
constexpr int kSensPairQ = 16;  // measurements per thread per pass
constexpr int kSensRow = 5;     // LDS entries per pair row of the pair form (4 used)

struct AdjArgs {
  int n, elanz, eanz, K, nimn, swrtr, mode;
  const int* el_nk;
  const int* el_nodes;
  const double* elbg;
  const double* kwnwi;
  const double2* kpot;    // [K][eanz][n]
  const double2* kpt;     // [K][n][eanz] (crt_kpot_transpose_kernel)
  const double2* Y;       // [K][n][eanz]
  const double2* sigma;
  double2* grad;          // elanz
  // wavenumbers k0 .. k0 + kc - 1 in kpt and Y (a pass of the recomputing
  // gradient; 0 .. K - 1 otherwise); acc carries the sums between passes
  int k0, kc, first, last;
  double2* acc;           // elanz
};

struct CgArgs {
  int nanz, manz, ltri, nx, maxnodes;
  double lam, eps;
  const double2* J;       // nanz x manz, column major
  const double* w;        // wmatd * wdfak
  const double* rdiag;    // diagonal of R used by the preconditioner
  const double* smatm;    // manz x (3 | maxnodes+1), column major
  const int* nachbar;     // manz x maxnodes (1-based, 0 = none)
  double* cgfac;
  double2* bvec;
  double2* rvec;
  double2* pvec;
  double2* ap;
  double2* dpar;
  double* cgres;          // ncgmax + 2
  double* scal;           // 8 device scalars of the recurrence
  double2* part;          // [nslice][nanz] partial sums of J (cgfac o p)
  int nslice, chunk;      // column slices of that product and their width
};

__global__ void crt_precal_area_kernel(PrecalArgs P);
__global__ void crt_precal_bnd_kernel(PrecalArgs P);
__global__ void crt_rtrafo_kernel(const double2* kpot, const double* kwnwi, int n, int eanz, int K,
                                  int swrtr, double2* hpot);
__global__ void crt_real_sigma_kernel(const double2* in, int n, double2* out);
__global__ void crt_real_part_kernel(const double2* in, size_t n, double* out);
__global__ void crt_rtrafo_elec_kernel(const double2* kpot, const double* kwnwi, const int* enr, int n, int eanz,
                                       int k0, int kc, int swrtr, int first, int last, double2* acc, double2* he);
__global__ void crt_volt_kernel(const double2* he, int eanz, int nanz, const int* strnr, const int* vnr,
                                int mode, const double* kfak, double2* volt, double2* sigmaa, int* err);
__global__ void crt_sens_kernel(SensArgs S);
// RUNS: each thread takes a contiguous run of measurements (else interleaved)
template <bool RUNS>
__global__ void crt_sens_pairs_kernel(SensArgs S, SensPairs Q);
// Gram form (FP64 matrix cores): kGramNT threads, MT tiles of 16 x 16 per
// wavefront; off[i] = the tile-buffer offsets of measurement i's four terms
// (crt_gram_offset), two 16-bit offsets per word; tiles = the nt tiles those
// offsets fall in, ascending; dynamic LDS crt_gram_lds_bytes(eanz)
constexpr int kGramMaxE = 128;
constexpr int kGramNT = 256;
constexpr int kGramQC = 4;  // measurements per thread per extraction pass
// REAL: the potentials are real (the DC twin): one real product per tile
template <int MT, bool REAL>
__global__ void crt_sens_gram_kernel(SensArgs S, const uint2* off, const int* tiles, int nt);
inline int crt_gram_tiles(int eanz) {
  const int tb = (eanz + 15) / 16;
  return tb * (tb + 1) / 2;
}
// Offset (in doubles) of G[a][b] in the kernel's buffer of lower-triangle
// tiles (tile x = ti (ti + 1) / 2 + tj row-major, 256 entries each) for
// 0-based electrodes a, b; the zero slot after the tiles where either is
// absent (-1).  G is symmetric: (a, b) in an upper tile reads (b, a).
inline int crt_gram_offset(int eanz, int a, int b) {
  if (a < 0 || b < 0) return crt_gram_tiles(eanz) * 256;
  if ((a >> 4) < (b >> 4)) {
    const int t = a;
    a = b;
    b = t;
  }
  const int ti = a >> 4, tj = b >> 4;
  return (ti * (ti + 1) / 2 + tj) * 256 + (a & 15) * 16 + (b & 15);
}
inline size_t crt_gram_lds_bytes(int eanz) {
  const size_t ep = 16 * (size_t)((eanz + 15) / 16);
  const size_t stage = 2 * 8 * ep * 16, tiles = (size_t)crt_gram_tiles(eanz) * 256 * 8 + 16;
  return stage > tiles ? stage : tiles;
}
__global__ void crt_kpot_transpose_kernel(const double2* kpot, int n, int E, double2* kpt);
__global__ void crt_adj_gamma_kernel(int nanz, int mode, const double* w, const double2* resid,
                                     const double2* volt, double2* gamma);
__global__ void crt_adj_cmat_kernel(int nanz, int eanz, const int* strnr, const int* vnr, const int* qptr,
                                    const int* qlist, const double2* gamma, double2* C);
__global__ void crt_adj_gemm_kernel(const double2* kpot, const double2* C, int n, int eanz, double2* Y);
__global__ void crt_adj_elem_kernel(AdjArgs A);
__global__ void crt_jhwr_kernel(const double2* J, int nanz, int manz, const double* w, const double2* r,
                                double2* g);
__global__ void crt_cg_rhs_kernel(CgArgs G, const double2* dat, const double2* sigmaa, const double2* rm);
__global__ void crt_cg_bap_kernel(CgArgs G);
__global__ void crt_cg_bap_sum_kernel(CgArgs G);
__global__ void crt_cg_bb_kernel(CgArgs G);
__global__ void crt_cg_scalar_kernel(CgArgs G, int stage, int iter);

constexpr int kAdjColsHost = 32;
constexpr int kAdjTilesHost = 8;  // electrodes up to 128 take the matrix-core GEMM
__global__ void crt_adj_gemm_mfma_kernel(const double2* kpot, const double2* C, int n, int eanz, double2* Y);
// C-stationary form: NS 4-electrode steps (eanz <= 4 NS), 64 ceil(eanz / 16)
// threads, dynamic LDS 2 x 64 NS complex values
template <int NS>
__global__ void crt_adj_gemm_cs_kernel(const double2* kpot, const double2* C, int n, int eanz, double2* Y);
