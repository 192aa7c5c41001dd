// Batched complex-symmetric band factorisation + forward sweep (gfx950).
//
// One workgroup factors one (wavenumber k, source electrode l) system at a
// time and loops over systems (persistent).  For each system it
//   1. computes the diagonal scaling f_r = 1/sqrt(|a_rr|) of every row
//      (scalab, src/scalab.f90:52-76),
//   2. assembles rows of the stiffness matrix on the fly from the per-element
//      matrices as the factor front reaches them (kompab, src/kompab.f90:75-183,
//      plus the Dirichlet rows of randb, src/randb.f90:43-82), already scaled,
//   3. factors the band L L^T without conjugation (chol, src/chol.f90:44-78)
//      with a right-looking panel of NB columns over an LDS-resident window,
//   4. runs the forward substitution L y = b inside the panel loop
//      (the first half of vre, src/vre.f90:45-56),
// and writes the finished rows of L (band-row layout, row i holds columns
// i-mb..i) and y to its slot of global scratch.  The back substitution runs
// in crt_backsolve_kernel.
//
// The window holds the not-yet-final lower triangle of rows/cols
// [j0, j0+W), W = mb + NB, as a packed triangle in LDS.  After each panel the
// trailing (mb x mb) triangle is updated and moved up by NB in the same pass.
// Every entry receives the subtractions -L_ik L_jk in ascending k, the order of
// the reference's dot product, and each L_ij is divided by L_jj exactly once.
#include "crt_device.h"
#include "crt_kernels.h"

namespace {

__device__ __forceinline__ int tri(int r, int c) { return (r * (r + 1)) / 2 + c; }

// Coefficient of one assembly term (the product added into the band entry).
__device__ __forceinline__ cplx term_value(const FactorArgs& A, int k, int l, int e, int imn) {
  if (e >= 0) {
    double dum = A.elbg[((size_t)k * A.elanz + e) * A.nimn + imn];
    return rmul(dum, ldc(A.sigma + e));
  }
  int rel = -e - 1;
  double dum = A.relbg[(size_t)rel * A.nimn + imn] * A.kg[((size_t)k * A.eanz + l) * A.relanz + rel];
  return rmul(dum, ldc(A.sigma + A.rnr[rel]));
}

// Sum of the diagonal terms of row r (in term order).
__device__ cplx row_diag(const FactorArgs& A, int k, int l, int r) {
  cplx acc = mkc(0.0, 0.0);
  const int p1 = A.rowptr[r + 1];
  for (int p = A.rowptr[r]; p < p1; ++p) {
    if (A.term_t[p] != A.mb) continue;
    int e = A.term_e[p];
    if (e == kDirichletUnit) return mkc(1.0, 0.0);
    acc = cadd(acc, term_value(A, k, l, e, A.term_imn[p]));
  }
  return acc;
}

// Assemble scaled row r (global) into window row rw of the packed triangle,
// window columns starting at global column j0w.  Entries not written here must
// already be zero.  Also returns the scaled RHS entry of the row.
__device__ void assemble_row(const FactorArgs& A, int k, int l, int r, int rw, int j0w,
                             const double* __restrict__ fak, cplx* T) {
  const int p0 = A.rowptr[r], p1 = A.rowptr[r + 1];
  if (p0 == p1) return;
  const double fr = fak[r];
  cplx acc = mkc(0.0, 0.0);
  int curt = A.term_t[p0];
  bool unit = false;
  for (int p = p0; p <= p1; ++p) {
    int t = (p < p1) ? A.term_t[p] : -1;
    if (t != curt) {
      int c = r - A.mb + curt;
      cplx v;
      if (curt == A.mb) {
        v = unit ? mkc(1.0, 0.0) : cdiv(acc, mkc(cabs_(acc), 0.0));
      } else {
        v = rmul(fr * fak[c], acc);
      }
      T[tri(rw, c - j0w)] = v;
      if (p == p1) break;
      acc = mkc(0.0, 0.0);
      curt = t;
    }
    int e = A.term_e[p];
    if (e == kDirichletUnit) {
      unit = true;
    } else {
      acc = cadd(acc, term_value(A, k, l, e, A.term_imn[p]));
    }
  }
}

__device__ __forceinline__ cplx rhs_value(const FactorArgs& A, int l, int r, const double* fak) {
  // b(enr(l)) = -1, b(nsink) = +1 (kompab :162,:183), Dirichlet rows zeroed
  // (randb :49), then scaled by f (scalab :65).
  cplx b = mkc(0.0, 0.0);
  if (r == A.enr[l]) b = mkc(-1.0, 0.0);
  if (A.lsink && r == A.nsink) b = mkc(1.0, 0.0);
  if (A.isdir && A.isdir[r]) b = mkc(0.0, 0.0);
  return rmul(fak[r], b);
}

}  // namespace

template <int NB>
__global__ __launch_bounds__(256) void crt_factor_kernel(FactorArgs A) {
  extern __shared__ double2 smem[];
  cplx* T = reinterpret_cast<cplx*>(smem);
  const int mb = A.mb, n = A.n;
  const int W = mb + NB;
  cplx* yw = T + (W * (W + 1)) / 2;
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;

  for (int s = blockIdx.x; s < A.nsys; s += gridDim.x) {
    const int sys = A.sys_list ? A.sys_list[s] : s;
    const int k = sys / A.eanz, l = sys % A.eanz;
    const int slot = blockIdx.x;
    double* fak = A.fak + (size_t)slot * n;
    double2* Lg = A.L + (size_t)slot * n * (mb + 1);
    double2* yg = A.y + (size_t)slot * n;

    // 1. scaling factors for all rows
    for (int r = tid; r < n; r += nthr) {
      cplx d = row_diag(A, k, l, r);
      double ad = cabs_(d);
      if (!(ad > 0.0)) {
        if (atomicCAS(A.err, 0, 27) == 0) A.err[1] = sys;
        ad = 1.0;
      }
      fak[r] = 1.0 / sqrt(ad);
    }
    __syncthreads();

    // 2. initial window: rows [0, min(W, n))
    const int ntri = (W * (W + 1)) / 2;
    for (int i = tid; i < ntri; i += nthr) T[i] = mkc(0.0, 0.0);
    __syncthreads();
    {
      const int rw_end = min(W, n);
      for (int r = tid; r < rw_end; r += nthr) {
        assemble_row(A, k, l, r, r, 0, fak, T);
        yw[r] = rhs_value(A, l, r, fak);
      }
    }
    __syncthreads();

    for (int j0 = 0; j0 < n; j0 += NB) {
      const int nbc = min(NB, n - j0);
      const int rows = min(W, n - j0);  // window rows present

      // 3a. panel factorisation, columns [0, nbc) of the window
      for (int c = 0; c < nbc; ++c) {
        const cplx s = T[tri(c, c)];
        if (!(cabs_(s) > 0.0)) {
          if (tid == 0 && atomicCAS(A.err, 0, 28) == 0) A.err[1] = sys;
        }
        const cplx lcc = csqrt_(s);
        const int rlast = min(rows - 1, c + mb);
        for (int r = c + 1 + tid; r <= rlast; r += nthr) T[tri(r, c)] = cdiv(T[tri(r, c)], lcc);
        if (tid == 0) {
          cplx yc = cdiv(yw[c], lcc);
          yw[c] = yc;
          yg[j0 + c] = make_double2(yc.re, yc.im);
        }
        __syncthreads();
        for (int r = c + 1 + tid; r <= rlast; r += nthr) {
          const cplx lrc = T[tri(r, c)];
          const int cend = min(nbc - 1, r);
          for (int c2 = c + 1; c2 <= cend; ++c2) T[tri(r, c2)] = cmsub(T[tri(r, c2)], lrc, T[tri(c2, c)]);
          yw[r] = cmsub(yw[r], lrc, yw[c]);
        }
        if (tid == 0) T[tri(c, c)] = lcc;
        __syncthreads();
      }

      // 3b. the panel's columns are final: store them into L (band-row layout)
      for (int idx = tid; idx < rows * NB; idx += nthr) {
        const int r = idx / NB, c = idx - r * NB;
        if (c < nbc && c <= r && r - c <= mb) {
          cplx v = T[tri(r, c)];
          Lg[(size_t)(j0 + r) * (mb + 1) + (c - r + mb)] = make_double2(v.re, v.im);
        }
      }
      if (j0 + NB >= n) break;  // last panel: nothing trails it

      // 3c. trailing update of the (rows-NB)^2 lower triangle, moved up by NB.
      const int m = rows - NB;               // trailing rows present
      const int nt = (m + 3) / 4;            // 4x4 tiles per side
      const int ntiles = (nt * (nt + 1)) / 2;
      constexpr int MAXT = 2;                // tiles per thread (mb <= 4*sqrt(2*512))
      cplx acc[MAXT][4][4];
      int tr0[MAXT], tc0[MAXT];
      cplx ysh[2];
#pragma unroll
      for (int q = 0; q < MAXT; ++q) {
        const int tile = tid + q * nthr;
        tr0[q] = -1;
        if (tile < ntiles) {
          // tile -> (ti, tj), tj <= ti, row-major over the lower triangle
          int ti = (int)((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
          while ((ti * (ti + 1)) / 2 > tile) --ti;
          while (((ti + 1) * (ti + 2)) / 2 <= tile) ++ti;
          const int tj = tile - (ti * (ti + 1)) / 2;
          tr0[q] = NB + 4 * ti;
          tc0[q] = NB + 4 * tj;
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int r = tr0[q] + i, c = tc0[q] + j;
              acc[q][i][j] = (r < rows && c <= r) ? T[tri(r, c)] : mkc(0.0, 0.0);
            }
          for (int kk = 0; kk < NB; ++kk) {
            cplx a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int r = min(tr0[q] + i, rows - 1);
              a[i] = T[tri(r, kk)];
              const int c = min(tc0[q] + i, rows - 1);
              b[i] = T[tri(c, kk)];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
              for (int j = 0; j < 4; ++j) acc[q][i][j] = cmsub(acc[q][i][j], a[i], b[j]);
          }
        }
      }
      for (int q = 0; q < 2; ++q) {
        const int r = NB + tid + q * nthr;
        ysh[q] = (r < rows) ? yw[r] : mkc(0.0, 0.0);
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < MAXT; ++q) {
        if (tr0[q] < 0) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int r = tr0[q] + i, c = tc0[q] + j;
            if (r < rows && c <= r) T[tri(r - NB, c - NB)] = acc[q][i][j];
          }
      }
      for (int q = 0; q < 2; ++q) {
        const int r = NB + tid + q * nthr;
        if (r < rows) yw[r - NB] = ysh[q];
      }
      // zero the rows that enter the window: window rows [W-NB, W)
      {
        const int base = tri(W - NB, 0);
        for (int i = base + tid; i < ntri; i += nthr) T[i] = mkc(0.0, 0.0);
      }
      __syncthreads();

      // 3d. assemble the rows entering the window
      {
        const int j0w = j0 + NB;
        const int r_first = j0w + W - NB;
        const int r_end = min(j0w + W, n);
        for (int r = r_first + tid; r < r_end; r += nthr) {
          assemble_row(A, k, l, r, r - j0w, j0w, fak, T);
          yw[r - j0w] = rhs_value(A, l, r, fak);
        }
      }
      __syncthreads();
    }
    __syncthreads();
    // the slot's factor is overwritten by its next system, so the back
    // substitution consumes it now
    crt_backsolve_block(A, sys, slot, T);
    __syncthreads();
  }
}

template __global__ void crt_factor_kernel<16>(FactorArgs);

// -------------------------------------------------------------------------
// Back substitution x = -L^{-T} y (the second half of vre, src/vre.f90:58-69)
// and kpot = x * f (src/fem.f90:385), done by the whole workgroup in blocks of
// BB rows, descending.  Within a block one wavefront solves the small
// triangle; the rows below the block are then updated by all threads, each
// applying the block's rows in descending order (the reference's order).
__device__ void crt_backsolve_block(const FactorArgs& A, int sys, int slot, cplx* smem) {
  constexpr int BB = 16;
  const int mb = A.mb, n = A.n, m1 = mb + 1;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int k = sys / A.eanz, l = sys % A.eanz;
  const double* fak = A.fak + (size_t)slot * n;
  const double2* yg = A.y + (size_t)slot * n;
  const double2* Lg = A.L + (size_t)slot * n * m1;
  double2* out = A.kpot + ((size_t)k * A.eanz + l) * n;
  int P = 1;
  while (P < mb + BB + 1) P <<= 1;
  cplx* Yc = smem;            // circular working vector, P entries
  cplx* Lb = smem + P;        // BB rows of L
  int loaded_lo = n;          // Yc holds y (or updated values) for rows >= loaded_lo
  for (int i1 = n; i1 > 0;) {
    const int i0 = max(0, i1 - BB);
    const int nb = i1 - i0;
    for (int idx = tid; idx < nb * m1; idx += nthr) Lb[idx] = ldc(Lg + (size_t)i0 * m1 + idx);
    // rows entering the working window [i0-mb, i1)
    {
      const int lo = max(0, i0 - mb);
      for (int j = lo + tid; j < loaded_lo; j += nthr) Yc[j & (P - 1)] = ldc(yg + j);
      loaded_lo = lo;
    }
    __syncthreads();
    if (tid < 64) {
      const int q = tid;
      cplx v = (q < nb) ? Yc[(i0 + q) & (P - 1)] : mkc(0.0, 0.0);
      for (int i = i1 - 1; i >= i0; --i) {
        const int qi = i - i0;
        cplx x = mkc(0.0, 0.0);
        if (q == qi) {
          cplx d = cdiv(v, Lb[qi * m1 + mb]);
          x = mkc(-d.re, -d.im);
          v = x;
        }
        x.re = __shfl(x.re, qi, 64);
        x.im = __shfl(x.im, qi, 64);
        if (q < qi && i - (i0 + q) <= mb) v = cmadd(v, Lb[qi * m1 + (i0 + q - i + mb)], x);
      }
      if (q < nb) {
        Yc[(i0 + q) & (P - 1)] = v;
        out[i0 + q] = make_double2(fak[i0 + q] * v.re, fak[i0 + q] * v.im);
      }
    }
    __syncthreads();
    for (int j = max(0, i0 - mb) + tid; j < i0; j += nthr) {
      cplx v = Yc[j & (P - 1)];
      for (int i = i1 - 1; i >= i0; --i) {
        if (i - j > mb) continue;
        v = cmadd(v, Lb[(i - i0) * m1 + (j - i + mb)], Yc[i & (P - 1)]);
      }
      Yc[j & (P - 1)] = v;
    }
    __syncthreads();
    i1 = i0;
  }
}
