// Batched complex-symmetric band factorisation with fused assembly, scaling
// and forward sweep (gfx950), plus the wavefront-per-system back substitution.
//
// crt_area_values_kernel: per wavenumber k and row r, the sums of the area
//   element contributions to each stored band entry of row r, in element order
//   (the first part of kompab's accumulation, src/kompab.f90:79-129).  They do
//   not depend on the source electrode.
//
// crt_factor_kernel: one workgroup factors one (k, l) system at a time and
//   loops over systems (persistent).  Rows enter an LDS window as the factor
//   front reaches them: the area sums are completed with the mixed-boundary
//   terms of the row (boundary elements follow all area elements in the
//   reference's element order, so the sum order is the reference's), then
//   scaled (scalab, src/scalab.f90:52-76).  The band L L^T without conjugation
//   (chol, src/chol.f90:44-78) is factored right-looking in panels of NB
//   columns; each L_ij is divided by L_jj exactly once and every entry receives
//   -L_ik L_jk in ascending k (the reference's dot-product order).  The forward
//   substitution L y = b (vre, src/vre.f90:45-56) runs inside the panel loop.
//   Finished rows of L (band-row layout: row i holds columns i-mb..i), y and the
//   scaling factors go to global memory; threads communicate only through LDS,
//   so the barriers order LDS alone and the stores stream out behind them.
//
// crt_backsolve_kernel: x = -L^{-T} y (vre, src/vre.f90:58-69) and
//   kpot = x * f (src/fem.f90:385), one wavefront per system.
#include "crt_device.h"
#include "crt_kernels.h"

namespace {

__device__ __forceinline__ int tri(int r, int c) { return (r * (r + 1)) / 2 + c; }

// Optional per-phase timing (thread 0 of each workgroup, constant-rate clock).
struct PhaseClock {
  unsigned long long* out;
  unsigned long long t, acc[8];
  __device__ void start() {
    if (!out) return;
    t = wall_clock64();
    for (int i = 0; i < 8; ++i) acc[i] = 0;
  }
  __device__ void mark(int phase) {
    if (!out) return;
    unsigned long long now = wall_clock64();
    acc[phase] += now - t;
    t = now;
  }
  __device__ void flush() {
    if (!out) return;
    for (int i = 0; i < 8; ++i) atomicAdd(out + i, acc[i]);
  }
};

__device__ __forceinline__ cplx bnd_term(const FactorArgs& A, int k, int l, int code, int imn) {
  const int rel = -code - 1;
  const double dum = A.relbg[(size_t)rel * A.nimn + imn] * A.kg[((size_t)k * A.eanz + l) * A.relanz + rel];
  return rmul(dum, ldc(A.sigma + A.rnr[rel]));
}

__device__ __forceinline__ cplx rhs_value(const FactorArgs& A, int l, int r, double fr) {
  // b(enr(l)) = -1, b(nsink) = +1 (kompab :162,:183), Dirichlet rows zeroed
  // (randb :49), then scaled by f (scalab :65).
  cplx b = mkc(0.0, 0.0);
  if (r == A.enr[l]) b = mkc(-1.0, 0.0);
  if (A.lsink && r == A.nsink) b = mkc(1.0, 0.0);
  if (A.isdir && A.isdir[r]) b = mkc(0.0, 0.0);
  return rmul(fr, b);
}

// Row r (window row rw, window columns from j0w): write the unscaled band
// entries -- the per-wavenumber area sums av (R slots) plus the row's
// mixed-boundary terms -- into the zeroed window row, and return the scaling
// factor 1/sqrt(|a_rr|) (scalab :56-63).
__device__ double put_row_unscaled(const FactorArgs& A, int k, int l, int r, int rw, int j0w,
                                   const double2* av, cplx* __restrict__ T, int sys) {
  cplx* row = T + tri(rw, 0);
  const int base = r - A.mb - j0w;
  for (int q = 0; q < A.R; ++q) {
    const int t = A.pat_t[(size_t)r * A.R + q];
    if (t >= 0) row[base + t] = ldc(av + (size_t)r * A.R + q);
  }
  const int b1 = A.bptr[r + 1];
  for (int p = A.bptr[r]; p < b1; ++p) {
    const int t = A.pat_t[(size_t)r * A.R + A.bslot[p]];
    row[base + t] = cadd(row[base + t], bnd_term(A, k, l, A.bcode[p], A.bimn[p]));
  }
  const double ad = cabs_(row[base + A.mb]);
  if (!(ad > 0.0)) {
    if (atomicCAS(A.err, 0, 27) == 0) A.err[1] = sys;
    return 1.0;
  }
  return 1.0 / sqrt(ad);
}

// Scale the stored entries of row r in place: a_rr/|a_rr| on the diagonal,
// a_rc f_r f_c elsewhere (scalab :58-73).
__device__ void scale_row(const FactorArgs& A, int r, int rw, int j0w, double fr, const double* fw,
                          cplx* __restrict__ T) {
  cplx* row = T + tri(rw, 0);
  const int base = r - A.mb - j0w;
  for (int q = 0; q < A.R; ++q) {
    const int t = A.pat_t[(size_t)r * A.R + q];
    if (t < 0) continue;
    const cplx v = row[base + t];
    if (t == A.mb) row[base + t] = cdiv(v, mkc(cabs_(v), 0.0));
    else row[base + t] = rmul(fr * fw[base + t], v);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
__global__ void crt_area_values_kernel(FactorArgs A, int K, double2* __restrict__ av) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)K * A.n) return;
  const int k = (int)(idx / A.n), r = (int)(idx % A.n);
  double2* out = av + ((size_t)k * A.n + r) * A.R;
  const int p1 = A.rowptr[r + 1];
  int p = A.rowptr[r];
  for (int q = 0; q < A.R; ++q) {
    const int t = A.pat_t[(size_t)r * A.R + q];
    cplx acc = mkc(0.0, 0.0);
    for (; t >= 0 && p < p1 && A.term_t[p] == t; ++p) {
      const int e = A.term_e[p];
      if (e == kDirichletUnit) {
        acc = mkc(1.0, 0.0);
      } else if (e >= 0) {
        const double dum = A.elbg[((size_t)k * A.elanz + e) * A.nimn + A.term_imn[p]];
        acc = cadd(acc, rmul(dum, ldc(A.sigma + e)));
      }
    }
    out[q] = make_double2(acc.re, acc.im);
  }
}

// ---------------------------------------------------------------------------
template <int NB, int MAXT>
__global__ __launch_bounds__(256) void crt_factor_kernel(FactorArgs A) {
  extern __shared__ double2 smem[];
  const int mb = A.mb, n = A.n;
  const int W = mb + NB;
  const int ntri = (W * (W + 1)) / 2;
  cplx* T = reinterpret_cast<cplx*>(smem);
  cplx* yw = T + ntri;                              // W: right-hand side / y
  cplx* colb = yw + W;                              // 2 x NB: one L column of the panel
  cplx* diagb = colb + 2 * NB;                      // 2: next pivot value
  cplx* ycb = diagb + 2;                            // 2: y of the pivot row
  double* fw = reinterpret_cast<double*>(ycb + 2);  // W: scaling factors of window rows
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;

  PhaseClock clk;
  clk.out = (tid == 0) ? A.prof : nullptr;
  clk.start();

  for (int s = blockIdx.x; s < A.nsys; s += gridDim.x) {
    const int sys = A.sys_first + s;
    const int k = sys / A.eanz, l = sys % A.eanz;
    double* fak = A.fak + (size_t)s * n;
    double2* Lg = A.L + (size_t)s * n * (mb + 1);
    double2* yg = A.y + (size_t)s * n;
    const double2* av = A.av + (size_t)k * n * A.R;

    // ---- first window: rows [0, min(W, n)); unscaled entries and scaling
    // factors, then the scaled entries
    for (int i = tid; i < ntri; i += nthr) T[i] = mkc(0.0, 0.0);
    lds_barrier();
    {
      const int rend = min(W, n);
      double fr = 0.0;
      if (tid < rend) {
        fr = put_row_unscaled(A, k, l, tid, tid, 0, av, T, sys);
        fw[tid] = fr;
        fak[tid] = fr;
      }
      lds_barrier();
      if (tid < rend) {
        scale_row(A, tid, tid, 0, fr, fw, T);
        yw[tid] = rhs_value(A, l, tid, fr);
      }
    }
    lds_barrier();
    clk.mark(1);

    for (int j0 = 0; j0 < n; j0 += NB) {
      const int nbc = min(NB, n - j0);
      const int rows = min(W, n - j0);  // window rows present

      // ---- panel factorisation.  During the panel each thread owns its window
      // row; columns are broadcast through double-buffered LDS slots, one
      // barrier per column.
      const int r = tid;
      const bool live = r < rows;
      cplx* __restrict__ myrow = T + tri(min(r, W - 1), 0);
      cplx yr = live ? yw[r] : mkc(0.0, 0.0);
      if (tid == 0) diagb[0] = T[0];
      lds_barrier();
      for (int c = 0; c < nbc; ++c) {
        const int buf = c & 1;
        const cplx d = diagb[buf];
        if (tid == 0 && !(cabs_(d) > 0.0) && atomicCAS(A.err, 0, 28) == 0) A.err[1] = sys;
        const cplx lcc = csqrt_(d);
        const bool below = live && r > c && r - c <= mb;
        cplx lrc = mkc(0.0, 0.0);
        if (below) {
          lrc = cdiv(myrow[c], lcc);
          myrow[c] = lrc;
          if (r < nbc) colb[buf * NB + r] = lrc;
          if (r == c + 1 && r < nbc) {
            // the next pivot takes its -L L now so it travels with this column
            const cplx dn = cmsub(myrow[r], lrc, lrc);
            myrow[r] = dn;
            diagb[buf ^ 1] = dn;
          }
        }
        if (r == c) {
          myrow[c] = lcc;
          const cplx yc = cdiv(yr, lcc);
          yr = yc;
          ycb[buf] = yc;
          yg[j0 + c] = make_double2(yc.re, yc.im);
        }
        lds_barrier();
        if (below) {
          // the row's remaining panel entries: all loads first, then the
          // updates (static indices, registers), then the stores
          const cplx* cb = colb + buf * NB;
          const int cend = min(nbc - 1, r);
          const int cskip = (r == c + 1) ? c + 1 : -1;
          constexpr int H = NB / 2;
#pragma unroll
          for (int h0 = 0; h0 < NB; h0 += H) {
            cplx rv[H], cv[H];
#pragma unroll
            for (int u = 0; u < H; ++u) {
              const int c2 = h0 + u;
              const bool upd = c2 > c && c2 <= cend && c2 != cskip;
              rv[u] = upd ? myrow[c2] : mkc(0.0, 0.0);
              cv[u] = upd ? cb[c2] : mkc(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < H; ++u) {
              const int c2 = h0 + u;
              const bool upd = c2 > c && c2 <= cend && c2 != cskip;
              if (upd) myrow[c2] = cmsub(rv[u], lrc, cv[u]);
            }
          }
          yr = cmsub(yr, lrc, ycb[buf]);
        }
      }
      // the panel's columns are final: out to L (band-row layout)
      for (int idx = tid; idx < rows * NB; idx += nthr) {
        const int rr = idx / NB, cc = idx - rr * NB;
        if (cc < nbc && cc <= rr && rr - cc <= mb) {
          const cplx v = T[tri(rr, cc)];
          Lg[(size_t)(j0 + rr) * (mb + 1) + (cc - rr + mb)] = make_double2(v.re, v.im);
        }
      }
      if (live) yw[r] = yr;
      lds_barrier();
      clk.mark(2);
      if (j0 + NB >= n) break;  // last panel: nothing trails it

      const int j0w = j0 + NB;
      const int r_new = j0w + W - NB + tid;  // threads 0..NB-1: one entering row each
      const bool has_new = tid < NB && r_new < n;

      // ---- trailing update of the (rows-NB)^2 lower triangle, moved up by NB
      const int m = rows - NB;
      const int nt = (m + 3) / 4;
      const int ntiles = (nt * (nt + 1)) / 2;
      cplx acc[MAXT][4][4];
      int tr0[MAXT], tc0[MAXT];
#pragma unroll
      for (int q = 0; q < MAXT; ++q) {
        const int tile = tid + q * nthr;
        tr0[q] = -1;
        tc0[q] = 0;
        if (tile < ntiles) {
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
              const int rr = tr0[q] + i, cc = tc0[q] + j;
              acc[q][i][j] = (rr < rows && cc <= rr) ? T[tri(rr, cc)] : mkc(0.0, 0.0);
            }
          for (int kk = 0; kk < NB; ++kk) {
            cplx a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              a[i] = T[tri(min(tr0[q] + i, rows - 1), kk)];
              b[i] = T[tri(min(tc0[q] + i, rows - 1), kk)];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
              for (int j = 0; j < 4; ++j) acc[q][i][j] = cmsub(acc[q][i][j], a[i], b[j]);
          }
        }
      }
      cplx ysh = mkc(0.0, 0.0);
      double fsh = 0.0;
      if (NB + tid < rows) {
        ysh = yw[NB + tid];
        fsh = fw[NB + tid];
      }
      lds_barrier();
#pragma unroll
      for (int q = 0; q < MAXT; ++q) {
        if (tr0[q] < 0) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int rr = tr0[q] + i, cc = tc0[q] + j;
            if (rr < rows && cc <= rr) T[tri(rr - NB, cc - NB)] = acc[q][i][j];
          }
      }
      if (NB + tid < rows) {
        yw[tid] = ysh;
        fw[tid] = fsh;
      }
      for (int i = tri(W - NB, 0) + tid; i < ntri; i += nthr) T[i] = mkc(0.0, 0.0);
      lds_barrier();
      clk.mark(4);

      // ---- the entering rows: unscaled entries and scaling factors, then scaled
      double fr = 0.0;
      if (has_new) {
        fr = put_row_unscaled(A, k, l, r_new, r_new - j0w, j0w, av, T, sys);
        fw[r_new - j0w] = fr;
        fak[r_new] = fr;
      }
      lds_barrier();
      if (has_new) {
        scale_row(A, r_new, r_new - j0w, j0w, fr, fw, T);
        yw[r_new - j0w] = rhs_value(A, l, r_new, fr);
      }
      lds_barrier();
      clk.mark(5);
    }
    lds_barrier();
  }
  clk.flush();
}

template __global__ void crt_factor_kernel<16, 1>(FactorArgs);
template __global__ void crt_factor_kernel<16, 2>(FactorArgs);

// ---------------------------------------------------------------------------
// Back substitution, one wavefront per system.  Lane j (mod 64) keeps the
// working values of the rows j = q (mod 64*SLOTS) of the band window in SLOTS
// registers; row i's pivot lane divides, the value is broadcast and every lane
// adds L_ij x_i to its rows j in [i-mb, i) -- for each j the updates arrive in
// descending i, the reference's order.  A slot freed by x_i is refilled with
// y_{i - 64*SLOTS}, which enters the window 64*SLOTS - mb rows later.  L rows
// are fetched DEPTH rows ahead.
template <int SLOTS, int DEPTH>
__global__ __launch_bounds__(64) void crt_backsolve_kernel(BacksolveArgs B) {
  const int lane = threadIdx.x;
  const int s = blockIdx.x;
  if (s >= B.nsys) return;
  const int sys = B.sys_first + s;
  const int k = sys / B.eanz, l = sys % B.eanz;
  const int n = B.n, mb = B.mb, m1 = mb + 1;
  constexpr int P = 64 * SLOTS;
  const double2* Lg = B.L + (size_t)s * n * m1;
  const double2* yg = B.y + (size_t)s * n;
  const double* fak = B.fak + (size_t)s * n;
  double2* out = B.kpot + ((size_t)k * B.eanz + l) * n;

  // class q = lane + 64*sl holds row j(i, q) = i - ((i - q) mod P) at step i
  cplx ya[SLOTS];
#pragma unroll
  for (int sl = 0; sl < SLOTS; ++sl) {
    const int q = lane + 64 * sl;
    const int i = n - 1;
    const int j = i - (((i - q) % P) + P) % P;
    ya[sl] = (j >= 0) ? ldc(yg + j) : mkc(0.0, 0.0);
  }
  // rows fetched DEPTH steps ahead: this lane's L entries, and for the pivot
  // lane L_ii, f_i and the refill value y_{i-P}
  cplx lr[DEPTH][SLOTS];
  cplx lii[DEPTH], ynx[DEPTH];
  double fr[DEPTH];
  auto fetch = [&](int i, cplx* dst, cplx& dii, cplx& dyn, double& fdst) {
#pragma unroll
    for (int sl = 0; sl < SLOTS; ++sl) {
      const int q = lane + 64 * sl;
      const int j = i - (((i - q) % P) + P) % P;
      const int t = j - i + mb;
      dst[sl] = (i >= 0 && j >= 0 && t >= 0 && j < i) ? ldc(Lg + (size_t)i * m1 + t) : mkc(0.0, 0.0);
    }
    const bool piv = i >= 0 && lane == (i % P) % 64;
    dii = piv ? ldc(Lg + (size_t)i * m1 + mb) : mkc(1.0, 0.0);
    fdst = piv ? fak[i] : 0.0;
    dyn = (piv && i - P >= 0) ? ldc(yg + (i - P)) : mkc(0.0, 0.0);
  };
#pragma unroll
  for (int d = 0; d < DEPTH; ++d) fetch(n - 1 - d, lr[d], lii[d], ynx[d], fr[d]);

  for (int ib = n - 1; ib >= 0; ib -= DEPTH) {
#pragma unroll
    for (int d = 0; d < DEPTH; ++d) {
      const int i = ib - d;
      if (i >= 0) {
        const int qp = i % P;
        const int plane = qp % 64, pslot = qp / 64;
        cplx x = mkc(0.0, 0.0);
        if (lane == plane) {
          cplx yi = ya[0];
#pragma unroll
          for (int sl = 1; sl < SLOTS; ++sl)
            if (sl == pslot) yi = ya[sl];
          const cplx dq = cdiv(yi, lii[d]);
          x = mkc(-dq.re, -dq.im);
          out[i] = make_double2(fr[d] * x.re, fr[d] * x.im);
          // the freed slot takes y_{i-P}, first updated by row i-P+mb
#pragma unroll
          for (int sl = 0; sl < SLOTS; ++sl)
            if (sl == pslot) ya[sl] = ynx[d];
        }
        x.re = __shfl(x.re, plane, 64);
        x.im = __shfl(x.im, plane, 64);
#pragma unroll
        for (int sl = 0; sl < SLOTS; ++sl) {
          const int q = lane + 64 * sl;
          const int j = i - (((i - q) % P) + P) % P;
          if (j < i && i - j <= mb && j >= 0) ya[sl] = cmadd(ya[sl], lr[d][sl], x);
        }
        fetch(i - DEPTH, lr[d], lii[d], ynx[d], fr[d]);
      }
    }
  }
}

template __global__ void crt_backsolve_kernel<2, 8>(BacksolveArgs);
template __global__ void crt_backsolve_kernel<4, 4>(BacksolveArgs);
