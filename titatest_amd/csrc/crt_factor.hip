// Batched complex-symmetric band factorisation with fused forward sweep
// (gfx950), the assembly/scaling kernels that feed it, and the
// wavefront-per-system back substitution.
//
// crt_area_values_kernel: per wavenumber k and row r, the sums of the area
//   element contributions to each stored band entry of row r, in element order
//   (the first part of kompab's accumulation, src/kompab.f90:79-129).  They do
//   not depend on the source electrode.
//
// crt_scale_kernel / crt_rows_kernel: per system (k, l) and row, the area sums
//   completed with the row's mixed-boundary terms (boundary elements follow all
//   area elements in the reference's element order, so the sum order is the
//   reference's), the scaling factor f = 1/sqrt|a_rr| and the scaled entries
//   a_rr/|a_rr|, a_rc f_r f_c (scalab, src/scalab.f90:52-76).  One record of R
//   values per row (the row's distinct band offsets).
//
// crt_factor_kernel: band L L^T without conjugation (chol, src/chol.f90:44-78)
//   plus the forward substitution L y = b (vre, src/vre.f90:45-56), one
//   workgroup per system, persistent over systems.  The matrix is processed in
//   stages of kNB columns.  Stage J: waves 1-3 first apply panel J-1 to the
//   kNB columns of panel J (the look-ahead), then, while wave 0 factors panel J
//   in registers without any workgroup barrier, they apply panel J-1 to the
//   rest of the trailing window.  Two barriers per stage.  Every entry receives
//   -L_ik L_jk in ascending k and is divided by L_jj once (the reference's
//   arithmetic per entry).  Rows enter from the scaled-row records one batch of
//   kNB rows per stage, staged in LDS a stage ahead.
//
// crt_backsolve_kernel: x = -L^{-T} y (vre, src/vre.f90:58-69) and
//   kpot = x * f (src/fem.f90:385), one wavefront per system.
#include <type_traits>

#include "crt_device.h"
#include "crt_kernels.h"

namespace {

constexpr int kStg = 3;  // staged row batches: the stage's previous, current and next

// Optional per-phase timing (constant-rate clock), kept by thread 0 (wave 0:
// accumulators 0-3) and thread 64 (waves 1-3: accumulators 4-7).
struct PhaseClock {
  unsigned long long* out;
  int base;
  unsigned long long t, acc[4];
  __device__ void start(unsigned long long* o, int b) {
    out = o;
    base = b;
    if (!out) return;
    t = wall_clock64();
    for (int i = 0; i < 4; ++i) acc[i] = 0;
  }
  __device__ void mark(int phase) {
    if (!out) return;
    unsigned long long now = wall_clock64();
    acc[phase] += now - t;
    t = now;
  }
  __device__ void flush() {
    if (!out) return;
    for (int i = 0; i < 4; ++i) atomicAdd(out + base + i, acc[i]);
  }
};

__device__ __forceinline__ cplx bnd_term(const FactorArgs& A, int k, int l, int code, int imn) {
  const int rel = -code - 1;
  const double dum = A.relbg[(size_t)rel * A.nimn + imn] * A.kg[((size_t)k * A.eanz + l) * A.relanz + rel];
  return rmul(dum, ldc(A.sigma + A.rnr[rel]));
}

// Trailing window storage.  The window holds rows and columns [o, o+mb); its
// diagonal t = r - c (0 <= t < mb) has live entries for c in [o, o+mb-t).
// Diagonals t and mb-1-t share one ring of mb+1 slots: diagonal t at slot
// c mod (mb+1), its partner at slot (r+1) mod (mb+1).  For every origin o the
// two live ranges are complementary arcs of the ring, so advancing the window
// moves no data, and the slot of a departing column's entry is the slot of an
// entering row's entry.  Arguments are window coordinates rw = r - o, cw = c - o.
struct Ring {
  int mb, M1, half, om;  // om = o mod (mb+1)
  __device__ __forceinline__ int pos(int rw, int cw) const {
    const int t = rw - cw;
    if (t < half) {
      int p = cw + om;
      p = p >= M1 ? p - M1 : p;
      return t * M1 + p;
    }
    int p = rw + 1 + om;
    p = p >= M1 ? p - M1 : p;
    return (mb - 1 - t) * M1 + p;
  }
};

// Panel buffer: W rows of kNB columns, column index swizzled by the row's
// 4-row group so that tiles reading one column from different rows spread
// over the LDS banks.
__device__ __forceinline__ int pix(int rw, int k) { return rw * kNB + (k ^ ((rw >> 2) & (kNB - 1))); }

// Value at band offset t of row i of a staged batch (0 when t is not one of
// the row's stored offsets).
__device__ __forceinline__ cplx staged(const cplx* sv, const int16_t* so, int i, int t) {
  int qf = -1;
#pragma unroll
  for (int q = 0; q < kMaxR; ++q) qf = (so[i * kMaxR + q] == t) ? q : qf;
  return csel(qf >= 0, sv[i * kMaxR + max(qf, 0)], mkc(0.0, 0.0));
}

__device__ __forceinline__ double rl(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ cplx rlc(cplx z, int l) { return mkc(rl(z.re, l), rl(z.im, l)); }

// Scaled-row record slot q of row r (value and band offset), or (0, -1).
__device__ __forceinline__ void row_slot(const FactorArgs& A, const double2* Sg, int r, int q, cplx& v, int& t) {
  v = mkc(0.0, 0.0);
  t = -1;
  if (r >= 0 && r < A.n && q < A.R) {
    t = A.pat_t[(size_t)r * A.R + q];
    if (t >= 0) v = ldc(Sg + (size_t)r * A.R + q);
  }
}

struct Stage {
  int n, mb, W, o;
  Ring ring;
  bool apply;              // panel J-1 exists
  const cplx* Pp;          // panel J-1 (its window coordinates: stage rows + kNB)
  const cplx* sv_old;      // batch J-1 = stage rows [mb - kNB, mb)
  const int16_t* so_old;
};

// Entry (rw, cw) of a tile exists: lower triangle, inside the window and the
// matrix.
__device__ __forceinline__ bool tile_ok(const Stage& C, int rw, int cw) {
  return cw <= rw && rw < C.mb && C.o + rw < C.n;
}

// TR x TC tile of the stage window at (rw0, cw0): the current values (ring, or
// the staged batch for the window's last kNB rows) minus panel J-1's rank-kNB
// product, accumulated in ascending k.  For the 2x2 look-ahead tiles the
// current values are fetched without branches, in two rounds of LDS loads
// whose latencies overlap: the ring entries and each tile row's offset record
// of the staged batch (kMaxR int16 offsets, one 16-byte load), then the
// staged values at the matching slots; every index is clamped in range and the
// choice is a select.
template <int TR, int TC>
__device__ __forceinline__ void tile_update(const Stage& C, const cplx* Tt, int rw0, int cw0, cplx acc[TR][TC]) {
  if constexpr (TR * TC <= 4) {
    // look-ahead tiles (on the critical path): branch-free two-round fetch
    int slot[TR][TC];
#pragma unroll
    for (int i = 0; i < TR; ++i) {
      const int rw = rw0 + i;
      const int ibc = min(max(rw - (C.mb - kNB), 0), kNB - 1);
      const u32x4 rec = *reinterpret_cast<const u32x4*>(C.so_old + ibc * kMaxR);
      const int rwc = min(max(rw, 0), C.mb - 1);
#pragma unroll
      for (int j = 0; j < TC; ++j) {
        const int cw = cw0 + j;
        acc[i][j] = Tt[C.ring.pos(rwc, min(max(cw, 0), rwc))];
        const int t = cw - rw + C.mb;
        int q = -1;
#pragma unroll
        for (int sl = 0; sl < kMaxR; ++sl) {
          const int off = (int)(short)(unsigned short)(rec[sl >> 1] >> (16 * (sl & 1)));
          q = off == t ? sl : q;
        }
        slot[i][j] = q;
      }
    }
#pragma unroll
    for (int i = 0; i < TR; ++i) {
      const int rw = rw0 + i;
      const int ib = rw - (C.mb - kNB);
      const int ibc = min(max(ib, 0), kNB - 1);
#pragma unroll
      for (int j = 0; j < TC; ++j) {
        const cplx sv = C.sv_old[ibc * kMaxR + max(slot[i][j], 0)];
        const cplx v = ib >= 0 ? csel(slot[i][j] >= 0, sv, mkc(0.0, 0.0)) : acc[i][j];
        acc[i][j] = csel(tile_ok(C, rw, cw0 + j), v, mkc(0.0, 0.0));
      }
    }
  } else {
    // trailing tiles: only the few on the window's last kNB rows take the
    // staged batch's slot search
#pragma unroll
    for (int i = 0; i < TR; ++i)
#pragma unroll
      for (int j = 0; j < TC; ++j) {
        const int rw = rw0 + i, cw = cw0 + j;
        cplx v = mkc(0.0, 0.0);
        if (tile_ok(C, rw, cw)) {
          const int ib = rw - (C.mb - kNB);
          v = ib >= 0 ? staged(C.sv_old, C.so_old, ib, cw - rw + C.mb) : Tt[C.ring.pos(rw, cw)];
        }
        acc[i][j] = v;
      }
  }
  if (!C.apply) return;
  // operand rows of panel J-1 (stage rows + kNB)
  int ra[TR], ca[TC];
#pragma unroll
  for (int i = 0; i < TR; ++i) ra[i] = min(rw0 + i + kNB, C.W - 1);
#pragma unroll
  for (int j = 0; j < TC; ++j) ca[j] = min(cw0 + j + kNB, C.W - 1);
  // the 2x2 look-ahead tiles issue all their operand loads at once; the
  // 4x4 trailing tiles keep one k in registers at a time
#pragma unroll (TR * TC <= 4 ? kNB : 1)
  for (int k = 0; k < kNB; ++k) {
    cplx a[TR], b[TC];
#pragma unroll
    for (int i = 0; i < TR; ++i) a[i] = C.Pp[pix(ra[i], k)];
#pragma unroll
    for (int j = 0; j < TC; ++j) b[j] = C.Pp[pix(ca[j], k)];
#pragma unroll
    for (int i = 0; i < TR; ++i)
#pragma unroll
      for (int j = 0; j < TC; ++j) acc[i][j] = cmsub(acc[i][j], a[i], b[j]);
  }
}

typedef double v4d __attribute__((ext_vector_type(4)));

// One 16 x 16 tile of the stage window on the matrix cores: window rows
// 16 TI + (lane >> 4) + 4 v (v < 4) and column 16 TJ + (lane & 15) -- the
// v_mfma_f64_16x16x4f64 accumulator layout.  out[v] = S - P with S the
// current value (ring, or the staged batch for the window's last kNB rows; 0
// outside the window's lower triangle) and P = sum_k L_rk L_ck over panel
// J-1's kNB columns, formed from zero: the real part as the L_re L_re
// products over k, then the -L_im L_im ones, the imaginary part as L_re L_im
// then L_im L_re, each chain accumulated by the matrix core, and subtracted
// from S once.  (A product block formed first and subtracted once is as
// close to the exact solution as the reference's one-product-at-a-time
// update on the worst-conditioned systems; accumulating the products into S
// itself is not.)  Every lane takes part (the matrix cores need the whole
// wavefront); the caller masks the stores.
__device__ __forceinline__ void mfma_tile(const Stage& C, const cplx* Tt, int TI, int TJ, int lane, cplx out[4]) {
  const int col = lane & 15, q = lane >> 4;
  const int cw = 16 * TJ + col;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int rw = 16 * TI + q + 4 * v;
    cplx sv = mkc(0.0, 0.0);
    if (tile_ok(C, rw, cw)) {
      const int ib = rw - (C.mb - kNB);
      sv = ib >= 0 ? staged(C.sv_old, C.so_old, ib, cw - rw + C.mb) : Tt[C.ring.pos(rw, cw)];
    }
    out[v] = sv;
  }
  if (!C.apply) return;
  // operand rows of panel J-1 (stage rows + kNB), clamped for the padding
  // rows of the last tile row (their results are not stored)
  const int ra = min(16 * TI + col + kNB, C.W - 1), ca = min(cw + kNB, C.W - 1);
  const cplx a0 = C.Pp[pix(ra, q)], a1 = C.Pp[pix(ra, q + 4)];
  const cplx b0 = C.Pp[pix(ca, q)], b1 = C.Pp[pix(ca, q + 4)];
  v4d pr = {0.0, 0.0, 0.0, 0.0}, pi = {0.0, 0.0, 0.0, 0.0};
  pr = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.re, b0.re, pr, 0, 0, 0);
  pr = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.re, b1.re, pr, 0, 0, 0);
  pr = __builtin_amdgcn_mfma_f64_16x16x4f64(-a0.im, b0.im, pr, 0, 0, 0);
  pr = __builtin_amdgcn_mfma_f64_16x16x4f64(-a1.im, b1.im, pr, 0, 0, 0);
  pi = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.re, b0.im, pi, 0, 0, 0);
  pi = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.re, b1.im, pi, 0, 0, 0);
  pi = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.im, b0.re, pi, 0, 0, 0);
  pi = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.im, b1.re, pi, 0, 0, 0);
#pragma unroll
  for (int v = 0; v < 4; ++v) out[v] = mkc(out[v].re - pr[v], out[v].im - pi[v]);
}

// lower-triangle tile index x -> (ti, tj), x = ti (ti + 1) / 2 + tj
__device__ __forceinline__ void tri_tile(int x, int& ti, int& tj) {
  int i = (int)((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
  while ((i * (i + 1)) / 2 > x) --i;
  while (((i + 1) * (i + 2)) / 2 <= x) ++i;
  ti = i;
  tj = x - (i * (i + 1)) / 2;
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
// f_r = 1/sqrt(|a_rr|) of every row of every system of the batch (scalab
// :56-63; errnr 27 when |a_rr| = 0), or 1/sqrt(a_rr) for the DC twin (scaldc
// :53-61; errnr 27 when a_rr <= 0).
__global__ void crt_scale_kernel(FactorArgs A) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)A.nsys * A.n) return;
  const int s = (int)(idx / A.n), r = (int)(idx % A.n);
  const int sys = (A.sys_first + s) * A.sys_stride;
  const int k = sys / A.eanz, l = sys % A.eanz;
  const int qd = A.pat_diag[r];
  cplx v = ldc(A.av + ((size_t)k * A.n + r) * A.R + qd);
  const int b1 = A.bptr[r + 1];
  for (int p = A.bptr[r]; p < b1; ++p)
    if (A.bslot[p] == qd) v = cadd(v, bnd_term(A, k, l, A.bcode[p], A.bimn[p]));
  // DC (scaldc :53-59): a_rr is real and must be positive
  const double ad = A.dc ? v.re : cabs_(v);
  double f = 1.0 / sqrt(ad);
  if (ad <= 0.0) {  // scalab :56-61
    f = 1.0;
    if (atomicCAS(A.err, 0, 27) == 0) A.err[1] = sys;
  }
  A.fak[(size_t)s * A.n + r] = f;
}

// Scaled band entries of every row of every system of the batch: a_rr/|a_rr|
// on the diagonal, (f_r f_c) a_rc elsewhere (scalab :58-73).  One thread per
// entry (system, row, slot), so that consecutive threads read av and write S
// at consecutive addresses.
__global__ void crt_rows_kernel(FactorArgs A) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)A.nsys * A.n * A.R) return;
  const long rr = idx / A.R;
  const int q = (int)(idx - rr * A.R);
  const int s = (int)(rr / A.n), r = (int)(rr % A.n);
  const int sys = (A.sys_first + s) * A.sys_stride;
  const int k = sys / A.eanz, l = sys % A.eanz;
  const double* fk = A.fak + (size_t)s * A.n;
  const int t = A.pat_t[(size_t)r * A.R + q];
  cplx v = mkc(0.0, 0.0);
  if (t >= 0) {
    v = ldc(A.av + ((size_t)k * A.n + r) * A.R + q);
    const int b0 = A.bptr[r], b1 = A.bptr[r + 1];
    for (int p = b0; p < b1; ++p)
      if (A.bslot[p] == q) v = cadd(v, bnd_term(A, k, l, A.bcode[p], A.bimn[p]));
    if (t == A.mb) v = A.dc ? mkc(1.0, 0.0) : cdiv(v, mkc(cabs_(v), 0.0));  // scaldc :60
    else v = rmul(fk[r] * fk[r - A.mb + t], v);
  }
  A.S[idx] = make_double2(v.re, v.im);
}

// Right-hand side of every system of the batch when boundary values are
// prescribed: kompab's source and sink (:162, :183; the sink wins a tie), zero
// on type-13 rows (randb :49), then randb2 (src/randb2.f90:30-71): -value on a
// prescribed row, and on a free row the moved terms value_k * a(k,j) in file
// order, each a(k,j) summed from its element terms in element order; scaled by
// f (scalab :65).  The DC twin (randbdc2) takes the real parts of the values.
__global__ void crt_rhs_kernel(FactorArgs A) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)A.nsys * A.n) return;
  const int s = (int)(idx / A.n), r = (int)(idx % A.n);
  const int sys = (A.sys_first + s) * A.sys_stride;
  const int k = sys / A.eanz, l = sys % A.eanz;
  const int rel = A.enr[l], rsink = A.lsink ? A.nsink : -1;
  cplx b = mkc(r == rsink ? 1.0 : (r == rel ? -1.0 : 0.0), 0.0);
  if (A.isdir && A.isdir[r]) b = mkc(0.0, 0.0);
  auto value = [&](int ir) {
    const cplx v = ldc(A.rwd + ir);
    return A.dc ? mkc(v.re, 0.0) : v;
  };
  const int last = A.rb_last[r];
  if (last >= 0) {
    const cplx v = value(last);
    b = mkc(-v.re, -v.im);
  } else {
    int p = A.mv_ptr[r];
    const int p1 = A.mv_ptr[r + 1];
    while (p < p1) {
      const int ir = A.mv_ir[p];
      cplx a = mkc(0.0, 0.0);
      for (; p < p1 && A.mv_ir[p] == ir; ++p) {
        const int e = A.mv_e[p];
        if (e >= 0) a = cadd(a, rmul(A.elbg[((size_t)k * A.elanz + e) * A.nimn + A.mv_imn[p]], ldc(A.sigma + e)));
        else a = cadd(a, bnd_term(A, k, l, e, A.mv_imn[p]));
      }
      b = cadd(b, cmul(value(ir), a));
    }
  }
  stc(A.rhs + (size_t)s * A.n + r, rmul(A.fak[(size_t)s * A.n + r], b));
}

// Stage barrier of the factor kernel: LDS only when the trailing window lives
// in LDS; with the window in global memory (GR) the workgroup-scope fences
// also order the ring's global loads and stores between the waves.
template <bool GR>
__device__ __forceinline__ void stage_barrier() {
  if constexpr (GR) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  } else {
    lds_barrier();
  }
}

// ---------------------------------------------------------------------------
// Lookahead band factorisation.  ROWS: panel rows per lane of wave 0
// (W = mb + kNB <= 64 ROWS); MAXT: trailing tiles per thread of the other
// waves; NT threads per workgroup: 256 (two workgroups per CU) or, when the
// LDS of one workgroup leaves no room for a second (mb above ~98), 512, so
// that the CU still holds 8 wavefronts (7 for the look-ahead and trailing
// update).  GR: the trailing-window ring lives in a global scratch of half x
// (mb + 1) entries per workgroup (A.ring) instead of LDS, for bands whose
// ring exceeds the CU's LDS (mb above 124); the panels and staged rows stay
// in LDS.  MF: the look-ahead and trailing updates run on the matrix cores in
// 16 x 16 tiles (mfma_tile): in phase A the tiles of window columns [0, 16)
// (panel J's columns to the panel buffer, columns [kNB, 16) back to the
// ring), in phase B the rest; MAXT is unused.
// The trailing 4x4 tiles of this thread for band mb (window rows and columns
// from 8 on; tile x of the lower triangle goes to thread 64 + x mod (NT - 64)),
// -1 where the thread has none.
template <int MAXT, int NT>
__device__ __forceinline__ void trailing_tiles(int mb, int (&rs_r0)[MAXT], int (&rs_c0)[MAXT]) {
  const int tid = threadIdx.x, wave = tid >> 6, x3 = tid - 64;
  const int tw = (mb + 3) / 4;
  const int nrest = tw > 2 ? ((tw - 2) * (tw - 1)) / 2 : 0;  // 4x4 tiles, tile columns 2..
#pragma unroll
  for (int q = 0; q < MAXT; ++q) {
    rs_r0[q] = -1;
    rs_c0[q] = 0;
    const int x = x3 + (NT - 64) * q;
    if (wave > 0 && x < nrest) {
      int ti, tj;
      tri_tile(x, ti, tj);
      rs_r0[q] = 4 * (ti + 2);
      rs_c0[q] = 4 * (tj + 2);
    }
  }
}

template <int ROWS, int MAXT, int NT, bool GR, bool MF>
__device__ __forceinline__ void factor_system(const FactorArgs& A, int s, double2* smem, PhaseClock& clk) {
  const int n = A.n, mb = A.mb, R = A.R;
  const int W = mb + kNB, M1 = mb + 1, half = (mb + 1) / 2;
  // half x M1 ring slots
  cplx* Tt = GR ? reinterpret_cast<cplx*>(A.ring) + (size_t)blockIdx.x * A.ring_stride : reinterpret_cast<cplx*>(smem);
  cplx* P0 = GR ? reinterpret_cast<cplx*>(smem) : Tt + half * M1;  // 2 panels of W x kNB
  cplx* sv = P0 + 2 * W * kNB;                // kStg batches of kNB x kMaxR values
  cplx* xb = sv + kStg * kNB * kMaxR;         // kNB: column broadcast of the panel
  cplx* xd = xb + kNB;                        // kNB: the panel's pivots d_c
  cplx* xi = xd + kNB;                        // kNB: 1/d_c
  cplx* xl = xi + kNB;                        // kNB: L_cc = sqrt(d_c)
  int16_t* so = reinterpret_cast<int16_t*>(xl + kNB);  // kStg x kNB x kMaxR offsets
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int x3 = tid - 64;                    // thread index within waves 1-3
  const int nst = (n + kNB - 1) / kNB;
  const int tw = (mb + 3) / 4;                          // 4x4 tiles per window side
  // look-ahead: 2x2 tiles over window rows [0, mb) x columns [0, kNB), column
  // by column of tiles; tile column j holds the tile rows i >= j
  const int t2 = (mb + 1) / 2;
  int nla = 0;
#pragma unroll
  for (int j = 0; j < kNB / 2; ++j) nla += max(t2 - j, 0);
  // trailing 4x4 tiles of this thread
  int rs_r0[MAXT], rs_c0[MAXT];
  trailing_tiles<MAXT, NT>(mb, rs_r0, rs_c0);

  {
    const int sys = (A.sys_first + s) * A.sys_stride;
    const int l = sys % A.eanz;
    double2* Lg = A.L + (size_t)s * n * M1;
    double2* yg = A.y + (size_t)s * n;
    const double2* Sg = A.S + (size_t)s * n * R;
    const double* fk = A.fak + (size_t)s * n;

    // ---- prologue: ring zeroed, rows [0, mb - kNB) in place, batches
    // [mb - kNB, mb) and [mb, mb + kNB) staged, wave 0's right-hand side
    for (int i = tid; i < half * M1; i += NT) Tt[i] = mkc(0.0, 0.0);
    stage_barrier<GR>();
    {
      Ring ring0{mb, M1, half, 0};
      for (int r = tid; r < min(mb - kNB, n); r += NT)
        for (int q = 0; q < R; ++q) {
          cplx v;
          int t;
          row_slot(A, Sg, r, q, v, t);
          if (t >= 0) Tt[ring0.pos(r, r - mb + t)] = v;
        }
      for (int x = tid; x < 2 * kNB * kMaxR; x += NT) {
        const int b = x / (kNB * kMaxR), i = (x / kMaxR) % kNB, q = x % kMaxR;
        const int r = (b == 0 ? mb - kNB : mb) + i;
        const int slot = (b == 0 ? kStg - 1 : 0) * kNB * kMaxR + i * kMaxR + q;
        cplx v;
        int t;
        row_slot(A, Sg, r, q, v, t);
        sv[slot] = v;
        so[slot] = (int16_t)t;
      }
    }
    // right-hand side b(enr(l)) = -1, b(nsink) = +1 (kompab :162,:183), zero on
    // Dirichlet rows (randb :49), scaled by f (scalab :65)
    const int rel = A.enr[l];
    const int rsink = A.lsink ? A.nsink : -1;
    double be = 0.0, bs = 0.0;
    if (wave == 0) {
      if (!(A.isdir && A.isdir[rel])) be = -fk[rel];
      if (rsink >= 0 && !(A.isdir && A.isdir[rsink])) bs = fk[rsink];
    }
    // prescribed boundary values: the right-hand side comes from crt_rhs_kernel
    const double2* Bs = A.rhs ? A.rhs + (size_t)s * n : nullptr;
    auto rhs = [&](int r) -> cplx {
      if (Bs) return ldc(Bs + min(max(r, 0), n - 1));
      return mkc(r == rsink ? bs : (r == rel ? be : 0.0), 0.0);
    };
    cplx yv[ROWS];
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
      const int rw = lane + 64 * q;
      yv[q] = csel(wave == 0 && rw < W && rw < n, rhs(rw), mkc(0.0, 0.0));
    }
    stage_barrier<GR>();
    clk.mark(0);

    for (int J = 0; J < nst; ++J) {
      const int o = J * kNB;
      const int nbc = min(kNB, n - o);
      Stage C;
      C.n = n;
      C.mb = mb;
      C.W = W;
      C.o = o;
      C.ring = Ring{mb, M1, half, o % M1};
      C.apply = J > 0;
      C.Pp = P0 + ((J + 1) & 1) * W * kNB;
      const int bo = (J + kStg - 1) % kStg, bc = J % kStg, bn = (J + 1) % kStg;
      C.sv_old = sv + bo * kNB * kMaxR;
      C.so_old = so + bo * kNB * kMaxR;
      cplx* Pc = P0 + (J & 1) * W * kNB;

      // The tile coordinates are loop invariant; re-declaring them as changed
      // each stage keeps the position arithmetic of 16 entries per tile from
      // being hoisted out of the stage loop (it would hold ~100 registers).
#pragma unroll
      for (int q = 0; q < MAXT; ++q) asm volatile("" : "+v"(rs_r0[q]), "+v"(rs_c0[q]));

      // ---- phase A: waves 1-3 apply panel J-1 to panel J's columns; wave 0
      // copies the entering batch J into panel J's last kNB rows
      if (MF && wave > 0) {
        __builtin_amdgcn_s_setprio(2);
        const int T = (mb + 15) / 16;
        for (int TI = wave - 1; TI < T; TI += NT / 64 - 1) {
          cplx out[4];
          mfma_tile(C, Tt, TI, 0, lane, out);
          const int cw = lane & 15;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int rw = 16 * TI + (lane >> 4) + 4 * v;
            if (tile_ok(C, rw, cw)) {
              if (cw < kNB) Pc[pix(rw, cw)] = out[v];
              else Tt[C.ring.pos(rw, cw)] = out[v];
            }
          }
        }
      } else if (wave > 0) {
        // the look-ahead precedes the panel on the critical path as well
        __builtin_amdgcn_s_setprio(2);
        for (int x = x3; x < nla; x += NT - 64) {
          int j = 0, i = x;
#pragma unroll
          for (int jj = 0; jj < kNB / 2 - 1; ++jj)
            if (i >= t2 - j) {
              i -= t2 - j;
              ++j;
            }
          const int rw0 = 2 * (i + j), cw0 = 2 * j;
          cplx acc[2][2];
          tile_update<2, 2>(C, Tt, rw0, cw0, acc);
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
              if (tile_ok(C, rw0 + a, cw0 + b)) Pc[pix(rw0 + a, cw0 + b)] = acc[a][b];
        }
      } else {
        const int i = lane / kNB, cw = lane % kNB, rw = mb + i;
        const bool ok = o + rw < n && cw >= i;
        Pc[pix(rw, cw)] = csel(ok, staged(sv + bc * kNB * kMaxR, so + bc * kNB * kMaxR, i, cw - i), mkc(0.0, 0.0));
      }
      __builtin_amdgcn_s_setprio(0);
      clk.mark(0);
      stage_barrier<GR>();
      clk.mark(1);

      // ---- phase B
      if (wave == 0) {
        // the panel is the stage's critical path: its wave takes issue
        // precedence over the other workgroup's trailing wave on its SIMD
        __builtin_amdgcn_s_setprio(3);
        // panel J in registers: lane holds panel rows lane + 64 q
        const int rows = min(W, n - o);
        cplx v[ROWS][kNB];
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
          const int rw = lane + 64 * q;
#pragma unroll
          for (int k = 0; k < kNB; ++k) {
            const bool ok = rw < rows && k < nbc && k <= rw && rw - k <= mb;
            v[q][k] = csel(ok, Pc[pix(min(rw, W - 1), k)], mkc(0.0, 0.0));
          }
        }
        // Column c, in the L D L^T form of the same elimination: with d the
        // pivot and s_rc the column's updated entries, l_rc = s_rc / d and the
        // later columns take -l_rc s_c2c (= -L_rc L_c2c).  The next pivot,
        // d_{c+1} = s_{c+1,c+1} - l_{c+1,c} s_{c+1,c}, needs no square root, so
        // the column-to-column dependency is a reciprocal and two products;
        // L_cc = sqrt(d), L_rc = l_rc L_cc and y_c = y'_c / L_cc follow off it.
        // lrc is 0 for rows at or above c, so every update is unconditional
        // (a - 0 * x leaves a unchanged).
        // Columns past the matrix end (the last stage only) are skipped; their
        // pivot slots hold 1 so the square roots below stay finite.  The loop leaves
        // l_rc in place (0 at and above the diagonal); the square roots of all
        // kNB pivots are taken at once afterwards (lane c) and turn l into L
        // and y' into y.  Broadcasts go through LDS (one wavefront: its LDS
        // operations complete in issue order), except the next pivot and y'_c.
        cplx piv = rlc(v[0][0], 0);
        bool zero_pivot = false;
#pragma unroll
        for (int c = 0; c < kNB; ++c) {
          if (c < nbc) {
            // chol :66 -- |s| <= 0 is an exact zero (a NaN pivot passes, as
            // there); choldc :66 -- s <= 0 for the real DC pivots
            zero_pivot |= A.dc ? piv.re <= 0.0 : (piv.re == 0.0 && piv.im == 0.0);
            const cplx inv = crecip(piv);
            xd[c] = piv;  // every lane stores the same value
            xi[c] = inv;
            cplx lrc[ROWS];
#pragma unroll
            for (int q = 0; q < ROWS; ++q) lrc[q] = csel(lane + 64 * q > c, cmul(v[q][c], inv), mkc(0.0, 0.0));
            const int cn = c + 1 < kNB ? c + 1 : kNB - 1;
            cplx pivn = mkc(1.0, 0.0);
            if (c + 1 < nbc) pivn = rlc(cmsub(v[0][cn], lrc[0], v[0][c]), cn);
            const cplx yp = rlc(yv[0], c);  // y'_c, final
            // s_{c2,c} of the panel's diagonal block from lanes c+1..kNB-1
            if (lane > c && lane < kNB) xb[lane] = v[0][c];
#pragma unroll
            for (int c2 = c + 1; c2 < kNB; ++c2) {
              if (c2 < nbc) {
                const cplx s2 = xb[c2];
#pragma unroll
                for (int q = 0; q < ROWS; ++q) v[q][c2] = cmsub(v[q][c2], lrc[q], s2);
              }
            }
#pragma unroll
            for (int q = 0; q < ROWS; ++q) {
              yv[q] = cmsub(yv[q], lrc[q], yp);
              v[q][c] = lrc[q];
            }
            piv = pivn;
          } else {
            xd[c] = mkc(1.0, 0.0);
            xi[c] = mkc(1.0, 0.0);
          }
        }
        const int lk = min(lane, kNB - 1);
        const cplx dd = xd[lk], invd = xi[lk];
        const cplx lcc = csqrt_(dd);
        if (lane < kNB) xl[lane] = lcc;
#pragma unroll
        for (int c = 0; c < kNB; ++c) {
          const cplx lc = xl[c];
#pragma unroll
          for (int q = 0; q < ROWS; ++q) v[q][c] = cmul(v[q][c], lc);
          v[0][c] = csel(lane == c, lc, v[0][c]);
        }
        yv[0] = csel(lane < kNB, cmul(cmul(yv[0], lcc), invd), yv[0]);
        if (zero_pivot && lane == 0 && atomicCAS(A.err, 0, 28) == 0) A.err[1] = sys;
        // L of panel J: back to the panel buffer (stage J+1's operand) and out
        // to the band rows; y of its columns is final
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
          const int rw = lane + 64 * q;
          if (rw < W) {
#pragma unroll
            for (int k = 0; k < kNB; ++k) {
              const bool ok = rw < rows && k < nbc && k <= rw && rw - k <= mb;
              Pc[pix(rw, k)] = csel(ok, v[q][k], mkc(0.0, 0.0));
              if (ok) Lg[(size_t)(o + rw) * M1 + (k - rw + mb)] = make_double2(v[q][k].re, v[q][k].im);
            }
          }
        }
        if (lane < nbc) yg[o + lane] = make_double2(yv[0].re, yv[0].im);
        // y of rows [o + kNB, o + W) moves up kNB rows; entering rows get b
        const int src = (lane + kNB) & 63;
        cplx ny[ROWS];
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
          cplx a, b = mkc(0.0, 0.0);
          a.re = __shfl(yv[q].re, src, 64);
          a.im = __shfl(yv[q].im, src, 64);
          if (q + 1 < ROWS) {
            b.re = __shfl(yv[q + 1].re, src, 64);
            b.im = __shfl(yv[q + 1].im, src, 64);
          }
          ny[q] = csel(lane + kNB < 64, a, b);
          const int rwn = lane + 64 * q;
          if (rwn >= mb) {
            const int r = o + kNB + rwn;
            ny[q] = csel(rwn < W && r < n, rhs(r), mkc(0.0, 0.0));
          }
        }
#pragma unroll
        for (int q = 0; q < ROWS; ++q) yv[q] = ny[q];
        __builtin_amdgcn_s_setprio(0);
        clk.mark(2);
      } else {
        // batch J+1 = rows [o + kNB + mb, o + 2 kNB + mb): loads issued first,
        // stored to LDS after the tiles
        cplx nv = mkc(0.0, 0.0);
        int nt = -1;
        if (x3 < kNB * kMaxR) row_slot(A, Sg, o + kNB + mb + x3 / kMaxR, x3 % kMaxR, nv, nt);
        // the rest of the window: panel J-1's update in place
        if constexpr (MF) {
          const int T = (mb + 15) / 16;
          int cnt = 0;
          for (int TJ = 1; TJ < T; ++TJ)
            for (int TI = TJ; TI < T; ++TI, ++cnt) {
              if (cnt % (NT / 64 - 1) != wave - 1) continue;
              cplx out[4];
              mfma_tile(C, Tt, TI, TJ, lane, out);
              const int cw = 16 * TJ + (lane & 15);
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const int rw = 16 * TI + (lane >> 4) + 4 * v;
                if (tile_ok(C, rw, cw)) Tt[C.ring.pos(rw, cw)] = out[v];
              }
            }
        }
#pragma unroll
        for (int q = 0; q < MAXT; ++q) {
          if (!MF && rs_r0[q] >= 0) {
            cplx acc[4][4];
            tile_update<4, 4>(C, Tt, rs_r0[q], rs_c0[q], acc);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
              for (int j = 0; j < 4; ++j)
                if (tile_ok(C, rs_r0[q] + i, rs_c0[q] + j)) Tt[C.ring.pos(rs_r0[q] + i, rs_c0[q] + j)] = acc[i][j];
          }
        }
        if (x3 < kNB * kMaxR) {
          sv[bn * kNB * kMaxR + x3] = nv;
          so[bn * kNB * kMaxR + x3] = (int16_t)nt;
        }
        clk.mark(2);
      }
      stage_barrier<GR>();
      clk.mark(3);
    }
  }
}

template <int ROWS, int MAXT, int NT, bool GR, bool MF>
__global__ __launch_bounds__(NT, 512 / NT) void crt_factor_kernel(FactorArgs A) {
  extern __shared__ double2 smem[];
  const int n = A.n, mb = A.mb, R = A.R;
  const int W = mb + kNB, M1 = mb + 1, half = (mb + 1) / 2;
  // half x M1 ring slots
  cplx* Tt = GR ? reinterpret_cast<cplx*>(A.ring) + (size_t)blockIdx.x * half * M1 : reinterpret_cast<cplx*>(smem);
  cplx* P0 = GR ? reinterpret_cast<cplx*>(smem) : Tt + half * M1;  // 2 panels of W x kNB
  cplx* sv = P0 + 2 * W * kNB;                // kStg batches of kNB x kMaxR values
  cplx* xb = sv + kStg * kNB * kMaxR;         // kNB: column broadcast of the panel
  cplx* xd = xb + kNB;                        // kNB: the panel's pivots d_c
  cplx* xi = xd + kNB;                        // kNB: 1/d_c
  cplx* xl = xi + kNB;                        // kNB: L_cc = sqrt(d_c)
  int16_t* so = reinterpret_cast<int16_t*>(xl + kNB);  // kStg x kNB x kMaxR offsets
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int x3 = tid - 64;                    // thread index within waves 1-3
  const int nst = (n + kNB - 1) / kNB;
  const int tw = (mb + 3) / 4;                          // 4x4 tiles per window side
  // look-ahead: 2x2 tiles over window rows [0, mb) x columns [0, kNB), column
  // by column of tiles; tile column j holds the tile rows i >= j
  const int t2 = (mb + 1) / 2;
  int nla = 0;
#pragma unroll
  for (int j = 0; j < kNB / 2; ++j) nla += max(t2 - j, 0);
  const int nrest = tw > 2 ? ((tw - 2) * (tw - 1)) / 2 : 0;  // 4x4 tiles, tile columns 2..

  // trailing 4x4 tiles of this thread (fixed for the kernel)
  int rs_r0[MAXT], rs_c0[MAXT];
#pragma unroll
  for (int q = 0; q < MAXT; ++q) {
    rs_r0[q] = -1;
    rs_c0[q] = 0;
    const int x = x3 + (NT - 64) * q;
    if (wave > 0 && x < nrest) {
      int ti, tj;
      tri_tile(x, ti, tj);
      rs_r0[q] = 4 * (ti + 2);
      rs_c0[q] = 4 * (tj + 2);
    }
  }

  PhaseClock clk;
  clk.start(tid == 0 ? A.prof : (tid == 64 ? A.prof : nullptr), tid == 0 ? 0 : 4);

  for (int s = blockIdx.x; s < A.nsys; s += gridDim.x) {
    const int sys = (A.sys_first + s) * A.sys_stride;
    const int l = sys % A.eanz;
    double2* Lg = A.L + (size_t)s * n * M1;
    double2* yg = A.y + (size_t)s * n;
    const double2* Sg = A.S + (size_t)s * n * R;
    const double* fk = A.fak + (size_t)s * n;

    // ---- prologue: ring zeroed, rows [0, mb - kNB) in place, batches
    // [mb - kNB, mb) and [mb, mb + kNB) staged, wave 0's right-hand side
    for (int i = tid; i < half * M1; i += NT) Tt[i] = mkc(0.0, 0.0);
    stage_barrier<GR>();
    {
      Ring ring0{mb, M1, half, 0};
      for (int r = tid; r < min(mb - kNB, n); r += NT)
        for (int q = 0; q < R; ++q) {
          cplx v;
          int t;
          row_slot(A, Sg, r, q, v, t);
          if (t >= 0) Tt[ring0.pos(r, r - mb + t)] = v;
        }
      for (int x = tid; x < 2 * kNB * kMaxR; x += NT) {
        const int b = x / (kNB * kMaxR), i = (x / kMaxR) % kNB, q = x % kMaxR;
        const int r = (b == 0 ? mb - kNB : mb) + i;
        const int slot = (b == 0 ? kStg - 1 : 0) * kNB * kMaxR + i * kMaxR + q;
        cplx v;
        int t;
        row_slot(A, Sg, r, q, v, t);
        sv[slot] = v;
        so[slot] = (int16_t)t;
      }
    }
    // right-hand side b(enr(l)) = -1, b(nsink) = +1 (kompab :162,:183), zero on
    // Dirichlet rows (randb :49), scaled by f (scalab :65)
    const int rel = A.enr[l];
    const int rsink = A.lsink ? A.nsink : -1;
    double be = 0.0, bs = 0.0;
    if (wave == 0) {
      if (!(A.isdir && A.isdir[rel])) be = -fk[rel];
      if (rsink >= 0 && !(A.isdir && A.isdir[rsink])) bs = fk[rsink];
    }
    // prescribed boundary values: the right-hand side comes from crt_rhs_kernel
    const double2* Bs = A.rhs ? A.rhs + (size_t)s * n : nullptr;
    auto rhs = [&](int r) -> cplx {
      if (Bs) return ldc(Bs + min(max(r, 0), n - 1));
      return mkc(r == rsink ? bs : (r == rel ? be : 0.0), 0.0);
    };
    cplx yv[ROWS];
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
      const int rw = lane + 64 * q;
      yv[q] = csel(wave == 0 && rw < W && rw < n, rhs(rw), mkc(0.0, 0.0));
    }
    stage_barrier<GR>();
    clk.mark(0);

    for (int J = 0; J < nst; ++J) {
      const int o = J * kNB;
      const int nbc = min(kNB, n - o);
      Stage C;
      C.n = n;
      C.mb = mb;
      C.W = W;
      C.o = o;
      C.ring = Ring{mb, M1, half, o % M1};
      C.apply = J > 0;
      C.Pp = P0 + ((J + 1) & 1) * W * kNB;
      const int bo = (J + kStg - 1) % kStg, bc = J % kStg, bn = (J + 1) % kStg;
      C.sv_old = sv + bo * kNB * kMaxR;
      C.so_old = so + bo * kNB * kMaxR;
      cplx* Pc = P0 + (J & 1) * W * kNB;

      // The tile coordinates are loop invariant; re-declaring them as changed
      // each stage keeps the position arithmetic of 16 entries per tile from
      // being hoisted out of the stage loop (it would hold ~100 registers).
#pragma unroll
      for (int q = 0; q < MAXT; ++q) asm volatile("" : "+v"(rs_r0[q]), "+v"(rs_c0[q]));

      // ---- phase A: waves 1-3 apply panel J-1 to panel J's columns; wave 0
      // copies the entering batch J into panel J's last kNB rows
      if (MF && wave > 0) {
        __builtin_amdgcn_s_setprio(2);
        const int T = (mb + 15) / 16;
        for (int TI = wave - 1; TI < T; TI += NT / 64 - 1) {
          cplx out[4];
          mfma_tile(C, Tt, TI, 0, lane, out);
          const int cw = lane & 15;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int rw = 16 * TI + (lane >> 4) + 4 * v;
            if (tile_ok(C, rw, cw)) {
              if (cw < kNB) Pc[pix(rw, cw)] = out[v];
              else Tt[C.ring.pos(rw, cw)] = out[v];
            }
          }
        }
      } else if (wave > 0) {
        // the look-ahead precedes the panel on the critical path as well
        __builtin_amdgcn_s_setprio(2);
        for (int x = x3; x < nla; x += NT - 64) {
          int j = 0, i = x;
#pragma unroll
          for (int jj = 0; jj < kNB / 2 - 1; ++jj)
            if (i >= t2 - j) {
              i -= t2 - j;
              ++j;
            }
          const int rw0 = 2 * (i + j), cw0 = 2 * j;
          cplx acc[2][2];
          tile_update<2, 2>(C, Tt, rw0, cw0, acc);
#pragma unroll
          for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
              if (tile_ok(C, rw0 + a, cw0 + b)) Pc[pix(rw0 + a, cw0 + b)] = acc[a][b];
        }
      } else {
        const int i = lane / kNB, cw = lane % kNB, rw = mb + i;
        const bool ok = o + rw < n && cw >= i;
        Pc[pix(rw, cw)] = csel(ok, staged(sv + bc * kNB * kMaxR, so + bc * kNB * kMaxR, i, cw - i), mkc(0.0, 0.0));
      }
      __builtin_amdgcn_s_setprio(0);
      clk.mark(0);
      stage_barrier<GR>();
      clk.mark(1);

      // ---- phase B
      if (wave == 0) {
        // the panel is the stage's critical path: its wave takes issue
        // precedence over the other workgroup's trailing wave on its SIMD
        __builtin_amdgcn_s_setprio(3);
        // panel J in registers: lane holds panel rows lane + 64 q
        const int rows = min(W, n - o);
        cplx v[ROWS][kNB];
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
          const int rw = lane + 64 * q;
#pragma unroll
          for (int k = 0; k < kNB; ++k) {
            const bool ok = rw < rows && k < nbc && k <= rw && rw - k <= mb;
            v[q][k] = csel(ok, Pc[pix(min(rw, W - 1), k)], mkc(0.0, 0.0));
          }
        }
        // Column c, in the L D L^T form of the same elimination: with d the
        // pivot and s_rc the column's updated entries, l_rc = s_rc / d and the
        // later columns take -l_rc s_c2c (= -L_rc L_c2c).  The next pivot,
        // d_{c+1} = s_{c+1,c+1} - l_{c+1,c} s_{c+1,c}, needs no square root, so
        // the column-to-column dependency is a reciprocal and two products;
        // L_cc = sqrt(d), L_rc = l_rc L_cc and y_c = y'_c / L_cc follow off it.
        // lrc is 0 for rows at or above c, so every update is unconditional
        // (a - 0 * x leaves a unchanged).
        // Columns past the matrix end (the last stage only) are skipped; their
        // pivot slots hold 1 so the square roots below stay finite.  The loop leaves
        // l_rc in place (0 at and above the diagonal); the square roots of all
        // kNB pivots are taken at once afterwards (lane c) and turn l into L
        // and y' into y.  Broadcasts go through LDS (one wavefront: its LDS
        // operations complete in issue order), except the next pivot and y'_c.
        cplx piv = rlc(v[0][0], 0);
        bool zero_pivot = false;
#pragma unroll
        for (int c = 0; c < kNB; ++c) {
          if (c < nbc) {
            // chol :66 -- |s| <= 0 is an exact zero (a NaN pivot passes, as
            // there); choldc :66 -- s <= 0 for the real DC pivots
            zero_pivot |= A.dc ? piv.re <= 0.0 : (piv.re == 0.0 && piv.im == 0.0);
            const cplx inv = crecip(piv);
            xd[c] = piv;  // every lane stores the same value
            xi[c] = inv;
            cplx lrc[ROWS];
#pragma unroll
            for (int q = 0; q < ROWS; ++q) lrc[q] = csel(lane + 64 * q > c, cmul(v[q][c], inv), mkc(0.0, 0.0));
            const int cn = c + 1 < kNB ? c + 1 : kNB - 1;
            cplx pivn = mkc(1.0, 0.0);
            if (c + 1 < nbc) pivn = rlc(cmsub(v[0][cn], lrc[0], v[0][c]), cn);
            const cplx yp = rlc(yv[0], c);  // y'_c, final
            // s_{c2,c} of the panel's diagonal block from lanes c+1..kNB-1
            if (lane > c && lane < kNB) xb[lane] = v[0][c];
#pragma unroll
            for (int c2 = c + 1; c2 < kNB; ++c2) {
              if (c2 < nbc) {
                const cplx s2 = xb[c2];
#pragma unroll
                for (int q = 0; q < ROWS; ++q) v[q][c2] = cmsub(v[q][c2], lrc[q], s2);
              }
            }
#pragma unroll
            for (int q = 0; q < ROWS; ++q) {
              yv[q] = cmsub(yv[q], lrc[q], yp);
              v[q][c] = lrc[q];
            }
            piv = pivn;
          } else {
            xd[c] = mkc(1.0, 0.0);
            xi[c] = mkc(1.0, 0.0);
          }
        }
        const int lk = min(lane, kNB - 1);
        const cplx dd = xd[lk], invd = xi[lk];
        const cplx lcc = csqrt_(dd);
        if (lane < kNB) xl[lane] = lcc;
#pragma unroll
        for (int c = 0; c < kNB; ++c) {
          const cplx lc = xl[c];
#pragma unroll
          for (int q = 0; q < ROWS; ++q) v[q][c] = cmul(v[q][c], lc);
          v[0][c] = csel(lane == c, lc, v[0][c]);
        }
        yv[0] = csel(lane < kNB, cmul(cmul(yv[0], lcc), invd), yv[0]);
        if (zero_pivot && lane == 0 && atomicCAS(A.err, 0, 28) == 0) A.err[1] = sys;
        // L of panel J: back to the panel buffer (stage J+1's operand) and out
        // to the band rows; y of its columns is final
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
          const int rw = lane + 64 * q;
          if (rw < W) {
#pragma unroll
            for (int k = 0; k < kNB; ++k) {
              const bool ok = rw < rows && k < nbc && k <= rw && rw - k <= mb;
              Pc[pix(rw, k)] = csel(ok, v[q][k], mkc(0.0, 0.0));
              if (ok) Lg[(size_t)(o + rw) * M1 + (k - rw + mb)] = make_double2(v[q][k].re, v[q][k].im);
            }
          }
        }
        if (lane < nbc) yg[o + lane] = make_double2(yv[0].re, yv[0].im);
        // y of rows [o + kNB, o + W) moves up kNB rows; entering rows get b
        const int src = (lane + kNB) & 63;
        cplx ny[ROWS];
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
          cplx a, b = mkc(0.0, 0.0);
          a.re = __shfl(yv[q].re, src, 64);
          a.im = __shfl(yv[q].im, src, 64);
          if (q + 1 < ROWS) {
            b.re = __shfl(yv[q + 1].re, src, 64);
            b.im = __shfl(yv[q + 1].im, src, 64);
          }
          ny[q] = csel(lane + kNB < 64, a, b);
          const int rwn = lane + 64 * q;
          if (rwn >= mb) {
            const int r = o + kNB + rwn;
            ny[q] = csel(rwn < W && r < n, rhs(r), mkc(0.0, 0.0));
          }
        }
#pragma unroll
        for (int q = 0; q < ROWS; ++q) yv[q] = ny[q];
        __builtin_amdgcn_s_setprio(0);
        clk.mark(2);
      } else {
        // batch J+1 = rows [o + kNB + mb, o + 2 kNB + mb): loads issued first,
        // stored to LDS after the tiles
        cplx nv = mkc(0.0, 0.0);
        int nt = -1;
        if (x3 < kNB * kMaxR) row_slot(A, Sg, o + kNB + mb + x3 / kMaxR, x3 % kMaxR, nv, nt);
        // the rest of the window: panel J-1's update in place
        if constexpr (MF) {
          const int T = (mb + 15) / 16;
          int cnt = 0;
          for (int TJ = 1; TJ < T; ++TJ)
            for (int TI = TJ; TI < T; ++TI, ++cnt) {
              if (cnt % (NT / 64 - 1) != wave - 1) continue;
              cplx out[4];
              mfma_tile(C, Tt, TI, TJ, lane, out);
              const int cw = 16 * TJ + (lane & 15);
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const int rw = 16 * TI + (lane >> 4) + 4 * v;
                if (tile_ok(C, rw, cw)) Tt[C.ring.pos(rw, cw)] = out[v];
              }
            }
        }
#pragma unroll
        for (int q = 0; q < MAXT; ++q) {
          if (!MF && rs_r0[q] >= 0) {
            cplx acc[4][4];
            tile_update<4, 4>(C, Tt, rs_r0[q], rs_c0[q], acc);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
              for (int j = 0; j < 4; ++j)
                if (tile_ok(C, rs_r0[q] + i, rs_c0[q] + j)) Tt[C.ring.pos(rs_r0[q] + i, rs_c0[q] + j)] = acc[i][j];
          }
        }
        if (x3 < kNB * kMaxR) {
          sv[bn * kNB * kMaxR + x3] = nv;
          so[bn * kNB * kMaxR + x3] = (int16_t)nt;
        }
        clk.mark(2);
      }
      stage_barrier<GR>();
      clk.mark(3);
    }
  }
  clk.flush();
}

// The same over the systems of several models whose bands share this form:
// work item w is system items[w].y of model items[w].x (its arguments
// args[items[w].x]); persistent workgroups take the items in list order from
// an atomic counter (zeroed before the launch) until the list is exhausted.
template <int ROWS, int MAXT, int NT, bool GR, bool MF>
__global__ __launch_bounds__(NT, 512 / NT) void crt_factor_many_kernel(const FactorArgs* __restrict__ args,
                                                                      const int2* __restrict__ items, int nitems,
                                                                      int* counter) {
  extern __shared__ double2 smem[];
  __shared__ int next;
  PhaseClock clk;
  clk.start(nullptr, 0);
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(counter, 1);
    __syncthreads();
    const int w = next;
    __syncthreads();
    if (w >= nitems) break;
    const int2 it = items[w];
    const FactorArgs A = args[it.x];
    factor_system<ROWS, MAXT, NT, GR, MF>(A, it.y, smem, clk);
  }
}


template __global__ void crt_factor_kernel<1, 1, 256, false, false>(FactorArgs);
template __global__ void crt_factor_kernel<2, 1, 256, false, false>(FactorArgs);
template __global__ void crt_factor_kernel<2, 2, 256, false, false>(FactorArgs);
template __global__ void crt_factor_kernel<2, 3, 256, false, false>(FactorArgs);
template __global__ void crt_factor_kernel<3, 3, 256, false, false>(FactorArgs);
template __global__ void crt_factor_kernel<2, 1, 512, false, false>(FactorArgs);
template __global__ void crt_factor_kernel<2, 2, 512, false, false>(FactorArgs);
template __global__ void crt_factor_kernel<3, 2, 512, false, false>(FactorArgs);
template __global__ void crt_factor_kernel<3, 5, 512, true, false>(FactorArgs);
template __global__ void crt_factor_kernel<4, 5, 512, true, false>(FactorArgs);
template __global__ void crt_factor_kernel<1, 1, 256, false, true>(FactorArgs);
template __global__ void crt_factor_kernel<2, 1, 256, false, true>(FactorArgs);
template __global__ void crt_factor_kernel<3, 1, 256, false, true>(FactorArgs);
template __global__ void crt_factor_kernel<2, 1, 512, false, true>(FactorArgs);
template __global__ void crt_factor_kernel<3, 1, 512, false, true>(FactorArgs);
template __global__ void crt_factor_kernel<3, 1, 512, true, true>(FactorArgs);
template __global__ void crt_factor_kernel<4, 1, 512, true, true>(FactorArgs);
template __global__ void crt_factor_many_kernel<1, 1, 256, false, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<2, 1, 256, false, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<2, 2, 256, false, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<2, 3, 256, false, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<3, 3, 256, false, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<2, 1, 512, false, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<2, 2, 512, false, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<3, 2, 512, false, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<3, 5, 512, true, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<4, 5, 512, true, false>(const FactorArgs*, const int2*, int, int*);

// ---------------------------------------------------------------------------
// Back substitution, one wavefront per right-hand side.  Lane q (q < 64) keeps
// the working values of the rows j = q + 64 s (mod 64 NS), s < NS, of the band
// window in NS slots (64 NS > mb): slot value y_j + sum_{i>j} L_ij x_i.  Row
// i's pivot lane forms x_i = -(its value) / L_ii, the value is broadcast and
// every lane adds L_ij x_i to its rows j in [i-mb, i) -- for each j the
// updates arrive in descending i, the reference's order.  Rows are taken in
// 64-row blocks.  Slot s of a lane holds the class of block b - s (rows
// lane + 64 ((b - s) mod NS)); after a block the slots rotate by one.  Row
// i's pivot is slot 0 of lane i mod 64, which holds -1/L_ii of its block row
// (formed once per block from a block-wide load, with f_i and the refill
// values y_{i-64 NS}).  The pivot lane's slot 0 takes no update at its own
// step (the load of L_ii is masked), so it keeps row i's final value; every G
// rows the group's lanes keep x (kpot is stored per block, one row per lane)
// and refill slot 0 with y_{i-64 NS}, which row i-64 NS+mb first updates --
// G < 64 NS + 1 - mb keeps the refill ahead of it.  Rows of L are fetched
// DEPTH rows ahead with buffer loads whose offsets for out-of-band, above-n and
// below-0 entries lie past the end of L, so they read zero: no lane-divergent
// branch and no per-value mask in the loop.
template <int DEPTH, int G, int NS>
__device__ __forceinline__ void backsolve_rhs(const BacksolveArgs& B, int r) {
  static_assert(DEPTH % G == 0 && 64 % DEPTH == 0, "row groups tile the block");
  static_assert(NS == 2 || NS == 4, "slot classes");
  constexpr int WIN = 64 * NS;                // rows of the slot window
  constexpr int RMASK = (WIN - 1) * 16;       // (x mod WIN) * 16 on 16-byte units
  const int lane = threadIdx.x;
  const int f = r / B.rhs_per_fac;
  const int sys = (B.fac_first + f) * B.fstride + r % B.rhs_per_fac;
  const int k = sys / B.eanz, l = sys % B.eanz;
  const int n = B.n, mb = B.mb, m1 = mb + 1;
  // n m1 16 < 2^31 (checked by the host): offsets of rows in [0, n + 64) fit,
  // offsets of rows below 0 wrap past the end
  const brsrc Lr = make_rsrc(B.L + (size_t)f * n * m1, (unsigned)(n * m1 * 16));
  const double2* yg = B.y + (size_t)r * n;
  const double* fak = B.fak + (size_t)f * n;
  double2* out = B.kpot + ((size_t)k * B.eanz + l) * n;

  // Entries of row t of L for every slot: slot s's row is j = t - rf, rf =
  // (t mod 64 - lane + 64 s) mod WIN, at band offset mb - rf, in range for 1
  // <= rf <= min(mb, t).  rf in 16-byte units.
  const int m16 = m1 * 16;
  auto fetch = [&](int t, cplx* dst) __attribute__((always_inline)) {
    const int base = t * m16 + mb * 16;
    const unsigned lim = (unsigned)max(min(mb, t), 0) * 16u;
    const int ra = (((t & 63) - lane) * 16) & RMASK;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const int rq = (ra + q * 0x400) & RMASK;
      dst[q] = bldc(Lr, (unsigned)(rq - 16) < lim ? (unsigned)(base - rq) : kBufOut);
    }
  };
  // f_i, L_ii and y_{i-WIN} of block b's row i = 64 b + lane (raw loads;
  // masked when the block becomes current)
  auto load_blk = [&](int b, double& fb, cplx& db, cplx& yb) __attribute__((always_inline)) {
    const int i = 64 * b + lane;
    fb = fak[min(max(i, 0), n - 1)];
    db = bldc(Lr, (unsigned)(i * m16 + mb * 16));
    yb = ldc(yg + min(max(i - WIN, 0), n - 1));
  };
  cplx ys[NS];  // slot s: the class of block b - s
  cplx lr[DEPTH][NS];
  // Rows are taken in 64-row blocks from the top; the top block is padded
  // above n with rows that change nothing (x = 0: their -1/L_ii is zero).
  const int btop = (n - 1) >> 6;
  double fcur, fnxt;
  cplx dcur, dnxt, ycur, ynxt, nl, xo = mkc(0.0, 0.0);
  auto enter_blk = [&](int b) __attribute__((always_inline)) {
    const int i = 64 * b + lane;
    fcur = i < n ? fcur : 0.0;
    ycur = csel(i - WIN >= 0 && i - WIN < n, ycur, mkc(0.0, 0.0));
    const cplx rc = crecip(dcur);
    nl = csel(i < n, mkc(-rc.re, -rc.im), mkc(0.0, 0.0));
  };
  load_blk(btop, fcur, dcur, ycur);
  load_blk(btop - 1, fnxt, dnxt, ynxt);
  const int itop = 64 * btop + 63;
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    // slot q's class at the top block: lane + 64 ((btop - q) mod NS)
    const int cls = lane + 64 * (((btop - q) % NS + NS) % NS);
    const int j = itop - (((itop - cls) % WIN) + WIN) % WIN;
    ys[q] = csel(j >= 0 && j < n, ldc(yg + min(max(j, 0), n - 1)), mkc(0.0, 0.0));
  }
  // in row order, as the loop refills them: the loop's wait for row i's
  // entries then counts the same loads behind them on entry as on the back edge
#pragma unroll
  for (int d = 0; d < DEPTH; ++d) {
    fetch(itop - d, lr[d]);
    __builtin_amdgcn_sched_barrier(0);
  }
  enter_blk(btop);

  for (int b = btop; b >= 0; --b) {
    if (b < btop) {
      fcur = fnxt;
      dcur = dnxt;
      ycur = ynxt;
      enter_blk(b);
      load_blk(b - 1, fnxt, dnxt, ynxt);
    }
#pragma unroll 1
    for (int c = 0; c < 64; c += DEPTH) {
#pragma unroll
      for (int d = 0; d < DEPTH; ++d) {
        const int i = 64 * b + 63 - c - d;
        const cplx xb = rlc(cmul(ys[0], nl), i & 63);
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          ys[q] = cmadd(ys[q], lr[d][q], xb);
          // the updates are formed before lr[d] is refilled (otherwise they
          // may sink below the refill, which then needs a copy and a wait)
          asm volatile("" : "+v"(ys[q].re), "+v"(ys[q].im));
        }
        fetch(i - DEPTH, lr[d]);
        // the fetch stays here, DEPTH rows ahead of its use (the scheduler
        // would otherwise sink it next to the use)
        __builtin_amdgcn_sched_barrier(0);
        if ((d + 1) % G == 0) {
          // rows i+G-1 .. i done: their lanes keep x and refill slot 0
          const bool g = (unsigned)(lane - (i & 63)) < (unsigned)G;
          xo = csel(g, cmul(ys[0], nl), xo);
          ys[0] = csel(g, ycur, ys[0]);
        }
      }
    }
    const int io = 64 * b + lane;
    if (io < n) out[io] = make_double2(fcur * xo.re, fcur * xo.im);
    // the next block's class is in slot 1; slot 0 (refilled with the class
    // WIN rows down) goes last
    const cplx t = ys[0];
#pragma unroll
    for (int q = 0; q + 1 < NS; ++q) ys[q] = ys[q + 1];
    ys[NS - 1] = t;
  }
}

template <int DEPTH, int G, int NS>
__global__ __launch_bounds__(64) void crt_backsolve_kernel(BacksolveArgs B) {
  if ((int)blockIdx.x < B.nrhs) backsolve_rhs<DEPTH, G, NS>(B, blockIdx.x);
}

// The same over the right-hand sides of several models: block w solves rhs
// items[w].y with the arguments of model items[w].x.
template <int DEPTH, int G, int NS>
__global__ __launch_bounds__(64) void crt_backsolve_many_kernel(const BacksolveArgs* __restrict__ args,
                                                               const int2* __restrict__ items, int nitems) {
  if ((int)blockIdx.x >= nitems) return;
  const int2 it = items[blockIdx.x];
  const BacksolveArgs B = args[it.x];
  backsolve_rhs<DEPTH, G, NS>(B, it.y);
}

template __global__ void crt_backsolve_kernel<8, 8, 2>(BacksolveArgs);
template __global__ void crt_backsolve_kernel<8, 4, 2>(BacksolveArgs);
template __global__ void crt_backsolve_kernel<8, 8, 4>(BacksolveArgs);
template __global__ void crt_backsolve_kernel<8, 4, 4>(BacksolveArgs);
template __global__ void crt_backsolve_many_kernel<8, 8, 2>(const BacksolveArgs*, const int2*, int);
template __global__ void crt_backsolve_many_kernel<8, 4, 2>(const BacksolveArgs*, const int2*, int);
template __global__ void crt_backsolve_many_kernel<8, 8, 4>(const BacksolveArgs*, const int2*, int);
template __global__ void crt_backsolve_kernel<16, 8, 2>(BacksolveArgs);
template __global__ void crt_backsolve_kernel<16, 8, 4>(BacksolveArgs);
template __global__ void crt_backsolve_many_kernel<16, 8, 2>(const BacksolveArgs*, const int2*, int);
template __global__ void crt_backsolve_many_kernel<16, 8, 4>(const BacksolveArgs*, const int2*, int);

// ---------------------------------------------------------------------------
// Column-band copy of shared factors for the forward sweep: row j of Lc holds
// column j of L from the diagonal down, Lc[j][s] = L_{j+s, j} for s <= mb
// (zero past row n-1).  One thread per entry of Lc, consecutive threads along
// a row of Lc (coalesced stores); the reads walk an anti-diagonal of L whose
// lines neighbouring rows of Lc share in cache.
__global__ void crt_lcol_kernel(const double2* __restrict__ L, double2* __restrict__ Lc, int n, int mb,
                                int nfac) {
  const int m1 = mb + 1;
  const size_t per = (size_t)n * m1;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= per * nfac) return;
  const size_t f = idx / per;
  const int rem = (int)(idx - f * per);
  const int j = rem / m1, s = rem - j * m1;
  double2 v = make_double2(0.0, 0.0);
  if (j + s < n) v = L[f * per + (size_t)(j + s) * m1 + (mb - s)];
  Lc[idx] = v;
}

// ---------------------------------------------------------------------------
// Forward sweep L y = b, one wavefront per right-hand side, for sources that
// share one factor (no mixed boundaries: the factor of source 1 serves all,
// src/fem.f90:372-376 with kompb, src/kompb.f90).  b has at most two non-zero
// rows (the source electrode, the sink), so y is zero up to the first of them
// and the sweep starts at that row's 64-row block.  The mirror image of the
// back substitution above, on the column-band copy Lc: lane q (q < 64) keeps
// the partial sums b_i - sum_{j'<j} L_ij' y_j' of the rows i = q + 64 s (mod
// 64 NS) ahead of step j in NS slots.  Row j's pivot lane forms w = -y_j =
// -(its sum) / L_jj, w is broadcast and every lane adds L_ij w to its rows i
// in (j, j+mb] -- for each i the terms arrive in ascending j, the order of
// vre's row sums (src/vre.f90).  Slot s of a lane holds block b + s's row
// 64 (b + s) + lane; the slots rotate by one between blocks.  The pivot lane's
// slot 0 takes no update at its own step (the load of its entry is masked), so
// it keeps row j's final sum; every G rows the group's lanes keep y and refill
// slot 0 with b_{j+64 NS}, which column j+64 NS-mb first updates -- G < 64 NS
// + 1 - mb keeps the refill ahead of it.  Columns of Lc are fetched DEPTH
// steps ahead with buffer loads whose offsets for entries past the band or
// past row n-1 lie beyond the end of Lc, so they read zero.
template <int DEPTH, int G, int NS>
__global__ __launch_bounds__(64) void crt_fwdsweep_kernel(FwdArgs F) {
  static_assert(DEPTH % G == 0 && 64 % DEPTH == 0, "row groups tile the block");
  static_assert(NS == 2 || NS == 4, "slot classes");
  constexpr int WIN = 64 * NS;
  constexpr int RMASK = (WIN - 1) * 16;
  const int lane = threadIdx.x;
  const int r = blockIdx.x;
  if (r >= F.nrhs) return;
  const int f = r / F.eanz, l = r % F.eanz;
  const int n = F.n, mb = F.mb, m1 = mb + 1;
  // n m1 16 < 2^31 (checked by the host): offsets of rows in [0, n + WIN) fit
  const brsrc Lr = make_rsrc(F.Lc + (size_t)f * n * m1, (unsigned)(n * m1 * 16));
  double2* yg = F.y + (size_t)r * n;
  // b = -f at the source electrode, +f at the sink (the sink wins a tie).
  // Source 0 is assembled with the matrix (kompab, randb: zero on Dirichlet
  // rows; with prescribed values the whole vector comes from crt_rhs_kernel);
  // the others take kompb's vector (src/kompb.f90:39-59), which leaves
  // Dirichlet rows as they are.
  const double2* B0 = (l == 0 && F.rhs) ? F.rhs + (size_t)f * n : nullptr;
  const brsrc Br = make_rsrc(B0 ? (const void*)B0 : (const void*)F.fak, B0 ? (unsigned)(n * 16) : 0u);
  const int re = F.enr[l], rs = F.lsink ? F.nsink : -1;
  const double* fak = F.fak + (size_t)f * n;
  const bool dz = l == 0 && F.isdir;
  const double bre = B0 || (dz && F.isdir[re]) ? 0.0 : -fak[re];
  const double brs = B0 || rs < 0 || (dz && F.isdir[rs]) ? 0.0 : fak[rs];
  // b_i, zero outside [0, n): one buffer load (empty without an assembled
  // vector) plus the electrode term, no branch.  Split in two so that the
  // refill values are combined only at the refill (a use next to the load
  // would wait for every fetch in flight).
  auto bload = [&](int i) __attribute__((always_inline)) -> cplx {
    return bldc(Br, (unsigned)i < (unsigned)n ? (unsigned)i * 16u : kBufOut);
  };
  auto bterm = [&](int i) __attribute__((always_inline)) -> double {
    return i == rs ? brs : (i == re ? bre : 0.0);
  };
  const int j0 = B0 ? 0 : (rs >= 0 ? min(re, rs) : re);
  const int b0 = j0 >> 6, btop = (n - 1) >> 6;
  for (int i = lane; i < 64 * b0; i += 64) yg[i] = make_double2(0.0, 0.0);

  const int m16 = m1 * 16;
  // Entries of column j for every slot: slot s's row is j + rf, rf = (lane +
  // 64 s - j mod 64) mod WIN; in range for 1 <= rf <= min(mb, n-1-j).  rf in
  // 16-byte units.
  auto fetch = [&](int j, cplx* dst) __attribute__((always_inline)) {
    const unsigned base = (unsigned)j * (unsigned)m16;
    const unsigned lim = (unsigned)max(min(mb, n - 1 - j), 0) * 16u;
    const int ra = ((lane - (j & 63)) * 16) & RMASK;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const int rq = (ra + q * 0x400) & RMASK;
      dst[q] = bldc(Lr, (unsigned)(rq - 16) < lim ? base + (unsigned)rq : kBufOut);
    }
  };
  // L_jj and b_{j+WIN} of block b's row j = 64 b + lane (L_jj reads zero for
  // rows past n-1; masked when the block becomes current)
  auto load_blk = [&](int b, cplx& db, cplx& bb) __attribute__((always_inline)) {
    const int j = 64 * b + lane;
    db = bldc(Lr, (unsigned)j * (unsigned)m16);
    bb = bload(j + WIN);
  };
  cplx dcur, dnxt, bcur, bnxt, nl, yo = mkc(0.0, 0.0);
  double ecur;
  // rows past n-1 (padding of the last block) change nothing: w = 0
  auto enter_blk = [&](int b) __attribute__((always_inline)) {
    const cplx rc = crecip(dcur);
    nl = csel(64 * b + lane < n, mkc(-rc.re, -rc.im), mkc(0.0, 0.0));
    ecur = bterm(64 * b + lane + WIN);
  };
  load_blk(b0, dcur, bcur);
  load_blk(b0 + 1, dnxt, bnxt);
  cplx ys[NS];  // slot s: block b + s
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    ys[q] = bload(64 * (b0 + q) + lane);
    ys[q].re += bterm(64 * (b0 + q) + lane);
  }
  cplx lr[DEPTH][NS];
  // in step order, as the loop refills them: the loop's wait for step j's
  // entries then counts the same loads behind them on entry as on the back edge
#pragma unroll
  for (int d = 0; d < DEPTH; ++d) {
    fetch(64 * b0 + d, lr[d]);
    __builtin_amdgcn_sched_barrier(0);
  }
  enter_blk(b0);

  for (int b = b0; b <= btop; ++b) {
    if (b > b0) {
      dcur = dnxt;
      bcur = bnxt;
      enter_blk(b);
      // the copy out of bnxt comes first, so the load can write its registers
      // directly (a load into a temporary needs a copy after it, and the copy
      // a wait for every fetch in flight)
      asm volatile("" : "+v"(bcur.re), "+v"(bcur.im));
      __builtin_amdgcn_sched_barrier(0);
      load_blk(b + 1, dnxt, bnxt);
    }
#pragma unroll 1
    for (int c = 0; c < 64; c += DEPTH) {
#pragma unroll
      for (int d = 0; d < DEPTH; ++d) {
        const int j = 64 * b + c + d;
        const cplx w = rlc(cmul(ys[0], nl), j & 63);
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          ys[q] = cmadd(ys[q], lr[d][q], w);
          // the updates are formed before lr[d] is refilled (otherwise they
          // may sink below the refill, which then needs a copy and a wait)
          asm volatile("" : "+v"(ys[q].re), "+v"(ys[q].im));
        }
        fetch(j + DEPTH, lr[d]);
        // the fetch stays here, DEPTH steps ahead of its use (the scheduler
        // would otherwise sink it next to the use)
        __builtin_amdgcn_sched_barrier(0);
        if ((d + 1) % G == 0) {
          // rows j-G+1 .. j done: their lanes keep y and refill slot 0
          const bool g = (unsigned)(lane - ((j & 63) - (G - 1))) < (unsigned)G;
          yo = csel(g, cmul(ys[0], nl), yo);
          ys[0] = csel(g, mkc(bcur.re + ecur, bcur.im), ys[0]);
        }
      }
    }
    const int jo = 64 * b + lane;
    if (jo < n) yg[jo] = make_double2(-yo.re, -yo.im);
    const cplx t = ys[0];
#pragma unroll
    for (int q = 0; q + 1 < NS; ++q) ys[q] = ys[q + 1];
    ys[NS - 1] = t;
  }
}

template __global__ void crt_fwdsweep_kernel<8, 8, 2>(FwdArgs);
template __global__ void crt_fwdsweep_kernel<8, 4, 2>(FwdArgs);
template __global__ void crt_fwdsweep_kernel<8, 8, 4>(FwdArgs);
template __global__ void crt_fwdsweep_kernel<8, 4, 4>(FwdArgs);
