// Batched complex-symmetric band factorisation with fused forward sweep
// (gfx950), the assembly/scaling kernels that feed it, and the
// wavefront-per-system substitutions; complex entries, or the real entries of
// the DC twin (scaldc, choldc, vredc).
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
// crt_factor_kernel: band L L^T without conjugation (chol, src/chol.f90:44-78;
//   choldc for DC) plus the forward substitution L y = b (vre, src/vre.f90:
//   45-56), one workgroup per system, persistent over systems (factor_system
//   below).  The matrix is processed in stages of kNB columns.  Stage J: the
//   update waves first apply panel J-1 to the kNB columns of panel J (the
//   look-ahead), then, while wave 0 factors panel J in registers without any
//   workgroup barrier, they apply panel J-1 to the rest of the trailing
//   window.  Two barriers per stage.  Every entry receives each earlier
//   panel's product sum_k L_ik L_jk (ascending k, formed from zero) as one
//   subtraction, and is divided by its pivot once.  The window lives in a
//   ring that rows enter from the scaled-row records one batch of kNB rows
//   per stage.
//
// crt_backsolve_kernel: x = -L^{-T} y (vre, src/vre.f90:58-69) and
//   kpot = x * f (src/fem.f90:385), one wavefront per system.
#include <type_traits>

#include "crt_device.h"
#include "crt_kernels.h"

namespace {

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

// Panel buffer: rows of kNB columns, one pad entry after every 4-row group,
// so that tiles reading one column from rows of different groups spread over
// the LDS banks, while the columns of a row stay at constant offsets from it
// (immediate offsets of the LDS loads).
__device__ __forceinline__ int pix(int rw, int k) { return rw * kNB + (rw >> 2) + k; }

// Panel rows of wave 0's lanes.  Replicated layout (REP): slot 0: lanes 0-7
// of every 16-lane row hold the panel's diagonal rows 0-7 (four copies, so
// that a DPP row broadcast reads diagonal row c from lane c), lanes 8-15 rows
// 8 + 8 g + (lane - 8) of 16-lane row g (rows 8-39); slot q >= 1: row 40 +
// 64 (q - 1) + lane; slots needed for the W = mb + kNB panel rows: 1 +
// ceil((W - 40) / 64).  Plain layout: row 64 q + lane, ceil(W / 64) slots, and
// the diagonal rows in lanes 0-7 of slot 0 are read by lane reads.
template <bool REP>
__device__ __forceinline__ int prow(int lane, int q) {
  if constexpr (!REP) return 64 * q + lane;
  const int j = lane & 15;
  return q == 0 ? (j < kNB ? j : kNB + kNB * (lane >> 4) + (j - kNB)) : 40 + 64 * (q - 1) + lane;
}

// Value at band offset t of row i of a staged batch (0 when t is not one of
// the row's stored offsets).
template <class T>
__device__ __forceinline__ T staged(const T* sv, const int16_t* so, int i, int t) {
  int qf = -1;
#pragma unroll
  for (int q = 0; q < kMaxR; ++q) qf = (so[i * kMaxR + q] == t) ? q : qf;
  return sel(qf >= 0, sv[i * kMaxR + max(qf, 0)], szero<T>());
}

__device__ __forceinline__ double rl(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ cplx rlc(cplx z, int l) { return mkc(rl(z.re, l), rl(z.im, l)); }
__device__ __forceinline__ double srl(double x, int l) { return rl(x, l); }
__device__ __forceinline__ cplx srl(cplx z, int l) { return rlc(z, l); }
// Lane L of every 16-lane row to the whole row (DPP row_newbcast, two 32-bit
// moves per double); rbcast takes L from a fully unrolled loop, where the
// switch folds to one case.
template <int L>
__device__ __forceinline__ double rbc(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)(b & 0xffffffffLL), 0x150 + L, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x150 + L, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int L>
__device__ __forceinline__ cplx rbc(cplx z) { return mkc(rbc<L>(z.re), rbc<L>(z.im)); }
template <class T>
__device__ __forceinline__ T rbcast(T x, int l) {
  switch (l) {
    case 0: return rbc<0>(x);
    case 1: return rbc<1>(x);
    case 2: return rbc<2>(x);
    case 3: return rbc<3>(x);
    case 4: return rbc<4>(x);
    case 5: return rbc<5>(x);
    case 6: return rbc<6>(x);
    default: return rbc<7>(x);
  }
}
__device__ __forceinline__ double sshfl(double x, int src) { return __shfl(x, src, 64); }
__device__ __forceinline__ cplx sshfl(cplx z, int src) { return mkc(__shfl(z.re, src, 64), __shfl(z.im, src, 64)); }

// Scaled-row record slot q of row r (value and band offset), or (0, -1).
template <class E, class T>
__device__ __forceinline__ void row_slot(const FactorArgs& A, const E* Sg, int r, int q, T& v, int& t) {
  v = szero<T>();
  t = -1;
  if (r >= 0 && r < A.n && q < A.R) {
    t = A.pat_t[(size_t)r * A.R + q];
    if (t >= 0) v = ld(Sg + (size_t)r * A.R + q);
  }
}

template <class T>
struct Stage {
  int n, mb, W, o;
  Ring ring;
  bool apply;              // panel J-1 exists
  const T* Pp;             // panel J-1 (its window coordinates: stage rows + kNB)
};

// Entry (rw, cw) of a tile exists: lower triangle, inside the window and the
// matrix.
template <class T>
__device__ __forceinline__ bool tile_ok(const Stage<T>& C, int rw, int cw) {
  return cw <= rw && rw < C.mb && C.o + rw < C.n;
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
// at consecutive addresses.  T = double (the DC twin) stores the real parts.
template <class T>
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
  st(reinterpret_cast<store_t<T>*>(A.S) + idx, sfrom<T>(v));
}
template __global__ void crt_rows_kernel<cplx>(FactorArgs);
template __global__ void crt_rows_kernel<double>(FactorArgs);

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

// TR x TC tile of the stage window at (rw0, cw0): the current values from the
// ring (every window entry lives there; 0 outside the window's lower triangle
// and the matrix) minus panel J-1's rank-kNB product, in ascending k.  slot
// receives each entry's ring position for the caller's store: the ring's
// discard slot where the entry does not exist, so that every store is
// unconditional.  rw0 and cw0 are multiples of TR = TC (2 or 4), so each of
// the tile's two sets of operand rows of panel J-1 (stage rows + kNB) lies in
// one 4-row group, its rows kNB entries apart (the panel buffer pads only
// between groups), and the panel buffer's 4 pad rows take the groups that
// reach past its W rows (their products fall on entries that are not
// stored).
template <int TR, int TC, class T>
__device__ __forceinline__ void ring_tile(const Stage<T>& C, const T* Tt, int rw0, int cw0, T acc[TR][TC],
                                          int slot[TR][TC]) {
  // The rank-kNB block product P = sum_k L_ik L_jk from zero (fused
  // multiply-adds, ascending k), subtracted from each entry once: as accurate
  // as the reference's entry-by-entry s - x*y (it rounds the products alone,
  // not the running entry), with 4 instead of 6 operations per complex term
  // and independent accumulator chains.  Small tiles read the window's
  // entries before the product (their LDS latency hides under it); 4x4 tiles
  // read them after it, which keeps the registers below the kernel's limit.
  constexpr bool early = TR * TC <= 4;
  auto window = [&]() __attribute__((always_inline)) {
    // Ring positions (Ring::pos) by parts: the tile's diagonals t = dt + i - j
    // take row t (below half) or mb - 1 - t of the ring; the slot within the
    // row is (c + om) mod (mb + 1) below half, (r + 1 + om) mod (mb + 1) for
    // the partners, one wrap per tile column and per tile row.  Absent
    // entries read the ring's zero slot and store to its discard slot.
    const int M1 = C.ring.M1, om = C.ring.om, half = C.ring.half;
    const int zs = half * M1;  // zero slot; zs + 1 discards
    const int dt = rw0 - cw0, t0 = dt * M1, q0 = (C.mb - 1 - dt) * M1;
    int cm[TC], rm[TR];
    bool okr[TR];
#pragma unroll
    for (int j = 0; j < TC; ++j) {
      const int p = cw0 + j + om;
      cm[j] = p >= M1 ? p - M1 : p;
    }
#pragma unroll
    for (int i = 0; i < TR; ++i) {
      const int p = rw0 + i + 1 + om;
      rm[i] = p >= M1 ? p - M1 : p;
      okr[i] = rw0 + i < C.mb && C.o + rw0 + i < C.n;
    }
#pragma unroll
    for (int i = 0; i < TR; ++i)
#pragma unroll
      for (int j = 0; j < TC; ++j) {
        const int d = i - j;
        const bool ok = okr[i] && dt + d >= 0;
        const int pos = dt + d < half ? t0 + d * M1 + cm[j] : q0 - d * M1 + rm[i];
        slot[i][j] = ok ? pos : zs + 1;
        acc[i][j] = Tt[ok ? pos : zs];
      }
  };
  if (early) window();
  T P[TR][TC];
#pragma unroll
  for (int i = 0; i < TR; ++i)
#pragma unroll
    for (int j = 0; j < TC; ++j) P[i][j] = szero<T>();
  if (C.apply) {
    const T* pa = C.Pp + pix(rw0 + kNB, 0);
    const T* pb = C.Pp + pix(cw0 + kNB, 0);
#pragma unroll (early ? kNB : 1)
    for (int k = 0; k < kNB; ++k) {
      T a[TR], b[TC];
#pragma unroll
      for (int i = 0; i < TR; ++i) a[i] = pa[i * kNB + k];
#pragma unroll
      for (int j = 0; j < TC; ++j) b[j] = pb[j * kNB + k];
#pragma unroll
      for (int i = 0; i < TR; ++i)
#pragma unroll
        for (int j = 0; j < TC; ++j) P[i][j] = fmacc(P[i][j], a[i], b[j]);
    }
  }
  if (!early) window();
#pragma unroll
  for (int i = 0; i < TR; ++i)
#pragma unroll
    for (int j = 0; j < TC; ++j) acc[i][j] = ssub(acc[i][j], P[i][j]);
}

// ---------------------------------------------------------------------------
// The factorisation of one system, on complex (T = cplx) or real (T = double,
// the DC twin) entries.  Phase A: the update waves apply panel J-1 to panel
// J's columns (2x2 tiles) while wave 0 copies the entering batch's entries in
// those columns into the panel's last kNB rows; phase B: wave 0 factors panel
// J in registers while the update waves apply panel J-1 to the rest of the
// window (4x4 tiles).
// * Every window entry is read from the ring.  The look-ahead moves panel
//   J's column entries out of the ring into the panel buffer and zeroes
//   their slots; in phase B the update waves scatter batch J's stored entries
//   right of panel J into the ring in window J+1's coordinates.  A departing
//   column's slots are the entering rows' slots (Ring), so the slots are free
//   and zero when the scatter writes them, and an entering row's unstored
//   band positions read 0.  Batches are staged in LDS for wave 0's copy and
//   the scatter only (two buffers, the current and the next).
// * The panel needs no column bounds: every panel position is written each
//   stage (0 outside the band, the matrix and the lower triangle), the
//   columns past the matrix end (last stage) are identity columns, and L goes
//   out through range-checked buffer stores.  Broadcasts within the panel
//   wave are DPP row broadcasts from the diagonal rows, which lanes 0-7 of
//   every 16-lane row hold, or lane reads.
// ROWS: panel row slots per lane of wave 0 (prow; W <= 40 + 64 (ROWS - 1)
// in the replicated layout of the 256-thread forms, W <= 64 ROWS in the plain
// layout of the 512-thread forms, which hold one workgroup per CU); MAXT:
// trailing tiles per thread of the update waves; NT threads per workgroup:
// 256 (one panel and three update waves; two workgroups per CU, four for
// the real entries) or, when one workgroup's LDS leaves no room for a second
// (complex, mb above ~98), 512 (seven update waves); GR: the ring lives in a
// global scratch of half x (mb + 1) entries per workgroup (A.ring) instead of
// LDS, for bands above 124; the panels and staged rows stay in LDS.
template <int ROWS, int MAXT, int NT, bool GR, class T>
__device__ __forceinline__ void factor_system(const FactorArgs& A, int s, double2* smem, PhaseClock& clk) {
  using E = store_t<T>;
  constexpr int ES = sizeof(E);
  constexpr int NB = kNB * kMaxR;             // records per staged batch
  const int n = A.n, mb = A.mb, R = A.R;
  const int W = mb + kNB, M1 = mb + 1, half = (mb + 1) / 2;
  T* Tt = GR ? reinterpret_cast<T*>(A.ring) + (size_t)blockIdx.x * A.ring_stride : reinterpret_cast<T*>(smem);
  const int PS = crt_panel_entries(mb);      // W rows, 4 pad rows, group pads
  T* P0 = GR ? reinterpret_cast<T*>(smem) : Tt + crt_ring_lds_entries(mb, ES);  // 2 panels
  T* sv = P0 + 2 * PS;                        // 2 batches of kNB x kMaxR values
  int16_t* so = reinterpret_cast<int16_t*>(sv + 2 * NB);  // their band offsets
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int x3 = tid - 64;                    // thread index within waves 1..
  // panel layout: replicated diagonal rows with DPP broadcasts for the
  // two-workgroup forms, plain rows with lane reads for the 512-thread forms
  // (one workgroup per CU, where the panel's slot count sets the pace)
  constexpr bool REP = NT == 256;
  const int nst = (n + kNB - 1) / kNB;
  // look-ahead: 2x2 tiles over window rows [0, mb) x columns [0, kNB), column
  // by column of tiles; tile column j holds the tile rows i >= j
  const int t2 = (mb + 1) / 2;
  int nla = 0;
#pragma unroll
  for (int j = 0; j < kNB / 2; ++j) nla += max(t2 - j, 0);
  int rs_r0[MAXT], rs_c0[MAXT];
  trailing_tiles<MAXT, NT>(mb, rs_r0, rs_c0);

  const int sys = (A.sys_first + s) * A.sys_stride;
  const int l = sys % A.eanz;
  E* Lg = reinterpret_cast<E*>(A.L) + (size_t)s * n * M1;
  const brsrc Lr = make_rsrc(Lg, (unsigned)(n * M1 * ES));
  E* yg = reinterpret_cast<E*>(A.y) + (size_t)s * n;
  const E* Sg = reinterpret_cast<const E*>(A.S) + (size_t)s * n * R;
  const double* fk = A.fak + (size_t)s * n;

  // ---- prologue: ring zeroed, window rows [0, mb) scattered into it, batch
  // 0 = rows [mb, mb + kNB) staged, wave 0's right-hand side
  for (int i = tid; i < crt_ring_entries(mb); i += NT) Tt[i] = szero<T>();
  stage_barrier<GR>();
  {
    Ring ring0{mb, M1, half, 0};
    for (int r = tid; r < min(mb, n); r += NT)
      for (int q = 0; q < R; ++q) {
        T v;
        int t;
        row_slot(A, Sg, r, q, v, t);
        if (t >= 0) Tt[ring0.pos(r, r - mb + t)] = v;
      }
    for (int x = tid; x < NB; x += NT) {
      T v;
      int t;
      row_slot(A, Sg, mb + x / kMaxR, x % kMaxR, v, t);
      sv[x] = v;
      so[x] = (int16_t)t;
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
  auto rhs = [&](int r) -> T {
    if (Bs) return sfrom<T>(ldc(Bs + min(max(r, 0), n - 1)));
    return sreal<T>(r == rsink ? bs : (r == rel ? be : 0.0));
  };
  // wave 0's forward-sweep values of its panel rows
  T yv[ROWS];
#pragma unroll
  for (int q = 0; q < ROWS; ++q) {
    const int rw = prow<REP>(lane, q);
    yv[q] = sel(wave == 0 && rw < W && rw < n, rhs(rw), szero<T>());
  }
  stage_barrier<GR>();
  clk.mark(0);

  for (int J = 0; J < nst; ++J) {
    const int o = J * kNB;
    const int nbc = min(kNB, n - o);
    Stage<T> C;
    C.n = n;
    C.mb = mb;
    C.W = W;
    C.o = o;
    C.ring = Ring{mb, M1, half, o % M1};
    C.apply = J > 0;
    C.Pp = P0 + ((J + 1) & 1) * PS;
    T* Pc = P0 + (J & 1) * PS;
    T* svc = sv + (J & 1) * NB;                 // batch J: rows [o + mb, o + mb + kNB)
    int16_t* soc = so + (J & 1) * NB;
    T* svn = sv + ((J + 1) & 1) * NB;           // batch J+1
    int16_t* son = so + ((J + 1) & 1) * NB;
    // the tile coordinates are loop invariant; re-declaring them as changed
    // each stage keeps the position arithmetic of 16 entries per tile from
    // being hoisted out of the stage loop (it would hold ~100 registers)
#pragma unroll
    for (int q = 0; q < MAXT; ++q) asm volatile("" : "+v"(rs_r0[q]), "+v"(rs_c0[q]));
    // the same for the lane-derived row and record indices below
    int ln = lane, lx = x3;
    asm volatile("" : "+v"(ln), "+v"(lx));

    // ---- phase A: waves 1.. apply panel J-1 to panel J's columns (out of the
    // ring into the panel buffer, their slots zeroed); wave 0 copies batch J's
    // entries in panel J's columns into the panel's last kNB rows
    if (wave > 0) {
      // the look-ahead is on the chain from panel to panel: it issues ahead
      // of the other workgroup's trailing update, as the panel does
      __builtin_amdgcn_s_setprio(3);
      for (int x = lx; x < nla; x += NT - 64) {
        int j = 0, i = x;
#pragma unroll
        for (int jj = 0; jj < kNB / 2 - 1; ++jj)
          if (i >= t2 - j) {
            i -= t2 - j;
            ++j;
          }
        const int rw0 = 2 * (i + j), cw0 = 2 * j;
        T acc[2][2];
        int slot[2][2];
        ring_tile<2, 2>(C, Tt, rw0, cw0, acc, slot);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            const int rw = rw0 + a, cw = cw0 + b;
            if (rw < mb) Pc[pix(rw, cw)] = acc[a][b];
            Tt[slot[a][b]] = szero<T>();
          }
      }
    } else {
      const int i = ln / kNB, cw = ln % kNB, rw = mb + i;
      const bool ok = o + rw < n && cw >= i;
      Pc[pix(rw, cw)] = sel(ok, staged(svc, soc, i, cw - i), szero<T>());
    }
    __builtin_amdgcn_s_setprio(0);
    clk.mark(0);
    stage_barrier<GR>();
    clk.mark(1);

    // ---- phase B
    if (wave == 0) {
      __builtin_amdgcn_s_setprio(3);
      const int rows = min(W, n - o);
      const int lj = REP ? (ln & 15) : ln;
      const bool dg = lj < kNB;  // the lane's slot-0 row is a diagonal row
      // lane c's value of a diagonal row to the whole wave
      auto bc = [&](T x, int c) __attribute__((always_inline)) {
        if constexpr (REP) return rbcast(x, c);
        else return srl(x, c);
      };
      int rwq[ROWS];
#pragma unroll
      for (int q = 0; q < ROWS; ++q) rwq[q] = prow<REP>(ln, q);
      T v[ROWS][kNB];
#pragma unroll
      for (int q = 0; q < ROWS; ++q) {
        const int rw = min(rwq[q], W + 3);  // rows past W read the pad rows; never stored
#pragma unroll
        for (int k = 0; k < kNB; ++k) v[q][k] = Pc[pix(rw, k)];
      }
      // identity columns past the matrix end
      if (nbc < kNB) {
#pragma unroll
        for (int k = 0; k < kNB; ++k) v[0][k] = sel(dg && lj == k && k >= nbc, sreal<T>(1.0), v[0][k]);
      }
      // Column c in the L D L^T form of the same elimination: with d the pivot
      // and s_rc the column's updated entries, l_rc = s_rc / d and the later
      // columns take -l_rc s_c2c (= -L_rc L_c2c).  The next pivot needs no
      // square root, so the column-to-column dependency is a reciprocal and
      // two products; L_cc = sqrt(d), L_rc = l_rc L_cc and y_c = y'_c / L_cc
      // follow off it.  lrc is 0 for the diagonal rows at or above c, so every
      // update is unconditional.  The pivot and y'_c, on the column-to-column
      // chain, come by DPP row broadcast from the diagonal rows; the other
      // entries of the diagonal block by lane reads into scalar registers,
      // which keeps them out of the vector registers.
      const bool dcz = A.dc != 0;
      T dd = sreal<T>(1.0);  // d of the lane's diagonal row
      bool zero_pivot = false;
#pragma unroll
      for (int c = 0; c < kNB; ++c) {
        const T piv = bc(v[0][c], c);
        // chol :66 -- |s| <= 0 is an exact zero (a NaN pivot passes, as
        // there); choldc :66 -- s <= 0 for the real DC pivots
        if constexpr (std::is_same<T, cplx>::value)
          zero_pivot |= (dcz & (piv.re <= 0.0)) | (!dcz & (piv.re == 0.0) & (piv.im == 0.0));
        else
          zero_pivot |= piv <= 0.0;
        const T inv = srecip(piv);
        dd = sel(lj == c, piv, dd);
        T lrc[ROWS];
        lrc[0] = sel(!dg || lj > c, smul(v[0][c], inv), szero<T>());
#pragma unroll
        for (int q = 1; q < ROWS; ++q) lrc[q] = smul(v[q][c], inv);
        const T yp = bc(yv[0], c);  // y'_c, final
        // With three or more row slots per lane each lane read is used
        // before the next one is issued (the updates pinned in order): the
        // compiler otherwise issues all of a column's reads at once and
        // spills them from the scalar registers (with two slots the
        // unpinned form measured faster).
#pragma unroll
        for (int c2 = c + 1; c2 < kNB; ++c2) {
          T src = v[0][c];
          if constexpr (ROWS >= 3 || !REP) spin(src);
          const T s2 = srl(src, c2);  // s_{c2,c} of the diagonal block
#pragma unroll
          for (int q = 0; q < ROWS; ++q) {
            v[q][c2] = msub(v[q][c2], lrc[q], s2);
            if constexpr (ROWS >= 3 || !REP) spin(v[q][c2]);
          }
        }
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
          yv[q] = msub(yv[q], lrc[q], yp);
          v[q][c] = lrc[q];
        }
      }
      const T lcc = ssqrt(dd);
#pragma unroll
      for (int c = 0; c < kNB; ++c) {
        const T lc = bc(lcc, c);
#pragma unroll
        for (int q = 0; q < ROWS; ++q) v[q][c] = smul(v[q][c], lc);
        v[0][c] = sel(dg && lj == c, lc, v[0][c]);
      }
      yv[0] = sel(dg, smul(yv[0], srecip(lcc)), yv[0]);
      if (zero_pivot && ln == 0 && atomicCAS(A.err, 0, 28) == 0) A.err[1] = sys;
      // L of panel J: rows kNB.. to the panel buffer (stage J+1's operands;
      // the diagonal rows are not read there) and every row to the band rows
      // (the diagonal rows from lanes 0-7)
#pragma unroll
      for (int q = 0; q < ROWS; ++q) {
        const int rw = rwq[q];
        if (rw >= kNB && rw < W) {
#pragma unroll
          for (int k = 0; k < kNB; ++k) Pc[pix(rw, k)] = v[q][k];
        }
        const bool mine = !REP || q > 0 || !dg || ln < 16;  // one copy of each row
        const int klo = max(rw - mb, 0), khi = min(rw, nbc - 1);
#pragma unroll
        for (int k = 0; k < kNB; ++k) {
          const bool ok = mine && rw < rows && k >= klo && k <= khi;
          bst(Lr, ok ? (unsigned)(((o + rw) * M1 + (k - rw + mb)) * ES) : kBufOut, v[q][k]);
        }
      }
      if (ln < nbc) st(yg + o + ln, yv[0]);
      // y of rows [o + kNB, o + W) moves up kNB rows (row r of the next
      // window is row r + kNB of this one: slot q or q + 1 of another lane);
      // entering rows get b
      T ny[ROWS];
#pragma unroll
      for (int q = 0; q < ROWS; ++q) {
        const int r1 = rwq[q] + kNB;  // this window's row
        int sl;
        bool hi;
        if (!REP) {
          sl = r1 & 63;
          hi = (r1 >> 6) > q;
        } else if (r1 < 40) {  // slot 0
          sl = 16 * ((r1 - kNB) >> 3) + kNB + ((r1 - kNB) & 7);
          hi = false;
        } else {
          sl = (r1 - 40) & 63;
          hi = 1 + (r1 - 40) / 64 > q;
        }
        const T a = sshfl(yv[q], sl);
        T b = szero<T>();
        if (q + 1 < ROWS) b = sshfl(yv[q + 1], sl);
        ny[q] = sel(hi, b, a);
        const int rwn = rwq[q];
        if (rwn >= mb) {
          const int r = o + kNB + rwn;
          ny[q] = sel(rwn < W && r < n, rhs(r), szero<T>());
        }
      }
#pragma unroll
      for (int q = 0; q < ROWS; ++q) yv[q] = ny[q];
      __builtin_amdgcn_s_setprio(0);
      clk.mark(2);
    } else {
      // real entries (four workgroups per CU): the trailing update at the
      // panel's priority as well (measured faster there, slower for complex)
      if constexpr (std::is_same<T, double>::value) __builtin_amdgcn_s_setprio(3);
      // batch J+1 = rows [o + kNB + mb, o + 2 kNB + mb): loads issued first,
      // staged after the tiles
      T nv = szero<T>();
      int nt = -1;
      if (lx < NB) row_slot(A, Sg, o + kNB + mb + lx / kMaxR, lx % kMaxR, nv, nt);
      // the rest of the window: panel J-1's update in place
#pragma unroll
      for (int q = 0; q < MAXT; ++q) {
        if (rs_r0[q] >= 0) {
          T acc[4][4];
          int slot[4][4];
          ring_tile<4, 4>(C, Tt, rs_r0[q], rs_c0[q], acc, slot);
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
              Tt[slot[i][j]] = acc[i][j];
        }
      }
      if (lx < NB) {
        // batch J's entries right of panel J into the ring (window J+1: row
        // mb - kNB + i, column i + t - kNB)
        const int i = lx / kMaxR, t = soc[lx], cw1 = i + t - kNB;
        if (t >= 0 && cw1 >= 0) {
          const Ring ring1{mb, M1, half, (o + kNB) % M1};
          Tt[ring1.pos(mb - kNB + i, cw1)] = svc[lx];
        }
        svn[lx] = nv;
        son[lx] = (int16_t)nt;
      }
      clk.mark(2);
    }
    stage_barrier<GR>();
    clk.mark(3);
  }
}

// Persistent over the batch's systems (workgroup b takes systems b, b + grid,
// ...).  T = double is the DC twin; its halved LDS (41 KB at mb = 82) and
// registers (128, a few values spilled outside the stage loop) leave room
// for four workgroups per CU at the narrower bands.
template <int ROWS, int MAXT, int NT, bool GR, class T>
__global__ __launch_bounds__(NT, (std::is_same<T, double>::value ? 1024 : 512) / NT) void crt_factor_kernel(
    FactorArgs A) {
  extern __shared__ double2 smem[];
  PhaseClock clk;
  clk.start(threadIdx.x == 0 ? A.prof : (threadIdx.x == 64 ? A.prof : nullptr), threadIdx.x == 0 ? 0 : 4);
  for (int s = blockIdx.x; s < A.nsys; s += gridDim.x) factor_system<ROWS, MAXT, NT, GR, T>(A, s, smem, clk);
  clk.flush();
}

// The same over the systems of several models whose bands share this form:
// work item w is system items[w].y of model items[w].x (its arguments
// args[items[w].x]); persistent workgroups take the items in list order from
// an atomic counter (zeroed before the launch) until the list is exhausted.
template <int ROWS, int MAXT, int NT, bool GR>
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
    factor_system<ROWS, MAXT, NT, GR, cplx>(A, it.y, smem, clk);
  }
}

template __global__ void crt_factor_kernel<1, 1, 256, false, cplx>(FactorArgs);
template __global__ void crt_factor_kernel<2, 1, 256, false, cplx>(FactorArgs);
template __global__ void crt_factor_kernel<2, 2, 256, false, cplx>(FactorArgs);
template __global__ void crt_factor_kernel<2, 3, 256, false, cplx>(FactorArgs);
template __global__ void crt_factor_kernel<3, 3, 256, false, cplx>(FactorArgs);
template __global__ void crt_factor_kernel<2, 1, 512, false, cplx>(FactorArgs);
template __global__ void crt_factor_kernel<2, 2, 512, false, cplx>(FactorArgs);
template __global__ void crt_factor_kernel<3, 2, 512, false, cplx>(FactorArgs);
template __global__ void crt_factor_kernel<3, 5, 512, true, cplx>(FactorArgs);
template __global__ void crt_factor_kernel<4, 5, 512, true, cplx>(FactorArgs);
template __global__ void crt_factor_kernel<5, 5, 512, true, cplx>(FactorArgs);
template __global__ void crt_factor_kernel<1, 1, 256, false, double>(FactorArgs);
template __global__ void crt_factor_kernel<2, 1, 256, false, double>(FactorArgs);
template __global__ void crt_factor_kernel<2, 2, 256, false, double>(FactorArgs);
template __global__ void crt_factor_kernel<2, 3, 256, false, double>(FactorArgs);
template __global__ void crt_factor_kernel<3, 3, 256, false, double>(FactorArgs);
template __global__ void crt_factor_kernel<2, 1, 512, false, double>(FactorArgs);
template __global__ void crt_factor_kernel<2, 2, 512, false, double>(FactorArgs);
template __global__ void crt_factor_kernel<3, 2, 512, false, double>(FactorArgs);
template __global__ void crt_factor_kernel<3, 5, 512, true, double>(FactorArgs);
template __global__ void crt_factor_kernel<4, 5, 512, true, double>(FactorArgs);
template __global__ void crt_factor_kernel<5, 5, 512, true, double>(FactorArgs);
template __global__ void crt_factor_many_kernel<1, 1, 256, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<2, 1, 256, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<2, 2, 256, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<2, 3, 256, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<3, 3, 256, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<2, 1, 512, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<2, 2, 512, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<3, 2, 512, false>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<3, 5, 512, true>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<4, 5, 512, true>(const FactorArgs*, const int2*, int, int*);
template __global__ void crt_factor_many_kernel<5, 5, 512, true>(const FactorArgs*, const int2*, int, int*);

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
// kpot = x f (complex; the DC twin's imaginary part is 0)
__device__ __forceinline__ double2 kpot_of(double f, cplx x) { return make_double2(f * x.re, f * x.im); }
__device__ __forceinline__ double2 kpot_of(double f, double x) { return make_double2(f * x, 0.0); }

template <int DEPTH, int G, int NS, class T>
__device__ __forceinline__ void backsolve_rhs(const BacksolveArgs& B, int r) {
  static_assert(DEPTH % G == 0 && 64 % DEPTH == 0, "row groups tile the block");
  static_assert(NS == 2 || NS == 4, "slot classes");
  using E = store_t<T>;
  constexpr int ES = sizeof(E);               // entry bytes: 16 complex, 8 real
  constexpr int WIN = 64 * NS;                // rows of the slot window
  constexpr int RMASK = (WIN - 1) * ES;       // (x mod WIN) * ES on entry units
  const int lane = threadIdx.x;
  const int f = r / B.rhs_per_fac;
  const int sys = (B.fac_first + f) * B.fstride + r % B.rhs_per_fac;
  const int k = sys / B.eanz, l = sys % B.eanz;
  const int n = B.n, mb = B.mb, m1 = mb + 1;
  // n m1 ES < 2^31 (checked by the host): offsets of rows in [0, n + 64) fit,
  // offsets of rows below 0 wrap past the end
  const brsrc Lr = make_rsrc(reinterpret_cast<const E*>(B.L) + (size_t)f * n * m1, (unsigned)(n * m1 * ES));
  const E* yg = reinterpret_cast<const E*>(B.y) + (size_t)r * n;
  const double* fak = B.fak + (size_t)f * n;
  double2* out = B.kpot + ((size_t)(k - B.kpot_k0) * B.eanz + l) * n;

  // Entries of row t of L for every slot: slot s's row is j = t - rf, rf =
  // (t mod 64 - lane + 64 s) mod WIN, at band offset mb - rf, in range for 1
  // <= rf <= min(mb, t).  rf in entry units (ES bytes).
  const int mES = m1 * ES;
  auto fetch = [&](int t, T* dst) __attribute__((always_inline)) {
    const int base = t * mES + mb * ES;
    const unsigned lim = (unsigned)max(min(mb, t), 0) * (unsigned)ES;
    const int ra = (((t & 63) - lane) * ES) & RMASK;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const int rq = (ra + q * 64 * ES) & RMASK;
      bld(Lr, (unsigned)(rq - ES) < lim ? (unsigned)(base - rq) : kBufOut, dst[q]);
    }
  };
  // f_i, L_ii and y_{i-WIN} of block b's row i = 64 b + lane (raw loads;
  // masked when the block becomes current)
  auto load_blk = [&](int b, double& fb, T& db, T& yb) __attribute__((always_inline)) {
    const int i = 64 * b + lane;
    fb = fak[min(max(i, 0), n - 1)];
    bld(Lr, (unsigned)(i * mES + mb * ES), db);
    yb = ld(yg + min(max(i - WIN, 0), n - 1));
  };
  T ys[NS];  // slot s: the class of block b - s
  T lr[DEPTH][NS];
  // Rows are taken in 64-row blocks from the top; the top block is padded
  // above n with rows that change nothing (x = 0: their -1/L_ii is zero).
  const int btop = (n - 1) >> 6;
  double fcur, fnxt;
  T dcur, dnxt, ycur, ynxt, nl, xo = szero<T>();
  auto enter_blk = [&](int b) __attribute__((always_inline)) {
    const int i = 64 * b + lane;
    fcur = i < n ? fcur : 0.0;
    ycur = sel(i - WIN >= 0 && i - WIN < n, ycur, szero<T>());
    nl = sel(i < n, sneg(srecip(dcur)), szero<T>());
  };
  load_blk(btop, fcur, dcur, ycur);
  load_blk(btop - 1, fnxt, dnxt, ynxt);
  const int itop = 64 * btop + 63;
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    // slot q's class at the top block: lane + 64 ((btop - q) mod NS)
    const int cls = lane + 64 * (((btop - q) % NS + NS) % NS);
    const int j = itop - (((itop - cls) % WIN) + WIN) % WIN;
    ys[q] = sel(j >= 0 && j < n, ld(yg + min(max(j, 0), n - 1)), szero<T>());
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
        const T xb = srl(smul(ys[0], nl), i & 63);
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          ys[q] = madd(ys[q], lr[d][q], xb);
          // the updates are formed before lr[d] is refilled (otherwise they
          // may sink below the refill, which then needs a copy and a wait)
          spin(ys[q]);
        }
        fetch(i - DEPTH, lr[d]);
        // the fetch stays here, DEPTH rows ahead of its use (the scheduler
        // would otherwise sink it next to the use)
        __builtin_amdgcn_sched_barrier(0);
        if ((d + 1) % G == 0) {
          // rows i+G-1 .. i done: their lanes keep x and refill slot 0
          const bool g = (unsigned)(lane - (i & 63)) < (unsigned)G;
          xo = sel(g, smul(ys[0], nl), xo);
          ys[0] = sel(g, ycur, ys[0]);
        }
      }
    }
    const int io = 64 * b + lane;
    if (io < n) out[io] = kpot_of(fcur, xo);
    // the next block's class is in slot 1; slot 0 (refilled with the class
    // WIN rows down) goes last
    const T t = ys[0];
#pragma unroll
    for (int q = 0; q + 1 < NS; ++q) ys[q] = ys[q + 1];
    ys[NS - 1] = t;
  }
}

template <int DEPTH, int G, int NS, class T>
__global__ __launch_bounds__(64) void crt_backsolve_kernel(BacksolveArgs B) {
  if ((int)blockIdx.x < B.nrhs) backsolve_rhs<DEPTH, G, NS, T>(B, blockIdx.x);
}

// The same over the right-hand sides of several models: block w solves rhs
// items[w].y with the arguments of model items[w].x.
template <int DEPTH, int G, int NS>
__global__ __launch_bounds__(64) void crt_backsolve_many_kernel(const BacksolveArgs* __restrict__ args,
                                                               const int2* __restrict__ items, int nitems) {
  if ((int)blockIdx.x >= nitems) return;
  const int2 it = items[blockIdx.x];
  const BacksolveArgs B = args[it.x];
  backsolve_rhs<DEPTH, G, NS, cplx>(B, it.y);
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
template __global__ void crt_backsolve_kernel<8, 8, 2, double>(BacksolveArgs);
template __global__ void crt_backsolve_kernel<8, 4, 2, double>(BacksolveArgs);
template __global__ void crt_backsolve_kernel<8, 8, 4, double>(BacksolveArgs);
template __global__ void crt_backsolve_kernel<16, 8, 2, double>(BacksolveArgs);
template __global__ void crt_backsolve_kernel<16, 8, 4, double>(BacksolveArgs);

// ---------------------------------------------------------------------------
// Column-band copy of shared factors for the forward sweep: row j of Lc holds
// column j of L from the diagonal down, Lc[j][s] = L_{j+s, j} for s <= mb
// (zero past row n-1).  One thread per entry of Lc, consecutive threads along
// a row of Lc (coalesced stores); the reads walk an anti-diagonal of L whose
// lines neighbouring rows of Lc share in cache.
template <class T>
__global__ void crt_lcol_kernel(const void* __restrict__ Lv, void* __restrict__ Lcv, int n, int mb, int nfac) {
  using E = store_t<T>;
  const E* L = reinterpret_cast<const E*>(Lv);
  E* Lc = reinterpret_cast<E*>(Lcv);
  const int m1 = mb + 1;
  const size_t per = (size_t)n * m1;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= per * nfac) return;
  const size_t f = idx / per;
  const int rem = (int)(idx - f * per);
  const int j = rem / m1, s = rem - j * m1;
  T v = szero<T>();
  if (j + s < n) v = ld(L + f * per + (size_t)(j + s) * m1 + (mb - s));
  st(Lc + idx, v);
}
template __global__ void crt_lcol_kernel<cplx>(const void*, void*, int, int, int);
template __global__ void crt_lcol_kernel<double>(const void*, void*, int, int, int);

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
template <int DEPTH, int G, int NS, class T>
__global__ __launch_bounds__(64) void crt_fwdsweep_kernel(FwdArgs F) {
  static_assert(DEPTH % G == 0 && 64 % DEPTH == 0, "row groups tile the block");
  static_assert(NS == 2 || NS == 4, "slot classes");
  using E = store_t<T>;
  constexpr int ES = sizeof(E);
  constexpr int WIN = 64 * NS;
  constexpr int RMASK = (WIN - 1) * ES;
  const int lane = threadIdx.x;
  const int r = blockIdx.x;
  if (r >= F.nrhs) return;
  const int f = r / F.eanz, l = r % F.eanz;
  const int n = F.n, mb = F.mb, m1 = mb + 1;
  // n m1 ES < 2^31 (checked by the host): offsets of rows in [0, n + WIN) fit
  const brsrc Lr = make_rsrc(reinterpret_cast<const E*>(F.Lc) + (size_t)f * n * m1, (unsigned)(n * m1 * ES));
  E* yg = reinterpret_cast<E*>(F.y) + (size_t)r * n;
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
  auto bload = [&](int i) __attribute__((always_inline)) -> T {
    return sfrom<T>(bldc(Br, (unsigned)i < (unsigned)n ? (unsigned)i * 16u : kBufOut));
  };
  auto bterm = [&](int i) __attribute__((always_inline)) -> double {
    return i == rs ? brs : (i == re ? bre : 0.0);
  };
  const int j0 = B0 ? 0 : (rs >= 0 ? min(re, rs) : re);
  const int b0 = j0 >> 6, btop = (n - 1) >> 6;
  for (int i = lane; i < 64 * b0; i += 64) st(yg + i, szero<T>());

  const int mES = m1 * ES;
  // Entries of column j for every slot: slot s's row is j + rf, rf = (lane +
  // 64 s - j mod 64) mod WIN; in range for 1 <= rf <= min(mb, n-1-j).  rf in
  // entry units (ES bytes).
  auto fetch = [&](int j, T* dst) __attribute__((always_inline)) {
    const unsigned base = (unsigned)j * (unsigned)mES;
    const unsigned lim = (unsigned)max(min(mb, n - 1 - j), 0) * (unsigned)ES;
    const int ra = ((lane - (j & 63)) * ES) & RMASK;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const int rq = (ra + q * 64 * ES) & RMASK;
      bld(Lr, (unsigned)(rq - ES) < lim ? base + (unsigned)rq : kBufOut, dst[q]);
    }
  };
  // L_jj and b_{j+WIN} of block b's row j = 64 b + lane (L_jj reads zero for
  // rows past n-1; masked when the block becomes current)
  auto load_blk = [&](int b, T& db, T& bb) __attribute__((always_inline)) {
    const int j = 64 * b + lane;
    bld(Lr, (unsigned)j * (unsigned)mES, db);
    bb = bload(j + WIN);
  };
  T dcur, dnxt, bcur, bnxt, nl, yo = szero<T>();
  double ecur;
  // rows past n-1 (padding of the last block) change nothing: w = 0
  auto enter_blk = [&](int b) __attribute__((always_inline)) {
    nl = sel(64 * b + lane < n, sneg(srecip(dcur)), szero<T>());
    ecur = bterm(64 * b + lane + WIN);
  };
  load_blk(b0, dcur, bcur);
  load_blk(b0 + 1, dnxt, bnxt);
  T ys[NS];  // slot s: block b + s
#pragma unroll
  for (int q = 0; q < NS; ++q) ys[q] = sadd_re(bload(64 * (b0 + q) + lane), bterm(64 * (b0 + q) + lane));
  T lr[DEPTH][NS];
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
      spin(bcur);
      __builtin_amdgcn_sched_barrier(0);
      load_blk(b + 1, dnxt, bnxt);
    }
#pragma unroll 1
    for (int c = 0; c < 64; c += DEPTH) {
#pragma unroll
      for (int d = 0; d < DEPTH; ++d) {
        const int j = 64 * b + c + d;
        const T w = srl(smul(ys[0], nl), j & 63);
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          ys[q] = madd(ys[q], lr[d][q], w);
          // the updates are formed before lr[d] is refilled (otherwise they
          // may sink below the refill, which then needs a copy and a wait)
          spin(ys[q]);
        }
        fetch(j + DEPTH, lr[d]);
        // the fetch stays here, DEPTH steps ahead of its use (the scheduler
        // would otherwise sink it next to the use)
        __builtin_amdgcn_sched_barrier(0);
        if ((d + 1) % G == 0) {
          // rows j-G+1 .. j done: their lanes keep y and refill slot 0
          const bool g = (unsigned)(lane - ((j & 63) - (G - 1))) < (unsigned)G;
          yo = sel(g, smul(ys[0], nl), yo);
          ys[0] = sel(g, sadd_re(bcur, ecur), ys[0]);
        }
      }
    }
    const int jo = 64 * b + lane;
    if (jo < n) st(yg + jo, sneg(yo));
    const T t = ys[0];
#pragma unroll
    for (int q = 0; q + 1 < NS; ++q) ys[q] = ys[q + 1];
    ys[NS - 1] = t;
  }
}

template __global__ void crt_fwdsweep_kernel<8, 8, 2>(FwdArgs);
template __global__ void crt_fwdsweep_kernel<8, 4, 2>(FwdArgs);
template __global__ void crt_fwdsweep_kernel<8, 8, 4>(FwdArgs);
template __global__ void crt_fwdsweep_kernel<8, 4, 4>(FwdArgs);
template __global__ void crt_fwdsweep_kernel<8, 8, 2, double>(FwdArgs);
template __global__ void crt_fwdsweep_kernel<8, 4, 2, double>(FwdArgs);
template __global__ void crt_fwdsweep_kernel<8, 8, 4, double>(FwdArgs);
