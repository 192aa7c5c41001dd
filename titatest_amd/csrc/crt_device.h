// Device-side helpers shared by the CRMod/CRTomo HIP kernels (gfx950).
//
// Complex values are interleaved (re, im) doubles, the layout of Fortran
// COMPLEX(KIND(0D0)) and C double _Complex.  Complex arithmetic follows the
// Fortran operator definitions; division uses the power-of-two scaled formula
// of C99 Annex G.5.1, which is what the reference's Fortran runtime evaluates,
// and the square root is the principal branch (src/chol.f90:73 uses cdsqrt).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

struct cplx {
  double re, im;
};

__device__ __forceinline__ cplx mkc(double r, double i) {
  cplx z;
  z.re = r;
  z.im = i;
  return z;
}
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return mkc(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return mkc(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return mkc(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// a - b*c as the reference writes it (s = s - x*y)
__device__ __forceinline__ cplx cmsub(cplx a, cplx b, cplx c) {
  return mkc(a.re - (b.re * c.re - b.im * c.im), a.im - (b.re * c.im + b.im * c.re));
}
// a - b*c as four fused multiply-adds.  The real part subtracts b.re c.re
// first, in one rounding with the exact product, so when a.re and b.re c.re
// nearly cancel (the usual case in an elimination update) the error is
// relative to the small difference, not to a.re.
__device__ __forceinline__ cplx cfms(cplx a, cplx b, cplx c) {
  return mkc(fma(b.im, c.im, fma(-b.re, c.re, a.re)), fma(-b.im, c.re, fma(-b.re, c.im, a.im)));
}
// a + b*c
__device__ __forceinline__ cplx cmadd(cplx a, cplx b, cplx c) {
  return mkc(a.re + (b.re * c.re - b.im * c.im), a.im + (b.re * c.im + b.im * c.re));
}
__device__ __forceinline__ cplx rmul(double r, cplx z) { return mkc(r * z.re, r * z.im); }
// Component-wise select.  A ?: between two cplx values may be lowered to a
// select of their addresses and a load, which keeps arrays of cplx in scratch
// memory instead of registers.
__device__ __forceinline__ cplx csel(bool c, cplx a, cplx b) { return mkc(c ? a.re : b.re, c ? a.im : b.im); }
__device__ __forceinline__ cplx cconj(cplx z) { return mkc(z.re, -z.im); }
__device__ __forceinline__ double cabs_(cplx z) { return hypot(z.re, z.im); }

// Division by b, split so that several numerators can share the divisor's
// scaling: b is scaled by 2^-il (il = logb(max(|b.re|, |b.im|)), exact, via
// v_frexp_exp_i32_f64 / v_ldexp_f64), den = |b 2^-il|^2, and each quotient is
// rescaled by 2^-il.
struct cdiv_by {
  double c, d, den;
  int il;
};
__device__ __forceinline__ cdiv_by cdiv_prep(cplx b) {
  cdiv_by p;
  const double m = fmax(fabs(b.re), fabs(b.im));
  // branch-free: il = 0 (no scaling) for a zero or non-finite divisor
  p.il = (m != 0.0 && isfinite(m)) ? __builtin_amdgcn_frexp_exp(m) - 1 : 0;
  p.c = __builtin_amdgcn_ldexp(b.re, -p.il);
  p.d = __builtin_amdgcn_ldexp(b.im, -p.il);
  p.den = p.c * p.c + p.d * p.d;
  return p;
}
__device__ __forceinline__ cplx cdiv_apply(cplx a, const cdiv_by& p) {
  return mkc(__builtin_amdgcn_ldexp((a.re * p.c + a.im * p.d) / p.den, -p.il),
             __builtin_amdgcn_ldexp((a.im * p.c - a.re * p.d) / p.den, -p.il));
}
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) { return cdiv_apply(a, cdiv_prep(b)); }

// 1/z = conj(z') / |z'|^2 * 2^-il with z' = z 2^-il (exact scaling, so |z'|^2
// lies in [1, 8)); the reciprocal of |z'|^2 is v_rcp_f64 refined by two Newton
// steps (within one ulp).
__device__ __forceinline__ cplx crecip(cplx z) {
  const cdiv_by p = cdiv_prep(z);
  double r = __builtin_amdgcn_rcp(p.den);
  double e = fma(-p.den, r, 1.0);
  r = fma(r, e, r);
  e = fma(-p.den, r, 1.0);
  r = fma(r, e, r);
  return mkc(__builtin_amdgcn_ldexp(p.c * r, -p.il), __builtin_amdgcn_ldexp(-p.d * r, -p.il));
}

// Principal square root: t = sqrt((|x| + |z|)/2), then (t, y/2t) for x >= 0
// and (|y|/2t, sign(y) t) otherwise -- one division either way.
__device__ __forceinline__ cplx csqrt_(cplx z) {
  const double x = z.re, y = z.im;
  const double t = sqrt((fabs(x) + hypot(x, y)) * 0.5);
  const bool pos = x >= 0.0;
  const double u = (pos ? y : fabs(y)) / (2.0 * t);
  // branch-free selects; z = 0 gives (0, y)
  const bool zero = x == 0.0 && y == 0.0;
  return mkc(zero ? 0.0 : (pos ? t : u), zero ? y : (pos ? u : copysign(t, y)));
}

__device__ __forceinline__ cplx clog_(cplx z) { return mkc(log(cabs_(z)), atan2(z.im, z.re)); }

// Entry-generic forms for the kernels instantiated on complex entries (cplx,
// stored in global memory as double2) and on the real entries of the DC twin
// (double, stored as double).  The complex forms are the functions above.
template <class T>
using store_t = std::conditional_t<std::is_same<T, cplx>::value, double2, double>;
template <class T>
__device__ __forceinline__ T sreal(double x);  // x + 0i
template <>
__device__ __forceinline__ cplx sreal<cplx>(double x) { return mkc(x, 0.0); }
template <>
__device__ __forceinline__ double sreal<double>(double x) { return x; }
template <class T>
__device__ __forceinline__ T szero() { return sreal<T>(0.0); }
template <class T>
__device__ __forceinline__ T sfrom(cplx z);  // the entry of a complex value (DC: its real part)
template <>
__device__ __forceinline__ cplx sfrom<cplx>(cplx z) { return z; }
template <>
__device__ __forceinline__ double sfrom<double>(cplx z) { return z.re; }
__device__ __forceinline__ double msub(double a, double b, double c) { return a - b * c; }
__device__ __forceinline__ cplx msub(cplx a, cplx b, cplx c) { return cmsub(a, b, c); }
__device__ __forceinline__ double madd(double a, double b, double c) { return a + b * c; }
__device__ __forceinline__ cplx madd(cplx a, cplx b, cplx c) { return cmadd(a, b, c); }
// acc + b*c as fused multiply-adds in a fixed order (terms of a block
// product accumulated from zero)
__device__ __forceinline__ double fmacc(double acc, double b, double c) { return fma(b, c, acc); }
__device__ __forceinline__ cplx fmacc(cplx acc, cplx b, cplx c) {
  return mkc(fma(-b.im, c.im, fma(b.re, c.re, acc.re)), fma(b.im, c.re, fma(b.re, c.im, acc.im)));
}
__device__ __forceinline__ double ssub(double a, double b) { return a - b; }
__device__ __forceinline__ cplx ssub(cplx a, cplx b) { return csub(a, b); }
__device__ __forceinline__ double smul(double a, double b) { return a * b; }
__device__ __forceinline__ cplx smul(cplx a, cplx b) { return cmul(a, b); }
__device__ __forceinline__ double sneg(double a) { return -a; }
__device__ __forceinline__ cplx sneg(cplx a) { return mkc(-a.re, -a.im); }
__device__ __forceinline__ double sadd_re(double a, double x) { return a + x; }
__device__ __forceinline__ cplx sadd_re(cplx a, double x) { return mkc(a.re + x, a.im); }
__device__ __forceinline__ double sel(bool c, double a, double b) { return c ? a : b; }
__device__ __forceinline__ cplx sel(bool c, cplx a, cplx b) { return csel(c, a, b); }
// 1/x as for crecip: x scaled by 2^-il into [1, 2) (exact), v_rcp_f64
// refined by two Newton steps, rescaled.
__device__ __forceinline__ double srecip(double x) {
  const double m = fabs(x);
  const int il = (m != 0.0 && isfinite(m)) ? __builtin_amdgcn_frexp_exp(m) - 1 : 0;
  const double xs = __builtin_amdgcn_ldexp(x, -il);
  double r = __builtin_amdgcn_rcp(xs);
  double e = fma(-xs, r, 1.0);
  r = fma(r, e, r);
  e = fma(-xs, r, 1.0);
  r = fma(r, e, r);
  return __builtin_amdgcn_ldexp(r, -il);
}
__device__ __forceinline__ cplx srecip(cplx z) { return crecip(z); }
__device__ __forceinline__ double ssqrt(double x) { return sqrt(x); }
__device__ __forceinline__ cplx ssqrt(cplx z) { return csqrt_(z); }
// Values the compiler must treat as changed here (keeps an update from being
// moved across the point).
__device__ __forceinline__ void spin(double& x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ void spin(cplx& z) { asm volatile("" : "+v"(z.re), "+v"(z.im)); }

// Workgroup barrier ordering LDS only.  __syncthreads() also waits for every
// outstanding global store of the calling wave; kernels whose threads only
// communicate through LDS use this instead so stores stream out behind it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Raw buffer access (no stride, no swizzle; dword 3 of the descriptor is the
// gfx9 data-format word).  A 16-byte load at a byte offset at or past the
// buffer's size returns zero -- a range check done by the memory pipeline, so
// a masked element costs one select of its offset instead of one per value.
typedef __amdgpu_buffer_rsrc_t brsrc;
typedef double dbl2v __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
constexpr unsigned kBufOut = 0xFFFFFFF0u;  // an offset past every buffer made here
__device__ __forceinline__ brsrc make_rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ cplx bldc(brsrc r, unsigned off) {
  const dbl2v v = __builtin_bit_cast(dbl2v, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
  return mkc(v.x, v.y);
}

__device__ __forceinline__ cplx ldc(const double2* p) {
  double2 v = *p;
  return mkc(v.x, v.y);
}
__device__ __forceinline__ void stc(double2* p, cplx z) { *p = make_double2(z.re, z.im); }

// Entry loads and stores for the entry-generic kernels.
__device__ __forceinline__ cplx ld(const double2* p) { return ldc(p); }
__device__ __forceinline__ double ld(const double* p) { return *p; }
__device__ __forceinline__ void st(double2* p, cplx z) { stc(p, z); }
__device__ __forceinline__ void st(double* p, double x) { *p = x; }
// An 8-byte buffer load (the range check of bldc).
__device__ __forceinline__ double bldd(brsrc r, unsigned off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
__device__ __forceinline__ void bld(brsrc r, unsigned off, cplx& z) { z = bldc(r, off); }
__device__ __forceinline__ void bld(brsrc r, unsigned off, double& x) { x = bldd(r, off); }
// Buffer stores: a store at a byte offset at or past the buffer's size is
// dropped by the memory pipeline (the store-side range check).
__device__ __forceinline__ void bst(brsrc r, unsigned off, cplx z) {
  const dbl2v v = {z.re, z.im};
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, (int)off, 0, 0);
}
__device__ __forceinline__ void bst(brsrc r, unsigned off, double x) {
  typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, (int)off, 0, 0);
}
