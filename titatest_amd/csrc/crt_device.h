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
// a + b*c
__device__ __forceinline__ cplx cmadd(cplx a, cplx b, cplx c) {
  return mkc(a.re + (b.re * c.re - b.im * c.im), a.im + (b.re * c.im + b.im * c.re));
}
__device__ __forceinline__ cplx rmul(double r, cplx z) { return mkc(r * z.re, r * z.im); }
__device__ __forceinline__ cplx cconj(cplx z) { return mkc(z.re, -z.im); }
__device__ __forceinline__ double cabs_(cplx z) { return hypot(z.re, z.im); }

// logb(m) for finite nonzero m is frexp's exponent minus one; the power-of-two
// scalings are exact (v_frexp_exp_i32_f64, v_ldexp_f64).
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
  double c = b.re, d = b.im;
  const double m = fmax(fabs(c), fabs(d));
  int il = 0;
  if (m != 0.0 && isfinite(m)) {
    il = __builtin_amdgcn_frexp_exp(m) - 1;
    c = __builtin_amdgcn_ldexp(c, -il);
    d = __builtin_amdgcn_ldexp(d, -il);
  }
  const double den = c * c + d * d;
  return mkc(__builtin_amdgcn_ldexp((a.re * c + a.im * d) / den, -il),
             __builtin_amdgcn_ldexp((a.im * c - a.re * d) / den, -il));
}

__device__ __forceinline__ cplx csqrt_(cplx z) {
  double x = z.re, y = z.im;
  if (x == 0.0 && y == 0.0) return mkc(0.0, y);
  double t = sqrt((fabs(x) + hypot(x, y)) * 0.5);
  if (x >= 0.0) return mkc(t, y / (2.0 * t));
  return mkc(fabs(y) / (2.0 * t), copysign(t, y));
}

__device__ __forceinline__ cplx clog_(cplx z) { return mkc(log(cabs_(z)), atan2(z.im, z.re)); }

// Workgroup barrier ordering LDS only.  __syncthreads() also waits for every
// outstanding global store of the calling wave; kernels whose threads only
// communicate through LDS use this instead so stores stream out behind it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ cplx ldc(const double2* p) {
  double2 v = *p;
  return mkc(v.x, v.y);
}
__device__ __forceinline__ void stc(double2* p, cplx z) { *p = make_double2(z.re, z.im); }
