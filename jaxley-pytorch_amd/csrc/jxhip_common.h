// Shared device helpers for the jaxhip kernels (gfx950 / CDNA4, wave64).
//
// Numerics follow jaxley-pytorch's gate solvers:
//   save_exp(x)       = exp(min(x, 20))                        solver_gate.py:8-11
//   x_over_expm1(x,y) = x / expm1(min(x/y, 20)), series y(1 - z/2 + z^2/12) for |z| < 1e-4
//                                                              solver_gate.py:14-30
//   exponential Euler x' = x e + x_inf (1 - e), e = save_exp(-dt / tau)
//                                                              solver_gate.py:45-64
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jxhip {

// float32 transcendentals: the hardware exp2 (v_exp_f32) after an exact-enough scaling,
// and expm1 from its Taylor series near 0 (|x| < 1/4: degree 7, relative error < 2e-8)
// where exp(x) - 1 would cancel. Define JXHIP_ACCURATE_F32 for the libm versions.
// float64 uses the correctly rounded libm functions.
#ifdef JXHIP_ACCURATE_F32
__device__ __forceinline__ float dexp(float x) { return expf(x); }
__device__ __forceinline__ float dexpm1(float x) { return expm1f(x); }
__device__ __forceinline__ float ddiv(float a, float b) { return a / b; }
#else
__device__ __forceinline__ float dexp(float x) { return __expf(x); }
__device__ __forceinline__ float dexpm1(float x) {
  if (fabsf(x) < 0.25f) {
    return x * (1.f + x * (0.5f + x * (1.f / 6 + x * (1.f / 24 + x * (1.f / 120 + x * (1.f / 720 + x * (1.f / 5040)))))));
  }
  return __expf(x) - 1.f;
}
__device__ __forceinline__ float ddiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
#endif
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ double dexpm1(double x) { return expm1(x); }
__device__ __forceinline__ double ddiv(double a, double b) { return a / b; }

template <typename T>
__device__ __forceinline__ T save_exp(T x) {
  return dexp(x < T(20) ? x : T(20));
}

// x / (exp(x / y) - 1) with the reference's clamp and series switch.
template <typename T>
__device__ __forceinline__ T x_over_expm1(T x, T y) {
  T z = x * (T(1) / y);
  z = z < T(20) ? z : T(20);
  if (fabs(z) < T(1e-4)) return y * (T(1) - z / T(2) + z * z / T(12));
  return ddiv(x, dexpm1(z));
}

template <typename T>
__device__ __forceinline__ T sigmoid(T x) {
  return ddiv(T(1), T(1) + dexp(-x));
}

// One HH gate by exponential Euler (channels/hh.py:37-50, solver_gate.py:45-64).
template <typename T>
__device__ __forceinline__ T gate_step(T x, T dt, T alpha, T beta) {
  T tau = ddiv(T(1), alpha + beta);
  T xinf = alpha * tau;
  T e = save_exp(-dt * (alpha + beta));
  return x * e + xinf * (T(1) - e);
}

// HH rates (channels/hh.py:87-108).
template <typename T>
__device__ __forceinline__ void hh_rates(T v, T& am, T& bm, T& ah, T& bh, T& an, T& bn) {
  am = T(0.1) * x_over_expm1(-(v + T(40)), T(10));
  bm = T(4.0) * save_exp(-(v + T(65)) * T(1.0 / 18));
  ah = T(0.07) * save_exp(-(v + T(65)) * T(1.0 / 20));
  bh = ddiv(T(1.0), save_exp(-(v + T(35)) * T(0.1)) + T(1));
  an = T(0.01) * x_over_expm1(-(v + T(55)), T(10));
  bn = T(0.125) * save_exp(-(v + T(65)) * T(1.0 / 80));
}

}  // namespace jxhip
