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
// float64: exp and x / expm1 below are branch-free and accurate to a few ulp (a wave's
// lanes never diverge on them); expm1 itself (adjoint and synapse slopes only) is libm's.
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
// A float64 constant materialized in an SGPR pair at its use (volatile asm: not hoisted
// into a loop-invariant VGPR, which would turn every Horner step into a copy and an fmac).
__device__ __forceinline__ double kc(double c) {
  asm volatile("" : "+s"(c));
  return c;
}
// float64 reciprocal: the hardware estimate refined by two Newton steps (within 2 ulp;
// the operands here are never 0, infinite or subnormal).
__device__ __forceinline__ double drcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
// float64 exp without branches: Cody-Waite reduction x = k ln2 + r (ln2 split in a
// 32-bit head and a tail, |r| <= ln2 / 2), the Taylor polynomial of degree 12 for e^r
// (truncation below 2e-16 relative) and a scaling by 2^k. The argument is clamped to
// [-746, 710], where the scaling gives 0 and infinity. Accurate to a few ulp.
__device__ __forceinline__ double exp_f64(double x) {
  x = fmin(fmax(x, -746.0), 710.0);
  const double k = rint(x * 1.4426950408889634074);          // log2(e)
  double r = fma(-k, 6.93147180369123816490e-01, x);           // ln2, head
  r = fma(-k, 1.90821492927058770002e-10, r);                  // ln2, tail
  double p = kc(2.08767569878680989792e-09);                   // 1/12!
  p = fma(p, r, kc(2.50521083854417187751e-08));               // 1/11!
  p = fma(p, r, kc(2.75573192239858906526e-07));               // 1/10!
  p = fma(p, r, kc(2.75573192239858906526e-06));               // 1/9!
  p = fma(p, r, kc(2.48015873015873015873e-05));               // 1/8!
  p = fma(p, r, kc(1.98412698412698412698e-04));               // 1/7!
  p = fma(p, r, kc(1.38888888888888888889e-03));               // 1/6!
  p = fma(p, r, kc(8.33333333333333333333e-03));               // 1/5!
  p = fma(p, r, kc(4.16666666666666666667e-02));               // 1/4!
  p = fma(p, r, kc(1.66666666666666666667e-01));               // 1/3!
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, static_cast<int>(k));
}
// z / expm1(z) for |z| < 1/4 from its Bernoulli series
//   z / (e^z - 1) = 1 - z/2 + z^2/12 - z^4/720 + z^6/30240 - z^8/1209600
//                   + z^10/47900160 - 691 z^12/1307674368000 + ...
// (the next term is below 6e-20 at |z| = 1/4).
__device__ __forceinline__ double z_over_expm1_series(double z) {
  const double y = z * z;
  double p = kc(-5.28419013868749318484e-10);                  // -691/1307674368000
  p = fma(p, y, kc(2.08767569878680989792e-08));               // 1/47900160
  p = fma(p, y, kc(-8.26719576719576719577e-07));              // -1/1209600
  p = fma(p, y, kc(3.30687830687830687831e-05));               // 1/30240
  p = fma(p, y, kc(-1.38888888888888888889e-03));              // -1/720
  p = fma(p, y, kc(8.33333333333333333333e-02));               // 1/12
  p = fma(p, y, 1.0);
  return fma(-0.5, z, p);
}
__device__ __forceinline__ double dexp(double x) { return exp_f64(x); }
__device__ __forceinline__ double dexpm1(double x) { return expm1(x); }
__device__ __forceinline__ double ddiv(double a, double b) { return a * drcp(b); }

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
// float64: y z / expm1(z) from the series for |z| < 1/4 (which also covers the
// reference's three-term switch at 1e-4 to within 2e-19 relative), x / (e^z - 1)
// elsewhere (cancellation at most a factor 4.5 there); both are evaluated and one is
// selected, so no lane diverges.
__device__ __forceinline__ double x_over_expm1(double x, double y) {
  double z = x * (1.0 / y);
  z = z < 20.0 ? z : 20.0;
  const double ser = y * z_over_expm1_series(z);
  const double dir = x * drcp(exp_f64(z) - 1.0);
  return fabs(z) < 0.25 ? ser : dir;
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
  return fma(x - xinf, e, xinf);  // x e + x_inf (1 - e)
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

// float64 HH rates from one exponential. With u = -(v + 65) every exponent of the six
// rates is a multiple of u / 720 plus a constant:
//   exp(u/80) = E^9, exp(u/20) = E^36, exp(u/18) = E^40, exp(u/10) = E^72, E = exp(u/720),
//   and -(v+35)/10 = u/10 + 3, -(v+40)/10 = u/10 + 5/2, -(v+55)/10 = u/10 + 1,
// so the powers replace five exponentials (relative error below 72 ulp of E, < 1e-14).
// x / expm1(z) is the series near z = 0 and x / (e^z - 1) elsewhere, selected without a
// branch. The clamps of save_exp / x_over_expm1 at 20 are inactive for u <= 170
// (v >= -235 mV); other voltages take the direct formulas (a wave-uniform branch in
// practice).
__device__ __forceinline__ void hh_rates_f64(double v, double& am, double& bm, double& ah,
                                             double& bh, double& an, double& bn) {
  const double u = -(v + 65.0);
  if (!(u <= 170.0)) {
    hh_rates(v, am, bm, ah, bh, an, bn);
    return;
  }
  const double e1 = exp_f64(u * (1.0 / 720));
  const double e2 = e1 * e1, e4 = e2 * e2, e8 = e4 * e4, e9 = e8 * e1;
  const double e18 = e9 * e9, e36 = e18 * e18, e40 = e36 * e4, e72 = e36 * e36;
  constexpr double kExp1 = 2.718281828459045235;    // e
  constexpr double kExp25 = 12.18249396070347343;   // e^(5/2)
  constexpr double kExp3 = 20.08553692318766774;    // e^3
  bm = 4.0 * e40;
  ah = 0.07 * e36;
  bn = 0.125 * e9;
  bh = drcp(e72 * kExp3 + 1.0);
  // -(v + 40) = u + 25, -(v + 55) = u + 10; z = x / 10.
  const double xm = u + 25.0, xn = u + 10.0;
  const double zm = xm * 0.1, zn = xn * 0.1;
  const double am_s = z_over_expm1_series(zm);           // 0.1 * 10 * z / expm1(z)
  const double an_s = 0.1 * z_over_expm1_series(zn);     // 0.01 * 10 * z / expm1(z)
  const double am_d = 0.1 * xm * drcp(e72 * kExp25 - 1.0);
  const double an_d = 0.01 * xn * drcp(e72 * kExp1 - 1.0);
  am = fabs(zm) < 0.25 ? am_s : am_d;
  an = fabs(zn) < 0.25 ? an_s : an_d;
}

// dx/dv of x / expm1(x / 10) (the HH alpha_m / alpha_n shape, solver_gate.py:14-30) at
// z = x / 10 with q = e^z - 1 given: 1 / q - x e^z / (10 q^2), or for |z| < 1/4 the
// derivative of its Bernoulli series, -1/2 + z/6 - z^3/180 + z^5/5040 - z^7/151200
// + z^9/4790016 - 691 z^11/108972864000 (next term below 3e-18), selected without a
// branch.
__device__ __forceinline__ double d_x_over_expm1_10(double x, double z, double q) {
  const double y = z * z;
  double p = kc(-6.34101567625469022e-09);                     // -691/108972864000
  p = fma(p, y, kc(2.08767569878680989792e-07));               // 1/4790016
  p = fma(p, y, kc(-6.61375661375661375661e-06));              // -1/151200
  p = fma(p, y, kc(1.98412698412698412698e-04));               // 1/5040
  p = fma(p, y, kc(-5.55555555555555555556e-03));              // -1/180
  p = fma(p, y, kc(1.66666666666666666667e-01));               // 1/6
  const double ser = fma(p, z, -0.5);
  const double rq = drcp(q);
  const double dir = rq - x * (q + 1.0) * rq * rq * 0.1;
  return fabs(z) < 0.25 ? ser : dir;
}

// HH rates and their derivatives with respect to v (channels/hh.py:87-108 as autograd
// differentiates them), float64, from one exponential as hh_rates_f64; other voltages
// (v < -235 mV) take the direct formulas.
__device__ __forceinline__ void hh_rates_dv_f64(double v, double& am, double& bm, double& ah,
                                                double& bh, double& an, double& bn, double& dam,
                                                double& dbm, double& dah, double& dbh,
                                                double& dan, double& dbn);

// The rates as the fused kernels evaluate them: the shared-exponential form in float64,
// the direct form in float32 (where powers of E would lose accuracy).
__device__ __forceinline__ void hh_rates_step(float v, float& am, float& bm, float& ah, float& bh,
                                              float& an, float& bn) {
  hh_rates(v, am, bm, ah, bh, an, bn);
}
__device__ __forceinline__ void hh_rates_step(double v, double& am, double& bm, double& ah,
                                              double& bh, double& an, double& bn) {
  hh_rates_f64(v, am, bm, ah, bh, an, bn);
}

// d/dx of x_over_expm1(x, y) as autograd differentiates solver_gate.py:14-30.
template <typename T>
__device__ __forceinline__ T d_x_over_expm1(T x, T y) {
  const T z = x / y;
  const T zc = z < T(20) ? z : T(20);
  const T pass = z <= T(20) ? T(1) : T(0);
  if (fabs(zc) < T(1e-4)) return pass * (T(-0.5) + zc / T(6));
  const T em1 = dexpm1(zc);
  return T(1) / em1 - x * (em1 + T(1)) / (em1 * em1) * pass / y;
}

template <typename T>
__device__ __forceinline__ T d_save_exp(T u) {
  return u <= T(20) ? dexp(u) : T(0);
}

// HH rates and their derivatives with respect to v by the reference's formulas (any
// type, any voltage).
template <typename T>
__device__ __forceinline__ void hh_rates_dv(T v, T& am, T& bm, T& ah, T& bh, T& an, T& bn, T& dam,
                                            T& dbm, T& dah, T& dbh, T& dan, T& dbn) {
  const T xm = -(v + T(40)), xn = -(v + T(55));
  const T um = -(v + T(65)) / T(18), uh = -(v + T(65)) / T(20);
  const T ub = -(v + T(35)) / T(10), un = -(v + T(65)) / T(80);
  am = T(0.1) * x_over_expm1(xm, T(10));
  dam = -T(0.1) * d_x_over_expm1(xm, T(10));
  bm = T(4) * save_exp(um);
  dbm = -T(4) / T(18) * d_save_exp(um);
  ah = T(0.07) * save_exp(uh);
  dah = -T(0.07) / T(20) * d_save_exp(uh);
  const T Eb = save_exp(ub);
  bh = T(1) / (Eb + T(1));
  dbh = d_save_exp(ub) / (T(10) * (Eb + T(1)) * (Eb + T(1)));
  an = T(0.01) * x_over_expm1(xn, T(10));
  dan = -T(0.01) * d_x_over_expm1(xn, T(10));
  bn = T(0.125) * save_exp(un);
  dbn = -T(0.125) / T(80) * d_save_exp(un);
}

__device__ __forceinline__ void hh_rates_dv_f64(double v, double& am, double& bm, double& ah,
                                                double& bh, double& an, double& bn, double& dam,
                                                double& dbm, double& dah, double& dbh,
                                                double& dan, double& dbn) {
  const double u = -(v + 65.0);
  if (!(u <= 170.0)) {
    hh_rates_dv<double>(v, am, bm, ah, bh, an, bn, dam, dbm, dah, dbh, dan, dbn);
    return;
  }
  const double e1 = exp_f64(u * (1.0 / 720));
  const double e2 = e1 * e1, e4 = e2 * e2, e8 = e4 * e4, e9 = e8 * e1;
  const double e18 = e9 * e9, e36 = e18 * e18, e40 = e36 * e4, e72 = e36 * e36;
  constexpr double kExp1 = 2.718281828459045235;    // e
  constexpr double kExp25 = 12.18249396070347343;   // e^(5/2)
  constexpr double kExp3 = 20.08553692318766774;    // e^3
  bm = 4.0 * e40;
  dbm = (-4.0 / 18.0) * e40;
  ah = 0.07 * e36;
  dah = (-0.07 / 20.0) * e36;
  bn = 0.125 * e9;
  dbn = (-0.125 / 80.0) * e9;
  const double Eb = e72 * kExp3;                    // exp(-(v + 35) / 10)
  bh = drcp(Eb + 1.0);
  dbh = bh * bh * Eb * 0.1;
  const double xm = u + 25.0, xn = u + 10.0;        // -(v + 40), -(v + 55)
  const double zm = xm * 0.1, zn = xn * 0.1;
  const double qm = e72 * kExp25 - 1.0, qn = e72 * kExp1 - 1.0;
  const double am_s = z_over_expm1_series(zm);
  const double an_s = 0.1 * z_over_expm1_series(zn);
  const double am_d = 0.1 * xm * drcp(qm);
  const double an_d = 0.01 * xn * drcp(qn);
  am = fabs(zm) < 0.25 ? am_s : am_d;
  an = fabs(zn) < 0.25 ? an_s : an_d;
  dam = -0.1 * d_x_over_expm1_10(xm, zm, qm);
  dan = -0.01 * d_x_over_expm1_10(xn, zn, qn);
}

// The rates and derivatives the reverse pass evaluates (the forward's forms).
__device__ __forceinline__ void hh_rates_dv_step(float v, float& am, float& bm, float& ah, float& bh,
                                                 float& an, float& bn, float& dam, float& dbm,
                                                 float& dah, float& dbh, float& dan, float& dbn) {
  hh_rates_dv<float>(v, am, bm, ah, bh, an, bn, dam, dbm, dah, dbh, dan, dbn);
}
__device__ __forceinline__ void hh_rates_dv_step(double v, double& am, double& bm, double& ah,
                                                 double& bh, double& an, double& bn, double& dam,
                                                 double& dbm, double& dah, double& dbh, double& dan,
                                                 double& dbn) {
  hh_rates_dv_f64(v, am, bm, ah, bh, an, bn, dam, dbm, dah, dbh, dan, dbn);
}

}  // namespace jxhip
