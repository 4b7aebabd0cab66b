// Standalone HH channel step (the eager `_step_channels` of one HH channel,
// modules/base.py:3414-3542 with channels/hh.py:37-67) for the per-step path.
//
// The reference linearizes the current by a finite difference with diff = 1e-3
// (modules/base.py:3520-3542). HH currents are linear in v at fixed gates, so the
// slope is exactly the total conductance; it is computed analytically here (identical
// up to round-off in float64, and free of the float32 cancellation of the difference).
#include "jxhip_common.h"

namespace jxhip {

template <typename T>
struct HHParams {
  const T* p[6];
  int64_t stride[6];
};

template <typename T>
__global__ void hh_step_kernel(const T* __restrict__ v, int64_t n_nodes,
                               const int32_t* __restrict__ inds, int64_t n_ch,
                               const T* __restrict__ m, const T* __restrict__ h,
                               const T* __restrict__ n, HHParams<T> P, T dt, T* m_out, T* h_out,
                               T* n_out, T* i_out, T* lin_out, T* cst_out, int64_t n_rows) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_rows * n_ch) return;
  const int64_t row = i / n_ch;
  const int64_t j = i - row * n_ch;
  const T vv = v[row * n_nodes + inds[j]];
  T am, bm, ah, bh, an, bn;
  hh_rates_step(vv, am, bm, ah, bh, an, bn);
  const T mm = gate_step(m[i], dt, am, bm);
  const T hh = gate_step(h[i], dt, ah, bh);
  const T nn = gate_step(n[i], dt, an, bn);
  const T gNa = P.p[0][row * P.stride[0] + j] * (mm * mm * mm) * hh;
  const T nn2 = nn * nn;
  const T gK = P.p[1][row * P.stride[1] + j] * (nn2 * nn2);
  const T gL = P.p[2][row * P.stride[2] + j];
  const T eNa = P.p[3][row * P.stride[3] + j];
  const T eK = P.p[4][row * P.stride[4] + j];
  const T eL = P.p[5][row * P.stride[5] + j];
  const T cur = gNa * (vv - eNa) + gK * (vv - eK) + gL * (vv - eL);
  const T g = gNa + gK + gL;
  m_out[i] = mm;
  h_out[i] = hh;
  n_out[i] = nn;
  if (i_out) i_out[i] = cur;
  lin_out[i] = g;
  cst_out[i] = gNa * eNa + gK * eK + gL * eL;
}

template <typename T>
int launch_hh(const T* v, int64_t n_nodes, const int32_t* inds, int64_t n_ch, const T* m,
              const T* h, const T* n, const T* const* params, const int64_t* strides, T dt,
              T* m_out, T* h_out, T* n_out, T* i_out, T* lin_out, T* cst_out, int64_t n_rows,
              hipStream_t stream) {
  const int64_t total = n_rows * n_ch;
  if (total <= 0) return hipSuccess;
  if (params == nullptr || strides == nullptr) return hipErrorInvalidValue;
  HHParams<T> P;
  for (int q = 0; q < 6; ++q) {
    if (params[q] == nullptr) return hipErrorInvalidValue;
    P.p[q] = params[q];
    P.stride[q] = strides[q];
  }
  const int64_t blocks = (total + 255) / 256;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((hh_step_kernel<T>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                     stream, v, n_nodes, inds, n_ch, m, h, n, P, dt, m_out, h_out, n_out, i_out,
                     lin_out, cst_out, n_rows);
  return hipGetLastError();
}

}  // namespace jxhip

extern "C" {

int jxhip_hh_step_f32(const float* v, int64_t n_nodes, const int32_t* inds, int64_t n_ch,
                      const float* m, const float* h, const float* n,
                      const float* const* params, const int64_t* p_row_stride, float dt,
                      float* m_out, float* h_out, float* n_out, float* i_out, float* lin_out,
                      float* cst_out, int64_t n_rows, void* stream) {
  return jxhip::launch_hh<float>(v, n_nodes, inds, n_ch, m, h, n, params, p_row_stride, dt,
                                 m_out, h_out, n_out, i_out, lin_out, cst_out, n_rows,
                                 static_cast<hipStream_t>(stream));
}

int jxhip_hh_step_f64(const double* v, int64_t n_nodes, const int32_t* inds, int64_t n_ch,
                      const double* m, const double* h, const double* n,
                      const double* const* params, const int64_t* p_row_stride, double dt,
                      double* m_out, double* h_out, double* n_out, double* i_out,
                      double* lin_out, double* cst_out, int64_t n_rows, void* stream) {
  return jxhip::launch_hh<double>(v, n_nodes, inds, n_ch, m, h, n, params, p_row_stride, dt,
                                  m_out, h_out, n_out, i_out, lin_out, cst_out, n_rows,
                                  static_cast<hipStream_t>(stream));
}

}  // extern "C"
