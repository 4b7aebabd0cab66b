// Standalone synapse step of one synapse type in a regular layout (k synapses per
// postsynaptic target, targets sorted): the drop-in for the reference's fused
// `synapse_step` / `synapse_step_backward` custom ops (kernels/synapse_step.py:121-421).
//
// Forward (per simulation b and target t, summing its k synapses without atomics):
//   Kinetic: T = sigmoid((v_pre - v_th)/delta), a = alpha T, r = a + beta,
//            s' = a/r + (s - a/r) exp(-dt r), i = gS s' (v_post - e_syn)
//   SigmoidRate: i = -gS sigmoid((v_pre - x_off) slope)
//   slope / offset of i / area * 1e5 by the finite difference with both voltages
//   perturbed by `diff` (modules/network.py:452-478).
// Backward: one thread per synapse loops over the batch (parameter gradients summed in
// registers, deterministic); presynaptic gradients go to a (synapse, batch) buffer and
// are summed per presynaptic node through the transposed index; postsynaptic ones are
// summed per target. No atomics.
#include "jxhip_common.h"

namespace jxhip {

template <typename T>
struct SynParams {
  const T* p[6];
};

template <typename T>
__device__ __forceinline__ void syn_forward_one(int model, const SynParams<T>& P, int64_t syn,
                                                T vpre, T vp, T s, T area, T dt, T diff,
                                                T& s_new, T& i0, T& vt, T& ct) {
  const T gs = P.p[0][syn];
  if (model == 0) {
    const T e_syn = P.p[1][syn], alpha = P.p[2][syn], beta = P.p[3][syn];
    const T v_th = P.p[4][syn], delta = P.p[5][syn];
    const T tr = sigmoid((vpre - v_th) / delta);
    const T a = alpha * tr;
    const T rate = a + beta;
    const T sinf = a / rate;
    s_new = sinf + (s - sinf) * dexp(-dt * rate);
    i0 = gs * s_new * (vp - e_syn);
    const T i1 = gs * s_new * ((vp + diff) - e_syn);
    const T d0 = i0 / area * T(100000);
    const T d1 = i1 / area * T(100000);
    vt = (d1 - d0) / diff;
    ct = d0 - vt * vp;
  } else {
    const T xo = P.p[1][syn], slope = P.p[2][syn];
    i0 = -gs * sigmoid((vpre - xo) * slope);
    const T i1 = -gs * sigmoid(((vpre + diff) - xo) * slope);
    const T d0 = i0 / area * T(100000);
    const T d1 = i1 / area * T(100000);
    vt = (d1 - d0) / diff;
    ct = d0 - vt * vp;
    s_new = s;
  }
}

template <typename T>
__global__ void syn_forward_kernel(const T* __restrict__ v, const T* __restrict__ s,
                                   const int32_t* __restrict__ pre,
                                   const int32_t* __restrict__ targets, SynParams<T> P,
                                   const T* __restrict__ area, T dt, T diff, T* s_new, T* i_syn,
                                   T* lin, T* cst, int64_t B, int64_t n_nodes, int64_t E,
                                   int64_t k, int32_t model, int32_t targets_are_nodes) {
  const int64_t T_ = E / k;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= B * T_) return;
  const int64_t b = i / T_;
  const int64_t t = i - b * T_;
  const int64_t node = targets_are_nodes ? t : targets[t];
  const T vp = v[b * n_nodes + node];
  T sl = T(0), sc = T(0);
  for (int64_t j = 0; j < k; ++j) {
    const int64_t syn = t * k + j;
    const T vpre = v[b * n_nodes + pre[syn]];
    const T st = model == 0 ? s[b * E + syn] : T(0);
    T sn, i0, vt, ct;
    syn_forward_one(model, P, syn, vpre, vp, st, area[syn], dt, diff, sn, i0, vt, ct);
    if (model == 0) s_new[b * E + syn] = sn;
    if (i_syn) i_syn[b * E + syn] = i0;
    sl += vt;
    sc += ct;
  }
  lin[b * T_ + t] = sl;
  cst[b * T_ + t] = sc;
}

template <typename T>
__global__ void fill_kernel(T* p, T val, int64_t count) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) p[i] = val;
}

// Thread per synapse: the reference's `_backward_kernel` arithmetic
// (synapse_step.py:184-305), with the batch loop inside the thread.
template <typename T>
__global__ void syn_backward_kernel(const T* __restrict__ v, const T* __restrict__ s,
                                    const int32_t* __restrict__ pre,
                                    const int32_t* __restrict__ targets, SynParams<T> P,
                                    const T* __restrict__ area, T dt, T diff,
                                    const T* __restrict__ g_sn_in, const T* __restrict__ g_i_in,
                                    const T* __restrict__ g_lin, const T* __restrict__ g_cst,
                                    T* g_s, T* g_params, T* gpre_work, T* gpost_work, int64_t B,
                                    int64_t n_nodes, int64_t E, int64_t k, int32_t model,
                                    int32_t targets_are_nodes) {
  const int64_t syn = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (syn >= E) return;
  const int64_t T_ = E / k;
  const int64_t t = syn / k;
  const int64_t node = targets_are_nodes ? t : targets[t];
  const int64_t pnode = pre[syn];
  const T ar = area[syn];
  const T gs = P.p[0][syn];
  T acc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  for (int64_t b = 0; b < B; ++b) {
    const T vp = v[b * n_nodes + node];
    const T vpre = v[b * n_nodes + pnode];
    const T glin = g_lin[b * T_ + t];
    const T gconst = g_cst[b * T_ + t];
    T g_vpre, g_vp;
    if (model == 0) {
      const T e_syn = P.p[1][syn], alpha = P.p[2][syn], beta = P.p[3][syn];
      const T v_th = P.p[4][syn], delta = P.p[5][syn];
      const T st = s[b * E + syn];
      const T x = (vpre - v_th) / delta;
      const T tr = sigmoid(x);
      const T a = alpha * tr;
      const T rate = a + beta;
      const T sinf = a / rate;
      const T decay = dexp(-dt * rate);
      const T sn = sinf + (st - sinf) * decay;
      const T w0 = vp - e_syn;
      const T w1 = (vp + diff) - e_syn;
      const T i0 = gs * sn * w0, i1 = gs * sn * w1;
      const T d0 = i0 / ar * T(100000), d1 = i1 / ar * T(100000);
      const T vt = (d1 - d0) / diff;
      const T g_vt = glin - gconst * vp;
      g_vp = -gconst * vt;
      const T g_d1 = g_vt / diff;
      const T g_d0 = gconst - g_d1;
      T g_i0 = g_d0 * T(100000) / ar;
      const T g_i1 = g_d1 * T(100000) / ar;
      if (g_i_in) g_i0 += g_i_in[b * E + syn];
      const T g_w = g_i0 * w0 + g_i1 * w1;
      T g_sn = gs * g_w;
      if (g_sn_in) g_sn += g_sn_in[b * E + syn];
      g_vp += gs * sn * (g_i0 + g_i1);
      const T g_sinf = g_sn * (T(1) - decay);
      const T g_rate = -dt * decay * g_sn * (st - sinf) - g_sinf * sinf / rate;
      const T g_a = g_sinf / rate + g_rate;
      const T g_x = g_a * alpha * tr * (T(1) - tr);
      g_vpre = g_x / delta;
      g_s[b * E + syn] = g_sn * decay;
      acc[0] += sn * g_w;
      acc[1] += -gs * sn * (g_i0 + g_i1);
      acc[2] += g_a * tr;
      acc[3] += g_rate;
      acc[4] += -g_vpre;
      acc[5] += -g_vpre * x;
    } else {
      const T xo = P.p[1][syn], slope = P.p[2][syn];
      const T u0 = vpre - xo, u1 = (vpre + diff) - xo;
      const T sig0 = sigmoid(u0 * slope), sig1 = sigmoid(u1 * slope);
      const T i0 = -gs * sig0, i1 = -gs * sig1;
      const T d0 = i0 / ar * T(100000), d1 = i1 / ar * T(100000);
      const T vt = (d1 - d0) / diff;
      const T g_vt = glin - gconst * vp;
      g_vp = -gconst * vt;
      const T g_d1 = g_vt / diff;
      const T g_d0 = gconst - g_d1;
      T g_i0 = g_d0 * T(100000) / ar;
      const T g_i1 = g_d1 * T(100000) / ar;
      if (g_i_in) g_i0 += g_i_in[b * E + syn];
      const T g_z0 = -gs * g_i0 * sig0 * (T(1) - sig0);
      const T g_z1 = -gs * g_i1 * sig1 * (T(1) - sig1);
      g_vpre = (g_z0 + g_z1) * slope;
      acc[0] += -(g_i0 * sig0 + g_i1 * sig1);
      acc[1] += -g_vpre;
      acc[2] += g_z0 * u0 + g_z1 * u1;
    }
    gpre_work[syn * B + b] = g_vpre;
    gpost_work[b * E + syn] = g_vp;
  }
  const int np = model == 0 ? 6 : 3;
  for (int q = 0; q < np; ++q) g_params[q * E + syn] = acc[q];
}

template <typename T>
__global__ void syn_presum_kernel(const T* __restrict__ gpre_work, const int32_t* __restrict__ inv,
                                  const int32_t* __restrict__ pre_nodes, T* g_v, int64_t B,
                                  int64_t n_nodes, int64_t n_pre, int64_t width) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_pre * B) return;
  const int64_t j = i / B;
  const int64_t b = i - j * B;
  T acc = T(0);
  for (int64_t q = 0; q < width; ++q) {
    const int32_t syn = inv[j * width + q];
    if (syn >= 0) acc += gpre_work[static_cast<int64_t>(syn) * B + b];
  }
  g_v[b * n_nodes + pre_nodes[j]] += acc;
}

template <typename T>
__global__ void syn_postsum_kernel(const T* __restrict__ gpost_work,
                                   const int32_t* __restrict__ targets, T* g_v, int64_t B,
                                   int64_t n_nodes, int64_t E, int64_t k,
                                   int32_t targets_are_nodes) {
  const int64_t T_ = E / k;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= B * T_) return;
  const int64_t b = i / T_;
  const int64_t t = i - b * T_;
  T acc = T(0);
  for (int64_t j = 0; j < k; ++j) acc += gpost_work[b * E + t * k + j];
  const int64_t node = targets_are_nodes ? t : targets[t];
  g_v[b * n_nodes + node] += acc;
}

inline unsigned blocks_for(int64_t total) { return static_cast<unsigned>((total + 255) / 256); }

template <typename T>
int launch_syn_fwd(const T* v, const T* s, const int32_t* pre, const int32_t* targets,
                   const T* const* params, const T* area, T dt, T diff, T* s_new, T* i_syn,
                   T* lin, T* cst, int64_t B, int64_t n_nodes, int64_t E, int64_t k,
                   int32_t model, int32_t targets_are_nodes, hipStream_t stream) {
  if (k <= 0 || E % k != 0 || (model != 0 && model != 1) || params == nullptr)
    return hipErrorInvalidValue;
  SynParams<T> P;
  const int np = model == 0 ? 6 : 3;
  for (int q = 0; q < 6; ++q) P.p[q] = q < np ? params[q] : params[0];
  const int64_t total = B * (E / k);
  if (total <= 0) return hipSuccess;
  if ((total + 255) / 256 > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((syn_forward_kernel<T>), dim3(blocks_for(total)), dim3(256), 0, stream, v,
                     s, pre, targets, P, area, dt, diff, s_new, i_syn, lin, cst, B, n_nodes, E, k,
                     model, targets_are_nodes);
  return hipGetLastError();
}

template <typename T>
int launch_syn_bwd(const T* v, const T* s, const int32_t* pre, const int32_t* targets,
                   const T* const* params, const T* area, T dt, T diff, const T* g_s_new,
                   const T* g_i, const T* g_lin, const T* g_cst, T* g_v, T* g_s, T* g_params,
                   T* work, const int32_t* pre_nodes, const int32_t* inv, int64_t n_pre,
                   int64_t width, int64_t B, int64_t n_nodes, int64_t E, int64_t k,
                   int32_t model, int32_t targets_are_nodes, hipStream_t stream) {
  if (k <= 0 || E % k != 0 || (model != 0 && model != 1) || params == nullptr || work == nullptr)
    return hipErrorInvalidValue;
  SynParams<T> P;
  const int np = model == 0 ? 6 : 3;
  for (int q = 0; q < 6; ++q) P.p[q] = q < np ? params[q] : params[0];
  if (B * n_nodes > 0) {
    hipLaunchKernelGGL((fill_kernel<T>), dim3(blocks_for(B * n_nodes)), dim3(256), 0, stream, g_v,
                       T(0), B * n_nodes);
  }
  if (E <= 0 || B <= 0) return hipGetLastError();
  T* gpre_work = work;
  T* gpost_work = work + E * B;
  hipLaunchKernelGGL((syn_backward_kernel<T>), dim3(blocks_for(E)), dim3(256), 0, stream, v, s,
                     pre, targets, P, area, dt, diff, g_s_new, g_i, g_lin, g_cst, g_s, g_params,
                     gpre_work, gpost_work, B, n_nodes, E, k, model, targets_are_nodes);
  if (n_pre > 0) {
    hipLaunchKernelGGL((syn_presum_kernel<T>), dim3(blocks_for(n_pre * B)), dim3(256), 0, stream,
                       gpre_work, inv, pre_nodes, g_v, B, n_nodes, n_pre, width);
  }
  hipLaunchKernelGGL((syn_postsum_kernel<T>), dim3(blocks_for(B * (E / k))), dim3(256), 0, stream,
                     gpost_work, targets, g_v, B, n_nodes, E, k, targets_are_nodes);
  return hipGetLastError();
}

}  // namespace jxhip

extern "C" {

int jxhip_synapse_step_f32(const float* v, const float* s, const int32_t* pre,
                           const int32_t* targets, const float* const* params, const float* area,
                           float dt, float diff, float* s_new, float* i_syn, float* lin,
                           float* cst, int64_t B, int64_t n_nodes, int64_t E, int64_t k,
                           int32_t model, int32_t targets_are_nodes, void* stream) {
  return jxhip::launch_syn_fwd<float>(v, s, pre, targets, params, area, dt, diff, s_new, i_syn,
                                      lin, cst, B, n_nodes, E, k, model, targets_are_nodes,
                                      static_cast<hipStream_t>(stream));
}

int jxhip_synapse_step_f64(const double* v, const double* s, const int32_t* pre,
                           const int32_t* targets, const double* const* params,
                           const double* area, double dt, double diff, double* s_new,
                           double* i_syn, double* lin, double* cst, int64_t B, int64_t n_nodes,
                           int64_t E, int64_t k, int32_t model, int32_t targets_are_nodes,
                           void* stream) {
  return jxhip::launch_syn_fwd<double>(v, s, pre, targets, params, area, dt, diff, s_new, i_syn,
                                       lin, cst, B, n_nodes, E, k, model, targets_are_nodes,
                                       static_cast<hipStream_t>(stream));
}

int jxhip_synapse_step_backward_f32(const float* v, const float* s, const int32_t* pre,
                                    const int32_t* targets, const float* const* params,
                                    const float* area, float dt, float diff,
                                    const float* g_s_new, const float* g_i, const float* g_lin,
                                    const float* g_cst, float* g_v, float* g_s, float* g_params,
                                    float* work, const int32_t* pre_nodes, const int32_t* inv,
                                    int64_t n_pre, int64_t inv_width, int64_t B, int64_t n_nodes,
                                    int64_t E, int64_t k, int32_t model, int32_t targets_are_nodes,
                                    void* stream) {
  return jxhip::launch_syn_bwd<float>(v, s, pre, targets, params, area, dt, diff, g_s_new, g_i,
                                      g_lin, g_cst, g_v, g_s, g_params, work, pre_nodes, inv,
                                      n_pre, inv_width, B, n_nodes, E, k, model,
                                      targets_are_nodes, static_cast<hipStream_t>(stream));
}

int jxhip_synapse_step_backward_f64(const double* v, const double* s, const int32_t* pre,
                                    const int32_t* targets, const double* const* params,
                                    const double* area, double dt, double diff,
                                    const double* g_s_new, const double* g_i,
                                    const double* g_lin, const double* g_cst, double* g_v,
                                    double* g_s, double* g_params, double* work,
                                    const int32_t* pre_nodes, const int32_t* inv, int64_t n_pre,
                                    int64_t inv_width, int64_t B, int64_t n_nodes, int64_t E,
                                    int64_t k, int32_t model, int32_t targets_are_nodes,
                                    void* stream) {
  return jxhip::launch_syn_bwd<double>(v, s, pre, targets, params, area, dt, diff, g_s_new, g_i,
                                       g_lin, g_cst, g_v, g_s, g_params, work, pre_nodes, inv,
                                       n_pre, inv_width, B, n_nodes, E, k, model,
                                       targets_are_nodes, static_cast<hipStream_t>(stream));
}

}  // extern "C"
