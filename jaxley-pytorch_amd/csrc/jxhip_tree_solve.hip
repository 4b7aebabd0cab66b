// Standalone tree (Hines / DHS) solves for gfx950: the drop-ins for jaxley-pytorch's
// `tree_solve` and `tree_solve_levels` custom ops (kernels/tree_solve.py:228-375, 518-630).
//
// System, per row: A[i,i] = d[i], A[p,c] = u[c], A[c,p] = l[c]; forest in solve order.
// Triangularization leaves first (a node of height h gathers its children, which all
// have smaller heights: no atomics, deterministic), then back-substitution.
//
// * levels variant: one workgroup per (block, row); the block's diagonal and rhs are
//   staged in LDS (or, for blocks too large for LDS, in the caller's workspace), and
//   every height level touches only its own nodes, with one workgroup barrier per level.
// * regs variant: one wavefront per (block, row); lane j holds nodes j, j+64, ... in
//   VGPRs; children's elimination terms are exchanged through LDS inside the wave and
//   the back-substitution is by recursive doubling over parent jumps.
#include "jxhip_common.h"

namespace jxhip {

constexpr int kLevelThreads = 256;

template <typename T, bool kLds>
__global__ void __launch_bounds__(kLevelThreads)
tree_solve_levels_kernel(const T* __restrict__ d, const T* __restrict__ l,
                         const T* __restrict__ u, const T* __restrict__ s, T* __restrict__ x,
                         T* __restrict__ work, int64_t n, int32_t n_blocks,
                         const int32_t* __restrict__ block_info,
                         const int32_t* __restrict__ level_ptr,
                         const int32_t* __restrict__ order, const int32_t* __restrict__ parent,
                         const int32_t* __restrict__ child_rng,
                         const int32_t* __restrict__ child_idx) {
  extern __shared__ unsigned char smem_raw[];
  const int64_t flat = blockIdx.x;
  const int32_t blk = static_cast<int32_t>(flat % n_blocks);
  const int64_t row = flat / n_blocks;
  const int32_t start = block_info[4 * blk + 0];
  const int32_t len = block_info[4 * blk + 1];
  const int32_t H = block_info[4 * blk + 2];
  const int32_t lp = block_info[4 * blk + 3];
  const int64_t base = row * n + start;
  const T* L = l + base;
  const T* U = u + base;
  T* dd;
  T* xx;
  if constexpr (kLds) {
    dd = reinterpret_cast<T*>(smem_raw);
    xx = dd + len;
  } else {
    dd = work + base;
    xx = x + base;
  }
  const int tid = threadIdx.x;
  for (int i = tid; i < len; i += blockDim.x) {
    dd[i] = d[base + i];
    xx[i] = s[base + i];
  }
  __syncthreads();
  const int32_t* ord = order + start;
  const int32_t* par = parent + start;
  const int32_t* crng = child_rng + 2 * static_cast<int64_t>(start);
  for (int h = 1; h <= H; ++h) {
    const int lo = level_ptr[lp + h];
    const int hi = level_ptr[lp + h + 1];
    for (int pos = lo + tid; pos < hi; pos += blockDim.x) {
      const int k = ord[pos];
      T dk = dd[k];
      T xk = xx[k];
      const int cb = crng[2 * k], ce = crng[2 * k + 1];
      for (int j = cb; j < ce; ++j) {
        const int c = child_idx[j];
        const T f = U[c] / dd[c];
        dk -= L[c] * f;
        xk -= xx[c] * f;
      }
      dd[k] = dk;
      xx[k] = xk;
    }
    __syncthreads();
  }
  for (int h = H; h >= 0; --h) {
    const int lo = level_ptr[lp + h];
    const int hi = level_ptr[lp + h + 1];
    for (int pos = lo + tid; pos < hi; pos += blockDim.x) {
      const int k = ord[pos];
      const int p = par[k];
      T xk = xx[k];
      if (p >= 0) xk -= L[k] * xx[p];
      xx[k] = xk / dd[k];
    }
    __syncthreads();
  }
  if constexpr (kLds) {
    for (int i = tid; i < len; i += blockDim.x) x[base + i] = xx[i];
  }
}

template <typename T, int NPL>
__global__ void __launch_bounds__(64)
tree_solve_regs_kernel(const T* __restrict__ d, const T* __restrict__ l,
                       const T* __restrict__ u, const T* __restrict__ s, T* __restrict__ x,
                       int64_t n, int32_t n_blocks, const int32_t* __restrict__ block_info,
                       const int32_t* __restrict__ level_ptr, const int32_t* __restrict__ order,
                       const int32_t* __restrict__ parent, const int32_t* __restrict__ child_rng,
                       const int32_t* __restrict__ child_idx) {
  constexpr int kBlock = 64 * NPL;
  __shared__ T sa[kBlock];
  __shared__ T sb[kBlock];
  __shared__ int32_t sj[kBlock];
  const int64_t flat = blockIdx.x;
  const int32_t blk = static_cast<int32_t>(flat % n_blocks);
  const int64_t row = flat / n_blocks;
  const int32_t start = block_info[4 * blk + 0];
  const int32_t len = block_info[4 * blk + 1];
  const int32_t H = block_info[4 * blk + 2];
  const int32_t lp = block_info[4 * blk + 3];
  const int64_t base = row * n + start;
  const int lane = threadIdx.x;
  // Heights (block-relative) through LDS.
  for (int h = 0; h <= H; ++h) {
    const int lo = level_ptr[lp + h], hi = level_ptr[lp + h + 1];
    for (int pos = lo + lane; pos < hi; pos += 64) sj[order[start + pos]] = h;
  }
  __syncthreads();
  T dr[NPL], sr[NPL], lr[NPL], ur[NPL];
  int hr[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int k = lane + 64 * i;
    if (k < len) {
      dr[i] = d[base + k];
      sr[i] = s[base + k];
      lr[i] = l[base + k];
      ur[i] = u[base + k];
      hr[i] = sj[k];
    } else {
      dr[i] = T(1);
      sr[i] = T(0);
      lr[i] = T(0);
      ur[i] = T(0);
      hr[i] = -1;
    }
  }
  const int32_t* crng = child_rng + 2 * static_cast<int64_t>(start);
  for (int h = 1; h <= H; ++h) {
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      if (hr[i] == h - 1) {
        const int k = lane + 64 * i;
        const T f = ur[i] / dr[i];
        sa[k] = lr[i] * f;
        sb[k] = sr[i] * f;
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      if (hr[i] == h) {
        const int k = lane + 64 * i;
        const int cb = crng[2 * k], ce = crng[2 * k + 1];
        T dk = dr[i], sk = sr[i];
        for (int j = cb; j < ce; ++j) {
          const int c = child_idx[j];
          dk -= sa[c];
          sk -= sb[c];
        }
        dr[i] = dk;
        sr[i] = sk;
      }
    }
    __syncthreads();
  }
  // Back-substitution by recursive doubling (jaxley.dhs.gpu; solver_voltage.py:113-140).
  T le[NPL], se[NPL];
  int jr[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int k = lane + 64 * i;
    const int p = (k < len) ? parent[start + k] : -1;
    jr[i] = p >= 0 ? p : k;
    le[i] = p >= 0 ? -lr[i] / dr[i] : T(0);
    se[i] = sr[i] / dr[i];
  }
  int doublings = 0;
  while ((1 << doublings) < H + 1) ++doublings;
  for (int it = 0; it < doublings; ++it) {
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      const int k = lane + 64 * i;
      if (k < kBlock) {
        sa[k] = se[i];
        sb[k] = le[i];
        sj[k] = jr[i];
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      const int j = jr[i];
      const T sej = sa[j], lej = sb[j];
      const int jj = sj[j];
      se[i] = le[i] * sej + se[i];
      le[i] = le[i] * lej;
      jr[i] = jj;
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int k = lane + 64 * i;
    if (k < len) x[base + k] = se[i];
  }
}

template <typename T>
__global__ void tree_solve_grads_kernel(const T* __restrict__ lam, const T* __restrict__ x,
                                        const int32_t* __restrict__ parent, T* g_d, T* g_l,
                                        T* g_u, int64_t n_rows, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_rows * n) return;
  const int64_t row = i / n;
  const int64_t k = i - row * n;
  const int32_t p = parent[k];
  const T lk = lam[i], xk = x[i];
  if (g_d) g_d[i] = -lk * xk;
  if (g_u) g_u[i] = p >= 0 ? -lam[row * n + p] * xk : T(0);
  if (g_l) g_l[i] = p >= 0 ? -lk * x[row * n + p] : T(0);
}

template <typename T>
int launch_levels(const T* d, const T* l, const T* u, const T* s, T* x, T* work, int64_t n_rows,
                  int64_t n, const int32_t* block_info, int32_t n_blocks, int32_t max_block,
                  const int32_t* level_ptr, const int32_t* order, const int32_t* parent,
                  const int32_t* child_rng, const int32_t* child_idx, hipStream_t stream) {
  if (n_rows <= 0 || n_blocks <= 0) return hipSuccess;
  const int64_t grid = n_rows * n_blocks;
  if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
  const size_t lds = 2 * sizeof(T) * static_cast<size_t>(max_block);
  if (lds <= 64 * 1024) {
    hipLaunchKernelGGL((tree_solve_levels_kernel<T, true>), dim3(static_cast<unsigned>(grid)),
                       dim3(kLevelThreads), lds, stream, d, l, u, s, x, work, n, n_blocks,
                       block_info, level_ptr, order, parent, child_rng, child_idx);
  } else {
    if (work == nullptr) return hipErrorInvalidValue;
    hipLaunchKernelGGL((tree_solve_levels_kernel<T, false>), dim3(static_cast<unsigned>(grid)),
                       dim3(kLevelThreads), 0, stream, d, l, u, s, x, work, n, n_blocks,
                       block_info, level_ptr, order, parent, child_rng, child_idx);
  }
  return hipGetLastError();
}

template <typename T>
int launch_regs(const T* d, const T* l, const T* u, const T* s, T* x, int64_t n_rows, int64_t n,
                const int32_t* block_info, int32_t n_blocks, int32_t npl,
                const int32_t* level_ptr, const int32_t* order, const int32_t* parent,
                const int32_t* child_rng, const int32_t* child_idx, hipStream_t stream) {
  if (n_rows <= 0 || n_blocks <= 0) return hipSuccess;
  const int64_t grid64 = n_rows * n_blocks;
  if (grid64 > 0x7fffffffLL) return hipErrorInvalidValue;
  const dim3 grid(static_cast<unsigned>(grid64));
#define JXHIP_REGS_CASE(NPL)                                                                \
  case NPL:                                                                                 \
    hipLaunchKernelGGL((tree_solve_regs_kernel<T, NPL>), grid, dim3(64), 0, stream, d, l, u, \
                       s, x, n, n_blocks, block_info, level_ptr, order, parent, child_rng,  \
                       child_idx);                                                          \
    break;
  switch (npl) {
    JXHIP_REGS_CASE(1)
    JXHIP_REGS_CASE(2)
    JXHIP_REGS_CASE(4)
    JXHIP_REGS_CASE(8)
    JXHIP_REGS_CASE(16)
    JXHIP_REGS_CASE(32)
    default:
      return hipErrorInvalidValue;
  }
#undef JXHIP_REGS_CASE
  return hipGetLastError();
}

template <typename T>
int launch_grads(const T* lam, const T* x, const int32_t* parent, T* g_d, T* g_l, T* g_u,
                 int64_t n_rows, int64_t n, hipStream_t stream) {
  const int64_t total = n_rows * n;
  if (total <= 0) return hipSuccess;
  const int64_t blocks = (total + 255) / 256;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((tree_solve_grads_kernel<T>), dim3(static_cast<unsigned>(blocks)), dim3(256),
                     0, stream, lam, x, parent, g_d, g_l, g_u, n_rows, n);
  return hipGetLastError();
}

}  // namespace jxhip

extern "C" {

int jxhip_tree_solve_levels_f32(const float* d, const float* l, const float* u, const float* s,
                                float* x, float* work, int64_t n_rows, int64_t n,
                                const int32_t* block_info, int32_t n_blocks, int32_t max_block,
                                const int32_t* level_ptr, const int32_t* order,
                                const int32_t* parent, const int32_t* child_rng,
                                const int32_t* child_idx, void* stream) {
  return jxhip::launch_levels<float>(d, l, u, s, x, work, n_rows, n, block_info, n_blocks,
                                     max_block, level_ptr, order, parent, child_rng, child_idx,
                                     static_cast<hipStream_t>(stream));
}

int jxhip_tree_solve_levels_f64(const double* d, const double* l, const double* u,
                                const double* s, double* x, double* work, int64_t n_rows,
                                int64_t n, const int32_t* block_info, int32_t n_blocks,
                                int32_t max_block, const int32_t* level_ptr,
                                const int32_t* order, const int32_t* parent,
                                const int32_t* child_rng, const int32_t* child_idx,
                                void* stream) {
  return jxhip::launch_levels<double>(d, l, u, s, x, work, n_rows, n, block_info, n_blocks,
                                      max_block, level_ptr, order, parent, child_rng, child_idx,
                                      static_cast<hipStream_t>(stream));
}

int jxhip_tree_solve_regs_f32(const float* d, const float* l, const float* u, const float* s,
                              float* x, int64_t n_rows, int64_t n, const int32_t* block_info,
                              int32_t n_blocks, int32_t npl, const int32_t* level_ptr,
                              const int32_t* order, const int32_t* parent,
                              const int32_t* child_rng, const int32_t* child_idx, void* stream) {
  return jxhip::launch_regs<float>(d, l, u, s, x, n_rows, n, block_info, n_blocks, npl,
                                   level_ptr, order, parent, child_rng, child_idx,
                                   static_cast<hipStream_t>(stream));
}

int jxhip_tree_solve_regs_f64(const double* d, const double* l, const double* u, const double* s,
                              double* x, int64_t n_rows, int64_t n, const int32_t* block_info,
                              int32_t n_blocks, int32_t npl, const int32_t* level_ptr,
                              const int32_t* order, const int32_t* parent,
                              const int32_t* child_rng, const int32_t* child_idx, void* stream) {
  return jxhip::launch_regs<double>(d, l, u, s, x, n_rows, n, block_info, n_blocks, npl,
                                    level_ptr, order, parent, child_rng, child_idx,
                                    static_cast<hipStream_t>(stream));
}

int jxhip_tree_solve_grads_f32(const float* lam, const float* x, const int32_t* parent_glob,
                               float* g_d, float* g_l, float* g_u, int64_t n_rows, int64_t n,
                               void* stream) {
  return jxhip::launch_grads<float>(lam, x, parent_glob, g_d, g_l, g_u, n_rows, n,
                                    static_cast<hipStream_t>(stream));
}

int jxhip_tree_solve_grads_f64(const double* lam, const double* x, const int32_t* parent_glob,
                               double* g_d, double* g_l, double* g_u, int64_t n_rows, int64_t n,
                               void* stream) {
  return jxhip::launch_grads<double>(lam, x, parent_glob, g_d, g_l, g_u, n_rows, n,
                                     static_cast<hipStream_t>(stream));
}

}  // extern "C"
