// C entry points of the fused forward and adjoint for one element type pair
// (gates / parameters: float, tree solve: float); the kernels are in jxhip_fused.h.
#include "jxhip_fused.h"

extern "C" {

int jxhip_fused_forward_f32(const jxhip_fused_args* args, void* stream) {
  return jxhip::launch_forward<float, float>(args, static_cast<hipStream_t>(stream));
}

int jxhip_fused_adjoint_f32(const jxhip_fused_args* args, const jxhip_adjoint_args* adj,
                              void* stream) {
  return jxhip::launch_adjoint<float, float>(args, adj, static_cast<hipStream_t>(stream));
}

}  // extern "C"
