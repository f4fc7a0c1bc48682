// Stencil kernel instantiations for F32 storage.
#include "wl_dslash_launch.h"

namespace wl {

hipError_t launch_dslash_f32(int nc, int recon, int op, int epi, const DslashArgs& a, hipStream_t stream) {
  return launch_dslash_prec<F32>(nc, recon, op, epi, a, stream);
}

hipError_t launch_eo_f32(int nc, int recon, int dag, int mode, int epi, const DslashArgs& a, hipStream_t stream) {
  return launch_eo_prec<F32>(nc, recon, dag, mode, epi, a, stream);
}

}  // namespace wl
