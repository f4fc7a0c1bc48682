// Stencil kernel instantiations for F16 storage.
#include "wl_dslash_launch.h"

namespace wl {

hipError_t launch_dslash_f16(int nc, int recon, int op, int epi, const DslashArgs& a, hipStream_t stream) {
  return launch_dslash_prec<F16>(nc, recon, op, epi, a, stream);
}

hipError_t launch_eo_f16(int nc, int recon, int dag, int mode, int epi, const DslashArgs& a, hipStream_t stream) {
  return launch_eo_prec<F16>(nc, recon, dag, mode, epi, a, stream);
}

}  // namespace wl
