// Stencil kernel instantiations for F32 storage (one translation unit per
// precision so the builds run in parallel).
#include "wl_dslash_kernels.h"

namespace wl {

template <int P, int NC, int RECON, bool DAG, bool DIAG, bool G5, int EPI>
static void launch_k(const DslashArgs& a, dim3 grid, hipStream_t stream) {
  if (NC == 3 && a.mrhs == 4)
    hipLaunchKernelGGL((k_dslash_mrhs<P, NC, RECON, DAG, DIAG, G5, EPI, 4>), dim3(a.ntiles * (a.ncol / 4)), dim3(NT), 0,
                       stream, a);
  else if (NC == 3 && a.mrhs == 2)
    hipLaunchKernelGGL((k_dslash_mrhs<P, NC, RECON, DAG, DIAG, G5, EPI, 2>), dim3(a.ntiles * (a.ncol / 2)), dim3(NT), 0,
                       stream, a);
  else if (a.march == 1)
    hipLaunchKernelGGL((k_dslash_site<P, NC, RECON, DAG, DIAG, G5, EPI>), grid, dim3(NT), 0, stream, a);
  else
    hipLaunchKernelGGL((k_dslash<P, NC, RECON, DAG, DIAG, G5, EPI>), grid, dim3(NT), 0, stream, a);
}

template <int P, int NC, int RECON>
static hipError_t launch_p(int op, int epi, const DslashArgs& a, hipStream_t stream) {
  dim3 grid((unsigned)(a.ntiles * a.ncol));
  // op: 0 D, 1 Ddag, 2 g5 D, 3 K, 4 Kdag
  if (epi == 0) {
    switch (op) {
      case 0: launch_k<P, NC, RECON, false, true, false, 0>(a, grid, stream); break;
      case 1: launch_k<P, NC, RECON, true, true, false, 0>(a, grid, stream); break;
      case 2: launch_k<P, NC, RECON, false, true, true, 0>(a, grid, stream); break;
      case 3: launch_k<P, NC, RECON, false, false, false, 0>(a, grid, stream); break;
      case 4: launch_k<P, NC, RECON, true, false, false, 0>(a, grid, stream); break;
      default: return hipErrorInvalidValue;
    }
  } else if (epi == 1 && op == 1) {
    launch_k<P, NC, RECON, true, true, false, 1>(a, grid, stream);
  } else if (epi == 2 && op == 1) {
    launch_k<P, NC, RECON, true, true, false, 2>(a, grid, stream);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_dslash_f32(int nc, int recon, int op, int epi, const DslashArgs& a, hipStream_t stream) {
  if (nc == 3) return recon == 12 ? launch_p<F32, 3, 12>(op, epi, a, stream) : launch_p<F32, 3, 18>(op, epi, a, stream);
  if (nc == 1) return launch_p<F32, 1, 18>(op, epi, a, stream);
  return hipErrorInvalidValue;
}

}  // namespace wl
