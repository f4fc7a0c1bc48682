// Wilson operator dispatch: D, D^dag, gamma5 D, the hopping term K and K^dag,
// with the epilogues of the fused CG iteration (Epilogue in wl_kernels.h).
// The kernels are in wl_dslash_kernels.h, instantiated per storage precision
// in wl_dslash_{f64,f32,f16}.hip; a reducing epilogue is followed by the
// per-column finish kernel.
#include "wl_device.h"
#include "wl_kernels.h"

namespace wl {

hipError_t launch_dslash_f64(int nc, int recon, int op, int epi, const DslashArgs& a, hipStream_t stream);
hipError_t launch_dslash_f32(int nc, int recon, int op, int epi, const DslashArgs& a, hipStream_t stream);
hipError_t launch_dslash_f16(int nc, int recon, int op, int epi, const DslashArgs& a, hipStream_t stream);
hipError_t launch_reduce_finish(const Reduction& red, int ncol, int nblk, int gate, hipStream_t stream);
hipError_t launch_eo_f64(int nc, int recon, int dag, int mode, int epi, const DslashArgs& a, hipStream_t stream);
hipError_t launch_eo_f32(int nc, int recon, int dag, int mode, int epi, const DslashArgs& a, hipStream_t stream);
hipError_t launch_eo_f16(int nc, int recon, int dag, int mode, int epi, const DslashArgs& a, hipStream_t stream);

hipError_t launch_dslash(int prec, int nc, int recon, int op, int epi, const DslashArgs& a, hipStream_t stream) {
  hipError_t e;
  if (prec == F64) e = launch_dslash_f64(nc, recon, op, epi, a, stream);
  else if (prec == F32) e = launch_dslash_f32(nc, recon, op, epi, a, stream);
  else if (prec == F16) e = launch_dslash_f16(nc, recon, op, epi, a, stream);
  else return hipErrorInvalidValue;
  if (e != hipSuccess || epi == EPI_STORE) return e;
  return launch_reduce_finish(a.red, a.ncol, a.ntiles, a.gate, stream);
}

// Even-odd hop / Schur combine (wl_eo_kernels.h), followed by the finish
// kernel for a reducing epilogue.
hipError_t launch_eo(int prec, int nc, int recon, int dag, int mode, int epi, const DslashArgs& a, hipStream_t stream) {
  hipError_t e;
  if (prec == F64) e = launch_eo_f64(nc, recon, dag, mode, epi, a, stream);
  else if (prec == F32) e = launch_eo_f32(nc, recon, dag, mode, epi, a, stream);
  else if (prec == F16) e = launch_eo_f16(nc, recon, dag, mode, epi, a, stream);
  else return hipErrorInvalidValue;
  if (e != hipSuccess || epi == EPI_STORE) return e;
  return launch_reduce_finish(a.red, a.ncol, a.ntiles, a.gate, stream);
}

}  // namespace wl
