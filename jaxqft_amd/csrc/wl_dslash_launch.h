// Kernel selection for one storage precision; included by the per-precision
// translation units wl_dslash_{f64,f32,f16}.hip so they compile in parallel.
#pragma once
#include "wl_dslash_kernels.h"
#include "wl_eo_kernels.h"

namespace wl {

template <int P, int NC, int RECON, bool DAG, bool DIAG, bool G5, int EPI>
static void launch_k(const DslashArgs& a, dim3 grid, hipStream_t stream) {
  // A launch gated on pending reliable updates (the FP64 true residual) is
  // idle in most iterations, and an idle launch still dispatches every
  // workgroup: with an even number of columns it takes the two-RHS kernel,
  // which halves the workgroups at about the one-RHS kernel's active speed.
  if constexpr (P == F64 && NC == 3) {
    if (a.gate == 2 && a.mrhs == 1 && a.march == 1 && a.cols_per_cfg % 2 == 0) {
      hipLaunchKernelGGL((k_dslash_mrhs<P, NC, RECON, DAG, DIAG, G5, EPI, 2>), dim3(a.ntiles * (a.ncol / 2)), dim3(NT),
                         a.lds_bytes, stream, a);
      return;
    }
  }
  if constexpr (P == F16 && NC == 3 && RECON == 12) {
    if (a.lbox) {
      // one workgroup per (configuration, box), all columns of the configuration inside it
      hipLaunchKernelGGL((k_dslash_box<RECON, DAG, DIAG, G5, EPI>), dim3(a.ntiles * (a.ncol / a.cols_per_cfg)),
                         dim3(NT), 0, stream, a);
      return;
    }
  }
  if constexpr (P == F16) {
    if (a.plane) {
      hipLaunchKernelGGL((k_dslash_plane<NC, RECON, DAG, DIAG, G5, EPI>), grid, dim3(NT), a.lds_bytes, stream, a);
      return;
    }
  }
  if constexpr (P == F16 && NC == 3) {
    if (a.mrhs == 2 && a.march > 1) {
      hipLaunchKernelGGL((k_dslash_mrhs_march<NC, RECON, DAG, DIAG, G5, EPI, 2>), dim3(a.ntiles * (a.ncol / 2)),
                         dim3(NT), a.lds_bytes, stream, a);
      return;
    }
  }
  if (NC == 3 && a.mrhs == 4)
    hipLaunchKernelGGL((k_dslash_mrhs<P, NC, RECON, DAG, DIAG, G5, EPI, 4>), dim3(a.ntiles * (a.ncol / 4)), dim3(NT),
                       a.lds_bytes, stream, a);
  else if (NC == 3 && a.mrhs == 2)
    hipLaunchKernelGGL((k_dslash_mrhs<P, NC, RECON, DAG, DIAG, G5, EPI, 2>), dim3(a.ntiles * (a.ncol / 2)), dim3(NT),
                       a.lds_bytes, stream, a);
  else if (a.march == 1)
    hipLaunchKernelGGL((k_dslash_site<P, NC, RECON, DAG, DIAG, G5, EPI>), grid, dim3(NT), a.lds_bytes, stream, a);
  else
    hipLaunchKernelGGL((k_dslash<P, NC, RECON, DAG, DIAG, G5, EPI>), grid, dim3(NT), a.lds_bytes, stream, a);
}

// op: 0 D, 1 D^dag, 2 gamma5 D, 3 K, 4 K^dag.  The reducing epilogues exist for
// the operator the solver uses them with: EPI_NORM on D, EPI_RESID[32|16] (FP64 only)
// and the CG steps on D^dag (an FP64 solve accumulates into FP64).
template <int P, int NC, int RECON>
static hipError_t launch_p(int op, int epi, const DslashArgs& a, hipStream_t stream) {
  dim3 grid((unsigned)(a.ntiles * a.ncol));
  if (epi == EPI_STORE) {
    switch (op) {
      case 0: launch_k<P, NC, RECON, false, true, false, EPI_STORE>(a, grid, stream); break;
      case 1: launch_k<P, NC, RECON, true, true, false, EPI_STORE>(a, grid, stream); break;
      case 2: launch_k<P, NC, RECON, false, true, true, EPI_STORE>(a, grid, stream); break;
      case 3: launch_k<P, NC, RECON, false, false, false, EPI_STORE>(a, grid, stream); break;
      case 4: launch_k<P, NC, RECON, true, false, false, EPI_STORE>(a, grid, stream); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  if (epi == EPI_NORM && op == 0) {
    launch_k<P, NC, RECON, false, true, false, EPI_NORM>(a, grid, stream);
    return hipGetLastError();
  }
  if (op != 1) return hipErrorInvalidValue;
  if constexpr (P == F64) {
    if (epi == EPI_RESID) launch_k<P, NC, RECON, true, true, false, EPI_RESID>(a, grid, stream);
    else if (epi == EPI_RESID32) launch_k<P, NC, RECON, true, true, false, EPI_RESID32>(a, grid, stream);
    else if (epi == EPI_RESID16) launch_k<P, NC, RECON, true, true, false, EPI_RESID16>(a, grid, stream);
    else if (epi == EPI_CG64) launch_k<P, NC, RECON, true, true, false, EPI_CG64>(a, grid, stream);
    else return hipErrorInvalidValue;
  } else {
    if (epi == EPI_CG64) launch_k<P, NC, RECON, true, true, false, EPI_CG64>(a, grid, stream);
    else if (epi == EPI_CG32) launch_k<P, NC, RECON, true, true, false, EPI_CG32>(a, grid, stream);
    else if (epi == EPI_CG16) launch_k<P, NC, RECON, true, true, false, EPI_CG16>(a, grid, stream);
    else return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Even-odd hop: dag (K or K^dag), mode 0 (hop) / 1 (Schur combine), epilogue
// EPI_STORE, or for the combine EPI_NORM (forward), EPI_RESID (FP64) and the CG
// steps (dagger).
template <int P, int NC, int RECON, bool D, int M, int E>
static void launch_eo_k(const DslashArgs& a, hipStream_t stream) {
  if (NC == 3 && a.mrhs == 2)
    hipLaunchKernelGGL((k_hop_eo<P, NC, RECON, D, M, E, 2>), dim3((unsigned)(a.ntiles * (a.ncol / 2))), dim3(NT), 0,
                       stream, a);
  else
    hipLaunchKernelGGL((k_hop_eo<P, NC, RECON, D, M, E, 1>), dim3((unsigned)(a.ntiles * a.ncol)), dim3(NT), 0, stream,
                       a);
}

template <int P, int NC, int RECON>
static hipError_t launch_eo_p(int dag, int mode, int epi, const DslashArgs& a, hipStream_t stream) {
#define WL_EOL(D, M, E) launch_eo_k<P, NC, RECON, D, M, E>(a, stream)
  if (epi == EPI_STORE) {
    if (mode == 0) {
      if (dag) WL_EOL(true, 0, EPI_STORE);
      else WL_EOL(false, 0, EPI_STORE);
    } else {
      if (dag) WL_EOL(true, 1, EPI_STORE);
      else WL_EOL(false, 1, EPI_STORE);
    }
    return hipGetLastError();
  }
  if (mode != 1) return hipErrorInvalidValue;
  if (epi == EPI_NORM && !dag) {
    WL_EOL(false, 1, EPI_NORM);
    return hipGetLastError();
  }
  if (!dag) return hipErrorInvalidValue;
  if constexpr (P == F64) {
    if (epi == EPI_RESID) WL_EOL(true, 1, EPI_RESID);
    else if (epi == EPI_RESID32) WL_EOL(true, 1, EPI_RESID32);
    else if (epi == EPI_RESID16) WL_EOL(true, 1, EPI_RESID16);
    else if (epi == EPI_CG64) WL_EOL(true, 1, EPI_CG64);
    else return hipErrorInvalidValue;
  } else {
    if (epi == EPI_CG64) WL_EOL(true, 1, EPI_CG64);
    else if (epi == EPI_CG32) WL_EOL(true, 1, EPI_CG32);
    else if (epi == EPI_CG16) WL_EOL(true, 1, EPI_CG16);
    else return hipErrorInvalidValue;
  }
#undef WL_EOL
  return hipGetLastError();
}

template <int P>
static hipError_t launch_eo_prec(int nc, int recon, int dag, int mode, int epi, const DslashArgs& a,
                                 hipStream_t stream) {
  if (nc == 3)
    return recon == 12 ? launch_eo_p<P, 3, 12>(dag, mode, epi, a, stream) : launch_eo_p<P, 3, 18>(dag, mode, epi, a, stream);
  if (nc == 1) return launch_eo_p<P, 1, 18>(dag, mode, epi, a, stream);
  return hipErrorInvalidValue;
}

template <int P>
static hipError_t launch_dslash_prec(int nc, int recon, int op, int epi, const DslashArgs& a, hipStream_t stream) {
  if (nc == 3) return recon == 12 ? launch_p<P, 3, 12>(op, epi, a, stream) : launch_p<P, 3, 18>(op, epi, a, stream);
  if (nc == 1) return launch_p<P, 1, 18>(op, epi, a, stream);
  return hipErrorInvalidValue;
}

}  // namespace wl
