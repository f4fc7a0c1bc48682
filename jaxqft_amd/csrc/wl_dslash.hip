// Wilson operator kernels: D, D^dag, gamma5 D, the hopping term K and K^dag,
// with optional fused reductions for CG (Re<p, A p> and |b - A x|^2).
//
// One site per thread, one spinor column per workgroup row; the 1D grid is
// ntiles x ncol workgroups ordered column-fastest and remapped so that each
// XCD walks one contiguous range of tiles: the k columns of one tile run
// back to back on the same XCD and share its link lines in L2.
#include "wl_kernels.h"
#include "wl_reduce.h"
#include "wl_stencil.h"

namespace wl {

// gamma5 = sigma1 x sigma3 in this basis: (g5 psi)_s = (psi2, -psi3, psi0, -psi1)
template <class T, int NC>
__device__ __forceinline__ void apply_g5(cx<T> (&v)[4 * NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    cx<T> a0 = v[0 * NC + c], a1 = v[1 * NC + c], a2 = v[2 * NC + c], a3 = v[3 * NC + c];
    v[0 * NC + c] = a2;
    v[1 * NC + c] = {-a3.re, -a3.im};
    v[2 * NC + c] = a0;
    v[3 * NC + c] = {-a1.re, -a1.im};
  }
}

// EPI: 0 store, 1 store + Re<w, out>, 2 out = w - result and |out|^2
template <int P, int NC, int RECON, bool DAG, bool DIAG, bool G5, int EPI>
__global__ __launch_bounds__(NT) void k_dslash(DslashArgs a) {
  using T = typename PT<P>::acc;
  const unsigned logical = xcd_tile(blockIdx.x, gridDim.x);
  const int col = logical % a.ncol;
  const int tile = logical / a.ncol;
  const ColState* st = a.red.cols;
  if (a.skip_inactive && !st[col].active) return;
  if (a.only_update && (!st[col].need_update || st[col].done)) return;

  const size_t V = (size_t)a.g.V;
  const size_t site = (size_t)tile * NT + threadIdx.x;
  const bool valid = site < V;
  const size_t s = valid ? site : 0;

  const int cfg = col / a.cols_per_cfg;
  const char* Ucfg = reinterpret_cast<const char*>(a.U) + (size_t)cfg * a.cfg_bytes;
  const char* inb = reinterpret_cast<const char*>(a.in) + (size_t)col * a.col_bytes;
  char* outb = reinterpret_cast<char*>(a.out) + (size_t)col * a.col_bytes;

  // coordinates (last axis fastest)
  int x[4];
  {
    size_t rem = s;
    x[3] = rem % a.g.L[3];
    rem /= a.g.L[3];
    x[2] = rem % a.g.L[2];
    rem /= a.g.L[2];
    x[1] = rem % a.g.L[1];
    x[0] = rem / a.g.L[1];
  }

  cx<T> acc[4 * NC];
#pragma unroll
  for (int c = 0; c < 4 * NC; ++c) acc[c] = {T(0), T(0)};

  constexpr int SF = DAG ? +1 : -1;
#define WL_HOP(MU)                                                                                        \
  {                                                                                                       \
    const int L = a.g.L[MU];                                                                              \
    const size_t str = (size_t)a.g.stride[MU];                                                            \
    const bool top = x[MU] == L - 1, bot = x[MU] == 0;                                                    \
    const size_t sf = top ? s - (size_t)(L - 1) * str : s + str;                                          \
    const size_t sb = bot ? s + (size_t)(L - 1) * str : s - str;                                          \
    cx<T> ph = {T(a.bc_re[MU]), T(a.bc_im[MU])};                                                          \
    hop_dir<P, NC, RECON, MU, SF>(Ucfg, a.dir_bytes, inb, V, s, sf, sb, top, bot, ph, acc);               \
  }
  WL_HOP(0)
  WL_HOP(1)
  WL_HOP(2)
  WL_HOP(3)
#undef WL_HOP

  cx<T> res[4 * NC];
  if (DIAG) {
    const T m0 = a.m0_from_state ? T(st[col].m0) : T(a.m0);
    cx<T> p[4 * NC];
    SpinorIO<P, NC>::load(inb, V, s, p);
#pragma unroll
    for (int c = 0; c < 4 * NC; ++c) res[c] = {m0 * p[c].re - T(0.5) * acc[c].re, m0 * p[c].im - T(0.5) * acc[c].im};
  } else {
#pragma unroll
    for (int c = 0; c < 4 * NC; ++c) res[c] = {T(-0.5) * acc[c].re, T(-0.5) * acc[c].im};
  }
  if (G5) apply_g5<T, NC>(res);

  double part = 0.0;
  if (EPI == 2) {
    const char* wb = reinterpret_cast<const char*>(a.w) + (size_t)col * a.col_bytes;
    cx<T> w[4 * NC];
    SpinorIO<P, NC>::load(wb, V, s, w);
#pragma unroll
    for (int c = 0; c < 4 * NC; ++c) {
      res[c] = csub(w[c], res[c]);
      part += (double)res[c].re * (double)res[c].re + (double)res[c].im * (double)res[c].im;
    }
  } else if (EPI == 1) {
    const char* wb = reinterpret_cast<const char*>(a.w) + (size_t)col * a.col_bytes;
    cx<T> w[4 * NC];
    SpinorIO<P, NC>::load(wb, V, s, w);
#pragma unroll
    for (int c = 0; c < 4 * NC; ++c) part += (double)w[c].re * (double)res[c].re + (double)w[c].im * (double)res[c].im;
  }
  if (valid) SpinorIO<P, NC>::store(outb, V, s, res);
  if (EPI != 0) {
    if (!valid) part = 0.0;
    finish_reduction(a.red, col, tile, a.ntiles, part);
  }
}

// ---------------------------------------------------------------------------
// host-side dispatch
// ---------------------------------------------------------------------------
template <int P, int NC, int RECON>
static hipError_t launch_p(int op, int epi, const DslashArgs& a, hipStream_t stream) {
  dim3 grid((unsigned)(a.ntiles * a.ncol)), block(NT);
  // op: 0 D, 1 Ddag, 2 g5 D, 3 K, 4 Kdag
  if (epi == 0) {
    switch (op) {
      case 0: hipLaunchKernelGGL((k_dslash<P, NC, RECON, false, true, false, 0>), grid, block, 0, stream, a); break;
      case 1: hipLaunchKernelGGL((k_dslash<P, NC, RECON, true, true, false, 0>), grid, block, 0, stream, a); break;
      case 2: hipLaunchKernelGGL((k_dslash<P, NC, RECON, false, true, true, 0>), grid, block, 0, stream, a); break;
      case 3: hipLaunchKernelGGL((k_dslash<P, NC, RECON, false, false, false, 0>), grid, block, 0, stream, a); break;
      case 4: hipLaunchKernelGGL((k_dslash<P, NC, RECON, true, false, false, 0>), grid, block, 0, stream, a); break;
      default: return hipErrorInvalidValue;
    }
  } else if (epi == 1 && op == 1) {
    hipLaunchKernelGGL((k_dslash<P, NC, RECON, true, true, false, 1>), grid, block, 0, stream, a);
  } else if (epi == 2 && op == 1) {
    hipLaunchKernelGGL((k_dslash<P, NC, RECON, true, true, false, 2>), grid, block, 0, stream, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_dslash(int prec, int nc, int recon, int op, int epi, const DslashArgs& a, hipStream_t stream) {
  if (nc == 3) {
    if (prec == F64) return recon == 12 ? launch_p<F64, 3, 12>(op, epi, a, stream) : launch_p<F64, 3, 18>(op, epi, a, stream);
    if (prec == F32) return recon == 12 ? launch_p<F32, 3, 12>(op, epi, a, stream) : launch_p<F32, 3, 18>(op, epi, a, stream);
    if (prec == F16) return recon == 12 ? launch_p<F16, 3, 12>(op, epi, a, stream) : launch_p<F16, 3, 18>(op, epi, a, stream);
  } else if (nc == 1) {
    if (prec == F64) return launch_p<F64, 1, 18>(op, epi, a, stream);
    if (prec == F32) return launch_p<F32, 1, 18>(op, epi, a, stream);
    if (prec == F16) return launch_p<F16, 1, 18>(op, epi, a, stream);
  }
  return hipErrorInvalidValue;
}

}  // namespace wl
