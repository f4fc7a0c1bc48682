// Analytic pseudofermion force of the Nf=2 Wilson action from a solve
// (su3_wilson_nf2.py:1556-1576; U(1): u1_wilson_nf2.py:1004-1025).  With
// X = (D^dag D)^-1 phi and Y = D X, for every link
//   c1[a][c] = sum_{s,t} conj(Y(x)[s][a]) (1 - g_mu)[s][t] (U~_mu(x) X(x+mu))[t][c]
//   c2[d][c] = sum_{s,t,a} conj(Y(x+mu)[s][a]) (1 + g_mu)[s][t] conj(U~_mu(x)[d][a]) X(x)[t][c]
//   F_mu(x) = TA(k),  k[i][j] = c1[j][i] - c2[j][i],
// TA the traceless anti-Hermitian part (U(1): (k - conj k) / 2), U~ the link
// with the fermion boundary phase (Wilson r = 1).  One thread per link; F is
// written in the reference layout, X and Y are read from the reference layout
// (a site's spinor and a link's matrix are contiguous, so a wavefront's
// accesses are contiguous runs) or from resident FP64 fields.
#include "wl_device.h"
#include "wl_kernels.h"
#include "wl_stencil.h"

namespace wl {

template <class S, int NC>
__device__ __forceinline__ void load_aos(const S* __restrict__ base, size_t idx, cx<double> (&v)[4][NC]) {
  const S* p = base + idx * (4 * NC) * 2;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < NC; ++c) v[s][c] = {(double)p[2 * (s * NC + c)], (double)p[2 * (s * NC + c) + 1]};
}

// The spinor of configuration cfg at site idx: from the reference layout, or
// (SOA) from a resident FP64 field, whose planes give each wavefront's load
// one contiguous run.  Same values either way.
template <bool SOA, class S, int NC>
__device__ __forceinline__ void load_spinor(const ForceArgs& a, const void* base, int cfg, size_t idx,
                                            cx<double> (&v)[4][NC]) {
  if constexpr (SOA) {
    const char* col = reinterpret_cast<const char*>(base) + (size_t)cfg * a.col_bytes;
    SpinorIO<F64, NC>::load(col, (size_t)a.g.Vp, idx, reinterpret_cast<cx<double>(&)[4 * NC]>(v));
  } else {
    const S* b = reinterpret_cast<const S*>(base) + (size_t)cfg * (size_t)a.g.V * (4 * NC) * 2;
    load_aos<S, NC>(b, idx, v);
  }
}

// One link's force through the projected rows.  With r = 1, (1 + S g_mu) =
// 2 P (P a Hermitian projector), and every spin row rebuilt from the two
// independent rows h0, h1 of spin_project carries a unimodular coefficient,
// so for spinors a, b and any colour indices
//   a^dag (1 + S g_mu) b = w_mu sum_k conj(ha_k) hb_k,   w_mu = 1 (mu < 3), 1/2 (mu = 3)
// (for mu = 3 the rows are 2 p).  The colour multiply commutes with the spin
// projection, so
//   c1[a][c] = w sum_k conj(hY_k[a]) (U hX_k)[c],   hY = rows of (1 - g) Y(x),   hX = rows of (1 - g) X(x+mu)
//   c2[d][c] = w sum_k conj((U hY'_k)[d]) hX'_k[c], hY' = rows of (1 + g) Y(x+mu), hX' = rows of (1 + g) X(x)
// Each spinor is projected as soon as it is loaded, so only the link, the
// accumulator and one pair of projected rows are live at a time.
template <int MU, int NC, class S, bool SOA>
__device__ __forceinline__ void force_link(const ForceArgs& a, int cfg, size_t site, const int (&x)[4]) {
  const size_t V = (size_t)a.g.V;
  const int L = a.g.L[MU];
  const size_t str = (size_t)a.g.stride[MU];
  const bool top = x[MU] == L - 1;
  const size_t sp = top ? site - (size_t)(L - 1) * str : site + str;
  const char* Ucfg = reinterpret_cast<const char*>(a.U) + (size_t)cfg * a.cfg_bytes;
  cx<double> u[NC][NC];
  LinkIO<F64, NC, 18>::load(Ucfg + MU * a.dir_bytes, (size_t)a.g.Vp, site, u);
  if (top) {
    const cx<double> ph = {a.bc_re[MU], a.bc_im[MU]};
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) u[i][j] = cmul(ph, u[i][j]);
  }
  constexpr double w = (MU == 3) ? 0.5 : 1.0;
  cx<double> k[NC][NC];  // k[i][j] = c1[j][i] - c2[j][i]
  {
    cx<double> hY[2][NC], uh[2][NC];
    {
      cx<double> t[4][NC];
      load_spinor<SOA, S, NC>(a, a.Y, cfg, site, t);
      spin_project<MU, -1>(reinterpret_cast<const cx<double>(&)[4 * NC]>(t), hY);
    }
    {
      cx<double> t[4][NC], hX[2][NC];
      load_spinor<SOA, S, NC>(a, a.X, cfg, sp, t);
      spin_project<MU, -1>(reinterpret_cast<const cx<double>(&)[4 * NC]>(t), hX);
      color_mul<false>(u, hX, uh);
    }
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const cx<double> v = cadd(cmulc(hY[0][j], uh[0][i]), cmulc(hY[1][j], uh[1][i]));
        k[i][j] = {w * v.re, w * v.im};
      }
  }
  {
    cx<double> uy[2][NC], hX0[2][NC];
    {
      cx<double> t[4][NC], hYp[2][NC];
      load_spinor<SOA, S, NC>(a, a.Y, cfg, sp, t);
      spin_project<MU, +1>(reinterpret_cast<const cx<double>(&)[4 * NC]>(t), hYp);
      color_mul<false>(u, hYp, uy);
    }
    {
      cx<double> t[4][NC];
      load_spinor<SOA, S, NC>(a, a.X, cfg, site, t);
      spin_project<MU, +1>(reinterpret_cast<const cx<double>(&)[4 * NC]>(t), hX0);
    }
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const cx<double> v = cadd(cmulc(uy[0][j], hX0[0][i]), cmulc(uy[1][j], hX0[1][i]));
        k[i][j] = {k[i][j].re - w * v.re, k[i][j].im - w * v.im};
      }
  }
  // traceless anti-Hermitian part
  cx<double> f[NC][NC];
  double tri = 0.0;
#pragma unroll
  for (int i = 0; i < NC; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      f[i][j] = {0.5 * (k[i][j].re - k[j][i].re), 0.5 * (k[i][j].im + k[j][i].im)};
      if (i == j) tri += f[i][j].im;
    }
  if (NC == 3) {
#pragma unroll
    for (int i = 0; i < NC; ++i) f[i][i].im -= tri / 3.0;
  }
  const size_t li = (a.layout == 0) ? (((size_t)cfg * 4 + MU) * V + site) : (((size_t)cfg * V + site) * 4 + MU);
  S* out = reinterpret_cast<S*>(a.F) + li * NC * NC * 2;
#pragma unroll
  for (int i = 0; i < NC; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      out[2 * (i * NC + j)] = (S)f[i][j].re;
      out[2 * (i * NC + j) + 1] = (S)f[i][j].im;
    }
}

// One thread per link: 64 sites per workgroup, each wavefront one direction
// (so the direction is uniform within a wavefront and selects a compile-time
// path); the four wavefronts read the same 64 sites' X and Y (L1/L2 reuse).
constexpr int kForceSites = NT / 4;

template <int NC, class S, bool SOA>
__global__ __launch_bounds__(NT) void k_pf_force(ForceArgs a) {
  const int mu = (int)(threadIdx.x / kForceSites);
  const size_t site = (size_t)blockIdx.x * kForceSites + (threadIdx.x % kForceSites);
  const int cfg = blockIdx.y;
  if (site >= (size_t)a.g.V) return;
  int x[4];
  {
    size_t rem = site;
    x[3] = (int)(rem % a.g.L[3]);
    rem /= a.g.L[3];
    x[2] = (int)(rem % a.g.L[2]);
    rem /= a.g.L[2];
    x[1] = (int)(rem % a.g.L[1]);
    x[0] = (int)(rem / a.g.L[1]);
  }
  switch (mu) {
    case 0: force_link<0, NC, S, SOA>(a, cfg, site, x); break;
    case 1: force_link<1, NC, S, SOA>(a, cfg, site, x); break;
    case 2: force_link<2, NC, S, SOA>(a, cfg, site, x); break;
    default: force_link<3, NC, S, SOA>(a, cfg, site, x); break;
  }
}

template <bool SOA>
static void launch_pf_force_t(int nc, int dtype, const ForceArgs& a, dim3 grid, dim3 block, hipStream_t st) {
  if (nc == 3) {
    if (dtype == 0) hipLaunchKernelGGL((k_pf_force<3, double, SOA>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((k_pf_force<3, float, SOA>), grid, block, 0, st, a);
  } else {
    if (dtype == 0) hipLaunchKernelGGL((k_pf_force<1, double, SOA>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((k_pf_force<1, float, SOA>), grid, block, 0, st, a);
  }
}

// X, Y in the reference layout (a.col_bytes == 0) or as resident FP64 fields
hipError_t launch_pf_force(int nc, int dtype, const ForceArgs& a, int ncfg, hipStream_t st) {
  dim3 grid((unsigned)((a.g.V + kForceSites - 1) / kForceSites), (unsigned)ncfg), block(NT);
  if (a.col_bytes) launch_pf_force_t<true>(nc, dtype, a, grid, block, st);
  else launch_pf_force_t<false>(nc, dtype, a, grid, block, st);
  return hipGetLastError();
}

}  // namespace wl
