// Even-odd (checkerboard) stencil: the hopping term between the two parities
// of the lattice and the Schur complement on one parity
// (su3_wilson_nf2.py:934-988 compact blocks, :1057-1079 K_oe/K_eo/K^dag
// blocks, :1112-1156 Schur operators).
//
// The parity of site x is (x0 + x1 + x2 + x3) mod 2.  A compact field of
// parity q holds the sites of that parity in ascending full-lattice order
// (su3_wilson_nf2.py:831-840).  With L3 even each pair x3 = (2j, 2j+1) of a
// row holds one site of each parity, so the compact index of any site is its
// full index / 2, and compact index i of parity q sits at
// x3 = 2 (i mod L3/2) + ((x0 + x1 + x2 + q) mod 2).  Links are stored per
// parity in the same compact order, so both the forward links U_mu(x) of the
// targets and the backward links U_mu(x - mu) of the sources are contiguous.
//
// Hop at target x of parity q from sources of parity 1 - q:
//   K psi(x) = -1/2 sum_mu [(1 - s g_mu) U~_mu(x) psi(x+mu) + (1 + s g_mu) U~_mu(x-mu)^dag psi(x-mu)]
// (s = -1: K, s = +1: K^dag).
//   MODE 0: out = K psi
//   MODE 1: Schur combine on the target parity, psi = K c computed by a MODE 0
//           launch: out = c - K psi / m0^2 (normalized) or m0 c - K psi / m0,
//           i.e. out = dc c + hc acc with acc = sum of the hop terms,
//           (dc, hc) = (1, 1/(2 m0^2)) or (m0, 1/(2 m0)).
#pragma once
#include "wl_dslash_kernels.h"

namespace wl {

// K right-hand sides of the same configuration per thread (columns
// col0 .. col0 + K - 1, K = a.mrhs): each link is loaded once for all K.
template <int P, int NC, int RECON, bool DAG, int MODE, int EPI, int K>
__global__ __launch_bounds__(NT) void k_hop_eo(DslashArgs a) {
  using T = typename PT<P>::acc;
  const int ngrp = a.ncol / K;
  const unsigned logical = xcd_tile(blockIdx.x, gridDim.x, a.xcd_remap);
  const int grp = (int)(logical % ngrp);
  const int tile = (int)(logical / ngrp);
  const int col0 = grp * K;
  const ColState* st = a.red.cols;
  bool on[K];
  bool any = false;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    on[k] = !(a.gate && gate_skip_col(a.gate, a.red, col0 + k));
    any |= on[k];
  }
  if (!any) return;

  const size_t V = (size_t)a.g.Vp;  // plane stride of the compact fields
  // compact target index: flat, or slabs of a.slab consecutive x1 planes walked
  // x0-fastest (each (x0, x1) plane holds L2 L3 / 2 compact sites, contiguous)
  size_t i0;
  bool valid;
  if (a.slab == 0) {
    i0 = (size_t)tile * NT + threadIdx.x;
    valid = i0 < (size_t)a.Vh;
  } else {
    const int sl = tile / a.slab_tiles;
    const int sub = tile - sl * a.slab_tiles;
    const unsigned plane = (unsigned)a.g.stride[1] >> 1;  // L2 * L3 / 2
    const unsigned ii = (unsigned)sub * NT + threadIdx.x;
    valid = ii < (unsigned)a.slab * plane;
    i0 = ((size_t)(sl % a.g.L[0]) * a.g.L[1] + (size_t)(sl / a.g.L[0]) * a.slab) * plane + ii;
  }
  const size_t i = valid ? i0 : 0;

  int x[4];
  const unsigned L3h = (unsigned)a.g.L[3] >> 1;
  const size_t row = i / L3h;
  const unsigned j = (unsigned)(i - row * L3h);
  x[2] = (int)(row % (unsigned)a.g.L[2]);
  const size_t r2 = row / (unsigned)a.g.L[2];
  x[1] = (int)(r2 % (unsigned)a.g.L[1]);
  x[0] = (int)(r2 / (unsigned)a.g.L[1]);
  x[3] = (int)(2 * j) + ((x[0] + x[1] + x[2] + a.pt) & 1);
  const size_t flat = row * (size_t)a.g.L[3] + (size_t)x[3];

  const int cfg = col0 / a.cols_per_cfg;
  const char* Ucfg = reinterpret_cast<const char*>(a.U) + (size_t)cfg * a.cfg_bytes;
  const char* Ut = Ucfg + (size_t)a.pt * 4 * a.dir_bytes;        // links at target-parity sites
  const char* Us = Ucfg + (size_t)(1 - a.pt) * 4 * a.dir_bytes;  // links at source-parity sites

  T dc[K], hc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const T m0 = a.m0_from_state ? T(st[col0 + k].m0) : T(a.m0);
    if (MODE == 0) {
      dc[k] = T(0);
      hc[k] = T(-0.5);
    } else if (a.normalized) {
      dc[k] = T(1);
      hc[k] = T(0.5) / (m0 * m0);
    } else {
      dc[k] = m0;
      hc[k] = T(0.5) / m0;
    }
  }

  cx<T> acc[K][4 * NC];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int q = 0; q < 4 * NC; ++q) acc[k][q] = {T(0), T(0)};
  constexpr int SF = DAG ? +1 : -1;
  constexpr bool FOLD = P == F16;
#define WL_EOHOP(MU)                                                                                      \
  {                                                                                                       \
    const int L = a.g.L[MU];                                                                              \
    const size_t str = (size_t)a.g.stride[MU];                                                            \
    const bool top = x[MU] == L - 1, bot = x[MU] == 0;                                                    \
    const size_t ff = top ? flat - (size_t)(L - 1) * str : flat + str;                                    \
    const size_t fb = bot ? flat + (size_t)(L - 1) * str : flat - str;                                    \
    const cx<T> ph = {T(a.bc_re[MU]), T(a.bc_im[MU])};                                                    \
    {                                                                                                     \
      typename LinkRegs<P, NC>::type u;                                                                   \
      load_link<P, NC, RECON>(Ut + MU * a.dir_bytes, V, i, u);                                            \
      _Pragma("unroll") for (int k = 0; k < K; ++k) {                                                     \
        if (!on[k]) continue;                                                                             \
        hop_one<P, NC, MU, SF, false, FOLD>(u, reinterpret_cast<const char*>(a.in) + (size_t)(col0 + k) * a.col_bytes, \
                                            V, ff >> 1, top, ph, acc[k], (float)hc[k]);                   \
      }                                                                                                   \
    }                                                                                                     \
    {                                                                                                     \
      typename LinkRegs<P, NC>::type u;                                                                   \
      load_link<P, NC, RECON>(Us + MU * a.dir_bytes, V, fb >> 1, u);                                      \
      _Pragma("unroll") for (int k = 0; k < K; ++k) {                                                     \
        if (!on[k]) continue;                                                                             \
        hop_one<P, NC, MU, -SF, true, FOLD>(u, reinterpret_cast<const char*>(a.in) + (size_t)(col0 + k) * a.col_bytes, \
                                            V, fb >> 1, bot, ph, acc[k], (float)hc[k]);                   \
      }                                                                                                   \
    }                                                                                                     \
  }
  WL_EOHOP(0)
  WL_EOHOP(1)
  WL_EOHOP(2)
  WL_EOHOP(3)
#undef WL_EOHOP

#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int col = col0 + k;
    double part = 0.0;
    if (on[k]) {
      cx<T> res[4 * NC];
      if (MODE == 0)
        site_result<P, NC, false, FOLD>(nullptr, V, i, T(0), acc[k], res, hc[k]);
      else
        site_result<P, NC, true, FOLD>(reinterpret_cast<const char*>(a.ctr) + (size_t)col * a.col_bytes, V, i, dc[k],
                                       acc[k], res, hc[k]);
      part = site_epilogue<P, NC, EPI>(a, col, V, i, valid, res);
    }
    if (EPI != EPI_STORE && on[k]) write_partial(a.red, col, tile, a.ntiles, part);
  }
}

}  // namespace wl
