// The Wilson hopping stencil on one site, with spin projection to half
// spinors before the colour multiply.
//
// Operator (wilson.py:228-256 and :258-284, sign = -1 for D, +1 for D^dag):
//   D psi(x) = m0 psi(x) - 1/2 sum_mu [ (1 - s g_mu) U~_mu(x) psi(x+mu)
//                                      + (1 + s g_mu) U~_mu(x-mu)^dag psi(x-mu) ]
// with s = +1 for D and s = -1 for D^dag (so D^dag swaps the projectors).
//
// Gamma basis (gamma.py:14-46, 4D): g0 = s1 x s1, g1 = s1 x s2, g2 = s2 x 1,
// g3 = s3 x 1.  Each (1 + S g_mu) has rank 2; the two independent rows h0, h1
// are formed from the spinor, multiplied by the link, and the four rows are
// rebuilt from them:
//   mu=0: h0 = p0 + S p3,   h1 = p1 + S p2;    rows (h0, h1, S h1, S h0)
//   mu=1: h0 = p0 - iS p3,  h1 = p1 + iS p2;   rows (h0, h1, -iS h1, iS h0)
//   mu=2: h0 = p0 - iS p2,  h1 = p1 - iS p3;   rows (h0, h1, iS h0, iS h1)
//   mu=3: S=+1: h = 2(p0, p1), rows (h0, h1, 0, 0);
//         S=-1: h = 2(p2, p3), rows (0, 0, h0, h1)
#pragma once
#include "wl_device.h"

namespace wl {

template <int MU, int S, class T, int NC>
__device__ __forceinline__ void spin_project(const cx<T> (&p)[4 * NC], cx<T> (&h)[2][NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const cx<T> p0 = p[0 * NC + c], p1 = p[1 * NC + c], p2 = p[2 * NC + c], p3 = p[3 * NC + c];
    if (MU == 0) {
      h[0][c] = (S > 0) ? cadd(p0, p3) : csub(p0, p3);
      h[1][c] = (S > 0) ? cadd(p1, p2) : csub(p1, p2);
    } else if (MU == 1) {
      // p0 - iS p3 ; p1 + iS p2
      h[0][c] = (S > 0) ? csub(p0, ci(p3)) : cadd(p0, ci(p3));
      h[1][c] = (S > 0) ? cadd(p1, ci(p2)) : csub(p1, ci(p2));
    } else if (MU == 2) {
      h[0][c] = (S > 0) ? csub(p0, ci(p2)) : cadd(p0, ci(p2));
      h[1][c] = (S > 0) ? csub(p1, ci(p3)) : cadd(p1, ci(p3));
    } else {
      if (S > 0) {
        h[0][c] = cadd(p0, p0);
        h[1][c] = cadd(p1, p1);
      } else {
        h[0][c] = cadd(p2, p2);
        h[1][c] = cadd(p3, p3);
      }
    }
  }
}

template <int MU, int S, class T, int NC>
__device__ __forceinline__ void spin_reconstruct_add(cx<T> (&acc)[4 * NC], const cx<T> (&h)[2][NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const cx<T> h0 = h[0][c], h1 = h[1][c];
    if (MU == 0) {
      acc[0 * NC + c] = cadd(acc[0 * NC + c], h0);
      acc[1 * NC + c] = cadd(acc[1 * NC + c], h1);
      acc[2 * NC + c] = (S > 0) ? cadd(acc[2 * NC + c], h1) : csub(acc[2 * NC + c], h1);
      acc[3 * NC + c] = (S > 0) ? cadd(acc[3 * NC + c], h0) : csub(acc[3 * NC + c], h0);
    } else if (MU == 1) {
      acc[0 * NC + c] = cadd(acc[0 * NC + c], h0);
      acc[1 * NC + c] = cadd(acc[1 * NC + c], h1);
      // -iS h1 ; iS h0
      acc[2 * NC + c] = (S > 0) ? csub(acc[2 * NC + c], ci(h1)) : cadd(acc[2 * NC + c], ci(h1));
      acc[3 * NC + c] = (S > 0) ? cadd(acc[3 * NC + c], ci(h0)) : csub(acc[3 * NC + c], ci(h0));
    } else if (MU == 2) {
      acc[0 * NC + c] = cadd(acc[0 * NC + c], h0);
      acc[1 * NC + c] = cadd(acc[1 * NC + c], h1);
      acc[2 * NC + c] = (S > 0) ? cadd(acc[2 * NC + c], ci(h0)) : csub(acc[2 * NC + c], ci(h0));
      acc[3 * NC + c] = (S > 0) ? cadd(acc[3 * NC + c], ci(h1)) : csub(acc[3 * NC + c], ci(h1));
    } else {
      if (S > 0) {
        acc[0 * NC + c] = cadd(acc[0 * NC + c], h0);
        acc[1 * NC + c] = cadd(acc[1 * NC + c], h1);
      } else {
        acc[2 * NC + c] = cadd(acc[2 * NC + c], h0);
        acc[3 * NC + c] = cadd(acc[3 * NC + c], h1);
      }
    }
  }
}

// uh[a] = U h[a]  (forward) or U^dag h[a]  (backward)
template <bool ADJ, class T, int NC>
__device__ __forceinline__ void color_mul(const cx<T> (&u)[NC][NC], const cx<T> (&h)[2][NC], cx<T> (&o)[2][NC]) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      cx<T> s = ADJ ? cmulc(u[0][i], h[a][0]) : cmul(u[i][0], h[a][0]);
#pragma unroll
      for (int j = 1; j < NC; ++j) s = cadd(s, ADJ ? cmulc(u[j][i], h[a][j]) : cmul(u[i][j], h[a][j]));
      o[a][i] = s;
    }
}

template <class T, int NC>
__device__ __forceinline__ void half_phase(cx<T> (&o)[2][NC], cx<T> ph) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int i = 0; i < NC; ++i) o[a][i] = cmul(ph, o[a][i]);
}

// One direction mu: forward hop from x+mu and backward hop from x-mu.
// SF is the sign S of the forward projector (1 + SF g_mu): SF = -1 for D,
// +1 for D^dag; the backward projector uses -SF.
template <int P, int NC, int RECON, int MU, int SF>
__device__ __forceinline__ void hop_dir(const char* __restrict__ Ucfg, size_t dir_bytes, const void* __restrict__ in,
                                        size_t V, size_t site, size_t s_fwd, size_t s_bwd, bool at_top,
                                        bool at_bottom, cx<typename PT<P>::acc> ph,
                                        cx<typename PT<P>::acc> (&acc)[4 * NC]) {
  using T = typename PT<P>::acc;
  const char* Umu = Ucfg + MU * dir_bytes;
  {
    cx<T> p[4 * NC];
    SpinorIO<P, NC>::load(in, V, s_fwd, p);
    cx<T> u[NC][NC];
    LinkIO<P, NC, RECON>::load(Umu, V, site, u);
    cx<T> h[2][NC], o[2][NC];
    spin_project<MU, SF>(p, h);
    color_mul<false>(u, h, o);
    if (at_top) half_phase(o, ph);
    spin_reconstruct_add<MU, SF>(acc, o);
  }
  {
    cx<T> p[4 * NC];
    SpinorIO<P, NC>::load(in, V, s_bwd, p);
    cx<T> u[NC][NC];
    LinkIO<P, NC, RECON>::load(Umu, V, s_bwd, u);
    cx<T> h[2][NC], o[2][NC];
    spin_project<MU, -SF>(p, h);
    color_mul<true>(u, h, o);
    if (at_bottom) half_phase(o, cconj(ph));
    spin_reconstruct_add<MU, -SF>(acc, o);
  }
}

}  // namespace wl
