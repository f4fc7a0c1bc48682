// Analytic pseudofermion force of the Nf=2 Wilson action from a solve
// (su3_wilson_nf2.py:1556-1576; U(1): u1_wilson_nf2.py:1004-1025).  With
// X = (D^dag D)^-1 phi and Y = D X, for every link
//   c1[a][c] = sum_{s,t} conj(Y(x)[s][a]) (r - g_mu)[s][t] (U~_mu(x) X(x+mu))[t][c]
//   c2[d][c] = sum_{s,t,a} conj(Y(x+mu)[s][a]) (r + g_mu)[s][t] conj(U~_mu(x)[d][a]) X(x)[t][c]
//   F_mu(x) = TA(k),  k[i][j] = c1[j][i] - c2[j][i],
// TA the traceless anti-Hermitian part (U(1): (k - conj k) / 2), U~ the link
// with the fermion boundary phase.  One thread per site and configuration,
// all four directions: X(x), Y(x) are loaded once; X, Y are read from and F
// written to the reference layouts directly (a site's spinor and a link's
// matrix are contiguous, so a wavefront's accesses are contiguous runs).
// Gamma basis (gamma.py:14-46): (g_mu psi)_s = phase[mu][s] psi_perm[mu][s].
#include "wl_device.h"
#include "wl_kernels.h"

namespace wl {


__constant__ int kGPerm[4][4] = {{3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}, {0, 1, 2, 3}};
// phase = (re, im) in {+-1, +-i}
__constant__ int kGPhRe[4][4] = {{1, 1, 1, 1}, {0, 0, 0, 0}, {0, 0, 0, 0}, {1, 1, -1, -1}};
__constant__ int kGPhIm[4][4] = {{0, 0, 0, 0}, {-1, 1, -1, 1}, {-1, -1, 1, 1}, {0, 0, 0, 0}};

template <class S, int NC>
__device__ __forceinline__ void load_aos(const S* __restrict__ base, size_t idx, cx<double> (&v)[4][NC]) {
  const S* p = base + idx * (4 * NC) * 2;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int c = 0; c < NC; ++c) v[s][c] = {(double)p[2 * (s * NC + c)], (double)p[2 * (s * NC + c) + 1]};
}

// out[s] = r v[s] + sgn g_mu v[s]  (sgn = -1: r - g, +1: r + g)
template <int NC>
__device__ __forceinline__ void spin_mix(int mu, double r, double sgn, const cx<double> (&v)[4][NC],
                                         cx<double> (&out)[4][NC]) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int t = kGPerm[mu][s];
    const double pr = sgn * kGPhRe[mu][s], pi = sgn * kGPhIm[mu][s];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const cx<double> w = v[t][c];
      out[s][c] = {r * v[s][c].re + pr * w.re - pi * w.im, r * v[s][c].im + pr * w.im + pi * w.re};
    }
  }
}

template <int NC, class S>
__global__ __launch_bounds__(NT) void k_pf_force(ForceArgs a) {
  const size_t site = (size_t)blockIdx.x * NT + threadIdx.x;
  const int cfg = blockIdx.y;
  if (site >= (size_t)a.g.V) return;
  const size_t V = (size_t)a.g.V;
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
  const S* Xb = reinterpret_cast<const S*>(a.X);
  const S* Yb = reinterpret_cast<const S*>(a.Y);
  cx<double> X0[4][NC], Y0[4][NC];
  load_aos<S, NC>(Xb, (size_t)cfg * V + site, X0);
  load_aos<S, NC>(Yb, (size_t)cfg * V + site, Y0);
  // Z = (r + g_mu) X(x) is needed per direction; formed inside the loop
  const char* Ucfg = reinterpret_cast<const char*>(a.U) + (size_t)cfg * a.cfg_bytes;
  for (int mu = 0; mu < 4; ++mu) {
    const int L = a.g.L[mu];
    const size_t str = (size_t)a.g.stride[mu];
    const bool top = x[mu] == L - 1;
    const size_t sp = top ? site - (size_t)(L - 1) * str : site + str;
    cx<double> Xp[4][NC], Yp[4][NC];
    load_aos<S, NC>(Xb, (size_t)cfg * V + sp, Xp);
    load_aos<S, NC>(Yb, (size_t)cfg * V + sp, Yp);
    cx<double> u[NC][NC];
    LinkIO<F64, NC, 18>::load(Ucfg + mu * a.dir_bytes, (size_t)a.g.Vp, site, u);
    if (top) {
      const cx<double> ph = {a.bc_re[mu], a.bc_im[mu]};
#pragma unroll
      for (int i = 0; i < NC; ++i)
#pragma unroll
        for (int j = 0; j < NC; ++j) u[i][j] = cmul(ph, u[i][j]);
    }
    // c1 = Y(x)^dag (r - g) (U X(x+mu)), colour indices [a][c]
    cx<double> UX[4][NC], W[4][NC];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < NC; ++i) {
        cx<double> s = {0.0, 0.0};
#pragma unroll
        for (int j = 0; j < NC; ++j) s = cadd(s, cmul(u[i][j], Xp[t][j]));
        UX[t][i] = s;
      }
    spin_mix<NC>(mu, a.r, -1.0, UX, W);
    cx<double> c1[NC][NC], Bm[NC][NC];
#pragma unroll
    for (int ai = 0; ai < NC; ++ai)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        cx<double> s = {0.0, 0.0};
#pragma unroll
        for (int sp4 = 0; sp4 < 4; ++sp4) s = cadd(s, cmulc(Y0[sp4][ai], W[sp4][c]));
        c1[ai][c] = s;
      }
    // B = Y(x+mu)^dag (r + g) X(x), then c2 = U^dag B
    spin_mix<NC>(mu, a.r, +1.0, X0, W);
#pragma unroll
    for (int ai = 0; ai < NC; ++ai)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        cx<double> s = {0.0, 0.0};
#pragma unroll
        for (int sp4 = 0; sp4 < 4; ++sp4) s = cadd(s, cmulc(Yp[sp4][ai], W[sp4][c]));
        Bm[ai][c] = s;
      }
    cx<double> k[NC][NC];  // k[i][j] = c1[j][i] - c2[j][i]
#pragma unroll
    for (int d = 0; d < NC; ++d)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        cx<double> s = {0.0, 0.0};
#pragma unroll
        for (int ai = 0; ai < NC; ++ai) s = cadd(s, cmulc(u[d][ai], Bm[ai][c]));  // conj(U[d][a]) B[a][c]
        k[c][d] = csub(c1[d][c], s);
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
    const size_t li = (a.layout == 0) ? (((size_t)cfg * 4 + mu) * V + site) : (((size_t)cfg * V + site) * 4 + mu);
    S* out = reinterpret_cast<S*>(a.F) + li * NC * NC * 2;
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        out[2 * (i * NC + j)] = (S)f[i][j].re;
        out[2 * (i * NC + j) + 1] = (S)f[i][j].im;
      }
  }
}

hipError_t launch_pf_force(int nc, int dtype, const ForceArgs& a, int ncfg, hipStream_t st) {
  dim3 grid((unsigned)((a.g.V + NT - 1) / NT), (unsigned)ncfg), block(NT);
  if (nc == 3) {
    if (dtype == 0) hipLaunchKernelGGL((k_pf_force<3, double>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((k_pf_force<3, float>), grid, block, 0, st, a);
  } else {
    if (dtype == 0) hipLaunchKernelGGL((k_pf_force<1, double>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((k_pf_force<1, float>), grid, block, 0, st, a);
  }
  return hipGetLastError();
}

}  // namespace wl
