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

// ---------------------------------------------------------------------------
// F16 path.  Spinors and links stay in their stored half2 form: a projected
// half spinor is formed with packed half adds (the sign and i-rotation of the
// projector are operand modifiers), the colour multiply is v_dot2_f32_f16
// (half products, FP32 accumulation; two per complex multiply-add), and the
// real factor f = (per-site scale) x (real boundary sign on the boundary hop)
// x (-1/2 when the caller folds the hopping coefficient) is applied in the
// FP32 reconstruction as one FMA per value.  Values entering half arithmetic
// are the stored ones (|.| <= 1, the site scale is kept apart) and sums of
// two of them, so nothing overflows.
typedef _Float16 hf2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ hf2 as_hf2(uint32_t u) { return __builtin_bit_cast(hf2, u); }
__device__ __forceinline__ uint32_t as_u32(hf2 h) { return __builtin_bit_cast(uint32_t, h); }
__device__ __forceinline__ float dot2(hf2 a, hf2 b, float c) { return __builtin_amdgcn_fdot2(a, b, c, false); }
// a.x b.x + a.y b.y with a zero accumulator as an inline constant (the
// builtin with c = 0 costs a register move per chain)
__device__ __forceinline__ float dot2z(hf2 a, hf2 b) {
  float r;
  asm("v_dot2_f32_f16 %0, %1, %2, 0" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

template <int NC>
struct HalfSpinor {
  uint32_t w[4 * NC];  // complex (spin s, colour c) at w[s * NC + c] as a half2
  float s;             // per-site scale
};

template <int NC>
__device__ __forceinline__ void load_half_spinor(const void* __restrict__ base, size_t V, size_t site,
                                                 HalfSpinor<NC>& r) {
  const uint4* p = reinterpret_cast<const uint4*>(base);
  r.s = reinterpret_cast<const float*>(p + NC * V)[site];
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    const uint4 t = p[j * V + site];
    r.w[4 * j + 0] = t.x;
    r.w[4 * j + 1] = t.y;
    r.w[4 * j + 2] = t.z;
    r.w[4 * j + 3] = t.w;
  }
}

// Links as stored half2 (row a, column b); 12-real storage rebuilds row 2 =
// conj(row0 x row1) with FP32 dot products, rounded once to half.  An SU(3)
// link is stored as LinkWords<RECON>::n half2 words: rows 0 and 1 (12-real,
// one uint4 and one uint2 plane) or all nine entries (one word plane each).
template <int RECON>
struct LinkWords {
  static constexpr int n = RECON == 12 ? 6 : 9;
};

template <int RECON>
__device__ __forceinline__ void load_link_words(const void* __restrict__ base, size_t V, size_t site,
                                                uint32_t (&w)[LinkWords<RECON>::n]) {
  if constexpr (RECON == 12) {
    const uint4* pa = reinterpret_cast<const uint4*>(base);
    const uint2* pb = reinterpret_cast<const uint2*>(pa + V);
    const uint4 t = pa[site];
    const uint2 q = pb[site];
    w[0] = t.x;
    w[1] = t.y;
    w[2] = t.z;
    w[3] = t.w;
    w[4] = q.x;
    w[5] = q.y;
  } else {
    const uint32_t* p = reinterpret_cast<const uint32_t*>(base);
#pragma unroll
    for (int c = 0; c < 9; ++c) w[c] = p[c * V + site];
  }
}

template <int RECON>
__device__ __forceinline__ void half_link_from_words(const uint32_t (&w)[LinkWords<RECON>::n], hf2 (&u)[3][3]) {
  if constexpr (RECON == 12) {
#pragma unroll
    for (int c = 0; c < 6; ++c) u[c / 3][c % 3] = as_hf2(w[c]);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int k = (j + 1) % 3, l = (j + 2) % 3;
      const hf2 ak = u[0][k], al = u[0][l], bk = u[1][k], bl = u[1][l];
      // c = ak bl - al bk ; row2_j = conj(c)
      float re = dot2(ak, hf2{bl.x, -bl.y}, 0.0f);
      re = dot2(al, hf2{-bk.x, bk.y}, re);
      float im = dot2(ak, hf2{bl.y, bl.x}, 0.0f);
      im = dot2(al, hf2{-bk.y, -bk.x}, im);
      u[2][j] = hf2{(_Float16)re, (_Float16)(-im)};
    }
  } else {
#pragma unroll
    for (int c = 0; c < 9; ++c) u[c / 3][c % 3] = as_hf2(w[c]);
  }
}

template <int NC, int RECON>
__device__ __forceinline__ void load_half_link(const void* __restrict__ base, size_t V, size_t site,
                                               hf2 (&u)[NC][NC]) {
  if constexpr (NC == 1) {
    u[0][0] = as_hf2(reinterpret_cast<const uint32_t*>(base)[site]);
  } else {
    uint32_t w[LinkWords<RECON>::n];
    load_link_words<RECON>(base, V, site, w);
    half_link_from_words<RECON>(w, u);
  }
}

// The two half2 forms of each projected value (re, im) = f-free projection of
// the stored spinor for (1 + S g_mu), rows as in spin_project:
//   forward  (U h):       hA = (re, -im), hB = (im, re)
//   backward (U^dag h):   hA = (re,  im), hB = (im, -re)
// so that Re = sum dot2(u, hA) and Im = sum dot2(u, hB).
template <int MU, int S, bool ADJ, int NC>
__device__ __forceinline__ void project_forms(const HalfSpinor<NC>& r, hf2 (&hA)[2][NC], hf2 (&hB)[2][NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      hf2 pa, q;  // (re, im) = pa + q
      if (MU == 3) {
        pa = as_hf2(r.w[((S > 0) ? k : 2 + k) * NC + c]);
        q = pa;
      } else {
        const int b = (MU == 2) ? 2 + k : 3 - k;
        pa = as_hf2(r.w[k * NC + c]);
        const hf2 pb = as_hf2(r.w[b * NC + c]);
        const _Float16 sg = (S > 0) ? (_Float16)1 : (_Float16)-1;
        if (MU == 0) q = hf2{sg * pb.x, sg * pb.y};
        else if (MU == 1 && k == 1) q = hf2{-sg * pb.y, sg * pb.x};
        else q = hf2{sg * pb.y, -sg * pb.x};  // MU == 1, k == 0 and MU == 2
      }
      if (!ADJ) {
        hA[k][c] = hf2{pa.x, -pa.y} + hf2{q.x, -q.y};
        hB[k][c] = hf2{pa.y, pa.x} + hf2{q.y, q.x};
      } else {
        hA[k][c] = pa + q;
        hB[k][c] = hf2{pa.y, -pa.x} + hf2{q.y, -q.x};
      }
    }
}

// o[a][i] = sum_j U_ij h_j (forward) or sum_j conj(U_ji) h_j (backward), unscaled
template <bool ADJ, int NC>
__device__ __forceinline__ void color_mul_half(const hf2 (&u)[NC][NC], const hf2 (&hA)[2][NC], const hf2 (&hB)[2][NC],
                                               cx<float> (&o)[2][NC]) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      float re = dot2z(ADJ ? u[0][i] : u[i][0], hA[a][0]);
      float im = dot2z(ADJ ? u[0][i] : u[i][0], hB[a][0]);
#pragma unroll
      for (int j = 1; j < NC; ++j) {
        const hf2 uij = ADJ ? u[j][i] : u[i][j];
        re = dot2(uij, hA[a][j], re);
        im = dot2(uij, hB[a][j], im);
      }
      o[a][i] = {re, im};
    }
}

// acc += f * (rows rebuilt from o), one FMA per accumulated value
template <int MU, int S, int NC>
__device__ __forceinline__ void spin_reconstruct_fma(cx<float> (&acc)[4 * NC], const cx<float> (&o)[2][NC], float f) {
  const float g = (S > 0) ? f : -f;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const cx<float> h0 = o[0][c], h1 = o[1][c];
    auto fma2 = [&](cx<float>& x, float re, float im, float k) {
      x.re = fmaf(re, k, x.re);
      x.im = fmaf(im, k, x.im);
    };
    if (MU == 3) {
      if (S > 0) {
        fma2(acc[0 * NC + c], h0.re, h0.im, f);
        fma2(acc[1 * NC + c], h1.re, h1.im, f);
      } else {
        fma2(acc[2 * NC + c], h0.re, h0.im, f);
        fma2(acc[3 * NC + c], h1.re, h1.im, f);
      }
      continue;
    }
    fma2(acc[0 * NC + c], h0.re, h0.im, f);
    fma2(acc[1 * NC + c], h1.re, h1.im, f);
    if (MU == 0) {  // rows (h0, h1, S h1, S h0)
      fma2(acc[2 * NC + c], h1.re, h1.im, g);
      fma2(acc[3 * NC + c], h0.re, h0.im, g);
    } else if (MU == 1) {  // rows (h0, h1, -iS h1, iS h0); i (x + iy) = -y + ix
      fma2(acc[2 * NC + c], h1.im, -h1.re, g);
      fma2(acc[3 * NC + c], -h0.im, h0.re, g);
    } else {  // MU == 2: rows (h0, h1, iS h0, iS h1)
      fma2(acc[2 * NC + c], -h0.im, h0.re, g);
      fma2(acc[3 * NC + c], -h1.im, h1.re, g);
    }
  }
}

// One neighbour of one direction for one spinor column, given its link u:
// forward (ADJ = false, projector sign S, phase ph on the boundary hop) or
// backward (ADJ = true, conj(ph)).  FOLD: accumulate fk times the term (F16
// path only; other precisions always accumulate the plain term).
// Link registers of one direction: FP32/FP64 complex, or stored half2 (F16).
template <int P, int NC>
struct LinkRegs {
  using type = cx<typename PT<P>::acc>[NC][NC];
};
template <int NC>
struct LinkRegs<F16, NC> {
  using type = hf2[NC][NC];
};

template <int P, int NC, int RECON>
__device__ __forceinline__ void load_link(const void* __restrict__ base, size_t V, size_t site,
                                          typename LinkRegs<P, NC>::type& u) {
  if constexpr (P == F16) load_half_link<NC, RECON>(base, V, site, u);
  else LinkIO<P, NC, RECON>::load(base, V, site, u);
}

// The F16 hop on a neighbour spinor already in registers (from memory or LDS).
template <int NC, int MU, int S, bool ADJ, bool FOLD>
__device__ __forceinline__ void hop_half(const hf2 (&u)[NC][NC], const HalfSpinor<NC>& r, bool at_bnd, cx<float> ph,
                                         cx<float> (&acc)[4 * NC], float fk = -0.5f) {
  cx<float> o[2][NC];
  const bool realbc = ph.im == 0.0f;
  float f = (FOLD ? fk : 1.0f) * r.s;
  if (realbc && at_bnd) f *= ph.re;
  hf2 hA[2][NC], hB[2][NC];
  project_forms<MU, S, ADJ>(r, hA, hB);
  color_mul_half<ADJ>(u, hA, hB, o);
  if (!realbc && at_bnd) half_phase(o, ADJ ? cconj(ph) : ph);  // complex boundary phase, in FP32
  spin_reconstruct_fma<MU, S>(acc, o, f);
}

template <int P, int NC, int MU, int S, bool ADJ, bool FOLD>
__device__ __forceinline__ void hop_one(const typename LinkRegs<P, NC>::type& u, const void* __restrict__ in,
                                        size_t V, size_t s_nb, bool at_bnd, cx<typename PT<P>::acc> ph,
                                        cx<typename PT<P>::acc> (&acc)[4 * NC], float fk = -0.5f) {
  using T = typename PT<P>::acc;
  cx<T> h[2][NC], o[2][NC];
  const cx<T> phe = ADJ ? cconj(ph) : ph;
  if constexpr (P == F16) {
    HalfSpinor<NC> r;
    load_half_spinor<NC>(in, V, s_nb, r);
    hop_half<NC, MU, S, ADJ, FOLD>(u, r, at_bnd, ph, acc, fk);
    return;
  } else {
    cx<T> p[4 * NC];
    SpinorIO<P, NC>::load(in, V, s_nb, p);
    spin_project<MU, S>(p, h);
    color_mul<ADJ>(u, h, o);
    if (at_bnd) half_phase(o, phe);
  }
  spin_reconstruct_add<MU, S>(acc, o);
}

// One direction mu: forward hop from x+mu and backward hop from x-mu.
// SF is the sign S of the forward projector (1 + SF g_mu): SF = -1 for D,
// +1 for D^dag; the backward projector uses -SF.
template <int P, int NC, int RECON, int MU, int SF, bool FOLD = false>
__device__ __forceinline__ void hop_dir(const char* __restrict__ Ucfg, size_t dir_bytes, const void* __restrict__ in,
                                        size_t V, size_t site, size_t s_fwd, size_t s_bwd, bool at_top,
                                        bool at_bottom, cx<typename PT<P>::acc> ph,
                                        cx<typename PT<P>::acc> (&acc)[4 * NC]) {
  using T = typename PT<P>::acc;
  const char* Umu = Ucfg + MU * dir_bytes;
  {
    typename LinkRegs<P, NC>::type u;
    load_link<P, NC, RECON>(Umu, V, site, u);
    hop_one<P, NC, MU, SF, false, FOLD>(u, in, V, s_fwd, at_top, ph, acc);
  }
  {
    typename LinkRegs<P, NC>::type u;
    load_link<P, NC, RECON>(Umu, V, s_bwd, u);
    hop_one<P, NC, MU, -SF, true, FOLD>(u, in, V, s_bwd, at_bottom, ph, acc);
  }
}

// Diagonal and hopping terms at the site: res = m0 psi(x) + hc acc, where
// the hops were accumulated plain (or, with FOLD, already times hc, so that
// res = m0 psi(x) + acc); DIAG = false gives the hopping term alone.
// hc = -1/2 is the Wilson operator's hopping coefficient.
template <int P, int NC, bool DIAG, bool FOLD>
__device__ __forceinline__ void site_result(const void* __restrict__ in, size_t V, size_t site,
                                            typename PT<P>::acc m0, const cx<typename PT<P>::acc> (&acc)[4 * NC],
                                            cx<typename PT<P>::acc> (&res)[4 * NC],
                                            typename PT<P>::acc hc = typename PT<P>::acc(-0.5)) {
  using T = typename PT<P>::acc;
  const T c = FOLD ? T(1) : hc;
  if (!DIAG) {
#pragma unroll
    for (int q = 0; q < 4 * NC; ++q) res[q] = FOLD ? acc[q] : cx<T>{c * acc[q].re, c * acc[q].im};
    return;
  }
  if constexpr (P == F16) {
    HalfSpinor<NC> r;
    load_half_spinor<NC>(in, V, site, r);
    const float ms = m0 * r.s;
#pragma unroll
    for (int q = 0; q < 4 * NC; ++q) {
      const float2 p = h2f(r.w[q]);
      res[q] = FOLD ? cx<float>{fmaf(p.x, ms, acc[q].re), fmaf(p.y, ms, acc[q].im)}
                    : cx<float>{fmaf(p.x, ms, c * acc[q].re), fmaf(p.y, ms, c * acc[q].im)};
    }
  } else {
    cx<T> p[4 * NC];
    SpinorIO<P, NC>::load(in, V, site, p);
#pragma unroll
    for (int q = 0; q < 4 * NC; ++q) res[q] = {m0 * p[q].re + c * acc[q].re, m0 * p[q].im + c * acc[q].im};
  }
}

}  // namespace wl
