// Wilson operator kernels: D, D^dag, gamma5 D, the hopping term K and K^dag,
// with epilogues that fuse the CG vector work into the stencil (Epilogue in
// wl_kernels.h).  One CG iteration on A = D^dag D is two stencil launches:
//   t = D p         with |t|^2 = <p, A p>          -> alpha = rr / |t|^2
//   A p = D^dag t   with y += alpha p, r -= alpha A p, |r|^2 -> beta
// so A p is never written to memory.  With the search direction by
// recurrence (search_rec) the first launch takes r and forms
// t = D r + beta t_prev = D (r + beta p_prev), and the second forms
// p = r + beta p_prev beside its update of r, so p needs no pass of its own.
//
// The 1D grid is ntiles x ncol workgroups, ordered spinor-column-fastest (the
// columns of one tile run back to back and share its link lines in L2) or
// column-slowest (one column sweeps the lattice before the next, so its
// neighbour spinors stay in L2), and remapped so that each XCD walks one
// contiguous range of that order.
#pragma once
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

template <int P, int NC>
__device__ __forceinline__ void store_nt(void* base, size_t V, size_t site, const cx<typename PT<P>::acc> (&v)[4 * NC]) {
  SpinorIO<P, NC>::store(base, V, site, v);
}

// Site of this thread in `tile` for the one-site-per-thread kernels.  Flat
// order: tile t covers sites [t NT, (t+1) NT).  Slab order: a slab is the
// sites with one x0, `slab` consecutive x1 values and `slab2` consecutive x2
// values (slab * slab2 * L3 sites in x1, x2, x3 order, ceil(./NT) tiles);
// slabs are numbered x0-fastest, then by x2 block, then by x1 block, so a
// tile's x0 neighbours are read by the tiles of the adjacent slabs moments
// before or after it, and each XCD (contiguous tile range) sweeps x0 over its
// own (x1, x2) block.  slab2 = L2 gives slabs of whole x1 planes.  `valid` is
// false for the padding threads of a slab's last tile.
//
// Box order: tile t is the 4D box (t0, t1, t2, t3) of extents 2^bsh (t3
// fastest); thread (l0, l1, l2, l3) of it, l3 fastest, so a wavefront reads
// whole 128-byte lines (x3 extent >= 8) and about half of the neighbours of
// a box lie inside it.  site_of_tile also returns the site's coordinates.
__device__ __forceinline__ size_t site_of_tile_linear(const DslashArgs& a, int tile, bool& valid) {
  if (a.slab == 0) {
    const size_t site = (size_t)tile * NT + threadIdx.x;
    valid = site < (size_t)a.g.V;
    return valid ? site : 0;
  }
  const int sl = tile / a.slab_tiles;
  const int sub = tile - sl * a.slab_tiles;
  const int x0 = sl % a.g.L[0];
  const int r = sl / a.g.L[0];
  const int nb2 = a.g.L[2] / a.slab2;
  const int b1 = r / nb2;
  const int b2 = r - b1 * nb2;
  const unsigned L3 = (unsigned)a.g.L[3];
  const unsigned row = (unsigned)a.slab2 * L3;  // sites of one x1 value of the slab
  const unsigned i = (unsigned)sub * NT + threadIdx.x;
  valid = i < (unsigned)a.slab * row;
  const unsigned ii = valid ? i : 0u;
  const unsigned q = ii / row;
  const unsigned rem = ii - q * row;
  const unsigned l2 = rem / L3;
  return (size_t)x0 * (size_t)a.g.stride[0] + (size_t)(b1 * a.slab + (int)q) * (size_t)a.g.stride[1] +
         (size_t)((unsigned)(b2 * a.slab2) + l2) * L3 + (rem - l2 * L3);
}

__device__ __forceinline__ size_t site_of_tile(const DslashArgs& a, int tile, bool& valid, int (&x)[4]) {
  if (a.box) {
    int t = tile;
    const int t3 = t % a.bn[3];
    t /= a.bn[3];
    const int t2 = t % a.bn[2];
    t /= a.bn[2];
    const int t1 = t % a.bn[1];
    const int t0 = t / a.bn[1];
    unsigned l = threadIdx.x;
    const unsigned l3 = l & ((1u << a.bsh[3]) - 1u);
    l >>= a.bsh[3];
    const unsigned l2 = l & ((1u << a.bsh[2]) - 1u);
    l >>= a.bsh[2];
    const unsigned l1 = l & ((1u << a.bsh[1]) - 1u);
    const unsigned l0 = l >> a.bsh[1];
    const unsigned x0 = ((unsigned)t0 << a.bsh[0]) + l0, x1 = ((unsigned)t1 << a.bsh[1]) + l1;
    const unsigned x2 = ((unsigned)t2 << a.bsh[2]) + l2, x3 = ((unsigned)t3 << a.bsh[3]) + l3;
    valid = true;
    x[0] = (int)x0;
    x[1] = (int)x1;
    x[2] = (int)x2;
    x[3] = (int)x3;
    return (((size_t)x0 * a.g.L[1] + x1) * a.g.L[2] + x2) * a.g.L[3] + x3;
  }
  const size_t s = site_of_tile_linear(a, tile, valid);
  size_t rem = s;
  x[3] = rem % a.g.L[3];
  rem /= a.g.L[3];
  x[2] = rem % a.g.L[2];
  rem /= a.g.L[2];
  x[1] = rem % a.g.L[1];
  x[0] = rem / a.g.L[1];
  return s;
}


// A spinor held as stored: FP16 keeps its half2 words and the site's scale
// (13 registers instead of 24 expanded values), so that a spinor read once can
// be used twice without being held expanded; other precisions keep the values.
// get(q) returns component q in the arithmetic precision.  Streaming loads.
template <int P, int NC>
struct Stored {
  using T = typename PT<P>::acc;
  cx<T> v[4 * NC];
  __device__ __forceinline__ void load(const void* base, size_t V, size_t site) {
    SpinorIO<P, NC>::template load<true>(base, V, site, v);
  }
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int q = 0; q < 4 * NC; ++q) v[q] = {T(0), T(0)};
  }
  __device__ __forceinline__ cx<T> get(int q) const { return v[q]; }
};

template <int NC>
struct Stored<F16, NC> {
  uint32_t w[4 * NC];
  float s;
  __device__ __forceinline__ void load(const void* base, size_t V, size_t site) {
    const uint4* p = reinterpret_cast<const uint4*>(base);
    s = gld<true>(reinterpret_cast<const float*>(p + NC * V) + site);
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      const uint4 t = gld<true>(p + j * V + site);
      w[4 * j] = t.x;
      w[4 * j + 1] = t.y;
      w[4 * j + 2] = t.z;
      w[4 * j + 3] = t.w;
    }
  }
  __device__ __forceinline__ void zero() {
    s = 0.f;
#pragma unroll
    for (int q = 0; q < 4 * NC; ++q) w[q] = 0u;
  }
  __device__ __forceinline__ cx<float> get(int q) const {
    const float2 h = h2f(w[q]);
    return {h.x * s, h.y * s};
  }
};

template <int EPI>
struct EpiAccum {
  static constexpr int prec = EPI == EPI_CG64 ? F64 : (EPI == EPI_CG32 ? F32 : F16);
};

// Epilogue at site s of column col: res holds the operator result (modified
// in place for EPI_RESID).  The epilogue's own vectors (b, r, p, y) are read
// and written once per launch at the site itself: streaming accesses, which
// keep them from displacing the neighbour spinors in the caches.  Stores what the epilogue stores (only when
// `valid`) and returns the site's contribution to the column's reduction
// partial (0 for an invalid thread).
template <int P, int NC, int EPI>
__device__ __forceinline__ double site_epilogue(const DslashArgs& a, int col, size_t V, size_t s, bool valid,
                                                cx<typename PT<P>::acc> (&res)[4 * NC]) {
  using T = typename PT<P>::acc;
  auto outb = [&]() { return reinterpret_cast<char*>(a.out) + (size_t)col * a.col_bytes; };
  double part = 0.0;
  if (EPI == EPI_STORE || EPI == EPI_NORM) {
    if (EPI == EPI_NORM) {
      if (a.search_rec) {
        // t = D r + beta t_prev = D (r + beta p_prev), t_prev at the site
        // itself; beta = 0 starts a new direction, and t_prev is not read
        // (it may hold anything then)
        const T beta = (T)a.red.cols[col].beta;
        if (beta != T(0)) {
          cx<T> tp[4 * NC];
          SpinorIO<P, NC>::template load<true>(outb(), V, s, tp);
#pragma unroll
          for (int q = 0; q < 4 * NC; ++q) {
            res[q].re += beta * tp[q].re;
            res[q].im += beta * tp[q].im;
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 4 * NC; ++q) part += (double)res[q].re * (double)res[q].re + (double)res[q].im * (double)res[q].im;
    }
    if (valid) store_nt<P, NC>(outb(), V, s, res);
  } else if (EPI == EPI_RESID || EPI == EPI_RESID32 || EPI == EPI_RESID16) {
    const char* wb = reinterpret_cast<const char*>(a.w) + (size_t)col * a.col_bytes;
    cx<T> w[4 * NC];
    SpinorIO<P, NC>::template load<true>(wb, V, s, w);
#pragma unroll
    for (int q = 0; q < 4 * NC; ++q) {
      res[q] = csub(w[q], res[q]);
      part += (double)res[q].re * (double)res[q].re + (double)res[q].im * (double)res[q].im;
    }
    if (EPI == EPI_RESID) {
      if (valid) store_nt<P, NC>(outb(), V, s, res);
    } else {
      // the inner solver's residual, rounded as a separate conversion would
      constexpr int PR = EPI == EPI_RESID32 ? F32 : F16;
      cx<float> rr[4 * NC];
#pragma unroll
      for (int q = 0; q < 4 * NC; ++q) rr[q] = {(float)res[q].re, (float)res[q].im};
      if (valid) SpinorIO<PR, NC>::template store<true>(reinterpret_cast<char*>(a.res) + (size_t)col * a.col_bytes_y, V, s, rr);
    }
  } else {
    constexpr int PY = EpiAccum<EPI>::prec;
    using TY = typename PT<PY>::acc;
    const ColState& cs = a.red.cols[col];
    const double alpha = cs.alpha;
    char* rb = reinterpret_cast<char*>(a.res) + (size_t)col * a.col_bytes;
    const T ar = (T)alpha;
    if (a.search_rec) {
      // The site's vectors r, p_prev and y are read together and held as
      // stored (FP16: 13 registers each, so that the three reads are in
      // flight at once), then: r -= alpha A p, stored; p = r + beta p_prev
      // with the old r, stored in place of p_prev (out); y += alpha p every
      // iteration, or (pairwise) nothing in the first iteration of a pair and
      // alpha_prev p_prev + alpha p in the second.  p_prev is read when beta
      // is not 0 or when it still owes its alpha_prev to y.
      const double beta = cs.beta;
      const bool second = cs.pairwise && cs.pend;
      const bool prev = beta != 0.0 || second;
      const bool yupd = !(cs.pairwise && !cs.pend);
      char* yb = reinterpret_cast<char*>(a.y) + (size_t)col * a.col_bytes_y;
      Stored<P, NC> r, pp;
      r.load(rb, V, s);
      if (prev) {
        pp.load(outb(), V, s);
      } else {
        pp.zero();
      }
      Stored<PY, NC> ys;
      if (yupd) ys.load(yb, V, s);
      {
        cx<T> rn[4 * NC];
#pragma unroll
        for (int q = 0; q < 4 * NC; ++q) {
          const cx<T> h = r.get(q);
          rn[q].re = h.re - ar * res[q].re;
          rn[q].im = h.im - ar * res[q].im;
          part += (double)rn[q].re * (double)rn[q].re + (double)rn[q].im * (double)rn[q].im;
        }
        if (valid) SpinorIO<P, NC>::template store<true>(rb, V, s, rn);
      }
      // p = r + beta p_prev
      const T bt = (T)beta;
      cx<T> p[4 * NC];
#pragma unroll
      for (int q = 0; q < 4 * NC; ++q) {
        const cx<T> h = r.get(q), g = pp.get(q);
        p[q] = {h.re + bt * g.re, h.im + bt * g.im};
      }
      if (valid) SpinorIO<P, NC>::template store<true>(outb(), V, s, p);
      if (yupd) {
        const TY ay = (TY)alpha;
        const TY aq = (TY)(second ? cs.alpha_prev : 0.0);
        cx<TY> w[4 * NC];  // takes p's registers when y has the inner precision's arithmetic
#pragma unroll
        for (int q = 0; q < 4 * NC; ++q) {
          const cx<T> g = pp.get(q);
          const cx<TY> yv = ys.get(q);
          w[q] = {yv.re + (ay * (TY)p[q].re + aq * (TY)g.re), yv.im + (ay * (TY)p[q].im + aq * (TY)g.im)};
        }
        if (valid) SpinorIO<PY, NC>::template store<true>(yb, V, s, w);
      }
      return valid ? part : 0.0;
    }
    cx<T> r[4 * NC];
    SpinorIO<P, NC>::template load<true>(rb, V, s, r);
#pragma unroll
    for (int q = 0; q < 4 * NC; ++q) {
      r[q].re -= ar * res[q].re;
      r[q].im -= ar * res[q].im;
      part += (double)r[q].re * (double)r[q].re + (double)r[q].im * (double)r[q].im;
    }
    if (valid) SpinorIO<P, NC>::template store<true>(rb, V, s, r);
    // y += alpha p every iteration, or (pairwise) nothing in the first
    // iteration of a pair and alpha_prev p_prev + alpha p in the second
    if (!(cs.pairwise && !cs.pend)) {
      const char* pb = reinterpret_cast<const char*>(a.w) + (size_t)col * a.col_bytes;
      char* yb = reinterpret_cast<char*>(a.y) + (size_t)col * a.col_bytes_y;
      cx<T> p[4 * NC];
      cx<TY> y[4 * NC];
      SpinorIO<P, NC>::template load<true>(pb, V, s, p);
      SpinorIO<PY, NC>::template load<true>(yb, V, s, y);
      const TY ay = (TY)alpha;
#pragma unroll
      for (int q = 0; q < 4 * NC; ++q) {
        y[q].re += ay * (TY)p[q].re;
        y[q].im += ay * (TY)p[q].im;
      }
      if (cs.pairwise) {
        const char* qb = reinterpret_cast<const char*>(a.w2) + (size_t)col * a.col_bytes;
        SpinorIO<P, NC>::template load<true>(qb, V, s, p);
        const TY aq = (TY)cs.alpha_prev;
#pragma unroll
        for (int q = 0; q < 4 * NC; ++q) {
          y[q].re += aq * (TY)p[q].re;
          y[q].im += aq * (TY)p[q].im;
        }
      }
      if (valid) SpinorIO<PY, NC>::template store<true>(yb, V, s, y);
    }
  }
  return valid ? part : 0.0;
}

// One site per thread (march = 1): every neighbour comes from memory.
template <int P, int NC, int RECON, bool DAG, bool DIAG, bool G5, int EPI>
__global__ __launch_bounds__(NT) void k_dslash_site(DslashArgs a) {
  using T = typename PT<P>::acc;
  const unsigned logical = xcd_tile(blockIdx.x, gridDim.x, a.xcd_remap);
  int col, tile;
  split_units(logical, a.ntiles, a.ncol, a.cols_outer, col, tile);
  const ColState* st = a.red.cols;
  if (a.gate && gate_skip_col(a.gate, a.red, col)) return;

  const size_t V = (size_t)a.g.Vp;  // plane stride
  bool valid;
  int x[4];  // coordinates (last axis fastest)
  const size_t s = site_of_tile(a, tile, valid, x);

  const int cfg = col / a.cols_per_cfg;
  const char* Ucfg = reinterpret_cast<const char*>(a.U) + (size_t)cfg * a.cfg_bytes;
  const char* inb = reinterpret_cast<const char*>(a.in) + (size_t)col * a.col_bytes;

  cx<T> acc[4 * NC];
#pragma unroll
  for (int c = 0; c < 4 * NC; ++c) acc[c] = {T(0), T(0)};

  constexpr int SF = DAG ? +1 : -1;
  constexpr bool FOLD = P == F16;
#define WL_HOP(MU)                                                                                        \
  {                                                                                                       \
    const int L = a.g.L[MU];                                                                              \
    const size_t str = (size_t)a.g.stride[MU];                                                            \
    const bool top = x[MU] == L - 1, bot = x[MU] == 0;                                                    \
    const size_t sf = top ? s - (size_t)(L - 1) * str : s + str;                                          \
    const size_t sb = bot ? s + (size_t)(L - 1) * str : s - str;                                          \
    cx<T> ph = {T(a.bc_re[MU]), T(a.bc_im[MU])};                                                          \
    hop_dir<P, NC, RECON, MU, SF, FOLD>(Ucfg, a.dir_bytes, inb, V, s, sf, sb, top, bot, ph, acc);         \
  }
  WL_HOP(0)
  WL_HOP(1)
  WL_HOP(2)
  WL_HOP(3)
#undef WL_HOP

  cx<T> res[4 * NC];
  const T m0 = DIAG ? (a.m0_from_state ? T(st[col].m0) : T(a.m0)) : T(0);
  site_result<P, NC, DIAG, FOLD>(inb, V, s, m0, acc, res);
  if (G5) apply_g5<T, NC>(res);

  const double part = site_epilogue<P, NC, EPI>(a, col, V, s, valid, res);
  if (EPI != EPI_STORE) write_partial(a.red, col, tile, a.ntiles, part);
}

// Multi-RHS: one site per thread and K right-hand sides of the same gauge
// configuration per thread; each link is loaded once and applied to K
// spinors.  Columns col0 .. col0+K-1 share one workgroup; a column whose
// solve is inactive is skipped (its loads and stores are predicated off).
// One reduction partial per column.
template <int P, int NC, int RECON, bool DAG, bool DIAG, bool G5, int EPI, int K>
__global__ __launch_bounds__(NT) void k_dslash_mrhs(DslashArgs a) {
  using T = typename PT<P>::acc;
  const int ngrp = a.ncol / K;
  const unsigned logical = xcd_tile(blockIdx.x, gridDim.x, a.xcd_remap);
  int grp, tile;
  split_units(logical, a.ntiles, ngrp, a.cols_outer, grp, tile);
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

  const size_t V = (size_t)a.g.Vp;  // plane stride
  bool valid;
  int x[4];
  const size_t s = site_of_tile(a, tile, valid, x);
  const int cfg = col0 / a.cols_per_cfg;
  const char* Ucfg = reinterpret_cast<const char*>(a.U) + (size_t)cfg * a.cfg_bytes;
  cx<T> acc[K][4 * NC];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int q = 0; q < 4 * NC; ++q) acc[k][q] = {T(0), T(0)};

  constexpr int SF = DAG ? +1 : -1;
  constexpr bool FOLD = P == F16;
#define WL_MHOP(MU)                                                                                         \
  {                                                                                                         \
    const int L = a.g.L[MU];                                                                                \
    const size_t str = (size_t)a.g.stride[MU];                                                              \
    const bool top = x[MU] == L - 1, bot = x[MU] == 0;                                                      \
    const size_t sf = top ? s - (size_t)(L - 1) * str : s + str;                                            \
    const size_t sb = bot ? s + (size_t)(L - 1) * str : s - str;                                            \
    const cx<T> ph = {T(a.bc_re[MU]), T(a.bc_im[MU])};                                                      \
    const char* Umu = Ucfg + MU * a.dir_bytes;                                                              \
    {                                                                                                       \
      typename LinkRegs<P, NC>::type u;                                                                     \
      load_link<P, NC, RECON>(Umu, V, s, u);                                                                \
      _Pragma("unroll") for (int k = 0; k < K; ++k) {                                                       \
        if (!on[k]) continue;                                                                               \
        hop_one<P, NC, MU, SF, false, FOLD>(u, reinterpret_cast<const char*>(a.in) + (size_t)(col0 + k) * a.col_bytes, \
                                            V, sf, top, ph, acc[k]);                                        \
      }                                                                                                     \
    }                                                                                                       \
    {                                                                                                       \
      typename LinkRegs<P, NC>::type u;                                                                     \
      load_link<P, NC, RECON>(Umu, V, sb, u);                                                               \
      _Pragma("unroll") for (int k = 0; k < K; ++k) {                                                       \
        if (!on[k]) continue;                                                                               \
        hop_one<P, NC, MU, -SF, true, FOLD>(u, reinterpret_cast<const char*>(a.in) + (size_t)(col0 + k) * a.col_bytes, \
                                            V, sb, bot, ph, acc[k]);                                        \
      }                                                                                                     \
    }                                                                                                       \
  }
  WL_MHOP(0)
  WL_MHOP(1)
  WL_MHOP(2)
  WL_MHOP(3)
#undef WL_MHOP

#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int col = col0 + k;
    double part = 0.0;
    if (on[k]) {
      const char* inb = reinterpret_cast<const char*>(a.in) + (size_t)col * a.col_bytes;
      cx<T> res[4 * NC];
      const T m0 = DIAG ? (a.m0_from_state ? T(st[col].m0) : T(a.m0)) : T(0);
      site_result<P, NC, DIAG, FOLD>(inb, V, s, m0, acc[k], res);
      if (G5) apply_g5<T, NC>(res);
      part = site_epilogue<P, NC, EPI>(a, col, V, s, valid, res);
    }
    if (EPI != EPI_STORE && on[k]) write_partial(a.red, col, tile, a.ntiles, part);
  }
}

// One thread owns one (x1, x2, x3) column and marches over M consecutive x0
// sites (x0 is the largest-stride axis).  The x0 neighbours come from
// registers: psi(x+0) loaded for the forward hop is the next site's centre,
// and U_0(x)^dag P psi(x), formed with the link already loaded for the
// forward hop, is the next site's backward hop.  Directions 1..3 load their
// neighbours from memory (L1/L2).  Workgroup tiles are ordered
// segment-fastest, so the segment that reads a boundary plane and the one
// that owns it run back to back on the same XCD.
template <int P, int NC, int RECON, bool DAG, bool DIAG, bool G5, int EPI>
__global__ __launch_bounds__(NT) void k_dslash(DslashArgs a) {
  using T = typename PT<P>::acc;
  const unsigned logical = xcd_tile(blockIdx.x, gridDim.x, a.xcd_remap);
  int col, tile;
  split_units(logical, a.ntiles, a.ncol, a.cols_outer, col, tile);
  const ColState* st = a.red.cols;
  if (a.gate && gate_skip_col(a.gate, a.red, col)) return;

  const size_t V = (size_t)a.g.Vp;  // plane stride
  const size_t S0 = (size_t)a.g.stride[0];  // sites per x0 slice
  const int G = a.nseg;
  const int chunk = tile / G, seg = tile % G;
  const size_t c = (size_t)chunk * NT + threadIdx.x;
  const bool valid = c < S0;
  const size_t cc = valid ? c : 0;
  const int L0 = a.g.L[0];
  const int M = a.march;
  const int x0s = seg * M;

  const int cfg = col / a.cols_per_cfg;
  const char* Ucfg = reinterpret_cast<const char*>(a.U) + (size_t)cfg * a.cfg_bytes;
  const char* inb = reinterpret_cast<const char*>(a.in) + (size_t)col * a.col_bytes;

  // coordinates of the column (last axis fastest)
  int x[4];
  {
    size_t rem = cc;
    x[3] = rem % a.g.L[3];
    rem /= a.g.L[3];
    x[2] = rem % a.g.L[2];
    x[1] = rem / a.g.L[2];
  }
  // signed neighbour offsets of directions 1..3 (independent of x0)
  long long off_f[4], off_b[4];
  bool top[4], bot[4];
#pragma unroll
  for (int mu = 1; mu < 4; ++mu) {
    const int L = a.g.L[mu];
    const long long str = a.g.stride[mu];
    top[mu] = x[mu] == L - 1;
    bot[mu] = x[mu] == 0;
    off_f[mu] = top[mu] ? -(long long)(L - 1) * str : str;
    off_b[mu] = bot[mu] ? (long long)(L - 1) * str : -str;
  }

  constexpr int SF = DAG ? +1 : -1;
  const cx<T> ph0 = {T(a.bc_re[0]), T(a.bc_im[0])};
  const T m0 = DIAG ? (a.m0_from_state ? T(st[col].m0) : T(a.m0)) : T(0);

  cx<T> cur[4 * NC];
  SpinorIO<P, NC>::load(inb, V, (size_t)x0s * S0 + cc, cur);
  cx<T> carry[2][NC];
  {
    const int xm = (x0s == 0) ? L0 - 1 : x0s - 1;
    const size_t sm = (size_t)xm * S0 + cc;
    cx<T> p[4 * NC], u[NC][NC], h[2][NC];
    SpinorIO<P, NC>::load(inb, V, sm, p);
    LinkIO<P, NC, RECON>::load(Ucfg, V, sm, u);
    spin_project<0, -SF>(p, h);
    color_mul<true>(u, h, carry);
    if (xm == L0 - 1) half_phase(carry, cconj(ph0));
  }

  double part = 0.0;
  for (int j = 0; j < M; ++j) {
    const int x0 = x0s + j;
    const size_t s = (size_t)x0 * S0 + cc;
    const bool top0 = x0 == L0 - 1;
    cx<T> acc[4 * NC];
#pragma unroll
    for (int q = 0; q < 4 * NC; ++q) acc[q] = {T(0), T(0)};
    // direction 0, backward: carried from the previous site
    spin_reconstruct_add<0, -SF>(acc, carry);
    // direction 0, forward
    cx<T> nxt[4 * NC];
    SpinorIO<P, NC>::load(inb, V, top0 ? cc : s + S0, nxt);
    {
      cx<T> u[NC][NC], h[2][NC], o[2][NC];
      LinkIO<P, NC, RECON>::load(Ucfg, V, s, u);
      spin_project<0, SF>(nxt, h);
      color_mul<false>(u, h, o);
      if (top0) half_phase(o, ph0);
      spin_reconstruct_add<0, SF>(acc, o);
      // backward hop of the next site: U_0(x)^dag P psi(x)
      spin_project<0, -SF>(cur, h);
      color_mul<true>(u, h, carry);
      if (top0) half_phase(carry, cconj(ph0));
    }
#define WL_HOP(MU)                                                                                          \
    {                                                                                                       \
      const cx<T> ph = {T(a.bc_re[MU]), T(a.bc_im[MU])};                                                    \
      hop_dir<P, NC, RECON, MU, SF>(Ucfg, a.dir_bytes, inb, V, s, (size_t)((long long)s + off_f[MU]),       \
                                    (size_t)((long long)s + off_b[MU]), top[MU], bot[MU], ph, acc);         \
    }
    WL_HOP(1)
    WL_HOP(2)
    WL_HOP(3)
#undef WL_HOP

    cx<T> res[4 * NC];
#pragma unroll
    for (int q = 0; q < 4 * NC; ++q)
      res[q] = DIAG ? cx<T>{m0 * cur[q].re - T(0.5) * acc[q].re, m0 * cur[q].im - T(0.5) * acc[q].im}
                    : cx<T>{T(-0.5) * acc[q].re, T(-0.5) * acc[q].im};
    if (G5) apply_g5<T, NC>(res);
    part += site_epilogue<P, NC, EPI>(a, col, V, s, valid, res);
#pragma unroll
    for (int q = 0; q < 4 * NC; ++q) cur[q] = nxt[q];
  }
  if (EPI != EPI_STORE) write_partial(a.red, col, tile, a.ntiles, part);
}

// Multi-RHS march (FP16 storage): K right-hand sides of one configuration per
// thread, one (x1, x2, x3) column per thread marching over M consecutive x0
// sites.  Links are loaded once per site and applied to the K spinors, and the
// x0 neighbours come from registers: psi(x+0), loaded for the forward hop,
// becomes the next site's centre, and U_0(x)^dag P psi(x), formed with the link
// of the forward hop, is carried as the next site's backward hop.  So a
// site's input is fetched once for direction 0 and its centre, instead of
// three times through a cache that a multi-RHS working set overflows.
// Directions 1..3 use the half-precision hop of the one-site kernels.
// Workgroups are ordered column groups fastest, then x1/x2/x3 chunks, then x0
// segments, so that the chunks holding a chunk's x1 and x2 neighbours run at
// the same time (on the same XCD with the XCD-contiguous order).
template <int NC, int RECON, bool DAG, bool DIAG, bool G5, int EPI, int K>
__global__ __launch_bounds__(NT) void k_dslash_mrhs_march(DslashArgs a) {
  using T = float;
  const int ngrp = a.ncol / K;
  const unsigned logical = xcd_tile(blockIdx.x, gridDim.x, a.xcd_remap);
  int grp, tile;
  split_units(logical, a.ntiles, ngrp, a.cols_outer, grp, tile);
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

  const size_t V = (size_t)a.g.Vp;          // plane stride
  const size_t S0 = (size_t)a.g.stride[0];  // sites per x0 slice
  const int nchunk = a.ntiles / a.nseg;
  const int chunk = tile % nchunk, seg = tile / nchunk;
  const size_t c = (size_t)chunk * NT + threadIdx.x;
  const bool valid = c < S0;
  const size_t cc = valid ? c : 0;
  const int L0 = a.g.L[0];
  const int M = a.march;
  const int x0s = seg * M;

  const int cfg = col0 / a.cols_per_cfg;
  const char* Ucfg = reinterpret_cast<const char*>(a.U) + (size_t)cfg * a.cfg_bytes;
  const char* inb[K];
#pragma unroll
  for (int k = 0; k < K; ++k) inb[k] = reinterpret_cast<const char*>(a.in) + (size_t)(col0 + k) * a.col_bytes;

  int x[4];
  {
    size_t rem = cc;
    x[3] = rem % a.g.L[3];
    rem /= a.g.L[3];
    x[2] = rem % a.g.L[2];
    x[1] = rem / a.g.L[2];
  }
  long long off_f[4], off_b[4];
  bool top[4], bot[4];
#pragma unroll
  for (int mu = 1; mu < 4; ++mu) {
    const int L = a.g.L[mu];
    const long long str = a.g.stride[mu];
    top[mu] = x[mu] == L - 1;
    bot[mu] = x[mu] == 0;
    off_f[mu] = top[mu] ? -(long long)(L - 1) * str : str;
    off_b[mu] = bot[mu] ? (long long)(L - 1) * str : -str;
  }

  constexpr int SF = DAG ? +1 : -1;
  const cx<T> ph0 = {T(a.bc_re[0]), T(a.bc_im[0])};
  const bool real0 = ph0.im == T(0);
  const char* U0 = Ucfg;  // direction 0

  HalfSpinor<NC> cur[K];
  cx<T> carry[K][2][NC];
  float carry_f[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    carry_f[k] = 0.0f;
    if (on[k]) load_half_spinor<NC>(inb[k], V, (size_t)x0s * S0 + cc, cur[k]);
  }
  {
    // backward x0 hop of the first site, from x0s - 1
    const int xm = (x0s == 0) ? L0 - 1 : x0s - 1;
    const size_t sm = (size_t)xm * S0 + cc;
    const bool bnd = xm == L0 - 1;
    hf2 u[NC][NC];
    load_half_link<NC, RECON>(U0, V, sm, u);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (!on[k]) continue;
      HalfSpinor<NC> r;
      load_half_spinor<NC>(inb[k], V, sm, r);
      hf2 hA[2][NC], hB[2][NC];
      project_forms<0, -SF, true>(r, hA, hB);
      color_mul_half<true>(u, hA, hB, carry[k]);
      if (!real0 && bnd) half_phase(carry[k], cconj(ph0));
      carry_f[k] = -0.5f * r.s * ((real0 && bnd) ? ph0.re : 1.0f);
    }
  }

  double part[K];
#pragma unroll
  for (int k = 0; k < K; ++k) part[k] = 0.0;
  for (int j = 0; j < M; ++j) {
    const int x0 = x0s + j;
    const size_t s = (size_t)x0 * S0 + cc;
    const bool top0 = x0 == L0 - 1;
    // the accumulator starts from the diagonal term m0 psi(x), so that the
    // centre spinor is dead once the x0 hops have used it
    cx<T> acc[K][4 * NC];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (DIAG && on[k]) {
        const T m0 = a.m0_from_state ? T(st[col0 + k].m0) : T(a.m0);
        const float ms = m0 * cur[k].s;
#pragma unroll
        for (int q = 0; q < 4 * NC; ++q) {
          const float2 p = h2f(cur[k].w[q]);
          acc[k][q] = {p.x * ms, p.y * ms};
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4 * NC; ++q) acc[k][q] = {T(0), T(0)};
      }
      if (on[k]) spin_reconstruct_fma<0, -SF>(acc[k], carry[k], carry_f[k]);
    }
    // direction 0: forward hop, and the backward hop carried to the next site
    {
      HalfSpinor<NC> nxt[K];
      const size_t sn = top0 ? cc : s + S0;
      const float fb = (real0 && top0) ? ph0.re : 1.0f;
      hf2 u[NC][NC];
      load_half_link<NC, RECON>(U0, V, s, u);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (!on[k]) continue;
        load_half_spinor<NC>(inb[k], V, sn, nxt[k]);
        hf2 hA[2][NC], hB[2][NC];
        cx<T> o[2][NC];
        project_forms<0, SF, false>(nxt[k], hA, hB);
        color_mul_half<false>(u, hA, hB, o);
        if (!real0 && top0) half_phase(o, ph0);
        spin_reconstruct_fma<0, SF>(acc[k], o, -0.5f * nxt[k].s * fb);
        project_forms<0, -SF, true>(cur[k], hA, hB);
        color_mul_half<true>(u, hA, hB, carry[k]);
        if (!real0 && top0) half_phase(carry[k], cconj(ph0));
        carry_f[k] = -0.5f * cur[k].s * fb;
        cur[k] = nxt[k];
      }
    }
#define WL_MMHOP(MU)                                                                                          \
    {                                                                                                         \
      const cx<T> ph = {T(a.bc_re[MU]), T(a.bc_im[MU])};                                                      \
      const char* Umu = Ucfg + MU * a.dir_bytes;                                                              \
      const size_t sf = (size_t)((long long)s + off_f[MU]), sb = (size_t)((long long)s + off_b[MU]);          \
      {                                                                                                       \
        hf2 u[NC][NC];                                                                                        \
        load_half_link<NC, RECON>(Umu, V, s, u);                                                              \
        _Pragma("unroll") for (int k = 0; k < K; ++k) {                                                       \
          if (!on[k]) continue;                                                                               \
          hop_one<F16, NC, MU, SF, false, true>(u, inb[k], V, sf, top[MU], ph, acc[k]);                       \
        }                                                                                                     \
      }                                                                                                       \
      {                                                                                                       \
        hf2 u[NC][NC];                                                                                        \
        load_half_link<NC, RECON>(Umu, V, sb, u);                                                             \
        _Pragma("unroll") for (int k = 0; k < K; ++k) {                                                       \
          if (!on[k]) continue;                                                                               \
          hop_one<F16, NC, MU, -SF, true, true>(u, inb[k], V, sb, bot[MU], ph, acc[k]);                       \
        }                                                                                                     \
      }                                                                                                       \
    }
    WL_MMHOP(1)
    WL_MMHOP(2)
    WL_MMHOP(3)
#undef WL_MMHOP

#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (!on[k]) continue;
      cx<T> res[4 * NC];
#pragma unroll
      for (int q = 0; q < 4 * NC; ++q) res[q] = acc[k][q];
      if (G5) apply_g5<T, NC>(res);
      part[k] += site_epilogue<F16, NC, EPI>(a, col0 + k, V, s, valid, res);
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (EPI != EPI_STORE && on[k]) write_partial(a.red, col0 + k, tile, a.ntiles, part[k]);
}

// LDS-staged plane kernel (FP16 storage, one column per thread).  A workgroup
// owns a tile of prow full x3 rows of the (x2, x3) plane at one x0 and marches
// over `march` consecutive x1 values.  Each step's plane tile, with one halo
// row on each side in x2 (periodic), sits in LDS (double-buffered: the next
// plane is staged while the current one is read), so the x2 and x3 neighbours
// of every site are LDS reads; the x1 neighbours come from registers (psi(x+1)
// loaded for the forward hop is the next step's centre, and U_1(x)^dag P psi(x)
// is carried as the next step's backward hop); only the x0 neighbours and the
// links are read from memory.  Workgroups are ordered x0 fastest, so the tiles
// whose planes are a tile's x0 neighbours run beside it.  One barrier per step.
// Three wavefronts per SIMD (168 VGPRs with a few spilled) measured faster
// for 12 right-hand sides than the unconstrained 192 VGPRs at two.
template <int NC, int RECON, bool DAG, bool DIAG, bool G5, int EPI>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3))) void k_dslash_plane(DslashArgs a) {
  using T = float;
  extern __shared__ __align__(16) unsigned char plane_lds[];
  const unsigned logical = xcd_tile(blockIdx.x, gridDim.x, a.xcd_remap);
  int col, tile;
  split_units(logical, a.ntiles, a.ncol, a.cols_outer, col, tile);
  if (a.gate && gate_skip_col(a.gate, a.red, col)) return;
  const ColState* st = a.red.cols;

  const int L0 = a.g.L[0], L1 = a.g.L[1], L2 = a.g.L[2], L3 = a.g.L[3];
  const int prow = a.prow;
  const int ntr = L2 / prow;
  const int x0 = tile % L0;
  const int rest = tile / L0;
  const int x2t = rest % ntr;
  const int seg = rest / ntr;
  const int M = a.march;
  const int x1s = seg * M;
  const int nthr = prow * L3;  // threads owning a site; the rest only stage halo sites
  const bool own = (int)threadIdx.x < nthr;
  const int tid = own ? (int)threadIdx.x : 0;
  const int x3 = tid % L3;
  const int rr = tid / L3;
  const int x2 = x2t * prow + rr;
  const size_t V = (size_t)a.g.Vp;
  const size_t S0 = (size_t)a.g.stride[0], S1 = (size_t)a.g.stride[1];
  const size_t col_base = (size_t)x0 * S0 + (size_t)x2 * L3 + x3;  // site index without the x1 term
  const int R = prow + 2;
  const int RL = R * L3;
  uint4* lw = reinterpret_cast<uint4*>(plane_lds);         // [2][NC][RL]
  float* ls = reinterpret_cast<float*>(lw + 2 * NC * RL);  // [2][RL]
  const int li = (rr + 1) * L3 + x3;                        // this site's LDS slot

  const int cfg = col / a.cols_per_cfg;
  const char* Ucfg = reinterpret_cast<const char*>(a.U) + (size_t)cfg * a.cfg_bytes;
  const char* inb = reinterpret_cast<const char*>(a.in) + (size_t)col * a.col_bytes;

  // halo slots: threads 0 .. 2 L3 - 1 stage x2 rows x2t*prow - 1 and x2t*prow + prow
  const int hx = (int)threadIdx.x;
  const bool halo = hx < 2 * L3;
  const int hrow = halo ? hx / L3 : 0;
  const int hx3 = hx - hrow * L3;
  const int hx2 = hrow == 0 ? (x2t * prow + L2 - 1) % L2 : (x2t * prow + prow) % L2;
  const int hli = (hrow == 0 ? 0 : R - 1) * L3 + hx3;
  const size_t hcol_base = (size_t)x0 * S0 + (size_t)hx2 * L3 + hx3;

  auto put = [&](int buf, int slot, const HalfSpinor<NC>& h) {
#pragma unroll
    for (int j = 0; j < NC; ++j)
      lw[(buf * NC + j) * RL + slot] = make_uint4(h.w[4 * j], h.w[4 * j + 1], h.w[4 * j + 2], h.w[4 * j + 3]);
    ls[buf * RL + slot] = h.s;
  };
  auto get = [&](int buf, int slot, HalfSpinor<NC>& h) {
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      const uint4 t = lw[(buf * NC + j) * RL + slot];
      h.w[4 * j] = t.x;
      h.w[4 * j + 1] = t.y;
      h.w[4 * j + 2] = t.z;
      h.w[4 * j + 3] = t.w;
    }
    h.s = ls[buf * RL + slot];
  };
  auto stage_halo = [&](int buf, int x1p) {
    if (halo) {
      HalfSpinor<NC> h;
      load_half_spinor<NC>(inb, V, hcol_base + (size_t)x1p * S1, h);
      put(buf, hli, h);
    }
  };

  // boundary flags and neighbour offsets of directions 0, 2, 3 (independent of x1)
  const bool top0 = x0 == L0 - 1, bot0 = x0 == 0;
  const long long off0f = top0 ? -(long long)(L0 - 1) * (long long)S0 : (long long)S0;
  const long long off0b = bot0 ? (long long)(L0 - 1) * (long long)S0 : -(long long)S0;
  const bool top2 = x2 == L2 - 1, bot2 = x2 == 0;
  const bool top3 = x3 == L3 - 1, bot3 = x3 == 0;
  const long long off2b = bot2 ? (long long)(L2 - 1) * L3 : -(long long)L3;
  const long long off3b = bot3 ? (long long)(L3 - 1) : -1LL;
  // LDS slots of the in-plane neighbours
  const int li2f = (rr + 2) * L3 + x3, li2b = rr * L3 + x3;
  const int li3f = (rr + 1) * L3 + (top3 ? 0 : x3 + 1), li3b = (rr + 1) * L3 + (bot3 ? L3 - 1 : x3 - 1);

  constexpr int SF = DAG ? +1 : -1;
  const cx<T> ph0 = {T(a.bc_re[0]), T(a.bc_im[0])}, ph1 = {T(a.bc_re[1]), T(a.bc_im[1])};
  const cx<T> ph2 = {T(a.bc_re[2]), T(a.bc_im[2])}, ph3 = {T(a.bc_re[3]), T(a.bc_im[3])};
  const bool real1 = ph1.im == T(0);
  const char* U1 = Ucfg + a.dir_bytes;

  {
    HalfSpinor<NC> first;
    load_half_spinor<NC>(inb, V, col_base + (size_t)x1s * S1, first);
    if (own) put(0, li, first);
    stage_halo(0, x1s);
  }
  cx<T> carry[2][NC];
  float carry_f;
  {
    const int xm = (x1s == 0) ? L1 - 1 : x1s - 1;
    const size_t sm = col_base + (size_t)xm * S1;
    const bool bnd = xm == L1 - 1;
    hf2 u[NC][NC];
    load_half_link<NC, RECON>(U1, V, sm, u);
    HalfSpinor<NC> r;
    load_half_spinor<NC>(inb, V, sm, r);
    hf2 hA[2][NC], hB[2][NC];
    project_forms<1, -SF, true>(r, hA, hB);
    color_mul_half<true>(u, hA, hB, carry);
    if (!real1 && bnd) half_phase(carry, cconj(ph1));
    carry_f = -0.5f * r.s * ((real1 && bnd) ? ph1.re : 1.0f);
  }

  double part = 0.0;
  for (int j = 0; j < M; ++j) {
    const int x1 = x1s + j;
    const size_t s = col_base + (size_t)x1 * S1;
    const bool top1 = x1 == L1 - 1;
    const int cb = j & 1, nb = cb ^ 1;
    __syncthreads();  // plane x1 staged in buffer cb; buffer nb free
    HalfSpinor<NC> cur;  // the centre, from this site's own slot
    get(cb, li, cur);
    cx<T> acc[4 * NC];
    if (DIAG) {
      const T m0 = a.m0_from_state ? T(st[col].m0) : T(a.m0);
      const float ms = m0 * cur.s;
#pragma unroll
      for (int q = 0; q < 4 * NC; ++q) {
        const float2 p = h2f(cur.w[q]);
        acc[q] = {p.x * ms, p.y * ms};
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4 * NC; ++q) acc[q] = {T(0), T(0)};
    }
    spin_reconstruct_fma<1, -SF>(acc, carry, carry_f);
    // direction 1: forward hop; stage the next plane; carry the backward hop
    HalfSpinor<NC> nxt;
    {
      const int x1n = top1 ? 0 : x1 + 1;
      load_half_spinor<NC>(inb, V, col_base + (size_t)x1n * S1, nxt);
      if (j + 1 < M) {
        if (own) put(nb, li, nxt);
        stage_halo(nb, x1n);
      }
      const float fb = (real1 && top1) ? ph1.re : 1.0f;
      hf2 u[NC][NC];
      load_half_link<NC, RECON>(U1, V, s, u);
      hf2 hA[2][NC], hB[2][NC];
      cx<T> o[2][NC];
      project_forms<1, SF, false>(nxt, hA, hB);
      color_mul_half<false>(u, hA, hB, o);
      if (!real1 && top1) half_phase(o, ph1);
      spin_reconstruct_fma<1, SF>(acc, o, -0.5f * nxt.s * fb);
      project_forms<1, -SF, true>(cur, hA, hB);
      color_mul_half<true>(u, hA, hB, carry);
      if (!real1 && top1) half_phase(carry, cconj(ph1));
      carry_f = -0.5f * cur.s * fb;
    }
    // direction 0: neighbours from memory
    {
      const char* U0 = Ucfg;
      hf2 u[NC][NC];
      load_half_link<NC, RECON>(U0, V, s, u);
      hop_one<F16, NC, 0, SF, false, true>(u, inb, V, (size_t)((long long)s + off0f), top0, ph0, acc);
      load_half_link<NC, RECON>(U0, V, (size_t)((long long)s + off0b), u);
      hop_one<F16, NC, 0, -SF, true, true>(u, inb, V, (size_t)((long long)s + off0b), bot0, ph0, acc);
    }
    // directions 2 and 3: neighbours from the staged plane
    {
      const char* U2 = Ucfg + 2 * a.dir_bytes;
      hf2 u[NC][NC];
      HalfSpinor<NC> r;
      load_half_link<NC, RECON>(U2, V, s, u);
      get(cb, li2f, r);
      hop_half<NC, 2, SF, false, true>(u, r, top2, ph2, acc);
      load_half_link<NC, RECON>(U2, V, (size_t)((long long)s + off2b), u);
      get(cb, li2b, r);
      hop_half<NC, 2, -SF, true, true>(u, r, bot2, ph2, acc);
    }
    {
      const char* U3 = Ucfg + 3 * a.dir_bytes;
      hf2 u[NC][NC];
      HalfSpinor<NC> r;
      load_half_link<NC, RECON>(U3, V, s, u);
      get(cb, li3f, r);
      hop_half<NC, 3, SF, false, true>(u, r, top3, ph3, acc);
      load_half_link<NC, RECON>(U3, V, (size_t)((long long)s + off3b), u);
      get(cb, li3b, r);
      hop_half<NC, 3, -SF, true, true>(u, r, bot3, ph3, acc);
    }
    cx<T> res[4 * NC];
#pragma unroll
    for (int q = 0; q < 4 * NC; ++q) res[q] = acc[q];
    if (G5) apply_g5<T, NC>(res);
    part += site_epilogue<F16, NC, EPI>(a, col, V, s, own, res);
  }
  if (EPI != EPI_STORE) write_partial(a.red, col, tile, a.ntiles, part);
}

// LDS box kernel (FP16 storage, SU(3)).  A workgroup owns one box of 256
// sites (extents 2^bsh, one site per thread) of one configuration and walks
// all of that configuration's columns, two per pass.  The forward links
// U_mu(x) of the box's sites (12-real storage) are staged once in LDS: a
// backward link U_mu(x - mu)
// inside the box is the forward link of that site, and only backward links
// across the box's lower faces are read from memory.  Each pass stages the
// two columns' spinors at the box's own sites in LDS; a neighbour inside the
// box is then an LDS read, and only neighbours outside it are read from
// memory.  Each link is rebuilt once per pass and applied to both columns.
// One barrier after the staging and one after the last read of the pass (the
// reduction's own barriers serve for the reducing epilogues).  With one
// workgroup per box and all columns inside it, the links are read once per
// launch and a box's out-of-box neighbours are read by its neighbour boxes in
// the same pass of theirs; numbering boxes x3-fastest inside 8 regions of the
// box grid, one region per XCD (workgroup b runs on XCD b mod 8), keeps those
// neighbour boxes on the same L2.  51 KiB of LDS per workgroup: three
// workgroups (wavefronts per SIMD) per CU.
template <int RECON, bool DAG, bool DIAG, bool G5, int EPI>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3))) void k_dslash_box(DslashArgs a) {
  static_assert(RECON == 12, "the box kernel stages 12-real links");
  constexpr int NC = 3;
  using T = float;
  __shared__ uint4 bsp[2][NC][NT];  // [column of the pass][plane][site]
  __shared__ float bss[2][NT];      // per-site scales
  __shared__ uint4 blk4[4][NT];     // forward links, words 0..3 of the 12-real storage
  __shared__ uint2 blk2[4][NT];     // words 4..5
  const int nbox = a.ntiles;
  const int cfg = (int)blockIdx.x / nbox;
  const int jb = (int)blockIdx.x - cfg * nbox;
  int tt[4];
  {
    unsigned r = a.lregions ? ((unsigned)jb & 7u) : 0u;
    unsigned i = a.lregions ? ((unsigned)jb >> 3) : (unsigned)jb;
    int rc[4], rn[4];
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      rn[mu] = a.bn[mu] / a.lsplit[mu];
      rc[mu] = (int)(r % (unsigned)a.lsplit[mu]);
      r /= (unsigned)a.lsplit[mu];
    }
#pragma unroll
    for (int mu = 3; mu >= 0; --mu) {
      tt[mu] = rc[mu] * rn[mu] + (int)(i % (unsigned)rn[mu]);
      i /= (unsigned)rn[mu];
    }
  }
  const int tid = (int)threadIdx.x;
  const int sh3 = a.bsh[3], sh2 = a.bsh[2], sh1 = a.bsh[1];
  const int lstr[4] = {1 << (sh1 + sh2 + sh3), 1 << (sh2 + sh3), 1 << sh3, 1};
  int x[4], lb[4];
  lb[3] = tid & ((1 << sh3) - 1);
  lb[2] = (tid >> sh3) & ((1 << sh2) - 1);
  lb[1] = (tid >> (sh3 + sh2)) & ((1 << sh1) - 1);
  lb[0] = tid >> (sh3 + sh2 + sh1);
  unsigned s = 0;  // the volume is below 2^31
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    x[mu] = (tt[mu] << a.bsh[mu]) + lb[mu];
    s += (unsigned)x[mu] * (unsigned)a.g.stride[mu];
  }
  // neighbour n = 2 mu (forward) or 2 mu + 1 (backward): its site, and its
  // slot in the box (-1 outside; a box spanning a whole axis wraps in itself)
  auto nb_site = [&](int n) -> unsigned {
    const int mu = n >> 1;
    const unsigned L = (unsigned)a.g.L[mu], str = (unsigned)a.g.stride[mu];
    if (n & 1) return x[mu] == 0 ? s + (L - 1) * str : s - str;
    return x[mu] == (int)L - 1 ? s - (L - 1) * str : s + str;
  };
  auto nb_slot = [&](int n) -> int {
    const int mu = n >> 1;
    const int ext = 1 << a.bsh[mu];
    const bool full = ext == a.g.L[mu];
    if (n & 1) return lb[mu] > 0 ? tid - lstr[mu] : (full ? tid + (ext - 1) * lstr[mu] : -1);
    return lb[mu] + 1 < ext ? tid + lstr[mu] : (full ? tid - (ext - 1) * lstr[mu] : -1);
  };
  const size_t V = (size_t)a.g.Vp;
  const char* Ucfg = reinterpret_cast<const char*>(a.U) + (size_t)cfg * a.cfg_bytes;
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    const uint4* pa = reinterpret_cast<const uint4*>(Ucfg + mu * a.dir_bytes);
    blk4[mu][tid] = pa[s];
    blk2[mu][tid] = reinterpret_cast<const uint2*>(pa + V)[s];
  }
  // the link of neighbour n: U_mu(x) forward, U_mu(x - mu) backward
  auto link = [&](int n, hf2 (&u)[NC][NC]) {
    const int mu = n >> 1;
    const int slot = (n & 1) ? nb_slot(n) : tid;
    uint32_t w[6];
    if (slot >= 0) {
      const uint4 t = blk4[mu][slot];
      const uint2 q = blk2[mu][slot];
      w[0] = t.x;
      w[1] = t.y;
      w[2] = t.z;
      w[3] = t.w;
      w[4] = q.x;
      w[5] = q.y;
    } else {
      load_link_words<12>(Ucfg + mu * a.dir_bytes, V, nb_site(n), w);
    }
    half_link_from_words<12>(w, u);
  };

  const ColState* st = a.red.cols;
  const int cpc = a.cols_per_cfg;
  const int colb = cfg * cpc;
  const int npass = (cpc + 1) >> 1;
  constexpr int SF = DAG ? +1 : -1;
  for (int p = 0; p < npass; ++p) {
    bool on[2];
    const char* inb[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int cc = 2 * p + k;
      on[k] = cc < cpc && !(a.gate && gate_skip_col(a.gate, a.red, colb + cc));
      inb[k] = reinterpret_cast<const char*>(a.in) + (size_t)(colb + (cc < cpc ? cc : 0)) * a.col_bytes;
    }
    if (!on[0] && !on[1]) continue;  // the same for every thread of the workgroup
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (!on[k]) continue;
      HalfSpinor<NC> h;
      load_half_spinor<NC>(inb[k], V, s, h);
#pragma unroll
      for (int j = 0; j < NC; ++j) bsp[k][j][tid] = make_uint4(h.w[4 * j], h.w[4 * j + 1], h.w[4 * j + 2], h.w[4 * j + 3]);
      bss[k][tid] = h.s;
    }
    __syncthreads();
    auto get = [&](int k, int slot, HalfSpinor<NC>& h) {
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const uint4 t = bsp[k][j][slot];
        h.w[4 * j] = t.x;
        h.w[4 * j + 1] = t.y;
        h.w[4 * j + 2] = t.z;
        h.w[4 * j + 3] = t.w;
      }
      h.s = bss[k][slot];
    };
    cx<T> acc[2][4 * NC];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int q = 0; q < 4 * NC; ++q) acc[k][q] = {T(0), T(0)};
#define WL_BHOP(MU, N, S, ADJ, BND)                                    \
  {                                                                   \
    hf2 u[NC][NC];                                                    \
    link(N, u);                                                       \
    const cx<T> ph = {T(a.bc_re[MU]), T(a.bc_im[MU])};                \
    const int tn = nb_slot(N);                                        \
    _Pragma("unroll") for (int k = 0; k < 2; ++k) {                   \
      if (!on[k]) continue;                                           \
      HalfSpinor<NC> r;                                               \
      if (tn >= 0) get(k, tn, r);                                     \
      else load_half_spinor<NC>(inb[k], V, nb_site(N), r);            \
      hop_half<NC, MU, S, ADJ, true>(u, r, BND, ph, acc[k]);          \
    }                                                                 \
  }
    WL_BHOP(0, 0, SF, false, x[0] == a.g.L[0] - 1)
    WL_BHOP(0, 1, -SF, true, x[0] == 0)
    WL_BHOP(1, 2, SF, false, x[1] == a.g.L[1] - 1)
    WL_BHOP(1, 3, -SF, true, x[1] == 0)
    WL_BHOP(2, 4, SF, false, x[2] == a.g.L[2] - 1)
    WL_BHOP(2, 5, -SF, true, x[2] == 0)
    WL_BHOP(3, 6, SF, false, x[3] == a.g.L[3] - 1)
    WL_BHOP(3, 7, -SF, true, x[3] == 0)
#undef WL_BHOP
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (!on[k]) continue;
      const int col = colb + 2 * p + k;
      cx<T> res[4 * NC];
      if (DIAG) {
        // res = m0 psi(x) + acc, the hops already times -1/2
        HalfSpinor<NC> c;
        get(k, tid, c);
        const float ms = (a.m0_from_state ? T(st[col].m0) : T(a.m0)) * c.s;
#pragma unroll
        for (int q = 0; q < 4 * NC; ++q) {
          const float2 pc = h2f(c.w[q]);
          res[q] = {fmaf(pc.x, ms, acc[k][q].re), fmaf(pc.y, ms, acc[k][q].im)};
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4 * NC; ++q) res[q] = acc[k][q];
      }
      if (G5) apply_g5<T, NC>(res);
      const double part = site_epilogue<F16, NC, EPI>(a, col, V, (size_t)s, true, res);
      if (EPI != EPI_STORE) write_partial(a.red, col, jb, nbox, part);
    }
    // the staged spinors are read until here (a reducing epilogue has just
    // passed the reduction's barriers, after this pass's last read)
    if (EPI == EPI_STORE) __syncthreads();
  }
}

}  // namespace wl
