// Wilson operator kernels: D, D^dag, gamma5 D, the hopping term K and K^dag,
// with optional fused reductions for CG (Re<p, A p> and |b - A x|^2).
//
// The 1D grid is ntiles x ncol workgroups ordered spinor-column-fastest and
// remapped so that each XCD walks one contiguous range of tiles: the k
// right-hand sides of one tile run back to back on the same XCD and share
// its link lines in L2.
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

// One site per thread (march = 1): every neighbour comes from memory.
// EPI: 0 store, 1 store + Re<w, out>, 2 out = w - result and |out|^2
template <int P, int NC, int RECON, bool DAG, bool DIAG, bool G5, int EPI>
__global__ __launch_bounds__(NT) void k_dslash_site(DslashArgs a) {
  using T = typename PT<P>::acc;
  const unsigned logical = xcd_tile(blockIdx.x, gridDim.x);
  const int col = logical % a.ncol;
  const int tile = logical / a.ncol;
  const ColState* st = a.red.cols;
  if (a.gate && gate_skip(a.gate, st[col])) return;

  const size_t V = (size_t)a.g.Vp;  // plane stride
  const size_t site = (size_t)tile * NT + threadIdx.x;
  const bool valid = site < (size_t)a.g.V;
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
  if (valid) store_nt<P, NC>(outb, V, s, res);
  if (EPI != 0) {
    if (!valid) part = 0.0;
    write_partial(a.red, col, tile, a.ntiles, part);
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
//
// EPI: 0 store, 1 store + Re<w, out> partials, 2 out = w - result and |out|^2 partials
template <int P, int NC, int RECON, bool DAG, bool DIAG, bool G5, int EPI>
__global__ __launch_bounds__(NT) void k_dslash(DslashArgs a) {
  using T = typename PT<P>::acc;
  const unsigned logical = xcd_tile(blockIdx.x, gridDim.x);
  const int col = logical % a.ncol;
  const int tile = logical / a.ncol;
  const ColState* st = a.red.cols;
  if (a.gate && gate_skip(a.gate, st[col])) return;

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
  char* outb = reinterpret_cast<char*>(a.out) + (size_t)col * a.col_bytes;
  const char* wb = reinterpret_cast<const char*>(a.w) + (size_t)col * a.col_bytes;

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
    if (EPI == 2) {
      cx<T> w[4 * NC];
      SpinorIO<P, NC>::load(wb, V, s, w);
#pragma unroll
      for (int q = 0; q < 4 * NC; ++q) {
        res[q] = csub(w[q], res[q]);
        part += (double)res[q].re * (double)res[q].re + (double)res[q].im * (double)res[q].im;
      }
    } else if (EPI == 1) {
      cx<T> w[4 * NC];
      SpinorIO<P, NC>::load(wb, V, s, w);
#pragma unroll
      for (int q = 0; q < 4 * NC; ++q) part += (double)w[q].re * (double)res[q].re + (double)w[q].im * (double)res[q].im;
    }
    if (valid) store_nt<P, NC>(outb, V, s, res);
#pragma unroll
    for (int q = 0; q < 4 * NC; ++q) cur[q] = nxt[q];
  }
  if (EPI != 0) write_partial(a.red, col, tile, a.ntiles, valid ? part : 0.0);
}

// ---------------------------------------------------------------------------
// host-side dispatch
// ---------------------------------------------------------------------------
template <int P, int NC, int RECON, bool DAG, bool DIAG, bool G5, int EPI>
static void launch_k(const DslashArgs& a, dim3 grid, hipStream_t stream) {
  if (a.march == 1)
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

hipError_t launch_reduce_finish(const Reduction& red, int ncol, int nblk, int gate, hipStream_t stream);

static hipError_t dslash_main(int prec, int nc, int recon, int op, int epi, const DslashArgs& a, hipStream_t stream) {
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

hipError_t launch_dslash(int prec, int nc, int recon, int op, int epi, const DslashArgs& a, hipStream_t stream) {
  hipError_t e = dslash_main(prec, nc, recon, op, epi, a, stream);
  if (e != hipSuccess || epi == 0) return e;
  return launch_reduce_finish(a.red, a.ncol, a.ntiles, a.gate, stream);
}

}  // namespace wl
